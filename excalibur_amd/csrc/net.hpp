// Host TCP transport between peers: the framing and collectives of client.cpp:19-411, with
// receive timeouts and TCP_NODELAY added.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace exc {

class PeerNet {
 public:
  // Peer j listens on base_port + j at the address of `listen` with j added to its last octet
  // (client.cpp:238-245). recv_timeout_s bounds every blocking receive (0 = none).
  PeerNet(int cid, int num_client, const std::string& listen, int base_port, double recv_timeout_s);
  ~PeerNet();

  // Opens the listener, connects to every other peer and waits until all of them connected.
  void connect_all(double timeout_s);

  int cid() const { return cid_; }
  int size() const { return P_; }

  // Frame = 8-byte length + payload (client.cpp:265-277), on the connection this peer opened.
  void send_to(int peer, const void* buf, size_t bytes);
  // Next frame from `peer`, in order. Throws on timeout or disconnect.
  std::vector<char> receive(int peer);
  // receive() into a caller buffer of exactly `bytes` bytes.
  void receive_into(int peer, void* buf, size_t bytes);
  // Returns a consumed frame's buffer (from receive) to the pool of the peer it came from.
  void recycle(int peer, std::vector<char>&& m);
  // The next frame from `peer` (exactly `bytes` bytes) into buf. When no part of it has arrived
  // yet, buf is posted to the reader, which reads the payload from the socket straight into it
  // (no copy through a pool buffer: for large frames into pinned memory, that copy is the cost);
  // otherwise the queued frame is copied. Throws on a size mismatch, timeout or disconnect; after
  // a timeout the reader may still be filling buf, so buf must outlive this PeerNet.
  void receive_direct(int peer, void* buf, size_t bytes);
  void broadcast(const void* buf, size_t bytes);

  // client.cpp:313-356: below P*10000 values a broadcast and a sum of the others' buffers in
  // peer order added to this peer's own; above, reduce-scatter then all-gather on the ring.
  void ring_all_reduce_sum(double* buf, long n);
  // The same sum with bit-identical results on every peer: below P*10000 values every peer adds
  // all buffers in peer order 0..P-1; above, the ring (whose chunks are reduced once and then
  // copied) as ring_all_reduce_sum.
  void all_reduce_sum_same(double* buf, long n);
  // out = sum over the peers of their FP32 buffers `in`, each widened to FP64 and added in peer
  // order 0..P-1, bit-identical on every peer; every peer sends its n floats to every other.
  void all_reduce_sum_same_f32(const float* in, double* out, long n);
  // client.cpp:358-411: peer `src` of group `grp` distributes `bytes` bytes (buf is filled on
  // the receivers).
  void all_gather(std::vector<char>& buf, size_t bytes, const std::vector<int>& grp, int src);

  // True once the connection from any peer has closed (the peer ended or failed).
  bool any_closed();
  // Every peer has reached this point (one empty frame to and from each peer).
  void barrier();

  double comm_seconds() const { return comm_s_; }
  void add_comm_seconds(double s) { comm_s_ += s; }
  void reset_comm() { comm_s_ = 0.0; }

 private:
  struct Inbox {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::vector<char>> q;
    std::atomic<int> n{0};
    bool closed = false;
    // frame buffers handed back after use: the reader fills them again without zeroing or
    // faulting in fresh pages (large frames otherwise cost a memset and a page fault per 4 KB)
    std::vector<std::vector<char>> pool;
    // receive_direct: a caller buffer posted for the next frame, which the reader then reads
    // straight from the socket into it (posted_done when complete); reading: a frame is being
    // read into a pool buffer
    char* posted = nullptr;
    size_t posted_bytes = 0;
    bool posted_done = false;
    bool reading = false;
  };
  void accept_loop();
  void reader_loop(int fd, int from);

  int cid_, P_;
  std::string listen_;
  int base_port_;
  double timeout_s_;
  int listen_fd_ = -1;
  std::vector<int> out_fd_;
  std::vector<Inbox*> inbox_;
  std::thread acceptor_;
  std::vector<std::thread> readers_;
  std::vector<int> in_fds_;
  std::mutex mu_;
  std::atomic<int> connected_in_{0};
  std::atomic<bool> stop_{false};
  double comm_s_ = 0.0;
};

double now_seconds();

}  // namespace exc
