// Host TCP transport between peer processes (client.cpp:19-411 behaviour, rewritten): one
// listening socket per peer, one outgoing connection to every other peer, a reader thread per
// incoming connection that queues whole frames, and the reference's collectives on top.
#include "net.hpp"

#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <sys/socket.h>
#include <sys/uio.h>
#include <unistd.h>

#include <chrono>
#include <cstring>
#include <stdexcept>

namespace exc {

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

namespace {

constexpr size_t kMaxChunk = size_t(1) << 30;  // client.cpp:13
constexpr double kHotSeconds = 2e-3;            // spin this long after a frame before blocking

void write_all(int fd, const void* p, size_t n) {
  const char* c = (const char*)p;
  while (n > 0) {
    ssize_t k = ::send(fd, c, std::min(n, kMaxChunk), MSG_NOSIGNAL);
    if (k < 0) {
      if (errno == EINTR) continue;
      throw std::runtime_error(std::string("send failed: ") + strerror(errno));
    }
    c += k;
    n -= (size_t)k;
  }
}

bool read_all(int fd, void* p, size_t n) {
  char* c = (char*)p;
  while (n > 0) {
    ssize_t k = ::recv(fd, c, std::min(n, kMaxChunk), MSG_WAITALL);
    if (k == 0) return false;
    if (k < 0) {
      if (errno == EINTR) continue;
      return false;
    }
    c += k;
    n -= (size_t)k;
  }
  return true;
}

// One frame (8-byte length + payload) in as few sendmsg calls as the socket takes, so the
// receiver wakes once per frame instead of once for the length and again for the payload.
void write_frame(int fd, const void* p, size_t n) {
  uint64_t size = n;
  iovec iov[2] = {{&size, sizeof size}, {const_cast<void*>(p), n}};
  int first = 0, cnt = n ? 2 : 1;
  while (first < cnt) {
    msghdr mh{};
    mh.msg_iov = iov + first;
    mh.msg_iovlen = cnt - first;
    ssize_t k = ::sendmsg(fd, &mh, MSG_NOSIGNAL);
    if (k < 0) {
      if (errno == EINTR) continue;
      throw std::runtime_error(std::string("send failed: ") + strerror(errno));
    }
    size_t left = (size_t)k;
    while (first < cnt && left >= iov[first].iov_len) left -= iov[first++].iov_len;
    if (first < cnt) {
      iov[first].iov_base = (char*)iov[first].iov_base + left;
      iov[first].iov_len -= left;
    }
  }
}

// Reads exactly n bytes. While `hot_until` lies ahead the socket is polled without sleeping
// (frames of the per-column all-reduce follow each other within tens of microseconds, and a
// thread woken from a blocking recv adds several microseconds to each); after it, the thread
// blocks in poll() with a 100 ms period so that `stop` is noticed.
bool read_spin(int fd, void* p, size_t n, double hot_until, const std::atomic<bool>& stop) {
  char* c = (char*)p;
  while (n > 0) {
    ssize_t k = ::recv(fd, c, std::min(n, kMaxChunk), MSG_DONTWAIT);
    if (k > 0) {
      c += k;
      n -= (size_t)k;
      continue;
    }
    if (k == 0) return false;
    if (errno == EINTR) continue;
    if (errno != EAGAIN && errno != EWOULDBLOCK) return false;
    if (stop.load(std::memory_order_relaxed)) return false;
    if (now_seconds() < hot_until) continue;
    pollfd pf{fd, POLLIN, 0};
    ::poll(&pf, 1, 100);
  }
  return true;
}

void tune(int fd) {
  int one = 1;
  setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof one);
  int buf = 8 << 20;
  setsockopt(fd, SOL_SOCKET, SO_SNDBUF, &buf, sizeof buf);
  setsockopt(fd, SOL_SOCKET, SO_RCVBUF, &buf, sizeof buf);
}

std::string peer_ip(const std::string& base, int j) {
  size_t dot = base.rfind('.');
  int last = std::stoi(base.substr(dot + 1));
  return base.substr(0, dot + 1) + std::to_string(last + j);
}

}  // namespace

PeerNet::PeerNet(int cid, int P, const std::string& listen, int base_port, double timeout_s)
    : cid_(cid), P_(P), listen_(listen), base_port_(base_port), timeout_s_(timeout_s) {
  out_fd_.assign(P, -1);
  for (int i = 0; i < P; i++) inbox_.push_back(new Inbox());
}

PeerNet::~PeerNet() {
  stop_ = true;
  for (int fd : out_fd_)
    if (fd >= 0) ::shutdown(fd, SHUT_RDWR), ::close(fd);
  if (listen_fd_ >= 0) ::shutdown(listen_fd_, SHUT_RDWR), ::close(listen_fd_);
  {
    std::lock_guard<std::mutex> g(mu_);
    for (int fd : in_fds_) ::shutdown(fd, SHUT_RDWR);
  }
  if (acceptor_.joinable()) acceptor_.join();
  for (auto& t : readers_)
    if (t.joinable()) t.join();
  {
    std::lock_guard<std::mutex> g(mu_);
    for (int fd : in_fds_) ::close(fd);
  }
  for (auto* b : inbox_) delete b;
}

void PeerNet::accept_loop() {
  for (int k = 0; k < P_ - 1 && !stop_; k++) {
    sockaddr_in a{};
    socklen_t len = sizeof a;
    int fd = ::accept(listen_fd_, (sockaddr*)&a, &len);
    if (fd < 0) {
      if (stop_) return;
      k--;
      continue;
    }
    tune(fd);
    int from = -1;
    if (!read_all(fd, &from, sizeof from) || from < 0 || from >= P_) {
      ::close(fd);
      k--;
      continue;
    }
    {
      std::lock_guard<std::mutex> g(mu_);
      in_fds_.push_back(fd);
      readers_.emplace_back(&PeerNet::reader_loop, this, fd, from);
    }
    connected_in_++;
  }
}

void PeerNet::reader_loop(int fd, int from) {
  Inbox& box = *inbox_[from];
  double hot_until = 0.0;
  while (!stop_) {
    uint64_t size = 0;
    if (!read_spin(fd, &size, sizeof size, hot_until, stop_)) break;
    std::vector<char> msg;
    char* direct = nullptr;
    {
      std::lock_guard<std::mutex> g(box.mu);
      if (box.posted && box.q.empty() && box.posted_bytes == size) {  // the next frame, posted for
        direct = box.posted;
        box.posted = nullptr;
        // a receive that gives up (time-out) while this read is in flight leaves it running; the
        // flag then sends the next receive to the queue instead of posting another buffer that
        // this read's completion would be taken for
        box.reading = true;
      } else {
        box.reading = true;
        if (!box.pool.empty()) {
          msg = std::move(box.pool.back());
          box.pool.pop_back();
        }
      }
    }
    if (direct) {
      if (size && !read_spin(fd, direct, size, now_seconds() + 1.0, stop_)) break;
      hot_until = now_seconds() + kHotSeconds;
      {
        std::lock_guard<std::mutex> g(box.mu);
        box.posted_done = true;
        box.reading = false;
      }
      box.cv.notify_all();
      continue;
    }
    msg.resize(size);  // a pooled buffer at least this large only changes its length
    if (size && !read_spin(fd, msg.data(), size, now_seconds() + 1.0, stop_)) break;
    hot_until = now_seconds() + kHotSeconds;
    {
      std::lock_guard<std::mutex> g(box.mu);
      box.q.push_back(std::move(msg));
      box.n++;
      box.reading = false;
    }
    box.cv.notify_all();
  }
  std::lock_guard<std::mutex> g(box.mu);
  box.closed = true;
  box.cv.notify_all();
}

void PeerNet::connect_all(double timeout_s) {
  listen_fd_ = ::socket(AF_INET, SOCK_STREAM, 0);
  if (listen_fd_ < 0) throw std::runtime_error("socket failed");
  int one = 1;
  setsockopt(listen_fd_, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
  sockaddr_in a{};
  a.sin_family = AF_INET;
  a.sin_port = htons(base_port_ + cid_);
  a.sin_addr.s_addr = INADDR_ANY;
  if (::bind(listen_fd_, (sockaddr*)&a, sizeof a) != 0)
    throw std::runtime_error("bind failed on port " + std::to_string(base_port_ + cid_) + ": " + strerror(errno));
  if (::listen(listen_fd_, 128) != 0) throw std::runtime_error("listen failed");
  acceptor_ = std::thread(&PeerNet::accept_loop, this);
  const double t0 = now_seconds();
  for (int j = 0; j < P_; j++) {
    if (j == cid_) continue;
    sockaddr_in s{};
    s.sin_family = AF_INET;
    s.sin_port = htons(base_port_ + j);
    s.sin_addr.s_addr = inet_addr(peer_ip(listen_, j).c_str());
    while (true) {
      int fd = ::socket(AF_INET, SOCK_STREAM, 0);
      if (::connect(fd, (sockaddr*)&s, sizeof s) == 0) {
        tune(fd);
        out_fd_[j] = fd;
        write_all(fd, &cid_, sizeof cid_);
        break;
      }
      ::close(fd);
      if (now_seconds() - t0 > timeout_s) throw std::runtime_error("could not connect to peer " + std::to_string(j));
      usleep(2000);
    }
  }
  while (connected_in_.load() < P_ - 1) {
    if (now_seconds() - t0 > timeout_s) throw std::runtime_error("peers did not connect in time");
    usleep(200);
  }
}

void PeerNet::send_to(int peer, const void* buf, size_t bytes) {
  write_frame(out_fd_[peer], buf, bytes);
}

std::vector<char> PeerNet::receive(int peer) {
  Inbox& box = *inbox_[peer];
  // short spin first: per-column all-reduce messages arrive within microseconds
  const double t0 = now_seconds();
  while (box.n.load(std::memory_order_acquire) == 0 && now_seconds() - t0 < 2e-4) {
  }
  std::unique_lock<std::mutex> lk(box.mu);
  auto ready = [&] { return !box.q.empty() || box.closed; };
  if (timeout_s_ > 0) {
    if (!box.cv.wait_for(lk, std::chrono::duration<double>(timeout_s_), ready))
      throw std::runtime_error("receive from peer " + std::to_string(peer) + " timed out");
  } else {
    box.cv.wait(lk, ready);
  }
  if (box.q.empty()) throw std::runtime_error("peer " + std::to_string(peer) + " disconnected");
  std::vector<char> m = std::move(box.q.front());
  box.q.pop_front();
  box.n--;
  return m;
}

void PeerNet::recycle(int peer, std::vector<char>&& m) {
  if (m.capacity() < (size_t(1) << 16)) return;  // small frames are cheap to allocate
  Inbox& box = *inbox_[peer];
  std::lock_guard<std::mutex> g(box.mu);
  if (box.pool.size() < 4) box.pool.push_back(std::move(m));
}

void PeerNet::receive_direct(int peer, void* buf, size_t bytes) {
  Inbox& box = *inbox_[peer];
  std::unique_lock<std::mutex> lk(box.mu);
  if (!box.q.empty() || box.reading || box.closed) {  // the frame is (being) queued: copy it
    lk.unlock();
    receive_into(peer, buf, bytes);
    return;
  }
  box.posted = (char*)buf;
  box.posted_bytes = bytes;
  box.posted_done = false;
  // done when the reader filled buf; a frame that went to the queue instead (another size) or a
  // closed connection ends the wait too
  auto ready = [&] { return box.posted_done || !box.q.empty() || box.closed; };
  bool ok = true;
  if (timeout_s_ > 0) ok = box.cv.wait_for(lk, std::chrono::duration<double>(timeout_s_), ready);
  else box.cv.wait(lk, ready);
  if (box.posted_done) {
    box.posted_done = false;
    return;
  }
  box.posted = nullptr;  // not taken yet: withdrawn
  if (!ok) throw std::runtime_error("receive from peer " + std::to_string(peer) + " timed out");
  if (!box.q.empty()) {
    lk.unlock();
    receive_into(peer, buf, bytes);  // throws on the size mismatch that queued it
    return;
  }
  throw std::runtime_error("peer " + std::to_string(peer) + " disconnected");
}

void PeerNet::receive_into(int peer, void* buf, size_t bytes) {
  std::vector<char> m = receive(peer);
  if (m.size() != bytes)
    throw std::runtime_error("message from peer " + std::to_string(peer) + " has " + std::to_string(m.size()) +
                             " bytes, expected " + std::to_string(bytes));
  if (bytes) memcpy(buf, m.data(), bytes);
  recycle(peer, std::move(m));
}

bool PeerNet::any_closed() {
  for (int j = 0; j < P_; j++) {
    if (j == cid_) continue;
    std::lock_guard<std::mutex> g(inbox_[j]->mu);
    if (inbox_[j]->closed) return true;
  }
  return false;
}

void PeerNet::barrier() {
  broadcast(nullptr, 0);
  for (int j = 0; j < P_; j++)
    if (j != cid_) receive_into(j, nullptr, 0);
}

void PeerNet::broadcast(const void* buf, size_t bytes) {
  for (int j = 0; j < P_; j++)
    if (j != cid_) send_to(j, buf, bytes);
}

void PeerNet::ring_all_reduce_sum(double* buf, long n) {
  const double t0 = now_seconds();
  if (n < (long)P_ * 10000) {
    broadcast(buf, n * sizeof(double));
    for (int j = 0; j < P_; j++) {
      if (j == cid_) continue;
      std::vector<char> m = receive(j);
      const double* x = (const double*)m.data();
      for (long q = 0; q < n; q++) buf[q] = buf[q] + x[q];
      recycle(j, std::move(m));
    }
  } else {
    std::vector<long> split(P_ + 1, 0);
    for (int i = 0; i < P_; i++) split[i + 1] = n / P_;
    for (int i = 1; i <= (int)(n % P_); i++) split[i] += 1;
    for (int i = 2; i <= P_; i++) split[i] += split[i - 1];
    const int next = (cid_ + 1) % P_, front = (cid_ - 1 + P_) % P_;
    for (int s = 0; s < P_ - 1; s++) {
      int di = (cid_ - s + P_) % P_;
      send_to(next, buf + split[di], (split[di + 1] - split[di]) * sizeof(double));
      std::vector<char> m = receive(front);
      di = (di - 1 + P_) % P_;
      const double* x = (const double*)m.data();
      for (long q = 0; q < split[di + 1] - split[di]; q++) buf[split[di] + q] = buf[split[di] + q] + x[q];
      recycle(front, std::move(m));
    }
    for (int s = 0; s < P_ - 1; s++) {
      int di = (cid_ + 1 - s + P_) % P_;
      send_to(next, buf + split[di], (split[di + 1] - split[di]) * sizeof(double));
      std::vector<char> m = receive(front);
      di = (di - 1 + P_) % P_;
      memcpy(buf + split[di], m.data(), (split[di + 1] - split[di]) * sizeof(double));
      recycle(front, std::move(m));
    }
  }
  comm_s_ += now_seconds() - t0;
}

void PeerNet::all_reduce_sum_same(double* buf, long n) {
  if (n >= (long)P_ * 10000 || P_ == 1) {
    ring_all_reduce_sum(buf, n);
    return;
  }
  const double t0 = now_seconds();
  broadcast(buf, n * sizeof(double));
  std::vector<std::vector<char>> got(P_);
  for (int j = 0; j < P_; j++)
    if (j != cid_) got[j] = receive(j);
  std::vector<double> own(buf, buf + n);
  for (long q = 0; q < n; q++) {
    double acc = 0.0;
    for (int j = 0; j < P_; j++) acc = acc + (j == cid_ ? own[q] : ((const double*)got[j].data())[q]);
    buf[q] = acc;
  }
  for (int j = 0; j < P_; j++)
    if (j != cid_) recycle(j, std::move(got[j]));
  comm_s_ += now_seconds() - t0;
}

void PeerNet::all_reduce_sum_same_f32(const float* in, double* out, long n) {
  const double t0 = now_seconds();
  broadcast(in, n * sizeof(float));
  std::vector<std::vector<char>> got(P_);
  for (int j = 0; j < P_; j++)
    if (j != cid_) {
      got[j] = receive(j);
      if (got[j].size() != (size_t)n * sizeof(float))
        throw std::runtime_error("all_reduce_sum_same_f32: frame of the wrong size from peer " + std::to_string(j));
    }
  for (long q = 0; q < n; q++) {
    double acc = 0.0;
    for (int j = 0; j < P_; j++) acc = acc + (double)(j == cid_ ? in[q] : ((const float*)got[j].data())[q]);
    out[q] = acc;
  }
  for (int j = 0; j < P_; j++)
    if (j != cid_) recycle(j, std::move(got[j]));
  comm_s_ += now_seconds() - t0;
}

void PeerNet::all_gather(std::vector<char>& buf, size_t bytes, const std::vector<int>& grp, int src) {
  const double t0 = now_seconds();
  const int nc = (int)grp.size();
  if (bytes <= (size_t)nc || nc == 2) {
    if (cid_ == src) {
      for (int j : grp)
        if (j != cid_) send_to(j, buf.data(), bytes);
    } else {
      buf = receive(src);
    }
  } else {
    std::vector<size_t> split(nc, 0);
    for (int i = 0; i < nc - 1; i++) split[i + 1] = bytes / (nc - 1);
    for (int i = 1; i <= (int)(bytes % (nc - 1)); i++) split[i] += 1;
    for (int i = 2; i <= nc - 1; i++) split[i] += split[i - 1];
    if (cid_ != src) buf.assign(bytes, 0);
    int fake = 0, pre = -1, nxt = -1, counter = 0;
    for (int i = 0; i < nc; i++) {
      if (grp[i] == src) continue;
      if (cid_ == src) {
        send_to(grp[i], buf.data() + split[counter], split[counter + 1] - split[counter]);
      } else if (cid_ == grp[i]) {
        fake = counter;
        pre = grp[(i - 1 + nc) % nc] != src ? grp[(i - 1 + nc) % nc] : grp[(i - 2 + nc) % nc];
        nxt = grp[(i + 1) % nc] != src ? grp[(i + 1) % nc] : grp[(i + 2) % nc];
        std::vector<char> m = receive(src);
        memcpy(buf.data() + split[counter], m.data(), split[counter + 1] - split[counter]);
      }
      counter++;
    }
    if (cid_ != src) {
      for (int s = 0; s < nc - 2; s++) {
        int di = (fake - s + nc - 1) % (nc - 1);
        send_to(nxt, buf.data() + split[di], split[di + 1] - split[di]);
        std::vector<char> m = receive(pre);
        di = (di - 1 + nc - 1) % (nc - 1);
        memcpy(buf.data() + split[di], m.data(), split[di + 1] - split[di]);
      }
    }
  }
  comm_s_ += now_seconds() - t0;
}

}  // namespace exc
