// DecentralizedClient on the GPU: the stage order of client.cpp:447-1213 (meta and seeds, local
// QR, Givens masks, re-partition, global QR, Ralha bidiagonalization, bidiagonal SVD, combine and
// unmask), every matrix stage through libexcalibur_amd's C ABI on device-resident data, the peer
// exchanges on host TCP (net.cpp).
#include "client.hpp"

#include <hip/hip_runtime.h>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>
#include <tuple>
#include <thread>
#include <functional>
#include <exception>
#include <memory>

#include "exc_common.hpp"
#include "excalibur_hip.h"
#include "net.hpp"
#include "gemm.hpp"

namespace exc {

void ck(int rc, const char* what) {
  if (rc != EXC_OK) throw std::runtime_error(std::string(what) + ": " + exc_last_error());
}

void ckh(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer of doubles (or bytes), zero-initialized.
struct DBuf {
  double* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { free_all(); }
  size_t cap = 0;
  // Reuses the allocation when it is large enough (repeated runs allocate nothing); the first
  // `count` elements are zeroed unless the caller overwrites all of them before any read (zero =
  // false: a full output of a beta = 0 product, or a workspace whose users initialize it).
  void alloc(size_t count, hipStream_t s, bool zero = true) {
    if (!p || cap < count) {
      free_all();
      void* q = nullptr;
      ck(exc_malloc(&q, std::max<size_t>(count, 1) * sizeof(double)), "device allocation");
      p = (double*)q;
      cap = std::max<size_t>(count, 1);
    }
    n = count;
    if (zero) ckh(hipMemsetAsync(p, 0, std::max<size_t>(count, 1) * sizeof(double), s), "memset");
  }
  void release() { n = 0; }  // keeps the memory for the next run
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(cap, o.cap);
  }
  void free_all() {
    if (p) exc_free(p);
    p = nullptr;
    n = cap = 0;
  }
};

// Pinned host staging buffer (reused across runs when large enough).
struct HBuf {
  double* p = nullptr;
  size_t n = 0, cap = 0;
  HBuf() = default;
  HBuf(const HBuf&) = delete;
  HBuf& operator=(const HBuf&) = delete;
  ~HBuf() {
    if (p) exc_host_free(p);
  }
  void alloc(size_t count, bool zero = true) {
    if (!p || cap < count) {
      if (p) exc_host_free(p);
      void* q = nullptr;
      ck(exc_host_alloc(&q, std::max<size_t>(count, 1) * sizeof(double)), "pinned allocation");
      p = (double*)q;
      cap = std::max<size_t>(count, 1);
    }
    n = count;
    if (zero) memset(p, 0, std::max<size_t>(count, 1) * sizeof(double));
  }
};

struct ClientDevice {
  hipStream_t s = nullptr;
  DBuf X;   // raw data (m x local_n); Q_i^T of the local QR in place afterwards
  DBuf X0;  // pristine input kept on the device for repeated runs
  DBuf masked, P, Vt, tau_g, Qg, F, U, alpha, beta, house, tau_h, cs, bu, bv, sigma, flip, part, FinalU, tmp,
      FinalVT, msg, g, scal, work, plan, rt, msgd, b2, cnt2;
  HBuf h_masked, h_repart, h_msg, h_U;
  // the mixed-precision form: FP32 copies of the data, the exchanged Grams, the Cholesky factors
  // of the passes, and the host staging of the Grams and of the shared small SVD
  DBuf x32, y32, g32, li32, Gm, L1, L2, L3, Lt, Uf, idm, pk, pacc, pstage, L3i, work2, probe, Pm;
  HBuf h_gram, h_tri32, h_pstage, h_probe;
  hipStream_t s2 = nullptr;  // side stream of the mixed form (work that overlaps the main chain)
  // the results as the host members hold them (pinned, so the copies run at the link's rate)
  HBuf r_sigma, r_FinalU, r_FinalVT, r_alpha, r_beta;
  void* link = nullptr;  // message link of the one-launch Ralha loop (coherent pinned memory)
  int link_m = 0;
  // Device-memory message exchange among peers on one device: -1 undecided, 0 off, 1 on.
  int boxes_state = -1;
  // Whether this run uses the one-launch bidiagonalization loop, decided by all peers together.
  int loop_ok = 0;
  void* box = nullptr;                // this peer's box
  std::vector<void*> boxes;           // every peer's box, opened through IPC (own box at cid)
  std::vector<int> box_nwg;           // every peer's loop grid size
  // Bulk exchange among peers on one device: this peer's outbox (device memory, exported by IPC)
  // and every other peer's outbox as last opened here (by handle).
  int colocated = 0;
  int samedev = 0;  // every peer runs on this peer's device (whether or not they exchange there)
  double* outbox = nullptr;
  size_t outbox_cap = 0;  // doubles
  char outbox_handle[EXC_IPC_HANDLE_BYTES] = {0};
  std::vector<std::string> peer_handle;
  std::vector<void*> peer_outbox;
  ~ClientDevice() {
    for (void* q : peer_outbox)
      if (q) exc_ipc_close(q);
    if (outbox) exc_free(outbox);
    for (size_t q = 0; q < boxes.size(); q++)
      if (boxes[q] && boxes[q] != box) exc_ipc_close(boxes[q]);
    if (box) exc_free(box);
    if (link) exc_ralha_link_free(link);
    if (s2) hipStreamDestroy(s2);
    if (s) hipStreamDestroy(s);
  }
};

// Bulk transfers between peers on one device. The sender packs the data into its outbox and
// sends a note (outbox handle, offset, count) over the network; the receiver copies from the
// outbox, mapped into its process, device to device, then acknowledges, after which the sender
// may reuse its outbox.
struct DmaNote {
  char handle[EXC_IPC_HANDLE_BYTES];
  uint64_t offset, count;  // doubles
};

double* outbox_reserve(ClientDevice& D, size_t count) {
  if (D.outbox_cap < count) {
    // every earlier transfer from the outbox has been acknowledged
    if (D.outbox) ck(exc_free(D.outbox), "outbox");
    D.outbox = nullptr;
    void* q = nullptr;
    ck(exc_malloc(&q, count * sizeof(double)), "outbox");
    D.outbox = (double*)q;
    D.outbox_cap = count;
    ck(exc_ipc_handle(D.outbox, D.outbox_handle), "outbox handle");
  }
  return D.outbox;
}

void dma_offer(PeerNet& net, ClientDevice& D, int to, size_t offset, size_t count) {
  DmaNote n;
  memcpy(n.handle, D.outbox_handle, sizeof n.handle);
  n.offset = offset;
  n.count = count;
  net.send_to(to, &n, sizeof n);
}

const double* dma_take(PeerNet& net, ClientDevice& D, int from, size_t count) {
  DmaNote n;
  net.receive_into(from, &n, sizeof n);
  if (n.count != count) throw std::runtime_error("device transfer of unexpected size");
  std::string key(n.handle, sizeof n.handle);
  if (D.peer_handle[from] != key) {
    if (D.peer_outbox[from]) exc_ipc_close(D.peer_outbox[from]);
    D.peer_outbox[from] = nullptr;
    ck(exc_ipc_open(n.handle, &D.peer_outbox[from]), "open a peer's outbox");
    D.peer_handle[from] = key;
  }
  return (const double*)D.peer_outbox[from] + n.offset;
}

void dma_ack(PeerNet& net, int to) { net.send_to(to, nullptr, 0); }
void dma_wait_ack(PeerNet& net, int from) { net.receive_into(from, nullptr, 0); }

}  // namespace exc

using exc::ck;
using exc::ckh;
using exc::DBuf;
using exc::HBuf;
using exc::now_seconds;

namespace {

// The Givens plans of one run (bbtsvd.cpp:727-779). Each is a function of (K, mask, seed) alone,
// since the plan reseeds the C library's generator itself, so all of them are drawn on a host
// thread while the GPU runs the local QR, into pinned buffers that the sweeps upload without a
// stream synchronization. The thread is the only user of rand() until it is joined: run() draws
// both seeds before it starts, and nothing after.
struct PlanJob {
  long K = 0;
  int mask = 0;
  unsigned seed = 0;
  size_t pair_words = 0, count = 0;  // doubles: pairs (as int32) then c/s
  HBuf host;
  DBuf dev;
  bool uploaded = false;  // dev holds the plan (uploaded once per run, by the first sweep)
};

class RunPlans {
 public:
  ~RunPlans() { join(); }
  void start(const std::vector<std::tuple<long, int, unsigned>>& want) {
    for (const auto& w : want) {
      bool dup = false;
      for (auto& j : jobs_) dup = dup || (j->K == std::get<0>(w) && j->mask == std::get<1>(w) && j->seed == std::get<2>(w));
      if (dup || std::get<0>(w) <= 1 || std::get<0>(w) > 10000) continue;
      auto j = std::make_unique<PlanJob>();
      j->K = std::get<0>(w);
      j->mask = std::get<1>(w);
      j->seed = std::get<2>(w);
      const long half = exc_givens_half(j->K);
      j->pair_words = ((size_t)32 * half * 2 * sizeof(int32_t) + 7) / 8;
      j->count = j->pair_words + (size_t)32 * half * 2;
      j->host.alloc(j->count);
      jobs_.push_back(std::move(j));
    }
    th_ = std::thread([this] {
      for (auto& j : jobs_) {
        if (exc_givens_plan(j->K, 32, j->mask, j->seed, (int32_t*)j->host.p, j->host.p + j->pair_words) != EXC_OK)
          failed_ = true;
      }
    });
  }
  PlanJob& get(long K, int mask, unsigned seed) {
    join();
    if (failed_) throw std::runtime_error("givens plan: " + std::string(exc_last_error()));
    for (auto& j : jobs_)
      if (j->K == K && j->mask == mask && j->seed == seed) return *j;
    throw std::runtime_error("givens plan not prepared");
  }
  // The plan thread is the only user of the C library's generator while it runs; callers that draw
  // plans themselves wait for it first.
  void join() {
    if (th_.joinable()) th_.join();
  }

 private:
  std::vector<std::unique_ptr<PlanJob>> jobs_;
  std::thread th_;
  bool failed_ = false;
};

// The plan of (K, mask, seed) on the device, uploaded by the first sweep that needs it; the copy
// is ordered before that sweep on s and overwrites the whole buffer, so it is not zeroed first.
PlanJob& device_plan(RunPlans& plans, long K, int mask, unsigned seed, hipStream_t s) {
  PlanJob& p = plans.get(K, mask, seed);
  if (!p.uploaded) {
    p.dev.alloc(p.count, s, false);
    ckh(hipMemcpyAsync(p.dev.p, p.host.p, p.count * sizeof(double), hipMemcpyHostToDevice, s), "plan");
    p.uploaded = true;
  }
  return p;
}

// block_shift_mask(m, n, 32, x, mask, is_left, seed) for K <= 10000 (bbtsvd.cpp:533-546). The
// sweep is enqueued on s behind an asynchronous upload of the plan; the caller synchronizes.
void givens_mask(double* d_x, long m, long n, int mask, int is_left, unsigned seed, RunPlans& plans, hipStream_t s) {
  const long K = is_left ? m : n;
  if (K <= 1) return;
  if (K > 10000) {  // the virtual-matrix mask (bbtsvd.cpp:548-668); draws its own plans
    plans.join();
    DBuf w;
    w.alloc((exc_block_shift_mask_workspace(m, n, 32, is_left) + 7) / 8, s, false);
    ck(exc_block_shift_mask(d_x, m, n, 32, mask, is_left, seed, w.p, w.n * 8, s), "virtual-matrix mask");
    return;
  }
  PlanJob& p = device_plan(plans, K, mask, seed, s);
  ck(exc_givens_apply(d_x, m, n, (const int32_t*)p.dev.p, p.dev.p + p.pair_words, 32, mask, is_left, s),
     "givens apply");
}

// The same sweep on FP32 data (the rotation values stay FP64 in the plan, rounded per use).
void givens_mask_f32(float* d_x, long m, long n, int mask, int is_left, unsigned seed, RunPlans& plans,
                     hipStream_t s) {
  const long K = is_left ? m : n;
  if (K <= 1) return;
  if (K > 10000) throw std::runtime_error("Givens masks of more than 10000 lines are not available in FP32");
  PlanJob& p = device_plan(plans, K, mask, seed, s);
  ck(exc_givens_apply_f32(d_x, m, n, (const int32_t*)p.dev.p, p.dev.p + p.pair_words, 32, mask, is_left, s),
     "givens apply");
}

}  // namespace

DecentralizedClient::DecentralizedClient(int cid, int nclient, int mode, size_t topk, size_t drow, size_t dcol,
                                         int port, std::string listen_, std::string dpath_, bool memmap,
                                         bool evaluate_, std::string log_dir_, unsigned int seed, int opt)
    : DecentralizedClient(cid, nclient, mode, topk, drow, dcol, port, listen_, dpath_, memmap, evaluate_, log_dir_,
                          seed, opt, ClientOptions()) {}

DecentralizedClient::DecentralizedClient(int cid, int nclient, int mode, size_t topk, size_t drow, size_t dcol,
                                         int port, std::string listen_, std::string dpath_, bool memmap,
                                         bool evaluate_, std::string log_dir_, unsigned int seed, int opt,
                                         const ClientOptions& options)
    : svd_mode(mode),
      top_k((int)topk),
      client_id(cid),
      num_client(nclient),
      m(drow),
      local_n(dcol),
      opt_control(opt),
      evaluate(evaluate_),
      is_memmap(memmap),
      listen(listen_),
      dpath(dpath_),
      log_dir(log_dir_),
      base_server_port(port),
      opt_(options),
      seed_(seed) {
  if (evaluate) opt_.dump_results = true;
  mkdir(log_dir.c_str(), 0755);
  log("Parameters: cid=" + std::to_string(cid) + " num_client=" + std::to_string(nclient) + " mode=" +
      std::to_string(mode) + " topk=" + std::to_string((long)top_k) + " m=" + std::to_string(m) + " local_n=" +
      std::to_string(local_n) + " port=" + std::to_string(port) + " listen=" + listen + " seed=" +
      std::to_string(seed) + " opt=" + std::to_string(opt) + " form=" + opt_.form);
  dev_.reset(new exc::ClientDevice());
  ckh(hipStreamCreateWithFlags(&dev_->s, hipStreamNonBlocking), "stream");
  load_data();
  const int repeat = std::max(1, opt_.repeat);
  if (repeat > 1 || evaluate) {  // runs overwrite X; the evaluation needs the raw data (raw_x)
    dev_->X0.alloc(m * local_n, dev_->s);
    ckh(hipMemcpyAsync(dev_->X0.p, dev_->X.p, m * local_n * 8, hipMemcpyDeviceToDevice, dev_->s), "copy");
    ckh(hipStreamSynchronize(dev_->s), "copy");
  }
  init_server_and_client();
  const bool dump = opt_.dump_results, keep = opt_.keep_results;
  for (int r = 0; r < repeat; r++) {
    if (r > 0) {  // fresh input on the device, outside the timed region
      dev_->X.alloc(m * local_n, dev_->s);
      ckh(hipMemcpyAsync(dev_->X.p, dev_->X0.p, m * local_n * 8, hipMemcpyDeviceToDevice, dev_->s), "copy");
      ckh(hipStreamSynchronize(dev_->s), "copy");
    }
    double token = 0.0;  // start together
    net_->ring_all_reduce_sum(&token, 1);
    net_->reset_comm();
    exc_prof_enable(opt_.profile ? 1 : 0);
    opt_.dump_results = dump && r == repeat - 1;
    opt_.keep_results = keep && r == repeat - 1;
    run();
  }
}

// The network goes first: its reader threads are joined before the device state (whose pinned
// staging a timed-out direct receive may still be filling) is freed.
DecentralizedClient::~DecentralizedClient() { net_.reset(); }

void DecentralizedClient::log(const std::string& line) {
  std::ofstream f(log_dir + "/Client" + std::to_string(client_id) + ".log", std::ios::app);
  f << line << "\n";
}

void DecentralizedClient::decide_device_exchange(int mi, int ci, bool one_launch) {
  exc::ClientDevice& D = *dev_;
  exc::PeerNet& net = *net_;
  const int P = num_client, cid = client_id;
  // this peer's record: the loop applies, its grid size, and the identity of its device
  struct Rec {
    int usable, nwg, slots;
    char where[160];
  } me{};
  int dev = 0;
  char bus[64] = {0};
  ckh(hipGetDevice(&dev), "device");
  ckh(hipDeviceGetPCIBusId(bus, sizeof bus, dev), "device");
  char host[64] = {0};
  gethostname(host, sizeof host - 1);
  std::string boot;
  {
    std::ifstream f("/proc/sys/kernel/random/boot_id");
    std::getline(f, boot);
  }
  snprintf(me.where, sizeof me.where, "%s|%s|%s", host, boot.c_str(), bus);
  me.usable = (opt_.device_exchange && P <= 16) ? 1 : 0;
  me.nwg = one_launch ? exc_ralha_loop_grid(mi, ci, 0) : 0;
  me.slots = one_launch ? exc_ralha_loop_slots(mi, ci, 0) : 0;
  net.broadcast(&me, sizeof me);
  std::vector<Rec> all(P);
  all[cid] = me;
  for (int j = 0; j < P; j++)
    if (j != cid) net.receive_into(j, &all[j], sizeof(Rec));
  bool samedev = true, usable = true;
  std::vector<int> nwg(P), slots(P);
  for (int j = 0; j < P; j++) {
    samedev = samedev && strncmp(all[j].where, me.where, sizeof me.where) == 0;
    usable = usable && all[j].usable == 1;
    nwg[j] = all[j].nwg;
    slots[j] = all[j].slots;
  }
  // Every peer takes the same decision from the same records. The loops wait on each other's
  // messages, so on one device all of them must be resident together, on TCP as on the boxes.
  D.loop_ok = exc_ralha_loops_colocate(P, nwg.data(), slots.data(), samedev ? 1 : 0);
  if (one_launch && !D.loop_ok)
    log("the peers' bidiagonalization loops do not fit on the device together: one launch per step");
  const bool same = samedev && usable;
  D.samedev = samedev ? 1 : 0;
  D.colocated = (same && P > 1) ? 1 : 0;
  D.peer_handle.assign(P, std::string());
  D.peer_outbox.assign(P, nullptr);
  if (D.colocated) log("bulk exchanges among " + std::to_string(P) + " peers in device memory");
  if (!same || !D.loop_ok) {
    D.boxes_state = 0;
    return;
  }
  // every peer allocates its box and sends the IPC handle to the others
  ck(exc_malloc(&D.box, exc_ralha_box_bytes(mi)), "ralha box");
  char handle[EXC_IPC_HANDLE_BYTES];
  ck(exc_ipc_handle(D.box, handle), "ralha box handle");
  net.broadcast(handle, sizeof handle);
  D.boxes.assign(P, nullptr);
  D.box_nwg.assign(P, 0);
  D.boxes[cid] = D.box;
  int opened = 1;
  for (int j = 0; j < P; j++) {
    D.box_nwg[j] = all[j].nwg;
    if (j == cid) continue;
    char h[EXC_IPC_HANDLE_BYTES];
    net.receive_into(j, h, sizeof h);
    if (opened && exc_ipc_open(h, &D.boxes[j]) != EXC_OK) {
      log(std::string("could not open a peer's box: ") + exc_last_error());
      D.boxes[j] = nullptr;
      opened = 0;
    }
  }
  // all peers switch to the device exchange, or none does
  net.broadcast(&opened, sizeof opened);
  int all_opened = opened;
  for (int j = 0; j < P; j++) {
    if (j == cid) continue;
    int o = 0;
    net.receive_into(j, &o, sizeof o);
    all_opened &= o;
  }
  if (!all_opened) {
    for (int j = 0; j < P; j++)
      if (j != cid && D.boxes[j]) exc_ipc_close(D.boxes[j]);
    D.boxes.clear();
    exc_free(D.box);
    D.box = nullptr;
    D.boxes_state = 0;
    return;
  }
  D.boxes_state = 1;
  log("bidiagonalization messages exchanged in device memory among " + std::to_string(P) + " peers");
}

void DecentralizedClient::load_data() {
  const size_t count = m * local_n;
  hipStream_t s = dev_->s;
  dev_->X.alloc(count, s);
  if (opt_.data) {
    ckh(hipMemcpyAsync(dev_->X.p, opt_.data, count * sizeof(double), hipMemcpyHostToDevice, s), "upload");
  } else if (is_memmap) {
    // -f 1 (utils.cpp:229-252): the file is mapped, not read into a host buffer of its size; it
    // streams to the device through a 64 MiB pinned staging buffer, so host memory stays bounded
    // whatever the block's size
    const std::string path = dpath + "/Client" + std::to_string(client_id) + ".mat";
    const int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("cannot open " + path);
    const off_t size = lseek(fd, 0, SEEK_END);
    if (size < (off_t)(count * sizeof(double))) {
      close(fd);
      throw std::runtime_error(path + " holds fewer than m * local_n doubles");
    }
    log("Load array size " + std::to_string(size / sizeof(double)));
    void* map = mmap(nullptr, count * sizeof(double), PROT_READ, MAP_SHARED, fd, 0);
    close(fd);
    if (map == MAP_FAILED) throw std::runtime_error("cannot map " + path);
    const size_t chunk = (size_t)8 << 20;  // doubles
    HBuf st[2];
    st[0].alloc(std::min(chunk, count), false);
    st[1].alloc(std::min(chunk, count), false);
    for (size_t o = 0, q = 0; o < count; o += chunk, q ^= 1) {
      const size_t n = std::min(chunk, count - o);
      ckh(hipStreamSynchronize(s), "upload");  // the staging buffer q is free again (two in flight)
      memcpy(st[q].p, (const double*)map + o, n * sizeof(double));
      ckh(hipMemcpyAsync(dev_->X.p + o, st[q].p, n * sizeof(double), hipMemcpyHostToDevice, s), "upload");
    }
    ckh(hipStreamSynchronize(s), "upload");
    munmap(map, count * sizeof(double));
  } else {
    const std::string path = dpath + "/Client" + std::to_string(client_id) + ".mat";
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    HBuf h;
    h.alloc(count);
    const size_t got = fread(h.p, sizeof(double), count, f);
    fclose(f);
    if (got != count) throw std::runtime_error(path + " holds fewer than m * local_n doubles");
    ckh(hipMemcpyAsync(dev_->X.p, h.p, count * sizeof(double), hipMemcpyHostToDevice, s), "upload");
    ckh(hipStreamSynchronize(s), "upload");
  }
  ckh(hipStreamSynchronize(s), "upload");
}

void DecentralizedClient::init_server_and_client() {
  net_.reset(new exc::PeerNet(client_id, num_client, listen, base_server_port, opt_.recv_timeout_s));
  my_next_client = (client_id + 1) % num_client;  // client.cpp:163-164
  my_front_client = (client_id - 1 + num_client) % num_client;
  net_->connect_all(std::max(60.0, opt_.recv_timeout_s));
  // this peer's share of the bidiagonalization columns (client.cpp:613-620)
  const size_t c = m / num_client + (client_id < (int)(m % num_client) ? 1 : 0);
  decide_device_exchange((int)m, (int)c, opt_.ralha_one_launch && exc_ralha_loop_fits((int)m, (int)c, 0) == 1);
}

void DecentralizedClient::run() {
  hipStream_t s = dev_->s;
  exc::ClientDevice& D = *dev_;
  exc::PeerNet& net = *net_;
  const int P = num_client, cid = client_id;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  // srand(seed) of client.cpp:196. The reference makes no rand() call between the constructor and
  // this point, but the HIP runtime's initialization (stream, allocations, load) may touch the C
  // library's generator, so the seed is set here, after it, to keep the reference's draws.
  srand(seed_);
  log("All Clients Connected, DecFedSVD Begin");
  const double t_run = now_seconds();
  double t0 = t_run;

  bool dense_mask_opt = true;  // client.cpp:480-488
  if (opt_control == 0 || opt_control == 2) dense_mask_opt = false;
  if (!dense_mask_opt) throw std::runtime_error("opt 0/2 (dense random orthogonal masks) is not supported");
  if (svd_mode == 4) throw std::runtime_error("mode 4 (linear regression) is not supported");
  if (opt_.form != "fp64" && opt_.form != "mixed")
    throw std::runtime_error("form '" + opt_.form + "' is not available (fp64 or mixed)");
  const bool compute_u = svd_mode != 3;
  const bool compute_v = svd_mode != 2;
  bidiag_order = (int)m;  // global_n once the tall-and-skinny branch has run its bidiagonal SVD

  // 1) meta and seeds (client.cpp:494-531)
  ClientMeta cm{m, local_n, cid, 0};
  if (cid == 0) {
    public_seed = (unsigned)rand();
    cm.shared_seed = public_seed;
  }
  net.broadcast(&cm, sizeof cm);
  const double thr = 5.0;
  auto reduce_n = [&](size_t ln) { return ((double)(ln / m)) > thr ? std::min(ln, m) : ln; };
  global_n = local_n;
  reduced_global_n = reduce_n(local_n);
  for (int j = 0; j < P; j++) {
    if (j == cid) continue;
    ClientMeta rm;
    net.receive_into(j, &rm, sizeof rm);
    local_ns[rm.cid] = rm.local_n;
    reduced_local_ns[rm.cid] = reduce_n(rm.local_n);
    global_n += rm.local_n;
    reduced_global_n += reduced_local_ns[rm.cid];
    if (rm.cid == 0) public_seed = rm.shared_seed;
  }
  local_ns[cid] = local_n;
  reduced_local_ns[cid] = reduce_n(local_n);
  for (int i = 0; i <= P; i++) {
    n_pos[i] = 0;
    reduced_n_pos[i] = 0;
    for (int j = 0; j < std::min(i, P); j++) {
      n_pos[i] += local_ns[j];
      reduced_n_pos[i] += reduced_local_ns[j];
    }
  }
  private_seed = (unsigned)rand();
  time_init = now_seconds() - t0;
  if (m < 3 || global_n < 3) throw std::runtime_error("m and global_n must be at least 3");
  if (opt_.form == "mixed" && m > global_n)
    throw std::runtime_error("the mixed form needs m <= global_n (its m x m Gram matrix is singular otherwise)");
  if (m > global_n) {
    const int k = run_tall(compute_u, compute_v);
    finish_run(k, compute_u, compute_v, true, t_run);
    return;
  }
  if (opt_.form == "mixed") {
    const int k = run_mixed(compute_u, compute_v);
    finish_run(k, compute_u, compute_v, true, t_run);
    return;
  }

  // 2) local QR (client.cpp:572-589)
  const size_t rl = reduced_local_ns[cid];
  const bool lqr = m < local_n && ((double)(local_n / m)) > thr;
  // the four Givens plans of this run: masks (left K = m, right K = rl) and their removal
  RunPlans plans;
  plans.start({{(long)m, 1, public_seed}, {(long)rl, 1, private_seed}, {(long)m, 0, public_seed},
               {(long)rl, 0, private_seed}});
  double* masked = D.X.p;
  if (lqr) {
    t0 = now_seconds();
    D.masked.alloc(m * m, s);
    const size_t wb = exc_qr_t_workspace((int)m, (long)local_n, 1);
    D.work.alloc((wb + 7) / 8, s, false);  // exc_qr_t writes every part of it before reading it
    ck(exc_qr_t((int)m, (long)local_n, D.X.p, (long)local_n, D.masked.p, compute_v ? D.X.p : nullptr, (long)local_n,
                nullptr, 0, nullptr, nullptr, 0, D.work.p, D.work.n * 8, s),
       "local QR");
    sync();
    masked = D.masked.p;
    time_local_qr = now_seconds() - t0;
    log("Local QR Costs(1) (short-and-wide) " + std::to_string(time_local_qr));
  }

  // 3) masks (client.cpp:592-605)
  t0 = now_seconds();
  givens_mask(masked, (long)m, (long)rl, 1, 1, public_seed, plans, s);
  givens_mask(masked, (long)m, (long)rl, 1, 0, private_seed, plans, s);
  sync();
  time_apply_mask = now_seconds() - t0;
  log("Masking Costs(1) (short-and-wide) " + std::to_string(time_apply_mask));

  // 4) re-partition (client.cpp:613-703)
  std::vector<size_t> m_part(P, m / P), m_part_pos(P + 1, 0);
  for (int i = 0; i < (int)(m % P); i++) m_part[i] += 1;
  for (int i = 0; i < P; i++) m_part_pos[i + 1] = m_part_pos[i] + m_part[i];
  const size_t c = m_part[cid];
  const size_t R = reduced_global_n;
  t0 = now_seconds();
  if (D.colocated) {
    // row slab j of the masked block goes to peer j, device to device
    double* ob = exc::outbox_reserve(D, m * rl);
    ckh(hipMemcpyAsync(ob, masked, m * rl * sizeof(double), hipMemcpyDeviceToDevice, s), "pack");
    sync();
    for (int j = 0; j < P; j++)
      if (j != cid) exc::dma_offer(net, D, j, m_part_pos[j] * rl, m_part[j] * rl);
    D.P.alloc(c * R, s);
    for (int j = 0; j < P; j++) {
      const size_t rj = reduced_local_ns[j], off = reduced_n_pos[j];
      const double* src = j == cid ? masked + m_part_pos[cid] * rl : exc::dma_take(net, D, j, c * rj);
      ckh(hipMemcpy2DAsync(D.P.p + off, R * sizeof(double), src, rj * sizeof(double), rj * sizeof(double), c,
                           hipMemcpyDeviceToDevice, s),
          "re-partition");
    }
    sync();
    for (int j = 0; j < P; j++)
      if (j != cid) exc::dma_ack(net, j);
    for (int j = 0; j < P; j++)
      if (j != cid) exc::dma_wait_ack(net, j);
  } else {
    HBuf& hm = D.h_masked;
    HBuf& hr = D.h_repart;
    hm.alloc(m * rl);
    ckh(hipMemcpyAsync(hm.p, masked, m * rl * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    sync();
    for (int j = 0; j < P; j++)
      if (j != cid) net.send_to(j, hm.p + m_part_pos[j] * rl, m_part[j] * rl * sizeof(double));
    hr.alloc(c * R);
    for (int j = 0; j < P; j++) {
      const size_t rj = reduced_local_ns[j], off = reduced_n_pos[j];
      if (j == cid) {
        for (size_t r = 0; r < c; r++) memcpy(hr.p + r * R + off, hm.p + (m_part_pos[cid] + r) * rl, rj * 8);
      } else {
        std::vector<char> blk = net.receive(j);
        if (blk.size() != c * rj * 8) throw std::runtime_error("re-partition block of unexpected size");
        const double* b = (const double*)blk.data();
        for (size_t r = 0; r < c; r++) memcpy(hr.p + r * R + off, b + r * rj, rj * 8);
      }
    }
    D.P.alloc(c * R, s);
    ckh(hipMemcpyAsync(D.P.p, hr.p, c * R * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
    sync();
  }
  if (!lqr) D.X.release();  // the masked raw data has been re-partitioned
  if (lqr) D.masked.release();
  time_exchange_low_dim = now_seconds() - t0;
  net.add_comm_seconds(time_exchange_low_dim);

  // 5) global QR (client.cpp:706-798)
  t0 = now_seconds();
  D.F.alloc(m * c, s);
  if (R > m) {
    D.Vt.alloc(m * R, s);
    D.tau_g.alloc(m, s);
    DBuf& rt = D.rt;
    DBuf& msgd = D.msgd;
    for (int i = 0; i < P; i++) {
      const long ci = (long)m_part[i], pos = (long)m_part_pos[i], L = (long)R - pos;
      const long len = exc_reflector_message_len(ci, (long)R, pos);
      std::vector<int> grp;
      if (compute_v) for (int j = 0; j < P; j++) grp.push_back(j);
      else for (int j = i; j < P; j++) grp.push_back(j);
      if (i == cid) {
        rt.alloc((size_t)ci * ci, s);
        const size_t wb = exc_qr_t_workspace((int)ci, L, 0);
        D.work.alloc((wb + 7) / 8, s);
        ck(exc_qr_t((int)ci, L, D.P.p + pos, (long)R, rt.p, nullptr, 0, D.Vt.p + pos * (long)R + pos, (long)R,
                    D.tau_g.p + pos, nullptr, 1, D.work.p, D.work.n * 8, s),
           "global QR panel");
        ck(exc_copy2d(ci, ci, rt.p, ci, D.P.p + pos, (long)R, s), "panel R");
        ck(exc_fill2d(ci, L - ci, 0.0, D.P.p + pos + ci, (long)R, s), "panel zero");
        if (D.colocated) {
          double* ob = exc::outbox_reserve(D, len);
          ck(exc_pack_reflectors(ci, (long)R, pos, D.Vt.p, (long)R, D.tau_g.p, ob, s), "pack");
          sync();
          for (int j : grp)
            if (j != cid) exc::dma_offer(net, D, j, 0, len);
          for (int j : grp)
            if (j != cid) exc::dma_wait_ack(net, j);
        } else {
          msgd.alloc(len, s);
          ck(exc_pack_reflectors(ci, (long)R, pos, D.Vt.p, (long)R, D.tau_g.p, msgd.p, s), "pack");
          std::vector<char> hmsg(len * sizeof(double));
          ckh(hipMemcpyAsync(hmsg.data(), msgd.p, len * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
          sync();
          if ((int)grp.size() > 1) net.all_gather(hmsg, len * sizeof(double), grp, cid);
        }
      } else if (compute_v || cid > i) {
        msgd.alloc(len, s);
        if (D.colocated) {
          const double* src = exc::dma_take(net, D, i, len);
          ckh(hipMemcpyAsync(msgd.p, src, len * sizeof(double), hipMemcpyDeviceToDevice, s), "reflectors");
          sync();
          exc::dma_ack(net, i);
        } else {
          std::vector<char> hmsg;
          net.all_gather(hmsg, len * sizeof(double), grp, i);
          if ((long)hmsg.size() != len * 8) throw std::runtime_error("global QR message of unexpected size");
          ckh(hipMemcpyAsync(msgd.p, hmsg.data(), len * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
        }
        ck(exc_unpack_reflectors(ci, (long)R, pos, msgd.p, D.Vt.p, (long)R, D.tau_g.p, s), "unpack");
        if (cid > i) {
          const size_t wb = exc_wy_workspace((int)ci, L, (long)c);
          D.work.alloc((wb + 7) / 8, s);
          ck(exc_apply_q_right((long)c, L, (int)ci, D.P.p + pos, (long)R, D.Vt.p + pos * (long)R + pos, (long)R,
                               D.tau_g.p + pos, D.work.p, D.work.n * 8, s),
             "apply panel reflectors");
        }
        sync();
      }
    }
    if (compute_v) {
      D.Qg.alloc(m * R, s);
      const size_t wb = exc_wy_workspace((int)m, (long)R, (long)m);
      D.work.alloc((wb + 7) / 8, s);
      ck(exc_orgqr_t((int)m, (long)R, D.Vt.p, (long)R, D.tau_g.p, D.Qg.p, (long)R, D.work.p, D.work.n * 8, s),
         "global Q");
    }
    sync();
    D.Vt.release();
  }
  // F = R[:, own] (m x c) kept column-major for the Ralha loop: rows of P (c x R) are its columns
  ck(exc_copy2d((long)c, (long)m, D.P.p, (long)R, D.F.p, (long)m, s), "shrink");
  sync();
  D.P.release();
  time_global_qr = now_seconds() - t0;

  // 6) Ralha bidiagonalization (client.cpp:800-910) + lower to upper (947-954)
  ralha_stage((int)m, (int)c, D.loop_ok == 1);
  D.F.release();

  // 7) bidiagonal SVD (client.cpp:963-996)
  const int k = bsvd_stage((int)m);

  // 8) combine and unmask (client.cpp:998-1199)
  t0 = now_seconds();
  if (compute_u) {
    // this peer's rows of FinalU: (bv U)^T = U^T bv^T, with U^T stored row-major (c x m)
    D.FinalU.alloc(m * (size_t)k, s);
    // split-K partials: a c x k output over K = m is too few tiles to fill the device
    D.work.alloc((exc_gemm_workspace((long)c, k, (long)m) + 7) / 8, s, false);
    ck(exc_dgemm(0, 1, (long)c, k, (long)m, 1.0, D.U.p, (long)m, D.bv.p, (long)m, 0.0,
                 D.FinalU.p + m_part_pos[cid] * (size_t)k, k, D.work.p, D.work.n * 8, s),
       "part final U");
    // ring all-gather of the row blocks of FinalU (client.cpp:1062-1096)
    const double tc = now_seconds();
    if (D.colocated) {
      double* ob = exc::outbox_reserve(D, c * (size_t)k);
      ckh(hipMemcpyAsync(ob, D.FinalU.p + m_part_pos[cid] * k, c * (size_t)k * 8, hipMemcpyDeviceToDevice, s),
          "pack");
      sync();
      for (int j = 0; j < P; j++)
        if (j != cid) exc::dma_offer(net, D, j, 0, c * (size_t)k);
      for (int j = 0; j < P; j++) {
        if (j == cid) continue;
        const double* src = exc::dma_take(net, D, j, m_part[j] * (size_t)k);
        ckh(hipMemcpyAsync(D.FinalU.p + m_part_pos[j] * k, src, m_part[j] * (size_t)k * 8, hipMemcpyDeviceToDevice,
                           s),
            "gather");
      }
      sync();
      for (int j = 0; j < P; j++)
        if (j != cid) exc::dma_ack(net, j);
      for (int j = 0; j < P; j++)
        if (j != cid) exc::dma_wait_ack(net, j);
    } else {
      HBuf& hU = D.h_U;
      hU.alloc(m * (size_t)k);
      ckh(hipMemcpyAsync(hU.p + m_part_pos[cid] * k, D.FinalU.p + m_part_pos[cid] * k, c * (size_t)k * 8,
                         hipMemcpyDeviceToHost, s),
          "d2h");
      sync();
      const int next = (cid + 1) % P, front = (cid - 1 + P) % P;
      for (int st = 0; st < P - 1; st++) {
        int di = (cid - st + P) % P;
        net.send_to(next, hU.p + m_part_pos[di] * k, m_part[di] * (size_t)k * 8);
        di = (di - 1 + P) % P;
        net.receive_into(front, hU.p + m_part_pos[di] * k, m_part[di] * (size_t)k * 8);
      }
      ckh(hipMemcpyAsync(D.FinalU.p, hU.p, m * (size_t)k * 8, hipMemcpyHostToDevice, s), "h2d");
      sync();
    }
    net.add_comm_seconds(now_seconds() - tc);
    givens_mask(D.FinalU.p, (long)m, k, 0, 1, public_seed, plans, s);
  }
  D.U.release();
  if (compute_v && m == global_n) {
    // m >= global_n: the reference's tall-and-skinny combine of the right vectors
    // (client.cpp:1113-1154): rows n_pos .. of H (rotations back) bu, transposed and unmasked
    ck(exc_rot_back((int)m, k, D.bu.p, D.cs.p, s), "back rotation");
    const size_t wb = exc_ormqr_house_workspace((int)m, k);
    D.work.alloc((wb + 7) / 8, s);
    ck(exc_ormqr_house_ln((int)m, k, D.house.p, D.tau_h.p, D.bu.p, k, D.work.p, D.work.n * 8, s), "ormqr house");
    D.tmp.alloc((size_t)k * local_n, s);
    ck(exc_transpose((long)local_n, k, D.bu.p + n_pos[cid] * (size_t)k, k, D.tmp.p, (long)local_n, s), "FinalVT");
    givens_mask(D.tmp.p, k, (long)local_n, 0, 0, private_seed, plans, s);
  } else if (compute_v) {
    ck(exc_rot_back((int)m, k, D.bu.p, D.cs.p, s), "back rotation");
    {
      const size_t wb = exc_ormqr_house_workspace((int)m, k);
      D.work.alloc((wb + 7) / 8, s);
      ck(exc_ormqr_house_ln((int)m, k, D.house.p, D.tau_h.p, D.bu.p, k, D.work.p, D.work.n * 8, s), "ormqr house");
    }
    D.tmp.alloc((size_t)k * rl, s);
    D.work.alloc((exc_gemm_workspace(k, (long)rl, (long)m) + 7) / 8, s, false);  // split-K partials
    ck(exc_dgemm(1, 0, k, (long)rl, (long)m, 1.0, D.bu.p, k, D.Qg.p + reduced_n_pos[cid], (long)R, 0.0, D.tmp.p,
                 (long)rl, D.work.p, D.work.n * 8, s),
       "combine V");
    sync();
    D.Qg.release();
    givens_mask(D.tmp.p, k, (long)rl, 0, 0, private_seed, plans, s);
    if (lqr) {
      D.FinalVT.alloc((size_t)k * local_n, s, false);  // every element written by the beta = 0 product
      exc_prof::label_next("gemm_f64_final_v");  // the largest single product of the run
      ck(exc_dgemm(0, 0, k, (long)local_n, (long)rl, 1.0, D.tmp.p, (long)rl, D.X.p, (long)local_n, 0.0, D.FinalVT.p,
                   (long)local_n, nullptr, 0, s),
         "final V");
      sync();
      D.X.release();
      D.tmp.release();
    }
  }
  sync();
  time_combine_svd_remove_mask = now_seconds() - t0;
  finish_run(k, compute_u, compute_v, lqr, t_run);
}

// The tall-and-skinny branch (m > global_n) of run(): client.cpp:543-569 (masks), 627-678
// (re-partition, local QR), 800-996 with the bidiagonal of order global_n, and the combine of
// client.cpp:998-1154. The raw block D.X (m * local_n doubles) is read as a local_n x m row-major
// matrix, as the reference reads it in this branch; FinalU ends k x m row-major (the reference's
// layout here) and FinalVT k x local_n. Exchanges go over the network.
int DecentralizedClient::run_tall(bool compute_u, bool compute_v) {
  hipStream_t s = dev_->s;
  exc::ClientDevice& D = *dev_;
  exc::PeerNet& net = *net_;
  const int P = num_client, cid = client_id;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  const size_t gn = global_n, ln = local_n;
  // the four Givens plans: masks (left K = local_n private, right K = m public) and their removal
  RunPlans plans;
  plans.start({{(long)ln, 1, private_seed}, {(long)m, 1, public_seed}, {(long)m, 0, public_seed},
               {(long)ln, 0, private_seed}});

  // masks (client.cpp:543-569): local_k = local_n
  double t0 = now_seconds();
  givens_mask(D.X.p, (long)ln, (long)m, 1, 1, private_seed, plans, s);
  givens_mask(D.X.p, (long)ln, (long)m, 1, 0, public_seed, plans, s);
  sync();
  time_apply_mask = now_seconds() - t0;
  log("Masking Costs (tall-and-skinny) " + std::to_string(time_apply_mask));

  // re-partition (client.cpp:627-659): peer j gets columns m_part_pos[j] .. of every block, and
  // this peer's columns of all blocks form Pm = global_n x c (row-major), rows n_pos[i] .. from
  // peer i
  std::vector<size_t> m_part(P, m / P), m_part_pos(P + 1, 0);
  for (int i = 0; i < (int)(m % P); i++) m_part[i] += 1;
  for (int i = 0; i < P; i++) m_part_pos[i + 1] = m_part_pos[i] + m_part[i];
  const size_t c = m_part[cid];
  t0 = now_seconds();
  {
    HBuf& hm = D.h_masked;
    HBuf& hr = D.h_repart;
    hm.alloc(ln * m);
    ckh(hipMemcpyAsync(hm.p, D.X.p, ln * m * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
    sync();
    std::vector<double> blk;
    for (int j = 0; j < P; j++) {
      if (j == cid) continue;
      blk.resize(ln * m_part[j]);
      for (size_t a = 0; a < ln; a++)
        memcpy(blk.data() + a * m_part[j], hm.p + a * m + m_part_pos[j], m_part[j] * 8);
      net.send_to(j, blk.data(), blk.size() * 8);
    }
    hr.alloc(gn * c);
    for (int i = 0; i < P; i++) {
      const size_t li = local_ns[i], off = n_pos[i];
      if (i == cid) {
        for (size_t a = 0; a < li; a++) memcpy(hr.p + (off + a) * c, hm.p + a * m + m_part_pos[cid], c * 8);
      } else {
        std::vector<char> r = net.receive(i);
        if (r.size() != li * c * 8) throw std::runtime_error("re-partition block of unexpected size");
        memcpy(hr.p + off * c, r.data(), li * c * 8);
      }
    }
    D.P.alloc(gn * c, s);
    ckh(hipMemcpyAsync(D.P.p, hr.p, gn * c * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
    sync();
  }
  D.X.release();
  time_exchange_low_dim = now_seconds() - t0;
  net.add_comm_seconds(time_exchange_low_dim);

  // local QR (client.cpp:661-676) when m_part > global_n: Pm^T = Q R (c x global_n), R^T into
  // D.rt and Q^T in place of Pm (u_reduction); the Ralha block is R (global_n x global_n), else Pm
  const bool lqr = c > gn;
  const size_t bl = std::min(c, gn);
  D.F.alloc(gn * bl, s);
  t0 = now_seconds();
  if (lqr) {
    D.rt.alloc(gn * gn, s);
    const size_t wb = exc_qr_t_workspace((int)gn, (long)c, 1);
    D.work.alloc((wb + 7) / 8, s, false);
    ck(exc_qr_t((int)gn, (long)c, D.P.p, (long)c, D.rt.p, compute_u ? D.P.p : nullptr, (long)c, nullptr, 0, nullptr,
                nullptr, 0, D.work.p, D.work.n * 8, s),
       "local QR (tall-and-skinny)");
    // the Ralha block column-major: element (i, l) = R^T[i][l]
    ck(exc_transpose((long)gn, (long)gn, D.rt.p, (long)gn, D.F.p, (long)gn, s), "Ralha block");
    sync();
    D.rt.release();
    time_local_qr = now_seconds() - t0;
    log("Local QR Reduction (tall-and-skinny) Costs " + std::to_string(time_local_qr));
  } else {
    ck(exc_transpose((long)gn, (long)c, D.P.p, (long)c, D.F.p, (long)gn, s), "Ralha block");
    sync();
    D.P.release();
  }

  // bidiagonalization of order global_n (one launch per step: the loop's co-residency decision
  // is taken for the short-and-wide shape) and the bidiagonal SVD
  ralha_stage((int)gn, (int)bl, false);
  D.F.release();
  const int k = bsvd_stage((int)gn);

  // combine (client.cpp:998-1154)
  t0 = now_seconds();
  if (compute_u) {
    // this peer's columns of FinalU (k x m): part = bv U_r (u_reduction), gathered as the rows of
    // FinalU^T (m x k) and transposed at the end
    DBuf& Ut = D.tmp;
    Ut.alloc(m * (size_t)k, s);
    D.work.alloc((exc_gemm_workspace((long)gn, k, (long)gn) + 7) / 8, s, false);
    if (lqr) {
      DBuf& T2 = D.Qg;  // (U_r^T bv^T): global_n x k
      T2.alloc(gn * (size_t)k, s);
      ck(exc_dgemm(0, 1, (long)gn, k, (long)gn, 1.0, D.U.p, (long)gn, D.bv.p, (long)gn, 0.0, T2.p, k, D.work.p,
                   D.work.n * 8, s),
         "part final U");
      ck(exc_dgemm(1, 0, (long)c, k, (long)gn, 1.0, D.P.p, (long)c, T2.p, k, 0.0, Ut.p + m_part_pos[cid] * (size_t)k,
                   k, nullptr, 0, s),
         "part final U (reduction)");
      sync();
      T2.release();
      D.P.release();
    } else {
      ck(exc_dgemm(0, 1, (long)c, k, (long)gn, 1.0, D.U.p, (long)gn, D.bv.p, (long)gn, 0.0,
                   Ut.p + m_part_pos[cid] * (size_t)k, k, D.work.p, D.work.n * 8, s),
         "part final U");
    }
    // ring all-gather of the column blocks of FinalU (client.cpp:1062-1096); the wire carries
    // each block as the reference does, k x m_part row-major
    const double tc = now_seconds();
    {
      HBuf& hU = D.h_U;
      hU.alloc(m * (size_t)k);
      ckh(hipMemcpyAsync(hU.p + m_part_pos[cid] * k, Ut.p + m_part_pos[cid] * k, c * (size_t)k * 8,
                         hipMemcpyDeviceToHost, s),
          "d2h");
      sync();
      const int next = (cid + 1) % P, front = (cid - 1 + P) % P;
      std::vector<double> wire;
      for (int st = 0; st < P - 1; st++) {
        int di = (cid - st + P) % P;
        wire.resize(m_part[di] * (size_t)k);
        for (size_t r = 0; r < m_part[di]; r++)
          for (int a = 0; a < k; a++) wire[(size_t)a * m_part[di] + r] = hU.p[(m_part_pos[di] + r) * k + a];
        net.send_to(next, wire.data(), wire.size() * 8);
        di = (di - 1 + P) % P;
        wire.resize(m_part[di] * (size_t)k);
        net.receive_into(front, wire.data(), wire.size() * 8);
        for (size_t r = 0; r < m_part[di]; r++)
          for (int a = 0; a < k; a++) hU.p[(m_part_pos[di] + r) * k + a] = wire[(size_t)a * m_part[di] + r];
      }
      ckh(hipMemcpyAsync(Ut.p, hU.p, m * (size_t)k * 8, hipMemcpyHostToDevice, s), "h2d");
    }
    net.add_comm_seconds(now_seconds() - tc);
    D.FinalU.alloc(m * (size_t)k, s, false);
    ck(exc_transpose((long)m, k, Ut.p, k, D.FinalU.p, (long)m, s), "FinalU");
    givens_mask(D.FinalU.p, k, (long)m, 0, 0, public_seed, plans, s);
    sync();
    Ut.release();
  }
  D.U.release();
  if (compute_v) {
    // right vectors (client.cpp:1113-1154): the back rotation and the reflectors of order
    // global_n, then this peer's rows transposed and unmasked with the private plan
    ck(exc_rot_back((int)gn, k, D.bu.p, D.cs.p, s), "back rotation");
    const size_t wb = exc_ormqr_house_workspace((int)gn, k);
    D.work.alloc((wb + 7) / 8, s);
    ck(exc_ormqr_house_ln((int)gn, k, D.house.p, D.tau_h.p, D.bu.p, k, D.work.p, D.work.n * 8, s), "ormqr house");
    D.FinalVT.alloc((size_t)k * ln, s, false);
    ck(exc_transpose((long)ln, k, D.bu.p + n_pos[cid] * (size_t)k, k, D.FinalVT.p, (long)ln, s), "FinalVT");
    givens_mask(D.FinalVT.p, k, (long)ln, 0, 0, private_seed, plans, s);
  }
  sync();
  time_combine_svd_remove_mask = now_seconds() - t0;
  return k;
}

// Ralha bidiagonalization (client.cpp:800-910) and lower to upper (947-954) of the mi x ci
// block D.F (column-major) this peer holds, every peer's all-reduce over the network or the
// device boxes; one_launch: the loop kernel in one launch (a decision all peers share).
void DecentralizedClient::ralha_stage(int mi, int ci, bool one_launch) {
  hipStream_t s = dev_->s;
  exc::ClientDevice& D = *dev_;
  exc::PeerNet& net = *net_;
  const int P = num_client, cid = client_id;
  (void)cid;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  const double t0 = now_seconds();
  const size_t mm = (size_t)mi, cc = (size_t)ci;
  D.U.alloc(mm * cc, s);
  D.alpha.alloc(mm, s);
  D.beta.alloc(mm, s);
  D.house.alloc(mm * mm, s);
  D.tau_h.alloc(mm, s);
  D.cs.alloc(2 * mm, s);
  D.msg.alloc(mm + 3, s);
  D.g.alloc(mm + 3, s);
  D.scal.alloc(4, s);
  sync();
  {
    HBuf& hb = D.h_msg;
    hb.alloc(mm + 3);
    double ar = 0.0;
    auto allreduce_dev = [&](double* dsrc, double* ddst, long n) {
      ckh(hipMemcpyAsync(hb.p, dsrc, n * sizeof(double), hipMemcpyDeviceToHost, s), "d2h");
      sync();
      const double ta = now_seconds();
      net.ring_all_reduce_sum(hb.p, n);
      ar += now_seconds() - ta;
      ckh(hipMemcpyAsync(ddst, hb.p, n * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
    };
    ck(exc_ralha_tail_update(-1, mi, ci, D.F.p, D.U.p, nullptr, D.beta.p, D.scal.p, s), "ralha init");
    allreduce_dev(D.scal.p, D.scal.p + 1, 1);
    ck(exc_ralha_tail_scale(0, mi, ci, D.F.p, D.U.p, D.scal.p + 1, D.alpha.p, s), "ralha init");
    if (one_launch) {
      // steps -1 .. m-3 in one launch; the messages go through device-memory boxes when every
      // peer shares this device, else this thread all-reduces them over the network
      const double tmo = std::max(30.0, opt_.recv_timeout_s);
      D.work.alloc((exc_ralha_loop_workspace(mi, ci) + 7) / 8, s);
      if (!D.link || D.link_m != mi) {
        if (D.link) ck(exc_ralha_link_free(D.link), "ralha link");
        D.link = nullptr;
        ck(exc_ralha_link_alloc(mi, &D.link), "ralha link");
        D.link_m = mi;
      }
      if (D.boxes_state == 1) {
        ck(exc_ralha_box_clear(D.box, s), "ralha box");
        sync();
        net.barrier();  // every peer's box counter is zero before any loop starts
        ck(exc_ralha_loop_start_boxes(mi, ci, D.F.p, D.U.p, D.house.p, D.tau_h.p, D.alpha.p, D.beta.p, D.link,
                                      D.work.p, D.work.n * 8, 0, tmo, (int)P, (int)cid, D.boxes.data(),
                                      D.box_nwg.data(), s),
           "ralha loop");
      } else {
        ck(exc_ralha_loop_start(mi, ci, D.F.p, D.U.p, D.house.p, D.tau_h.p, D.alpha.p, D.beta.p, D.link, D.work.p,
                                D.work.n * 8, 0, tmo, s),
           "ralha loop");
      }
      try {
        if (D.boxes_state == 1) {
          const double t0w = now_seconds();
          while (true) {
            const int r = exc_ralha_loop_running(s);
            if (r < 0) ck(r, "ralha loop");
            if (r == 0) break;
            if (net.any_closed()) throw std::runtime_error("a peer disconnected during the bidiagonalization");
            if (now_seconds() - t0w > tmo) throw std::runtime_error("the bidiagonalization loop timed out");
            usleep(50);
          }
        } else {
          for (int j = 0; j + 2 < mi; j++) {
            double* msg = nullptr;
            ck(exc_ralha_loop_wait(D.link, j, tmo, &msg), "ralha loop");
            const double ta = now_seconds();
            net.ring_all_reduce_sum(msg, (long)mi + 3 - j);
            ar += now_seconds() - ta;
            ck(exc_ralha_loop_post(D.link, j), "ralha loop");
          }
        }
      } catch (...) {
        exc_ralha_loop_abort(D.link);
        exc_ralha_loop_finish(D.link, s);
        throw;
      }
      ck(exc_ralha_loop_finish(D.link, s), "ralha loop");
    } else {
      const size_t rwb = exc_ralha_workspace(mi, ci);
      D.work.alloc((rwb + 7) / 8, s);
      // the step kernels write the local message straight into pinned host memory
      ck(exc_ralha_step(-1, mi, ci, D.F.p, D.U.p, nullptr, D.house.p, D.tau_h.p, D.alpha.p, D.beta.p, hb.p,
                        D.work.p, D.work.n * 8, s),
         "ralha step");
      for (int i = 0; i + 2 < mi; i++) {
        sync();
        const double ta = now_seconds();
        net.ring_all_reduce_sum(hb.p, (long)mi + 3 - i);
        ar += now_seconds() - ta;
        ckh(hipMemcpyAsync(D.g.p, hb.p, ((long)mi + 3 - i) * sizeof(double), hipMemcpyHostToDevice, s), "h2d");
        ck(exc_ralha_step(i, mi, ci, D.F.p, D.U.p, D.g.p, D.house.p, D.tau_h.p, D.alpha.p, D.beta.p, hb.p,
                          D.work.p, D.work.n * 8, s),
           "ralha step");
      }
    }
    for (int i = mi - 2; i < mi; i++) {
      ck(exc_ralha_dot(i, mi, ci, D.F.p, D.U.p, D.scal.p, s), "ralha dot");
      allreduce_dev(D.scal.p, D.scal.p + 1, 1);
      ck(exc_ralha_tail_update(i, mi, ci, D.F.p, D.U.p, D.scal.p + 1, D.beta.p, D.scal.p + 2, s), "ralha tail");
      allreduce_dev(D.scal.p + 2, D.scal.p + 3, 1);
      ck(exc_ralha_tail_scale(i, mi, ci, D.F.p, D.U.p, D.scal.p + 3, D.alpha.p, s), "ralha tail");
    }
    ck(exc_lower_to_upper(mi, D.alpha.p, D.beta.p, D.cs.p, s), "lower to upper");
    sync();
    time_ralha_allreduce = ar;
  }
  time_bi_diag = now_seconds() - t0;
  log("One-side Householder reflector Costs " + std::to_string(time_bi_diag));

}

// Bidiagonal SVD of the upper bidiagonal (D.alpha, D.beta) of order n (client.cpp:963-996):
// the count of singular values, bisection + twisted factorization and the peers' sign rule.
// Returns the number of triplets kept (sigma_count).
int DecentralizedClient::bsvd_stage(int n) {
  hipStream_t s = dev_->s;
  exc::ClientDevice& D = *dev_;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  const double t0 = now_seconds();
  {
    DBuf& b2 = D.b2;
    b2.alloc(2, s);
    DBuf& cnt2 = D.cnt2;
    cnt2.alloc(1, s);
    ck(exc_bisection_bound(n, D.alpha.p, D.beta.p, b2.p, s), "bound");
    ck(exc_sturm_count(n, D.alpha.p, D.beta.p, b2.p, 2, (int*)cnt2.p, s), "count");
    int hc[2];
    ckh(hipMemcpyAsync(hc, cnt2.p, sizeof hc, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    sigma_count = hc[0] - hc[1];
    log("Ori Sigma Count = " + std::to_string(sigma_count));
    if (top_k > 0) sigma_count = std::min(top_k, sigma_count);
    log("Adj Sigma Count = " + std::to_string(sigma_count));
  }
  const int k = sigma_count;
  bidiag_order = n;
  D.bu.alloc((size_t)n * k, s);
  D.bv.alloc((size_t)n * k, s);
  D.sigma.alloc(std::max(k, 1), s);
  D.flip.alloc(std::max(k, 1), s);
  {
    const size_t wb = exc_bidiag_svd_workspace(n, k);
    D.work.alloc((wb + 7) / 8, s);
    const int rc = exc_bidiag_svd(n, D.alpha.p, D.beta.p, D.bu.p, D.bv.p, D.sigma.p, k, nullptr, D.work.p,
                                  D.work.n * 8, s);
    ck(rc, "bidiagonal SVD");
    ck(exc_sign_normalize(n, k, D.bu.p, D.bv.p, D.flip.p, s), "sign rule");
    sync();
  }
  time_bsvd = now_seconds() - t0;
  return k;
}

// Logs, host results and the result line of one run() (client.cpp:1200-1213).
void DecentralizedClient::finish_run(int k, bool compute_u, bool compute_v, bool lqr, double t_run) {
  exc::ClientDevice& D = *dev_;
  hipStream_t s = D.s;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  time_total = now_seconds() - t_run;
  time_communication = net_->comm_seconds();

  log("Finish DecFedSVD");
  log("DecFedSVD totally Costs " + std::to_string(time_total));
  log("Overall Communication Cost " + std::to_string(time_communication));
  log("time_init=" + std::to_string(time_init));
  log("time_local_qr=" + std::to_string(time_local_qr));
  log("time_global_qr=" + std::to_string(time_global_qr));
  log("time_apply_mask=" + std::to_string(time_apply_mask));
  log("time_exchange_low_dim=" + std::to_string(time_exchange_low_dim));
  log("time_bi_diag=" + std::to_string(time_bi_diag));
  log("time_bsvd=" + std::to_string(time_bsvd));
  log("time_combine_svd_remove_mask=" + std::to_string(time_combine_svd_remove_mask));
  log("time_ralha_allreduce=" + std::to_string(time_ralha_allreduce));
  if (opt_.form == "mixed") log("time_refine=" + std::to_string(time_refine));

  sigma = FinalU = FinalVT = x_diagonal = x_upper_diagonal = nullptr;
  time_results_d2h = 0;
  if (opt_.keep_results || opt_.dump_results) {
    const double td = now_seconds();
    copy_results_to_host(k, compute_u, compute_v, lqr);
    time_results_d2h = now_seconds() - td;
    log("time_results_d2h=" + std::to_string(time_results_d2h));
    if (opt_.dump_results) {
      auto dump = [&](const std::string& name, const double* v, size_t count) {
        FILE* f = fopen((log_dir + "/Client" + std::to_string(client_id) + "_" + name + ".bin").c_str(), "wb");
        if (f) {
          fwrite(v, sizeof(double), count, f);
          fclose(f);
        }
      };
      dump("sigma", sigma, k);
      if (compute_u) dump("FinalU", FinalU, m * (size_t)k);
      if (compute_v) dump("FinalVT", FinalVT, (size_t)k * local_n);
      // the bidiagonal and its vectors, for stage-by-stage comparison
      const size_t nb = bidiag_order > 0 ? (size_t)bidiag_order : m;
      dump("alpha", x_diagonal, nb);
      dump("beta", x_upper_diagonal, nb - 1);
      if (opt_.form != "mixed") {
        std::vector<double> v(nb * (size_t)k);
        ckh(hipMemcpyAsync(v.data(), D.bv.p, v.size() * 8, hipMemcpyDeviceToHost, s), "d2h");
        sync();
        dump("bidiag_v", v.data(), v.size());
      }
    }
  }
  if (evaluate && m > global_n) log("the evaluation report is computed for the short-and-wide branch only");
  else if (evaluate) evaluate_results(k, compute_u, compute_v, lqr);
  char json[1024];
  snprintf(json, sizeof json,
           "{\"cid\": %d, \"sigma_count\": %d, \"total\": %.6f, \"communication\": %.6f, \"time_init\": %.6f, "
           "\"time_local_qr\": %.6f, \"time_global_qr\": %.6f, \"time_apply_mask\": %.6f, "
           "\"time_exchange_low_dim\": %.6f, \"time_bi_diag\": %.6f, \"time_ralha_allreduce\": %.6f, "
           "\"time_bsvd\": %.6f, \"time_combine_svd_remove_mask\": %.6f, \"time_results_d2h\": %.6f, "
           "\"time_refine\": %.6f, \"form\": \"%s\"}",
           client_id, k, time_total, time_communication, time_init, time_local_qr, time_global_qr, time_apply_mask,
           time_exchange_low_dim, time_bi_diag, time_ralha_allreduce, time_bsvd, time_combine_svd_remove_mask,
           time_results_d2h, time_refine, opt_.form.c_str());
  printf("EXC_RESULT %s\n", json);
  if (opt_.profile) {
    std::vector<char> pb(1 << 16);
    if (exc_prof_report(pb.data(), pb.size()) == EXC_OK) printf("EXC_PROF %s\n", pb.data());
  }
  fflush(stdout);
}

void DecentralizedClient::copy_results_to_host(int k, bool compute_u, bool compute_v, bool lqr) {
  exc::ClientDevice& D = *dev_;
  hipStream_t s = D.s;
  D.r_sigma.alloc(std::max(k, 1), false);
  ckh(hipMemcpyAsync(D.r_sigma.p, D.sigma.p, k * 8, hipMemcpyDeviceToHost, s), "d2h");
  sigma = D.r_sigma.p;
  const size_t nb = bidiag_order > 0 ? (size_t)bidiag_order : m;  // global_n in the tall branch
  D.r_alpha.alloc(nb, false);
  D.r_beta.alloc(nb, false);
  ckh(hipMemcpyAsync(D.r_alpha.p, D.alpha.p, nb * 8, hipMemcpyDeviceToHost, s), "d2h");
  ckh(hipMemcpyAsync(D.r_beta.p, D.beta.p, (nb - 1) * 8, hipMemcpyDeviceToHost, s), "d2h");
  x_diagonal = D.r_alpha.p;
  x_upper_diagonal = D.r_beta.p;
  if (compute_u) {
    D.r_FinalU.alloc(m * (size_t)k, false);
    ckh(hipMemcpyAsync(D.r_FinalU.p, D.FinalU.p, m * (size_t)k * 8, hipMemcpyDeviceToHost, s), "d2h");
    FinalU = D.r_FinalU.p;
  }
  if (compute_v) {
    const double* src = lqr ? D.FinalVT.p : D.tmp.p;
    D.r_FinalVT.alloc((size_t)k * local_n, false);
    ckh(hipMemcpyAsync(D.r_FinalVT.p, src, (size_t)k * local_n * 8, hipMemcpyDeviceToHost, s), "d2h");
    FinalVT = D.r_FinalVT.p;
  }
  ckh(hipStreamSynchronize(s), "d2h");
}

void DecentralizedClient::remove_input_data() {  // client.cpp:215-218
  remove((dpath + "/Client" + std::to_string(client_id) + ".mat").c_str());
  if (client_id == 0) remove((dpath + "/command.txt").c_str());
}

std::string DecentralizedClient::filename_with_cid(std::string filename) {  // client.cpp:413-415
  return log_dir + "/Client" + std::to_string(client_id) + "_" + filename;
}

void DecentralizedClient::send_to(int cid, char* buffer, size_t size) { net_->send_to(cid, buffer, size); }

char* DecentralizedClient::receive_frame(int cid) {
  std::vector<char> f = net_->receive(cid);
  char* p = new char[std::max<size_t>(f.size(), 1)];
  if (!f.empty()) memcpy(p, f.data(), f.size());
  return p;
}

void DecentralizedClient::broadcast(char* buffer, size_t size) { net_->broadcast(buffer, size); }

void DecentralizedClient::ring_all_reduce_sum(double* buffer, int size) {
  const double t0 = now_seconds();
  net_->ring_all_reduce_sum(buffer, size);
  time_communication += now_seconds() - t0;
}

void DecentralizedClient::all_gather(char** buffer, size_t size, int* c_group, int nc, int src) {
  std::vector<int> grp(c_group, c_group + nc);
  std::vector<char> b;
  if (client_id == src) b.assign(*buffer, *buffer + size);
  const double t0 = now_seconds();
  net_->all_gather(b, size, grp, src);
  time_communication += now_seconds() - t0;
  if (client_id != src) {
    char* p = new char[std::max<size_t>(size, 1)];
    if (size) memcpy(p, b.data(), size);
    *buffer = p;
  }
}

void DecentralizedClient::local_qr_reduction(size_t mr, size_t n, double** raw_x, double* q, bool transpose_x,
                                             bool compute_q) {
  if (!(mr > n) || n < 1) throw std::invalid_argument("local_qr_reduction needs m > n >= 1");
  hipStream_t s = dev_->s;
  DBuf x, rt, work;
  x.alloc(mr * n, s, false);
  rt.alloc(n * n, s);
  const size_t wb = exc_qr_t_workspace((int)n, (long)mr, compute_q ? 1 : 0);
  work.alloc((wb + 7) / 8, s, false);
  // the column-major mr x n input is the row-major n x mr matrix exc_qr_t factorizes as X^T
  ckh(hipMemcpyAsync(x.p, *raw_x, mr * n * 8, hipMemcpyHostToDevice, s), "h2d");
  ck(exc_qr_t((int)n, (long)mr, x.p, (long)mr, rt.p, compute_q ? x.p : nullptr, (long)mr, nullptr, 0, nullptr,
              nullptr, 0, work.p, work.n * 8, s),
     "local QR");
  std::vector<double> hrt(n * n);
  ckh(hipMemcpyAsync(hrt.data(), rt.p, n * n * 8, hipMemcpyDeviceToHost, s), "d2h");
  // Q (mr x n column-major) is Q^T row-major, the layout exc_qr_t leaves
  if (compute_q) ckh(hipMemcpyAsync(q, x.p, mr * n * 8, hipMemcpyDeviceToHost, s), "d2h");
  ckh(hipStreamSynchronize(s), "local QR");
  double* out = new double[n * n]();
  // hrt is R^T row-major (lower); the reference's shrink (client.cpp:434-444) leaves R row-major
  // when transpose_x, else R^T row-major
  for (size_t i = 0; i < n; i++)
    for (size_t j = 0; j <= i; j++) {
      if (transpose_x) out[j * n + i] = hrt[i * n + j];
      else out[i * n + j] = hrt[i * n + j];
    }
  delete[] *raw_x;
  *raw_x = out;
}

// The mixed-precision form (ClientOptions::form = "mixed"). The m x m factor R of A (A = Q R,
// A^T = [X_0 ... X_{P-1}]) comes from a CholeskyQR over all peers' Gram matrices instead of the
// local QR, the masked exchange and the global QR of client.cpp:572-798, and its SVD is computed
// by every peer from the same bits instead of the per-column all-reduce of client.cpp:829-910:
//   pass 1 (FP32 MFMA): G = sum_i X_i X_i^T, masked with the public Givens plan on both sides
//     before the exchange (only the lower triangle of the symmetric masked matrix travels, as
//     FP32, summed in FP64), shifted by u32 * trace(G) until its Cholesky factor L1 exists;
//   pass 2 (FP32 MFMA): Y_i = L1^-1 X_i and the Cholesky factor L2 of sum_i Y_i Y_i^T: R1 = L1 L2
//     is the R^T of a shifted CholeskyQR2 in single precision;
//   pass 3 (FP64, the refinement of Q): Z_i = R1^-1 X_i by triangular solve, L3 from
//     sum_i Z_i Z_i^T, which is I to within FP32 error; Q_i^T = L3^-1 Z_i and R^T = Lt = R1 L3
//     (repeated, Z_i <- L3^-1 Z_i, while ||sum_i Z_i Z_i^T - I||_F > 1/2: at most 3 passes);
//   every peer: one-sided bidiagonalization of Lt^T (one owner, no exchange), bisection + twisted
//     factorization (FP64, as the plain form), then the FP64 Ogita-Aishima refinement of
//     Lt = U diag(sigma) W^T (exc_oa_refine), W' = L3^-T W and FinalVT_i = W'^T Z_i = W^T Q_i^T.
// Every exchange is a sum of Gram matrices, bit-identical on every peer, so every peer factors the
// same Lt; a checksum of sigma, U and W' exchanged at the end confirms that they agree.
int DecentralizedClient::run_mixed(bool compute_u, bool compute_v) {
  hipStream_t s = dev_->s;
  exc::ClientDevice& D = *dev_;
  exc::PeerNet& net = *net_;
  const int cid = client_id;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  const int mi = (int)m;
  const long r = (long)local_n;
  const size_t mm = m * m;
  const double u32 = std::ldexp(1.0, -24);
  RunPlans plans;
  plans.start({{(long)m, 1, public_seed}, {(long)m, 0, public_seed}});
  double t_mask = 0, t_comm = 0;
  double t0 = now_seconds();

  // workspaces: split-K partials of the Grams and the factorization workspace
  const size_t gws = std::max(exc_gemm_workspace(m, m, r), exc_gemm_workspace(m, m, m));
  const size_t fws = exc_fact_workspace(mi, std::max<long>(r, (long)m));
  D.work.alloc((std::max(gws, fws) + 7) / 8, s, false);
  void* work = D.work.p;
  const size_t work_bytes = D.work.n * 8;
  ExcGemmOpts gram;
  gram.lower_only = 1;
  gram.mirror = 1;
  gram.work = work;
  gram.work_bytes = work_bytes;
  ExcGemmOpts plain;
  plain.work = work;
  plain.work_bytes = work_bytes;

  D.x32.alloc((m * local_n + 1) / 2, s, false);
  D.y32.alloc((m * local_n + 1) / 2, s, false);
  D.g32.alloc((mm + 1) / 2, s, false);
  D.li32.alloc((mm + 1) / 2, s, false);
  float* x32 = (float*)D.x32.p;
  float* y32 = (float*)D.y32.p;
  float* g32 = (float*)D.g32.p;
  float* li32 = (float*)D.li32.p;
  D.Gm.alloc(mm, s, false);
  D.L1.alloc(mm, s, false);
  D.L2.alloc(mm, s, false);
  D.L3.alloc(mm, s, false);
  D.Lt.alloc(mm, s, false);
  D.idm.alloc(mm, s, false);
  D.scal.alloc(4, s);
  D.h_gram.alloc(mm, false);

  // The masked Gram matrix Q G Q^T (the public plan on both sides) is symmetric, so only its lower
  // triangle travels: n_tri = m (m + 1) / 2 values, packed by rows.
  const long n_tri = (long)m * ((long)m + 1) / 2;
  D.pk.alloc((size_t)n_tri, s, false);
  D.h_tri32.alloc(((size_t)n_tri + 1) / 2, false);
  D.pacc.alloc((size_t)n_tri, s, false);
  D.pstage.alloc((size_t)n_tri * 2, s, false);
  D.h_pstage.alloc((size_t)n_tri * 2, false);
  // Every peer's packed triangle (this peer's in D.pk, FP32 or FP64) summed in FP64 on the device
  // in peer order 0..P-1 (exc_acc_packed), so every peer holds the same bits in D.pacc. The own
  // triangle, already in pinned memory at h_own, goes to every peer over TCP; each other peer's
  // frame is read from the socket into a pinned staging slot (PeerNet::receive_direct) and
  // uploaded from there. The two slots alternate over the senders; a slot is refilled only after
  // its upload has completed.
  hipEvent_t slot_free[2] = {nullptr, nullptr};
  for (auto& e : slot_free) ckh(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard() {
      for (int i = 0; i < 2; i++)
        if (e[i]) hipEventDestroy(e[i]);
    }
  } slot_guard{slot_free};
  auto sum_over_peers = [&](bool f32, const void* h_own) {
    const size_t bytes = (size_t)n_tri * (f32 ? 4 : 8);
    const double tb = now_seconds();
    net.broadcast(h_own, bytes);
    int used = 0;
    for (int j = 0; j < (int)num_client; j++) {
      const void* part = D.pk.p;
      if (j != cid) {
        const int slot = used & 1;
        if (used >= 2) ckh(hipEventSynchronize(slot_free[slot]), "staging slot");
        char* hs = (char*)(D.h_pstage.p + (size_t)slot * n_tri);
        char* ds = (char*)(D.pstage.p + (size_t)slot * n_tri);
        net.receive_direct(j, hs, bytes);  // read from the socket straight into the pinned slot
        ckh(hipMemcpyAsync(ds, hs, bytes, hipMemcpyHostToDevice, s), "h2d");
        ckh(hipEventRecord(slot_free[slot], s), "event");
        part = ds;
        used++;
      }
      ck(exc_acc_packed(n_tri, part, f32 ? 1 : 0, j == 0 ? 1 : 0, D.pacc.p, s), "Gram sum");
    }
    const double dt = now_seconds() - tb;
    net.add_comm_seconds(dt);
    return dt;
  };
  // masked FP32 Gram in g32 -> its triangle summed over the peers in FP64 -> unmasked FP64 Gram
  // in D.Gm
  auto exchange_f32 = [&]() {
    double ta = now_seconds();
    givens_mask_f32(g32, (long)m, (long)m, 1, 1, public_seed, plans, s);
    givens_mask_f32(g32, (long)m, (long)m, 1, 0, public_seed, plans, s);
    ck(exc_pack_lower_f32(mi, g32, (long)m, (float*)D.pk.p, s), "pack");
    ckh(hipMemcpyAsync(D.h_tri32.p, D.pk.p, (size_t)n_tri * 4, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    double tb = now_seconds();
    const double tx = sum_over_peers(true, D.h_tri32.p);
    double tc = tb + tx;
    ck(exc_unpack_lower_f64(mi, D.pacc.p, D.Gm.p, (long)m, s), "unpack");
    givens_mask(D.Gm.p, (long)m, (long)m, 0, 1, public_seed, plans, s);
    givens_mask(D.Gm.p, (long)m, (long)m, 0, 0, public_seed, plans, s);
    sync();
    t_mask += (tb - ta) + (now_seconds() - tc);
    t_comm += tc - tb;
  };
  // Cholesky of D.Gm + shift I into L, the shift doubled from `shift` (0: first without one)
  auto shifted_chol = [&](DBuf& L, double shift, const char* what) {
    // the shift step u32 * trace(G), read back only when a shift is needed (a pass that starts
    // without one needs it only if that attempt fails)
    double step = -1.0;
    auto shift_step = [&]() {
      if (step < 0.0) {
        double tr = 0.0;
        ck(exc_trace(mi, D.Gm.p, (long)m, D.scal.p, s), "trace");
        ckh(hipMemcpyAsync(&tr, D.scal.p, 8, hipMemcpyDeviceToHost, s), "d2h");
        sync();
        step = u32 * tr;
      }
      return step;
    };
    if (shift < 0) shift = shift_step();  // a first pass in single precision always needs one
    for (int attempt = 0; attempt < 60; attempt++) {
      ckh(hipMemcpyAsync(L.p, D.Gm.p, mm * 8, hipMemcpyDeviceToDevice, s), "copy");
      if (shift > 0) ck(exc_add_diag(mi, L.p, (long)m, shift, s), "shift");
      int info = 0;
      ck(exc_potrf_lower(mi, L.p, (long)m, (int*)(D.scal.p + 1), work, work_bytes, s), what);
      ckh(hipMemcpyAsync(&info, D.scal.p + 1, sizeof(int), hipMemcpyDeviceToHost, s), "d2h");
      sync();
      if (info == 0) {
        if (shift > 0) log(std::string(what) + ": shift " + std::to_string(shift));
        return;
      }
      shift = shift > 0 ? 2 * shift : shift_step();
    }
    throw std::runtime_error(std::string(what) + ": no shift made the Gram matrix positive definite");
  };

  // ---- pass 1 (FP32): G = X X^T
  ck(exc_convert_f64_f32((long)m, r, D.X.p, r, x32, r, s), "convert");
  ck(exc_sgemm_opts(0, 1, (long)m, (long)m, r, 1.0, x32, r, x32, r, 0.0, g32, (long)m, gram, s), "gram 1");
  sync();
  double tf = now_seconds() - t0;
  exchange_f32();
  t0 = now_seconds();
  shifted_chol(D.L1, -1.0, "pass 1 Cholesky");
  // ---- pass 2 (FP32): Y = L1^-1 X through the explicit inverse, G2 = Y Y^T
  ck(exc_fill2d((long)m, (long)m, 0.0, D.idm.p, (long)m, s), "identity");
  ck(exc_add_diag(mi, D.idm.p, (long)m, 1.0, s), "identity");
  ck(exc_trsm_left(0, 0, mi, (long)m, D.L1.p, (long)m, D.idm.p, (long)m, work, work_bytes, s), "inverse");
  ck(exc_convert_f64_f32((long)m, (long)m, D.idm.p, (long)m, li32, (long)m, s), "convert");
  {
    ExcGemmOpts tri = plain;
    tri.tri = EXC_TRI_A_LOWER;
    ck(exc_sgemm_opts(0, 0, (long)m, r, (long)m, 1.0, li32, (long)m, x32, r, 0.0, y32, r, tri, s), "pass 2 solve");
  }
  ck(exc_sgemm_opts(0, 1, (long)m, (long)m, r, 1.0, y32, r, y32, r, 0.0, g32, (long)m, gram, s), "gram 2");
  sync();
  tf += now_seconds() - t0;
  exchange_f32();
  t0 = now_seconds();
  shifted_chol(D.L2, 0.0, "pass 2 Cholesky");
  // R1 = L1 L2 (lower)
  {
    ExcGemmOpts tri = plain;
    tri.tri = EXC_TRI_A_LOWER;
    ck(exc_gemm_opts(0, 0, (long)m, (long)m, (long)m, 1.0, D.L1.p, (long)m, D.L2.p, (long)m, 0.0, D.Lt.p, (long)m,
                     tri, s),
       "R1");
  }
  // ---- pass 3 (FP64): Z = R1^-1 X in place, G3 = Z Z^T
  ck(exc_trsm_left(0, 0, mi, r, D.Lt.p, (long)m, D.X.p, r, work, work_bytes, s), "pass 3 solve");
  // G3 summed over the peers, masked for the exchange like the FP32 Grams
  auto gram_f64 = [&]() {
    ck(exc_gemm_opts(0, 1, (long)m, (long)m, r, 1.0, D.X.p, r, D.X.p, r, 0.0, D.Gm.p, (long)m, gram, s), "gram 3");
    sync();
    tf += now_seconds() - t0;
    double ta = now_seconds();
    givens_mask(D.Gm.p, (long)m, (long)m, 1, 1, public_seed, plans, s);
    givens_mask(D.Gm.p, (long)m, (long)m, 1, 0, public_seed, plans, s);
    ck(exc_pack_lower_f64(mi, D.Gm.p, (long)m, D.pk.p, s), "pack");
    ckh(hipMemcpyAsync(D.h_gram.p, D.pk.p, (size_t)n_tri * 8, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    double tb = now_seconds();
    const double tx = sum_over_peers(false, D.h_gram.p);
    double tc = tb + tx;
    ck(exc_unpack_lower_f64(mi, D.pacc.p, D.Gm.p, (long)m, s), "unpack");
    givens_mask(D.Gm.p, (long)m, (long)m, 0, 1, public_seed, plans, s);
    givens_mask(D.Gm.p, (long)m, (long)m, 0, 0, public_seed, plans, s);
    sync();
    t_mask += (tb - ta) + (now_seconds() - tc);
    t_comm += tc - tb;
    t0 = now_seconds();
  };
  auto lt_times_l3 = [&]() {
    // Lt = Lt L3 (both lower); D.L1 is free after pass 2
    ExcGemmOpts tri = plain;
    tri.tri = EXC_TRI_A_LOWER;
    ckh(hipMemcpyAsync(D.L1.p, D.Lt.p, mm * 8, hipMemcpyDeviceToDevice, s), "copy");
    ck(exc_gemm_opts(0, 0, (long)m, (long)m, (long)m, 1.0, D.L1.p, (long)m, D.L3.p, (long)m, 0.0, D.Lt.p, (long)m,
                     tri, s),
       "Lt");
  };
  // One FP64 CholeskyQR pass leaves ||Q^T Q - I|| ~ u kappa(Z)^2, with kappa(Z)^2 <=
  // (1 + e) / (1 - e) for e = ||Z Z^T - I||_F. When the FP32 passes left e above 1/2 (a condition
  // number near the reach of single precision), Z is first replaced by L3^-1 Z and the pass
  // repeated. The Gram matrix is bit-identical on every peer, so every peer takes the same branch.
  int fp64_passes = 0;
  double last_dev = 0.0;
  for (;;) {
    gram_f64();
    fp64_passes++;
    ck(exc_dev_identity(mi, D.Gm.p, (long)m, D.scal.p + 2, work, work_bytes, s), "deviation");
    double dev[2] = {0.0, 0.0};
    ckh(hipMemcpyAsync(dev, D.scal.p + 2, 16, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    log("mixed form: FP64 pass " + std::to_string(fp64_passes) + ", max |Z Z^T - I| " + std::to_string(dev[0]) +
        ", ||Z Z^T - I||_F " + std::to_string(dev[1]));
    shifted_chol(D.L3, 0.0, "pass 3 Cholesky");
    lt_times_l3();
    last_dev = dev[1];
    if (dev[1] <= 0.5 || fp64_passes == 3) break;
    ck(exc_trsm_left(0, 0, mi, r, D.L3.p, (long)m, D.X.p, r, work, work_bytes, s), "extra pass solve");
  }
  // L3^-1 on the side stream while the bidiagonalization runs, so that W' = L3^-T W is one product
  // at the end instead of a triangular solve on the critical path. With ||G3 - I||_F <= 1/2 the
  // eigenvalues of G3 = L3 L3^T lie in [1/2, 3/2], so kappa(L3) <= sqrt(3) and the explicit inverse
  // is as accurate as the solve; otherwise (the third FP64 pass still short of that) W' keeps the
  // solve.
  const bool l3_inverse = last_dev <= 0.5;
  // events of the side stream's work (destroyed on every exit)
  struct Events {
    std::vector<hipEvent_t> e;
    hipEvent_t make() {
      hipEvent_t v = nullptr;
      ckh(hipEventCreateWithFlags(&v, hipEventDisableTiming), "event");
      e.push_back(v);
      return v;
    }
    ~Events() {
      for (hipEvent_t v : e) hipEventDestroy(v);
    }
  } events;
  if (!D.s2) ckh(hipStreamCreateWithFlags(&D.s2, hipStreamNonBlocking), "stream");
  // Side-stream work is enqueued from a host thread of its own: its tens of launches would
  // otherwise hold back the main stream's next launch by the time they take on the host. The
  // thread makes the main thread's device current first (the current device is per thread); a
  // failure in it is rethrown by join(), which comes before any wait on the thread's events.
  int device = 0;
  ckh(hipGetDevice(&device), "device");
  struct SideThread {
    std::thread th;
    std::exception_ptr err;
    int device = 0;
    void run(std::function<void()> f) {
      th = std::thread([this, f]() {
        try {
          ckh(hipSetDevice(device), "device");
          f();
        } catch (...) {
          err = std::current_exception();
        }
      });
    }
    void join() {
      if (th.joinable()) th.join();
      if (err) {
        std::exception_ptr e = err;
        err = nullptr;
        std::rethrow_exception(e);
      }
    }
    ~SideThread() {
      if (th.joinable()) th.join();
    }
  } side;
  side.device = device;
  // the side stream's workspace, sized for both of its jobs before either runs
  D.work2.alloc((std::max(exc_fact_workspace(mi, (long)m), exc_ormqr_house_workspace(mi, (long)m)) + 7) / 8, s,
                false);
  D.Pm.alloc(mm, s, false);
  hipEvent_t ev_inv = nullptr;
  if (l3_inverse) {
    hipEvent_t ev_l3 = events.make();
    ev_inv = events.make();
    D.L3i.alloc(mm, s, false);
    ckh(hipEventRecord(ev_l3, s), "event");
    side.run([&D, ev_l3, ev_inv, mi, m = (long)m]() {
      ckh(hipStreamWaitEvent(D.s2, ev_l3, 0), "event");
      ck(exc_fill2d(m, m, 0.0, D.L3i.p, m, D.s2), "identity");
      ck(exc_add_diag(mi, D.L3i.p, m, 1.0, D.s2), "identity");
      ck(exc_trsm_left(0, 0, mi, m, D.L3.p, m, D.L3i.p, m, D.work2.p, D.work2.n * 8, D.s2), "L3 inverse");
      ckh(hipEventRecord(ev_inv, D.s2), "event");
    });
  }
  sync();
  tf += now_seconds() - t0;
  time_local_qr = tf;
  time_apply_mask = t_mask;
  time_exchange_low_dim = t_comm;
  time_global_qr = 0;
  log("Mixed CholeskyQR (FP32 x 2, FP64 x " + std::to_string(fp64_passes) + ") Costs " + std::to_string(tf));

  // ---- the SVD of Lt, on every peer: Lt is bit-identical on all of them (every exchange is a sum
  // in a fixed order that all peers evaluate alike) and every step below is deterministic, so all
  // peers obtain the same sigma, U and W' without sending them; a checksum confirms it.
  int k = 0;
  D.alpha.alloc(m, s);
  D.beta.alloc(m, s);
  D.Uf.alloc(mm, s, false);
  D.bu.alloc(mm, s);
  D.sigma.alloc(m, s);
  {
    t0 = now_seconds();
    D.F.alloc(mm, s, false);
    D.U.alloc(mm, s);
    D.house.alloc(mm, s);
    D.tau_h.alloc(m, s);
    D.cs.alloc(2 * m, s);
    ckh(hipMemcpyAsync(D.F.p, D.Lt.p, mm * 8, hipMemcpyDeviceToDevice, s), "copy");  // F = Lt^T column-major
    if (!D.link || D.link_m != mi) {
      if (D.link) ck(exc_ralha_link_free(D.link), "ralha link");
      D.link = nullptr;
      ck(exc_ralha_link_alloc(mi, &D.link), "ralha link");
      D.link_m = mi;
    }
    D.msgd.alloc((exc_bidiag_local_workspace(mi, mi) + 7) / 8, s, false);
    // peers that share the device run their loops at the same time
    const int colocated = D.samedev ? std::max(1, (int)num_client) : 1;
    ck(exc_bidiag_local(mi, mi, D.F.p, D.U.p, D.house.p, D.tau_h.p, D.alpha.p, D.beta.p, D.cs.p,
                        opt_.ralha_one_launch ? D.link : nullptr, colocated, D.msgd.p, D.msgd.n * 8,
                        std::max(30.0, opt_.recv_timeout_s), s),
       "bidiagonalization");
    sync();
    time_bi_diag = now_seconds() - t0;
    t0 = now_seconds();
    // P = H R^T on the side stream while the bisection and the twisted factorization run: the back
    // rotation R^T (the rotations of the lower-to-upper step) and the Householder product H both
    // multiply the columns of bu from the left, so W = H (R^T bu) = P bu is then one product
    // instead of a chain of them after the twisted factorization.
    hipEvent_t ev_p = events.make();
    {
      hipEvent_t ev_bd = events.make();
      ckh(hipEventRecord(ev_bd, s), "event");
      side.join();  // the L3 inverse's launches are all on the side stream before these
      side.run([&D, ev_bd, ev_p, mi, m = (long)m]() {
        ckh(hipStreamWaitEvent(D.s2, ev_bd, 0), "event");
        ck(exc_fill2d(m, m, 0.0, D.Pm.p, m, D.s2), "identity");
        ck(exc_add_diag(mi, D.Pm.p, m, 1.0, D.s2), "identity");
        ck(exc_rot_back(mi, mi, D.Pm.p, D.cs.p, D.s2), "back rotation");
        ck(exc_ormqr_house_ln(mi, m, D.house.p, D.tau_h.p, D.Pm.p, m, D.work2.p, D.work2.n * 8, D.s2),
           "ormqr house");
        ckh(hipEventRecord(ev_p, D.s2), "event");
      });
    }
    DBuf& b2 = D.b2;
    b2.alloc(2, s);
    DBuf& cnt2 = D.cnt2;
    cnt2.alloc(1, s);
    ck(exc_bisection_bound(mi, D.alpha.p, D.beta.p, b2.p, s), "bound");
    ck(exc_sturm_count(mi, D.alpha.p, D.beta.p, b2.p, 2, (int*)cnt2.p, s), "count");
    int hc[2];
    ckh(hipMemcpyAsync(hc, cnt2.p, sizeof hc, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    k = std::min(hc[0] - hc[1], mi);
    D.bv.alloc(mm, s);
    D.flip.alloc(m, s);
    const size_t wb = exc_bidiag_svd_workspace(mi, k);
    D.work.alloc(std::max(D.work.n, (wb + 7) / 8), s, false);
    ck(exc_bidiag_svd(mi, D.alpha.p, D.beta.p, D.bu.p, D.bv.p, D.sigma.p, k, nullptr, D.work.p, D.work.n * 8, s),
       "bidiagonal SVD");
    ck(exc_sign_normalize(mi, k, D.bu.p, D.bv.p, D.flip.p, s), "sign rule");
    // U = (bv U_r)^T: the left vectors of Lt; W = H (rotations back) bu: the right vectors
    D.work.alloc(std::max(D.work.n, (exc_gemm_workspace((long)m, k, (long)m) + 7) / 8), s, false);
    ck(exc_dgemm(0, 1, (long)m, k, (long)m, 1.0, D.U.p, (long)m, D.bv.p, (long)m, 0.0, D.Uf.p, k, D.work.p,
                 D.work.n * 8, s),
       "U");
    // W = P bu, into D.idm, which then trades places with D.bu
    side.join();
    ckh(hipStreamWaitEvent(s, ev_p, 0), "event");
    ck(exc_dgemm(0, 0, (long)m, k, (long)m, 1.0, D.Pm.p, (long)m, D.bu.p, k, 0.0, D.idm.p, k, D.work.p,
                 D.work.n * 8, s),
       "W");
    D.bu.swap(D.idm);
    sync();
    time_bsvd = now_seconds() - t0;
    t0 = now_seconds();
    if (k == mi && opt_.refine_iters > 0) {
      const size_t rw = exc_oa_refine_workspace(mi);
      D.work.alloc(std::max(D.work.n, (rw + 7) / 8), s, false);
      ck(exc_oa_refine(mi, D.Lt.p, (long)m, 1, D.Uf.p, (long)m, D.bu.p, (long)m, D.sigma.p, opt_.refine_iters, 1e-6,
                       1e-4, D.work.p, D.work.n * 8, s),
         "refinement");
    }
    if (l3_inverse) {  // W' = (L3^-1)^T W, into D.idm, which then trades places with D.bu
      side.join();
      ckh(hipStreamWaitEvent(s, ev_inv, 0), "event");
      D.work.alloc(std::max(D.work.n, (exc_gemm_workspace((long)m, k, (long)m) + 7) / 8), s, false);
      ck(exc_dgemm(1, 0, (long)m, k, (long)m, 1.0, D.L3i.p, (long)m, D.bu.p, k, 0.0, D.idm.p, k, D.work.p,
                   D.work.n * 8, s),
         "W'");
      D.bu.swap(D.idm);
    } else {  // W' = L3^-T W through the upper triangle L3^T
      ck(exc_transpose((long)m, (long)m, D.L3.p, (long)m, D.idm.p, (long)m, s), "transpose");
      ck(exc_trsm_left(1, 0, mi, k, D.idm.p, (long)m, D.bu.p, k, D.work.p, D.work.n * 8, s), "W'");
    }
    sync();
    time_refine = now_seconds() - t0;
  }
  if (num_client > 1) {
    // FNV-1a over k, sigma and the first and last rows of U and W'
    // the five rows gathered on the device, then one copy into pinned memory
    const size_t np = (size_t)1 + 5 * k;
    D.probe.alloc(np, s, false);
    D.h_probe.alloc(np, false);
    const double* rows[5] = {D.sigma.p, D.Uf.p, D.Uf.p + (m - 1) * (size_t)k, D.bu.p, D.bu.p + (m - 1) * (size_t)k};
    for (int q = 0; q < 5; q++)
      ckh(hipMemcpyAsync(D.probe.p + 1 + (size_t)q * k, rows[q], (size_t)k * 8, hipMemcpyDeviceToDevice, s), "copy");
    ckh(hipMemcpyAsync(D.h_probe.p + 1, D.probe.p + 1, (size_t)5 * k * 8, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    D.h_probe.p[0] = k;
    uint64_t h = 1469598103934665603ull;
    const unsigned char* bytes = (const unsigned char*)D.h_probe.p;
    for (size_t q = 0; q < np * sizeof(double); q++) h = (h ^ bytes[q]) * 1099511628211ull;
    const double tb = now_seconds();
    net.broadcast(&h, sizeof h);
    for (int j = 0; j < (int)num_client; j++) {
      if (j == cid) continue;
      uint64_t o = 0;
      net.receive_into(j, &o, sizeof o);
      if (o != h)
        throw std::runtime_error("mixed form: peer " + std::to_string(j) +
                                 " computed a different SVD of the shared factor (checksums differ)");
    }
    t_comm += now_seconds() - tb;
  }
  time_exchange_low_dim = t_comm;

  // ---- results: the leading kk triplets
  t0 = now_seconds();
  const int kk = top_k > 0 ? std::min(top_k, k) : k;
  sigma_count = kk;
  if (compute_u) {
    D.FinalU.alloc(m * (size_t)kk, s, false);
    ck(exc_copy2d((long)m, kk, D.Uf.p, k, D.FinalU.p, kk, s), "FinalU");
  }
  if (compute_v) {
    D.FinalVT.alloc((size_t)kk * local_n, s, false);
    exc_prof::label_next("gemm_f64_final_v");
    ck(exc_dgemm(1, 0, kk, r, (long)m, 1.0, D.bu.p, k, D.X.p, r, 0.0, D.FinalVT.p, r, nullptr, 0, s), "final V");
  }
  sync();
  time_combine_svd_remove_mask = time_refine + (now_seconds() - t0);
  D.x32.release();
  D.y32.release();
  return kk;
}

// The reference's evaluation report (client.cpp:1215-1309, short-and-wide branch), computed on
// the device with the reductions of eval.hip; ground truth from dpath/u.mat (k x m, rows u_i) and
// dpath/vt.mat (k x global_n, rows v_i) as generate_data.cpp writes them. The values go to the log
// under the reference's texts and to one EXC_EVAL line.
extern "C" {
size_t exc_eval_workspace(void);
int exc_eval_reduce(int op, long rows, long cols, const double* d_x, long ldx, const double* d_y, long ldy, long row0,
                    double* d_out, void* d_work, size_t work_bytes, exc_stream_t stream);
}

namespace {
std::vector<double> read_doubles(const std::string& path, size_t count) {
  std::vector<double> v(count);
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t got = fread(v.data(), sizeof(double), count, f);
  fclose(f);
  if (got != count) throw std::runtime_error(path + " is shorter than expected");
  return v;
}
}  // namespace

void DecentralizedClient::evaluate_results(int k, bool compute_u, bool compute_v, bool lqr) {
  exc::ClientDevice& D = *dev_;
  hipStream_t s = D.s;
  auto sync = [&] { ckh(hipStreamSynchronize(s), "stream sync"); };
  DBuf ew, out, tmp, gt;
  ew.alloc((exc_eval_workspace() + 7) / 8, s, false);
  out.alloc(2, s);
  auto reduce = [&](int op, long rows, long cols, const double* x, long ldx, const double* y, long ldy, long row0) {
    ck(exc_eval_reduce(op, rows, cols, x, ldx, y, ldy, row0, out.p, ew.p, ew.n * 8, s), "evaluation");
    double h[2];
    ckh(hipMemcpyAsync(h, out.p, sizeof h, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    return std::make_pair(h[0], h[1]);
  };
  const double* fvt = compute_v ? (lqr ? D.FinalVT.p : D.tmp.p) : nullptr;
  std::string js = "{\"cid\": " + std::to_string(client_id);
  char num[64];
  auto add = [&](const char* key, double v) {
    snprintf(num, sizeof num, "%.17g", v);
    js += std::string(", \"") + key + "\": " + num;
  };
  if (compute_u && client_id == 0) {  // orthogonal_test(FinalU, m, k, false), utils.cpp:38-66
    tmp.alloc((size_t)k * k, s, false);
    ck(exc_dgemm(1, 0, k, k, (long)m, 1.0, D.FinalU.p, k, D.FinalU.p, k, 0.0, tmp.p, k, nullptr, 0, s), "U^T U");
    const auto r = reduce(3, k, k, tmp.p, k, nullptr, 0, 0);
    const double mean = k > 1 ? r.first / ((double)k * k - k) : 0.0;
    log("Orthogonal Test " + std::to_string(mean) + " Max " + std::to_string(r.second));
    log("Final U OrthTest " + std::to_string(mean));
    add("u_orth_mean", mean);
    add("u_orth_max", r.second);
  }
  if (svd_mode == 0 && compute_v && compute_u) {
    // singular vector error (RMSE of |v| entries) against vt.mat, and the SVD error
    const size_t gn = global_n;
    std::vector<double> h = read_doubles(dpath + "/vt.mat", (size_t)k * gn);
    gt.alloc((size_t)k * local_n, s, false);
    ckh(hipMemcpy2DAsync(gt.p, local_n * 8, h.data() + n_pos[client_id], gn * 8, local_n * 8, k,
                         hipMemcpyHostToDevice, s),
        "h2d");
    double err = reduce(2, k, (long)local_n, fvt, (long)local_n, gt.p, (long)local_n, 0).first;
    net_->ring_all_reduce_sum(&err, 1);
    const double rmse = std::sqrt(err / ((double)k * gn));
    add("singular_vector_error_vt", rmse);
    // mean |U diag(sigma) V^T - X| (utils.cpp:96-150)
    tmp.alloc(m * (size_t)k, s, false);
    ckh(hipMemcpyAsync(tmp.p, D.FinalU.p, m * (size_t)k * 8, hipMemcpyDeviceToDevice, s), "copy");
    DBuf sc;
    sc.alloc(m * (size_t)k, s, false);
    // U diag(sigma): scale the columns through a diagonal product
    DBuf dg;
    dg.alloc((size_t)k * k, s);
    std::vector<double> hs(k);
    ckh(hipMemcpyAsync(hs.data(), D.sigma.p, k * 8, hipMemcpyDeviceToHost, s), "d2h");
    sync();
    std::vector<double> hd((size_t)k * k, 0.0);
    for (int i = 0; i < k; i++) hd[(size_t)i * k + i] = hs[i];
    ckh(hipMemcpyAsync(dg.p, hd.data(), hd.size() * 8, hipMemcpyHostToDevice, s), "h2d");
    ck(exc_dgemm(0, 0, (long)m, k, k, 1.0, tmp.p, k, dg.p, k, 0.0, sc.p, k, nullptr, 0, s), "U S");
    DBuf rx;
    rx.alloc(m * local_n, s, false);
    ckh(hipMemcpyAsync(rx.p, D.X0.p, m * local_n * 8, hipMemcpyDeviceToDevice, s), "copy");
    ck(exc_dgemm(0, 0, (long)m, (long)local_n, k, 1.0, sc.p, k, fvt, (long)local_n, -1.0, rx.p, (long)local_n,
                 nullptr, 0, s),
       "U S V^T - X");
    const double e = reduce(0, (long)m, (long)local_n, rx.p, (long)local_n, nullptr, 0, 0).first / ((double)m * local_n);
    log("Final SVD Error " + std::to_string(e));
    add("svd_error", e);
  }
  if ((svd_mode == 1 || svd_mode == 2) && compute_u && client_id == 0) {
    // projection distance of U against u.mat (k x m): the upper triangle of U_gt U_gt^T - U U^T
    std::vector<double> h = read_doubles(dpath + "/u.mat", (size_t)k * m);
    gt.alloc((size_t)k * m, s, false);
    ckh(hipMemcpyAsync(gt.p, h.data(), (size_t)k * m * 8, hipMemcpyHostToDevice, s), "h2d");
    tmp.alloc(m * m, s, false);
    ck(exc_dgemm(1, 0, (long)m, (long)m, k, 1.0, gt.p, (long)m, gt.p, (long)m, 0.0, tmp.p, (long)m, nullptr, 0, s),
       "P_gt");
    ck(exc_dgemm(0, 1, (long)m, (long)m, k, 1.0, D.FinalU.p, k, D.FinalU.p, k, -1.0, tmp.p, (long)m, nullptr, 0, s),
       "P - P_gt");
    const double e = reduce(1, (long)m, (long)m, tmp.p, (long)m, nullptr, 0, 0).first / ((double)m * (m + 1) / 2);
    log("Projection Distance U " + std::to_string(e));
    add("projection_distance_u", e);
  }
  if ((svd_mode == 1 || svd_mode == 3) && compute_v) {
    // projection distance of V (this peer's columns of vt.mat), by row tiles of the
    // local_n x local_n difference
    const size_t gn = global_n;
    std::vector<double> h = read_doubles(dpath + "/vt.mat", (size_t)k * gn);
    gt.alloc((size_t)k * local_n, s, false);
    ckh(hipMemcpy2DAsync(gt.p, local_n * 8, h.data() + n_pos[client_id], gn * 8, local_n * 8, k,
                         hipMemcpyHostToDevice, s),
        "h2d");
    const long ln = (long)local_n, T = std::min<long>(ln, 2048);
    tmp.alloc((size_t)T * ln, s, false);
    double err = 0.0;
    for (long i0 = 0; i0 < ln; i0 += T) {
      const long rows = std::min(T, ln - i0);
      ck(exc_dgemm(1, 0, rows, ln, k, 1.0, gt.p + i0, ln, gt.p, ln, 0.0, tmp.p, ln, nullptr, 0, s), "P_gt");
      ck(exc_dgemm(1, 0, rows, ln, k, 1.0, fvt + i0, ln, fvt, ln, -1.0, tmp.p, ln, nullptr, 0, s), "P - P_gt");
      err += reduce(1, rows, ln, tmp.p, ln, nullptr, 0, i0).first;
    }
    const double e = err / ((double)ln * (ln + 1) / 2);
    log("Projection Distance VT " + std::to_string(e));
    add("projection_distance_vt", e);
  }
  log("Evaluation Done");
  printf("EXC_EVAL %s}\n", js.c_str());
  fflush(stdout);
}
