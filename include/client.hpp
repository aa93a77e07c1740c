// Drop-in for the reference's DecentralizedClient (include/client.hpp:31-101): same constructor
// parameters and stage order of run() (client.cpp:447-1213), with every matrix stage executed on
// the GPU through libexcalibur_amd's C ABI and the peer exchanges kept on host TCP.
//
// Supported: the short-and-wide branch (m <= global_n) and the tall-and-skinny branch
// (m > global_n, where the input block is read as local_n x m and FinalU is k x m, as the
// reference has them), modes 1 (two-sided), 2 (left) and 3 (right) with bisection + twisted
// factorization, opt 3 / 1 (Givens masks). Mode 0 (full SVD) is routed through the same
// bisection + twisted path instead of LAPACKE_dbdsdc. Mode 4 (linear regression) and the
// dense-mask ablation (opt 0 / 2) are rejected with an error message.
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <string>
#include <vector>

struct ClientMeta {  // client.hpp:16-21 (wire format of the first broadcast)
  size_t m;
  size_t local_n;
  int cid;
  unsigned int shared_seed;
};

namespace exc {
class PeerNet;
struct ClientDevice;
}  // namespace exc

// Knobs beyond the reference's constructor (defaults reproduce the reference).
struct ClientOptions {
  double recv_timeout_s = 600.0;  // a peer that stops answering ends the run instead of hanging
  bool keep_results = true;       // copy sigma / FinalU / FinalVT back into the host members
  bool dump_results = false;      // write them as raw float64 files into log_dir
  std::string form = "fp64";      // "fp64" (plain) or "mixed" (FP32 bulk + FP64 refinement)
  int refine_iters = 2;           // refinement sweeps of the mixed form
  const double* data = nullptr;   // optional in-memory input (m x local_n row-major) instead of dpath
  int repeat = 1;                 // run() this many times on the same input (benchmarking)
  bool profile = false;           // time the MFMA GEMM / Givens launches with HIP events
  bool ralha_one_launch = true;   // the bidiagonalization loop as one kernel launch where it fits
  // Peers that share one device exchange the protocol's data through device memory (IPC) instead
  // of host TCP. Off by default: the reference's transport is host TCP.
  bool device_exchange = false;
};

class DecentralizedClient {
 public:
  DecentralizedClient(int cid, int num_client, int mode, size_t topk, size_t drow, size_t dcol, int port,
                      std::string listen, std::string dpath, bool is_memmap, bool evaluate, std::string log_dir,
                      unsigned int seed, int opt);
  DecentralizedClient(int cid, int num_client, int mode, size_t topk, size_t drow, size_t dcol, int port,
                      std::string listen, std::string dpath, bool is_memmap, bool evaluate, std::string log_dir,
                      unsigned int seed, int opt, const ClientOptions& options);
  ~DecentralizedClient();

  // configuration (client.hpp:35-45)
  int svd_mode, top_k;
  int client_id, num_client;
  size_t m, local_n, global_n = 0, reduced_global_n = 0;
  std::map<int, size_t> n_pos, reduced_n_pos, local_ns, reduced_local_ns;
  unsigned int public_seed = 0, private_seed = 0;
  int opt_control;
  bool evaluate, is_memmap;
  std::string listen, dpath, log_dir;
  int base_server_port;

  // Results in host memory after construction, as the reference leaves them (client.hpp:51-55):
  // sigma (sigma_count), FinalU (m x sigma_count row-major, modes 0, 1, 2), FinalVT
  // (sigma_count x local_n row-major, modes 0, 1, 3), and the upper bidiagonal x_diagonal (m) /
  // x_upper_diagonal (m - 1). nullptr when not computed or when ClientOptions::keep_results is
  // off. The storage belongs to the client.
  int sigma_count = 0;
  double *sigma = nullptr, *FinalU = nullptr, *FinalVT = nullptr;
  double *x_diagonal = nullptr, *x_upper_diagonal = nullptr;
  // The reference's intermediate buffers (client.hpp:47-54: raw_x, masked_x, shrinked_x, local_v,
  // qr_u, mask_shared_v) and the linear-regression vectors (w_lr, y) live in device memory here and
  // are released by run(); these host pointers stay nullptr.
  double *raw_x = nullptr, *masked_x = nullptr, *shrinked_x = nullptr, *local_v = nullptr, *qr_u = nullptr,
         *mask_shared_v = nullptr, *w_lr = nullptr, *y = nullptr;
  // ring neighbours (client.hpp:66); the mask sizes (client.hpp:73) are declared but never set
  // by the reference
  int my_next_client = 0, my_front_client = 0;
  int global_mask_size = 0, local_mask_size = 0;

  // stage timers (client.hpp:83-84), seconds
  double time_init = 0, time_local_qr = 0, time_global_qr = 0, time_apply_mask = 0, time_exchange_low_dim = 0,
         time_bi_diag = 0, time_bsvd = 0, time_combine_svd_remove_mask = 0, time_remove_mask = 0,
         time_combine_svd = 0, time_communication = 0, time_total = 0;
  double time_ralha_allreduce = 0;  // part of time_bi_diag spent in the per-column all-reduce
  double time_results_d2h = 0;      // copy of the results into the host members, after "Finish"
  double time_refine = 0;           // mixed form: the FP64 refinement of the triplets

  void load_data();
  void remove_input_data();
  void init_server_and_client();
  void run();

  // Communicators (client.hpp:87-93), over the same TCP connections and framing as run().
  // receive() allocates the buffer with new char[] (free it with delete[]).
  void send_to(int cid, char* buffer, size_t size);
  template <typename T>
  void receive(T** buffer, int cid) {
    *buffer = (T*)receive_frame(cid);
  }
  void broadcast(char* buffer, size_t size);
  void ring_all_reduce_sum(double* buffer, int size);
  // Peer src of the nc peers in c_group sends `size` bytes at *buffer; the others get a new
  // buffer (new char[]) in *buffer.
  void all_gather(char** buffer, size_t size, int* c_group, int nc, int src);

  // client.hpp:96 on host memory: QR of the column-major m x n *raw_x (m > n) on the GPU; Q
  // (m x n column-major) into q when compute_q, and *raw_x replaced by a new n x n buffer holding
  // R in row-major order (transpose_x) or in column-major order (R^T row-major), zeros elsewhere,
  // as client.cpp:417-445 leaves it. The old *raw_x is freed with delete[] and the new one
  // allocated with new double[].
  void local_qr_reduction(size_t m, size_t n, double** raw_x, double* q, bool transpose_x, bool compute_q);
  std::string filename_with_cid(std::string filename);

 private:
  char* receive_frame(int cid);
  void copy_results_to_host(int k, bool compute_u, bool compute_v, bool lqr);
  void finish_run(int k, bool compute_u, bool compute_v, bool lqr, double t_run);
  // The reference's evaluation report (client.cpp:1215-1309) when `evaluate` is set.
  void evaluate_results(int k, bool compute_u, bool compute_v, bool lqr);
  // The mixed-precision form of run() after the seeds; returns the number of triplets kept.
  int run_mixed(bool compute_u, bool compute_v);
  // The tall-and-skinny branch (m > global_n) after the seeds; returns the triplets kept.
  int run_tall(bool compute_u, bool compute_v);
  // Stages shared by the branches: the bidiagonalization of this peer's mi x ci block D.F, and
  // the bidiagonal SVD of order n (returns sigma_count).
  void ralha_stage(int mi, int ci, bool one_launch);
  int bsvd_stage(int n);
  int bidiag_order = 0;  // order of the bidiagonal of the last run (m, or global_n when tall)
  std::vector<double> h_sigma_, h_FinalU_, h_FinalVT_, h_alpha_, h_beta_;
  void log(const std::string& line);
  // Once per client: whether all peers share this device (and each runs the one-launch
  // bidiagonalization), in which case they exchange IPC handles of their message boxes.
  void decide_device_exchange(int m, int c, bool one_launch);
  ClientOptions opt_;
  std::unique_ptr<exc::PeerNet> net_;
  std::unique_ptr<exc::ClientDevice> dev_;
  std::vector<double> raw_host_;
  unsigned int seed_;
};
