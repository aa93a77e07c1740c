// Drop-in for the reference's DecentralizedClient (include/client.hpp:31-101): same constructor
// parameters and stage order of run() (client.cpp:447-1213), with every matrix stage executed on
// the GPU through libexcalibur_amd's C ABI and the peer exchanges kept on host TCP.
//
// Supported: the short-and-wide branch (m <= global_n, the branch every benchmark configuration
// takes), modes 1 (two-sided), 2 (left) and 3 (right) with bisection + twisted factorization,
// opt 3 / 1 (Givens masks). Mode 0 (full SVD) is routed through the same bisection + twisted
// path instead of LAPACKE_dbdsdc. The tall-and-skinny branch, mode 4 (linear regression) and the
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
  bool keep_results = false;      // copy FinalU / FinalVT / sigma back into the host members
  bool dump_results = false;      // write them as raw float64 files into log_dir
  std::string form = "fp64";      // "fp64" (plain) or "mixed" (FP32 bulk + FP64 refinement)
  int refine_iters = 2;           // refinement sweeps of the mixed form
  const double* data = nullptr;   // optional in-memory input (m x local_n row-major) instead of dpath
  int repeat = 1;                 // run() this many times on the same input (benchmarking)
  bool profile = false;           // time the MFMA GEMM / Givens launches with HIP events
  bool ralha_one_launch = true;   // the bidiagonalization loop as one kernel launch where it fits
  bool device_exchange = true;  // peers sharing one device exchange data through device memory (IPC)
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

  // results: sigma (sigma_count), FinalU (m x sigma_count), FinalVT (sigma_count x local_n),
  // filled when ClientOptions::keep_results is set
  int sigma_count = 0;
  std::vector<double> sigma, FinalU, FinalVT;

  // stage timers (client.hpp:83-84), seconds
  double time_init = 0, time_local_qr = 0, time_global_qr = 0, time_apply_mask = 0, time_exchange_low_dim = 0,
         time_bi_diag = 0, time_bsvd = 0, time_combine_svd_remove_mask = 0, time_communication = 0, time_total = 0;
  double time_ralha_allreduce = 0;  // part of time_bi_diag spent in the per-column all-reduce

  void load_data();
  void init_server_and_client();
  void run();

 private:
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
