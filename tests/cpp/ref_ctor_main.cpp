// Test driver: constructs DecentralizedClient through the reference's 14-argument constructor
// (include/client.hpp:77 of the reference, main.cpp:104), as reference code does, then reads the
// results from the public host members and writes them as raw float64 files for the test to check.
//   ref_ctor_main cid nclient mode topk m local_n port dpath logdir seed outdir
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

#include "client.hpp"

static void dump(const std::string& path, const double* p, size_t n) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f || !p) {
    if (f) fclose(f);
    return;
  }
  fwrite(p, sizeof(double), n, f);
  fclose(f);
}

int main(int argc, char** argv) {
  if (argc != 12) {
    fprintf(stderr, "usage: %s cid nclient mode topk m local_n port dpath logdir seed outdir\n", argv[0]);
    return 2;
  }
  const int cid = atoi(argv[1]), nclient = atoi(argv[2]), mode = atoi(argv[3]);
  const long topk = atol(argv[4]);
  const size_t m = strtoul(argv[5], nullptr, 10), local_n = strtoul(argv[6], nullptr, 10);
  const int port = atoi(argv[7]);
  const std::string dpath = argv[8], logdir = argv[9], outdir = argv[11];
  const unsigned seed = (unsigned)strtoul(argv[10], nullptr, 10);
  try {
    DecentralizedClient client(cid, nclient, mode, (size_t)topk, m, local_n, port, "127.0.0.1", dpath, false, false,
                               logdir, seed, 3);
    const size_t k = (size_t)client.sigma_count;
    const std::string base = outdir + "/ctor" + std::to_string(cid) + "_";
    dump(base + "sigma.bin", client.sigma, k);
    if (client.FinalU) dump(base + "FinalU.bin", client.FinalU, m * k);
    if (client.FinalVT) dump(base + "FinalVT.bin", client.FinalVT, k * local_n);
    dump(base + "alpha.bin", client.x_diagonal, m);
    printf("ctor ok cid=%d k=%zu d2h=%.6f total=%.6f\n", cid, k, client.time_results_d2h, client.time_total);
  } catch (const std::exception& e) {
    fprintf(stderr, "peer %d failed: %s\n", cid, e.what());
    return 3;
  }
  return 0;
}
