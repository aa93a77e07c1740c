// Test driver of the host transport (excalibur_amd/csrc/net.cpp) without a GPU: peer `cid` of P
// fills buffers from a seeded generator, runs the collectives the mixed form uses and prints a
// digest of each result, so that the caller can check that every peer holds the same bits and
// that they equal the sum in peer order.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "net.hpp"

namespace {
uint64_t fnv(const void* p, size_t n) {
  uint64_t h = 1469598103934665603ull;
  const unsigned char* b = (const unsigned char*)p;
  for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}
// the value of entry q on peer j (exactly representable in FP32 and FP64)
double value(int j, long q) {
  const long long r = (long long)(((unsigned long long)q * 2654435761ull + 97ull * (j + 1)) % 2000003ull);
  return (double)(r - 1000001) / 1024.0;
}
}  // namespace

int main(int argc, char** argv) {
  if (argc < 5) return 2;
  const int cid = atoi(argv[1]), P = atoi(argv[2]), port = atoi(argv[3]);
  const long n = atol(argv[4]);
  exc::PeerNet net(cid, P, "127.0.0.1", port, 60.0);
  net.connect_all(30.0);
  std::vector<double> d(n);
  std::vector<float> f(n);
  for (long q = 0; q < n; q++) {
    d[q] = value(cid, q);
    f[q] = (float)value(cid, q);
  }
  net.all_reduce_sum_same(d.data(), n);
  std::vector<double> o(n);
  net.all_reduce_sum_same_f32(f.data(), o.data(), n);
  // the expected sums in peer order, from the same generator
  long bad = 0;
  for (long q = 0; q < n; q++) {
    double acc = 0.0;
    for (int j = 0; j < P; j++) acc = acc + value(j, q);
    if (o[q] != acc) bad++;
  }
  // receive_direct: every peer sends its values to every other peer twice; the first round is
  // received at once (the buffer is posted before the frame arrives, for large n), the second
  // after a pause (the frame is already queued and is copied)
  long bad_direct = 0;
  for (int round = 0; round < 2; round++) {
    std::vector<double> mine(n), got(n);
    for (long q = 0; q < n; q++) mine[q] = value(cid, q) + round;
    net.barrier();
    net.broadcast(mine.data(), n * sizeof(double));
    if (round == 1) std::this_thread::sleep_for(std::chrono::milliseconds(200));
    for (int j = 0; j < P; j++) {
      if (j == cid) continue;
      net.receive_direct(j, got.data(), n * sizeof(double));
      for (long q = 0; q < n; q++)
        if (got[q] != value(j, q) + round) bad_direct++;
    }
  }
  printf("%016llx %016llx %ld %ld\n", (unsigned long long)fnv(d.data(), n * 8),
         (unsigned long long)fnv(o.data(), n * 8), bad, bad_direct);
  net.barrier();
  return 0;
}
