// Shared helpers for the gfx950 kernels and the C ABI of libexcalibur_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "excalibur_hip.h"

#define EXC_CHECK_HIP(expr)                                                                  \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      exc_detail::set_last_error(#expr, _e, __FILE__, __LINE__);                             \
      return EXC_ERR_HIP;                                                                    \
    }                                                                                        \
  } while (0)

#define EXC_REQUIRE(cond, code, msg)                                                         \
  do {                                                                                       \
    if (!(cond)) {                                                                           \
      exc_detail::set_last_message(msg, __FILE__, __LINE__);                                 \
      return (code);                                                                         \
    }                                                                                        \
  } while (0)

namespace exc_detail {
void set_last_error(const char* what, hipError_t e, const char* file, int line);
void set_last_message(const char* what, const char* file, int line);
inline hipStream_t stream(exc_stream_t s) { return reinterpret_cast<hipStream_t>(s); }
inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }
}  // namespace exc_detail

// Wavefront helpers (CDNA: 64 lanes).
constexpr int kWave = 64;

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ inline float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// One DPP move of a double (both halves), lanes without a source or outside ROW_MASK reading 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The sum over the 64 lanes as a wave-uniform value, in a fixed order: prefix sums within each
// row of 16 lanes (row_shr 1, 2, 4, 8), then the row totals carried into lane 63 by row_bcast 15
// (rows 1 and 3) and row_bcast 31 (rows 2 and 3), and lane 63 read back. All register moves, no
// LDS traffic; several times shorter than the shuffle butterfly of wave_sum.
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_move_f64<0x111, 0xf>(v);
  v += dpp_move_f64<0x112, 0xf>(v);
  v += dpp_move_f64<0x114, 0xf>(v);
  v += dpp_move_f64<0x118, 0xf>(v);
  v += dpp_move_f64<0x142, 0xa>(v);
  v += dpp_move_f64<0x143, 0xc>(v);
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}

// Data shared between workgroups of one running grid without cache maintenance: relaxed
// device-scope atomics compile to sc1 loads and stores, coherent across the XCDs' L2s, so no
// L2 writeback or invalidate (what a device-scope fence costs on gfx950) is needed. A writer stores
// with coh_st, waits for its own stores (coh_stores_done, s_waitcnt vmcnt(0)) in every wavefront,
// meets the others at a workgroup barrier, then one thread raises a flag with a relaxed atomic; a
// reader spins on the flag with relaxed loads and then reads the data with coh_ld.
__device__ __forceinline__ double coh_ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coh_st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void coh_stores_done() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0)

// Workgroup barrier that orders LDS accesses only. __syncthreads also waits for every outstanding
// global load and store of the thread (vmcnt(0)); where a barrier only publishes LDS data, this one
// lets global stores that no other thread reads before the kernel ends (and prefetches) stay in
// flight across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Live kernel timing with HIP events on the launch stream (off unless exc_prof_enable(1)).
namespace exc_prof {
bool enabled();
// Returns an event recorded on `s` now (nullptr when profiling is off).
hipEvent_t mark(hipStream_t s);
// Books the interval between two marks under `name` with its algorithmic flops and bytes; the
// same interval is also booked under the label set by label_next() on this thread, if any.
void book(const char* name, hipEvent_t start, hipEvent_t end, double flops, double bytes);
// Names the next booked launch of this thread a second time (e.g. one particular product).
void label_next(const char* label);
}  // namespace exc_prof
