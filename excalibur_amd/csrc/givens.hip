// Givens-rotation masking sweep on gfx950 (the arithmetic of
// givens_mask_cache_optimized_parallel_with_simd, bbtsvd.cpp:781-875).
//
// A workgroup owns a strip of independent lines (columns for a left mask, rows for a right
// mask) and keeps the whole strip in LDS for all passes: one coalesced HBM read and one write
// per element per call, whatever the number of passes. Within a pass the rotation pairs are
// disjoint (a permutation) except, for odd K, the extra pair the reference appends; that pair
// is applied on its own after (mask) or before (unmask) the others, which is the order the
// reference's sequential loop gives it. Each update is c*x1 + s'*x2 with separately rounded
// products and sum, so the result is bit-identical to the reference's loop.
#include <algorithm>

#include "exc_common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ void rot(T* p1, T* p2, T c, T ss, T nss) {
  const T a = *p1, b = *p2;
  *p1 = a * c + b * ss;
  *p2 = a * nss + b * c;
}

// odd_mode: 0 none, 1 the special pair is pair 0 (applied first), 2 it is the last pair.
template <typename T, bool LEFT>
__device__ __forceinline__ void givens_strip(T* __restrict__ x, long m, long n, int W,
                                             const int32_t* __restrict__ pairs, const double* __restrict__ cs,
                                             int passes, int half, int mask, int odd_mode, T* buf) {
  const int tid = threadIdx.x;
  const long K = LEFT ? m : n;
  const long base = (long)blockIdx.x * W;
  const int Wc = (int)min((long)W, (LEFT ? n : m) - base);
  // load the strip: LEFT buf[r*W + cc] = x[r*n + base + cc]; RIGHT buf[rr*n + q] = x[(base+rr)*n + q]
  if (LEFT) {
    const long tot = K * Wc;
    for (long idx = tid; idx < tot; idx += kThreads) {
      const long r = idx / Wc;
      const int cc = (int)(idx - r * Wc);
      buf[r * W + cc] = x[r * n + base + cc];
    }
  } else {
    const long tot = (long)Wc * n;
    const T* src = x + base * n;
    for (long idx = tid; idx < tot; idx += kThreads) buf[idx] = src[idx];
  }
  __syncthreads();
  const T sgn = mask ? T(1) : T(-1);
  auto apply_range = [&](const int32_t* pp, const double* cp, int j0, int j1) {
    const int nj = j1 - j0;
    const long items = (long)nj * Wc;
    for (long it = tid; it < items; it += kThreads) {
      int j, cc;
      if (LEFT) { j = j0 + (int)(it / Wc); cc = (int)(it % Wc); }
      else { j = j0 + (int)(it % nj); cc = (int)(it / nj); }
      const int r1 = pp[2 * j], r2 = pp[2 * j + 1];
      if ((unsigned)r1 >= (unsigned long)K || (unsigned)r2 >= (unsigned long)K) continue;  // malformed plan
      const T c = (T)cp[2 * j];
      const T s = (T)cp[2 * j + 1];
      const T ss = sgn * s;
      const T nss = -sgn * s;
      if (LEFT) rot(&buf[(long)r1 * W + cc], &buf[(long)r2 * W + cc], c, ss, nss);
      else rot(&buf[(long)cc * n + r1], &buf[(long)cc * n + r2], c, ss, nss);
    }
  };
  for (int p = 0; p < passes; p++) {
    const int32_t* pp = pairs + (long)p * half * 2;
    const double* cp = cs + (long)p * half * 2;
    if (odd_mode == 1) {
      apply_range(pp, cp, 0, 1);
      __syncthreads();
      apply_range(pp, cp, 1, half);
    } else if (odd_mode == 2) {
      apply_range(pp, cp, 0, half - 1);
      __syncthreads();
      apply_range(pp, cp, half - 1, half);
    } else {
      apply_range(pp, cp, 0, half);
    }
    __syncthreads();
  }
  if (LEFT) {
    const long tot = K * Wc;
    for (long idx = tid; idx < tot; idx += kThreads) {
      const long r = idx / Wc;
      const int cc = (int)(idx - r * Wc);
      x[r * n + base + cc] = buf[r * W + cc];
    }
  } else {
    const long tot = (long)Wc * n;
    T* dst = x + base * n;
    for (long idx = tid; idx < tot; idx += kThreads) dst[idx] = buf[idx];
  }
}

template <typename T, bool LEFT>
__global__ __launch_bounds__(kThreads) void k_givens(T* __restrict__ x, long m, long n, int W,
                                                     const int32_t* __restrict__ pairs,
                                                     const double* __restrict__ cs, int passes, int half, int mask,
                                                     int odd_mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  givens_strip<T, LEFT>(x, m, n, W, pairs, cs, passes, half, mask, odd_mode, reinterpret_cast<T*>(smem_raw));
}

// The same sweep over a batch of matrices of one shape, each with its own plan: blockIdx.y picks
// the matrix.
template <typename T, bool LEFT>
__global__ __launch_bounds__(kThreads) void k_givens_batched(T* const* __restrict__ xs, long m, long n, int W,
                                                             const int32_t* const* __restrict__ pairs,
                                                             const double* const* __restrict__ cs, int passes,
                                                             int half, int mask, int odd_mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = blockIdx.y;
  givens_strip<T, LEFT>(xs[b], m, n, W, pairs[b], cs[b], passes, half, mask, odd_mode, reinterpret_cast<T*>(smem_raw));
}

constexpr size_t kLdsSoft = 64 * 1024;   // two workgroups per CU
constexpr size_t kLdsHard = 160 * 1024;  // one workgroup per CU

template <typename T>
int givens_launch(T* x, long m, long n, const int32_t* pairs, const double* cs, long passes, int mask, int is_left,
                  hipStream_t s) {
  EXC_REQUIRE(x && pairs && cs && m >= 1 && n >= 1 && passes >= 0, EXC_ERR_ARG, "exc_givens_apply: bad arguments");
  const long K = is_left ? m : n;
  const long lines = is_left ? n : m;
  if (K <= 1 || passes == 0) return EXC_OK;
  const int half = (int)((K % 2 == 0 ? K : K + 1) / 2);
  const int odd_mode = (K % 2 == 1) ? (mask ? 2 : 1) : 0;
  const size_t line_bytes = (size_t)K * sizeof(T);
  EXC_REQUIRE(line_bytes <= kLdsHard, EXC_ERR_ARG, "exc_givens_apply: a line does not fit in LDS (K too large)");
  long w_lds = (long)(kLdsSoft / line_bytes);
  if (w_lds < 1) w_lds = 1;
  // enough workgroups to spread over the 256 CUs, as long as the strip fits
  long w = lines / 256;
  if (w < 1) w = 1;
  if (w > w_lds) w = w_lds;
  if (w > 64) w = 64;
  const size_t lds = (size_t)w * line_bytes;
  const unsigned grid = exc_detail::ceil_div(lines, w);
  hipEvent_t e0 = exc_prof::mark(s);
  if (is_left) {
    if (lds > 64 * 1024)
      EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_givens<T, true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_givens<T, true>), dim3(grid), dim3(kThreads), lds, s, x, m, n, (int)w, pairs, cs,
                       (int)passes, half, mask, odd_mode);
  } else {
    if (lds > 64 * 1024)
      EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_givens<T, false>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_givens<T, false>), dim3(grid), dim3(kThreads), lds, s, x, m, n, (int)w, pairs, cs,
                       (int)passes, half, mask, odd_mode);
  }
  EXC_CHECK_HIP(hipGetLastError());
  if (e0) {
    // one read and one write of the matrix; 6 flops per element per rotation
    const double bytes = 2.0 * sizeof(T) * (double)m * (double)n;
    const double flops = 6.0 * (double)passes * half * (double)lines * 2.0 / 2.0 * 2.0 / 2.0;
    exc_prof::book(sizeof(T) == 8 ? "givens_f64" : "givens_f32", e0, exc_prof::mark(s), flops, bytes);
  }
  return EXC_OK;
}

template <typename T>
int givens_batched_launch(int nb, T* const* d_xs, long m, long n, const int32_t* const* d_pairs,
                          const double* const* d_cs, long passes, int mask, int is_left, hipStream_t s) {
  EXC_REQUIRE(nb >= 0 && m >= 1 && n >= 1 && passes >= 0 && (nb == 0 || (d_xs && d_pairs && d_cs)), EXC_ERR_ARG,
              "exc_givens_apply_batched: bad arguments");
  const long K = is_left ? m : n;
  const long lines = is_left ? n : m;
  if (nb == 0 || K <= 1 || passes == 0) return EXC_OK;
  const int half = (int)((K % 2 == 0 ? K : K + 1) / 2);
  const int odd_mode = (K % 2 == 1) ? (mask ? 2 : 1) : 0;
  const size_t line_bytes = (size_t)K * sizeof(T);
  EXC_REQUIRE(line_bytes <= kLdsHard, EXC_ERR_ARG, "exc_givens_apply_batched: K too large");
  long w_lds = (long)(kLdsSoft / line_bytes);
  if (w_lds < 1) w_lds = 1;
  // the batch already spreads the work: give each workgroup as many lines as fit in LDS
  long w = std::max<long>(1, std::min<long>({w_lds, 64L, std::max<long>(1, lines * nb / 512)}));
  const size_t lds = (size_t)w * line_bytes;
  const dim3 grid(exc_detail::ceil_div(lines, w), nb);
  if (is_left)
    hipLaunchKernelGGL((k_givens_batched<T, true>), grid, dim3(kThreads), lds, s, d_xs, m, n, (int)w, d_pairs, d_cs,
                       (int)passes, half, mask, odd_mode);
  else
    hipLaunchKernelGGL((k_givens_batched<T, false>), grid, dim3(kThreads), lds, s, d_xs, m, n, (int)w, d_pairs, d_cs,
                       (int)passes, half, mask, odd_mode);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // namespace

extern "C" {

int exc_givens_apply_batched(int nb, double* const* d_xs, long m, long n, const int32_t* const* d_pairs,
                             const double* const* d_cs, long passes, int mask, int is_left, exc_stream_t stream) {
  return givens_batched_launch<double>(nb, d_xs, m, n, d_pairs, d_cs, passes, mask, is_left,
                                       exc_detail::stream(stream));
}

int exc_givens_apply(double* d_x, long m, long n, const int32_t* d_pairs, const double* d_cs, long passes, int mask,
                     int is_left, exc_stream_t stream) {
  return givens_launch<double>(d_x, m, n, d_pairs, d_cs, passes, mask, is_left, exc_detail::stream(stream));
}

int exc_givens_apply_f32(float* d_x, long m, long n, const int32_t* d_pairs, const double* d_cs, long passes,
                         int mask, int is_left, exc_stream_t stream) {
  return givens_launch<float>(d_x, m, n, d_pairs, d_cs, passes, mask, is_left, exc_detail::stream(stream));
}

}  // extern "C"
