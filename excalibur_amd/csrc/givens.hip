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

// Plans of at most this many pairs per pass are staged in LDS (two or three pass buffers): every
// workgroup reads each pass's pairs and rotations from HBM or L2 once, one or two passes ahead of
// their use, instead of once per line it owns and in the shadow of nothing.
constexpr int kStageMaxHalf = 1024;
constexpr int kStagePerThread = kStageMaxHalf / 256;
size_t stage_bytes(int half, int nbuf) { return half <= kStageMaxHalf ? (size_t)half * 24 * nbuf : 0; }
// items (pair, line) per thread and pass up to which a pass takes the precomputed-index path
constexpr int kBatch = 8;

// odd_mode: 0 none, 1 the special pair is pair 0 (applied first), 2 it is the last pair.
// NBUF plan buffers in LDS (stage: NBUF x half x {2 int32, 2 double}); 0: no staging. With 2, the
// plan of pass p + 1 is loaded into registers before the rotations of pass p and stored into the
// other buffer after them; with 3, the plan of pass p + 2 is loaded before the rotations of pass p
// and stored after those of pass p + 1 (two register sets), so a load has two passes to arrive.
template <typename T, bool LEFT, int NBUF>
__device__ __forceinline__ void givens_strip(T* __restrict__ x, long m, long n, int W,
                                             const int32_t* __restrict__ pairs, const double* __restrict__ cs,
                                             int passes, int half, int mask, int odd_mode, T* buf,
                                             unsigned char* stage) {
  constexpr bool STAGE = NBUF > 0;
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
  // The pairs of one pass are disjoint, so the rotations of a range touch disjoint elements: a
  // thread takes several of its items at once (kBatchGeneral here, kBatch on the fast path), all
  // their loads before any of their stores, and the LDS latencies overlap instead of chaining item
  // after item. Every element still receives the same operations in the same order.
  // (LDS indices fit in 32 bits: the strip is at most 160 KB.)
  constexpr int kBatchGeneral = 4;
  auto apply_range = [&](const int32_t* pp, const double* cp, int j0, int j1) {
    const int nj = j1 - j0;
    const int items = nj * Wc;
    for (int it0 = tid; it0 < items; it0 += kThreads * kBatchGeneral) {
      int i1[kBatchGeneral], i2[kBatchGeneral];
      T c[kBatchGeneral], s[kBatchGeneral], a[kBatchGeneral], b[kBatchGeneral];
      bool ok[kBatchGeneral];
#pragma unroll
      for (int u = 0; u < kBatchGeneral; u++) {
        const int it = it0 + u * kThreads;
        ok[u] = false;
        if (it < items) {
          int j, cc;
          if (LEFT) { j = j0 + it / Wc; cc = it - (j - j0) * Wc; }
          else { cc = it / nj; j = j0 + (it - cc * nj); }
          const int r1 = pp[2 * j], r2 = pp[2 * j + 1];
          // a malformed plan's out-of-range pair is skipped
          ok[u] = (unsigned)r1 < (unsigned long)K && (unsigned)r2 < (unsigned long)K;
          i1[u] = LEFT ? r1 * W + cc : cc * (int)n + r1;
          i2[u] = LEFT ? r2 * W + cc : cc * (int)n + r2;
          c[u] = (T)cp[2 * j];
          s[u] = (T)cp[2 * j + 1];
        }
      }
#pragma unroll
      for (int u = 0; u < kBatchGeneral; u++)
        if (ok[u]) {
          a[u] = buf[i1[u]];
          b[u] = buf[i2[u]];
        }
#pragma unroll
      for (int u = 0; u < kBatchGeneral; u++)
        if (ok[u]) {
          const T ss = sgn * s[u];
          const T nss = -sgn * s[u];
          buf[i1[u]] = a[u] * c[u] + b[u] * ss;
          buf[i2[u]] = a[u] * nss + b[u] * c[u];
        }
    }
  };
  // Fast path, when a pass holds at most kBatch items per thread (the common shapes): thread tid
  // owns items tid + u * kThreads of every pass, and the item -> (pair j, line cc) split is the
  // same in every pass, so it is computed once here; per pass an item then costs its plan reads,
  // two 32-bit index products, two loads, the rotation and two stores. For odd K the special pair
  // goes in a sub-step of its own, before (unmask) or after (mask) the others, as in apply_range.
  const int all_items = half * Wc;
  const bool fast = all_items <= kThreads * kBatch;
  int fj[kBatch], fbase[kBatch];
  const int fmul = LEFT ? W : 1;
#pragma unroll
  for (int u = 0; u < kBatch; u++) {
    const int it = tid + u * kThreads;
    fj[u] = -1;
    fbase[u] = 0;
    if (fast && it < all_items) {
      int cc;
      if (LEFT) { fj[u] = it / Wc; cc = it - fj[u] * Wc; }
      else { cc = it / half; fj[u] = it - cc * half; }
      fbase[u] = LEFT ? cc : cc * (int)n;
    }
  }
  const int special = odd_mode == 1 ? 0 : (odd_mode == 2 ? half - 1 : -1);
  // sub = 0: every pair (no special pair); 1: the special pair only; 2: all but the special pair
  auto apply_fast = [&](const int32_t* pp, const double* cp, int sub) {
    int i1[kBatch], i2[kBatch];
    T c[kBatch], sv[kBatch], a[kBatch], b[kBatch];
    bool ok[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; u++) {
      const int j = fj[u];
      ok[u] = j >= 0 && (sub == 0 || ((j == special) == (sub == 1)));
      if (ok[u]) {
        const int r1 = pp[2 * j], r2 = pp[2 * j + 1];
        ok[u] = (unsigned)r1 < (unsigned long)K && (unsigned)r2 < (unsigned long)K;  // malformed plan
        i1[u] = r1 * fmul + fbase[u];
        i2[u] = r2 * fmul + fbase[u];
        c[u] = (T)cp[2 * j];
        sv[u] = (T)cp[2 * j + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; u++)
      if (ok[u]) {
        a[u] = buf[i1[u]];
        b[u] = buf[i2[u]];
      }
#pragma unroll
    for (int u = 0; u < kBatch; u++)
      if (ok[u]) {
        const T ss = sgn * sv[u];
        const T nss = -sgn * sv[u];
        buf[i1[u]] = a[u] * c[u] + b[u] * ss;
        buf[i2[u]] = a[u] * nss + b[u] * c[u];
      }
  };
  // staged plan buffers: b holds half int2 pairs then half double2 rotations
  int32_t* spp[3] = {nullptr, nullptr, nullptr};
  double* scp[3] = {nullptr, nullptr, nullptr};
  // a pass's pairs (r1, r2) and rotations (c, s) of up to kStagePerThread pairs per thread, in
  // registers between their load and their store into LDS; scalar loads, so no alignment beyond
  // the elements' own is assumed of the plan
  struct PlanRegs {
    int32_t a[kStagePerThread], b[kStagePerThread];
    double c[kStagePerThread], s[kStagePerThread];
  };
  PlanRegs R0, R1;
  auto issue = [&](PlanRegs& R, int q) {  // loads of pass q's plan
    const int32_t* gp = pairs + (long)q * half * 2;
    const double* gc = cs + (long)q * half * 2;
#pragma unroll
    for (int u = 0; u < kStagePerThread; u++) {
      const int idx = tid + u * kThreads;
      if (idx < half) {
        R.a[u] = gp[2 * idx];
        R.b[u] = gp[2 * idx + 1];
        R.c[u] = gc[2 * idx];
        R.s[u] = gc[2 * idx + 1];
      }
    }
  };
  auto commit = [&](const PlanRegs& R, int q) {  // pass q's plan into its LDS buffer
    int32_t* np = spp[q % (STAGE ? NBUF : 1)];
    double* nc = scp[q % (STAGE ? NBUF : 1)];
#pragma unroll
    for (int u = 0; u < kStagePerThread; u++) {
      const int idx = tid + u * kThreads;
      if (idx < half) {
        np[2 * idx] = R.a[u];
        np[2 * idx + 1] = R.b[u];
        nc[2 * idx] = R.c[u];
        nc[2 * idx + 1] = R.s[u];
      }
    }
  };
  if (STAGE) {
    for (int b = 0; b < NBUF; b++) {
      unsigned char* base_b = stage + (size_t)b * half * 24;
      spp[b] = reinterpret_cast<int32_t*>(base_b);
      scp[b] = reinterpret_cast<double*>(base_b + (size_t)half * 8);
    }
    for (int idx = tid; idx < 2 * half; idx += kThreads) {
      spp[0][idx] = pairs[idx];
      scp[0][idx] = cs[idx];
    }
    if (NBUF == 3 && passes > 1) issue(R1, 1);
    lds_barrier();
  }
  auto apply_pass = [&](const int32_t* pp, const double* cp) {
    if (fast) {
      if (odd_mode == 1) {
        apply_fast(pp, cp, 1);
        lds_barrier();
        apply_fast(pp, cp, 2);
      } else if (odd_mode == 2) {
        apply_fast(pp, cp, 2);
        lds_barrier();
        apply_fast(pp, cp, 1);
      } else {
        apply_fast(pp, cp, 0);
      }
    } else if (odd_mode == 1) {
      apply_range(pp, cp, 0, 1);
      lds_barrier();
      apply_range(pp, cp, 1, half);
    } else if (odd_mode == 2) {
      apply_range(pp, cp, 0, half - 1);
      lds_barrier();
      apply_range(pp, cp, half - 1, half);
    } else {
      apply_range(pp, cp, 0, half);
    }
  };
  // Pass p reads buffer p % NBUF. With 3 buffers the loads of pass p + 2 go out at its start into
  // the register set of p's parity (free: it held pass p's plan, committed at the end of pass
  // p - 1), and pass p + 1's plan is committed at its end into buffer (p + 1) % 3, last read by
  // pass p - 2; with 2, pass p + 1's plan is loaded at its start and committed at its end into the
  // buffer pass p - 1 read. Either buffer was released by the barrier that ended its last pass.
  auto one_pass = [&](int p, PlanRegs& Rp, PlanRegs& Rn) {
    const int32_t* pp = STAGE ? spp[p % (STAGE ? NBUF : 1)] : pairs + (long)p * half * 2;
    const double* cp = STAGE ? scp[p % (STAGE ? NBUF : 1)] : cs + (long)p * half * 2;
    if (NBUF == 3 && p + 2 < passes) issue(Rp, p + 2);
    if (NBUF == 2 && p + 1 < passes) issue(Rp, p + 1);
    apply_pass(pp, cp);
    if (NBUF == 3 && p + 1 < passes) commit(Rn, p + 1);
    if (NBUF == 2 && p + 1 < passes) commit(Rp, p + 1);
    lds_barrier();
  };
  for (int p = 0; p < passes; p += 2) {
    one_pass(p, R0, R1);
    if (p + 1 < passes) one_pass(p + 1, R1, R0);
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

// LDS: the strip (W lines, rounded up to 256 bytes), then the NBUF staged plan buffers.
template <typename T, bool LEFT, int NBUF>
__global__ __launch_bounds__(kThreads) void k_givens(T* __restrict__ x, long m, long n, int W,
                                                     const int32_t* __restrict__ pairs,
                                                     const double* __restrict__ cs, int passes, int half, int mask,
                                                     int odd_mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const size_t strip = ((size_t)W * (LEFT ? m : n) * sizeof(T) + 255) / 256 * 256;
  givens_strip<T, LEFT, NBUF>(x, m, n, W, pairs, cs, passes, half, mask, odd_mode, reinterpret_cast<T*>(smem_raw),
                               smem_raw + strip);
}

// The same sweep over a batch of matrices of one shape, each with its own plan: blockIdx.y picks
// the matrix.
template <typename T, bool LEFT, int NBUF>
__global__ __launch_bounds__(kThreads) void k_givens_batched(T* const* __restrict__ xs, long m, long n, int W,
                                                             const int32_t* const* __restrict__ pairs,
                                                             const double* const* __restrict__ cs, int passes,
                                                             int half, int mask, int odd_mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int b = blockIdx.y;
  const size_t strip = ((size_t)W * (LEFT ? m : n) * sizeof(T) + 255) / 256 * 256;
  givens_strip<T, LEFT, NBUF>(xs[b], m, n, W, pairs[b], cs[b], passes, half, mask, odd_mode,
                               reinterpret_cast<T*>(smem_raw), smem_raw + strip);
}

constexpr size_t kLdsSoft = 80 * 1024;   // two workgroups per CU
constexpr size_t kLdsHard = 160 * 1024;  // one workgroup per CU

// The staged plan (nbuf buffers, none when it does not fit beside one line) and the lines per
// workgroup: w_want, or as many as fit beside the plan within the soft budget.
void sweep_geom(int half, size_t line_bytes, int nbuf, long w_want, size_t& stage, long& w) {
  stage = stage_bytes(half, nbuf) + 256 <= kLdsHard - line_bytes ? stage_bytes(half, nbuf) : 0;
  const size_t budget = kLdsSoft > stage + line_bytes ? kLdsSoft - stage : line_bytes;
  const long w_lds = std::max<long>(1, (long)(budget / line_bytes));
  w = std::max<long>(1, std::min<long>(w_want, w_lds));
}

// Three plan buffers when the third costs no line per workgroup and the whole still leaves two
// workgroups per CU; else two (0: the plan is not staged).
int sweep_plan(int half, size_t line_bytes, long w_want, size_t& stage, long& w) {
  size_t s3, s2;
  long w3, w2;
  sweep_geom(half, line_bytes, 3, w_want, s3, w3);
  sweep_geom(half, line_bytes, 2, w_want, s2, w2);
  if (s3 && w3 == w2 && ((size_t)w3 * line_bytes + 255) / 256 * 256 + s3 <= kLdsSoft) {
    stage = s3;
    w = w3;
    return 3;
  }
  stage = s2;
  w = w2;
  return s2 ? 2 : 0;
}

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
  // enough workgroups to spread over the 256 CUs, as long as the strip fits and a pass keeps at
  // most kBatch items per thread (the precomputed-index path)
  const long w_want = std::max<long>(1, std::min<long>({lines / 256, 64L, (long)kThreads * kBatch / half}));
  size_t stage = 0;
  long w = 1;
  const int nbuf = sweep_plan(half, line_bytes, w_want, stage, w);
  const size_t lds = ((size_t)w * line_bytes + 255) / 256 * 256 + stage;
  const unsigned grid = exc_detail::ceil_div(lines, w);
  hipEvent_t e0 = exc_prof::mark(s);
  auto launch = [&](auto kernel) -> int {
    if (lds > 64 * 1024)
      EXC_CHECK_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), lds, s, x, m, n, (int)w, pairs, cs, (int)passes, half,
                       mask, odd_mode);
    return EXC_OK;
  };
  if (is_left) {
    if (nbuf == 3) launch(k_givens<T, true, 3>);
    else if (nbuf == 2) launch(k_givens<T, true, 2>);
    else launch(k_givens<T, true, 0>);
  } else {
    if (nbuf == 3) launch(k_givens<T, false, 3>);
    else if (nbuf == 2) launch(k_givens<T, false, 2>);
    else launch(k_givens<T, false, 0>);
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
  // the batch already spreads the work: give each workgroup as many lines as fit in LDS
  const long w_want = std::max<long>(1, std::min<long>(64L, lines * nb / 512));
  size_t stage = 0;
  long w = 1;
  const int nbuf = sweep_plan(half, line_bytes, w_want, stage, w);
  const size_t lds = ((size_t)w * line_bytes + 255) / 256 * 256 + stage;
  const dim3 grid(exc_detail::ceil_div(lines, w), nb);
  auto launch = [&](auto kernel) {
    if (lds > 64 * 1024)
      EXC_CHECK_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kernel, grid, dim3(kThreads), lds, s, d_xs, m, n, (int)w, d_pairs, d_cs, (int)passes, half,
                       mask, odd_mode);
    return EXC_OK;
  };
  if (is_left) {
    if (nbuf == 3) launch(k_givens_batched<T, true, 3>);
    else if (nbuf == 2) launch(k_givens_batched<T, true, 2>);
    else launch(k_givens_batched<T, true, 0>);
  } else {
    if (nbuf == 3) launch(k_givens_batched<T, false, 3>);
    else if (nbuf == 2) launch(k_givens_batched<T, false, 2>);
    else launch(k_givens_batched<T, false, 0>);
  }
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
