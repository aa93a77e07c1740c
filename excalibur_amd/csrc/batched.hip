// Batched local stage: the local QR and the Givens masks of many peers' blocks at once
// (client.cpp:572-605 for every peer), for a host that runs many peers on one device.
//
// Every block X_b (m x r_b row-major, r_b > 5 m so that the reference reduces it by a local QR)
// gets R_b^T (m x m, lower) from a CholeskyQR whose bulk pass runs in FP32 on the MFMA:
//   pass 1 (FP32): G_b = X_b X_b^T for all blocks in one grouped launch (the autotuner picks the K
//     chunk and the persistent grid), L1_b = chol(G_b + s_b I) with s_b = u32 trace(G_b), doubled
//     until the factor exists;
//   pass 2 (FP64): Z_b = L1_b^-1 X_b (one grouped launch, the inverses from one batched launch),
//     G3_b = Z_b Z_b^T (one grouped launch), L3_b = chol(G3_b); a pass repeats for blocks whose
//     G3_b is still far from I (ill-conditioned blocks);
//   R_b^T = L1_b L3_b (one grouped launch), Q_b^T = L3_b^-1 Z_b when asked.
// The masked blocks P R_b^T Q_b (public left plan, private right plan per block, bbtsvd.cpp:
// 727-876) come from two batched sweeps. The per-block m x m factorizations run one workgroup per
// block in one launch each.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "exc_common.hpp"
#include "fact.hpp"
#include "gemm.hpp"

namespace {

// Lower Cholesky of A_b + shift_b I for every block, in place (upper triangle zeroed); info_b = 0
// or 1 + the index of the first non-positive pivot. One workgroup per block: column k is scaled,
// staged in LDS, and each thread updates whole rows of the trailing lower triangle (orders above
// the tile kernels' limit; below it the tile form factors all blocks in one launch).
__global__ __launch_bounds__(256) void k_potrf_batched(int n, double* const* __restrict__ As, long lda,
                                                       const double* __restrict__ shift, int* __restrict__ info) {
  extern __shared__ double col[];
  __shared__ int bad;
  double* A = As[blockIdx.x];
  const double sh = shift ? shift[blockIdx.x] : 0.0;
  if (threadIdx.x == 0) bad = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) A[(long)i * lda + i] += sh;
  __syncthreads();
  for (int k = 0; k < n; k++) {
    if (threadIdx.x == 0) {
      const double d = A[(long)k * lda + k];
      if (!(d > 0.0)) bad = k + 1;
      else A[(long)k * lda + k] = sqrt(d);
    }
    __syncthreads();
    if (bad) break;
    const double piv = A[(long)k * lda + k];
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      const double v = A[(long)i * lda + k] / piv;
      A[(long)i * lda + k] = v;
      col[i] = v;
    }
    __syncthreads();
    for (int i = k + 1 + threadIdx.x; i < n; i += blockDim.x) {
      const double li = col[i];
      double* row = A + (long)i * lda;
      for (int j = k + 1; j <= i; j++) row[j] -= li * col[j];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) info[blockIdx.x] = bad;
  for (long e = threadIdx.x; e < (long)n * n; e += blockDim.x) {
    const long i = e / n, j = e - i * n;
    if (j > i) A[i * lda + j] = 0.0;
  }
}

// A_b += shift_b I for every block.
__global__ __launch_bounds__(256) void k_add_shift_batched(int n, double* const* __restrict__ As, long lda,
                                                           const double* __restrict__ shift) {
  double* A = As[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(long)i * lda + i] += shift[blockIdx.y];
}

// X_b = L_b^-1 (L_b lower, n x n) by forward substitution, one thread per column of X (orders
// above the tile kernels' limit).
__global__ __launch_bounds__(256) void k_trinv_batched(int n, const double* const* __restrict__ Ls, long ldl,
                                                       double* const* __restrict__ Xs, long ldx) {
  const double* L = Ls[blockIdx.y];
  double* X = Xs[blockIdx.y];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  for (int i = 0; i < j; i++) X[(long)i * ldx + j] = 0.0;
  for (int i = j; i < n; i++) {
    double s = i == j ? 1.0 : 0.0;
    for (int k = j; k < i; k++) s -= L[(long)i * ldl + k] * X[(long)k * ldx + j];
    X[(long)i * ldx + j] = s / L[(long)i * ldl + i];
  }
}

// out[2b] = trace(A_b), out[2b+1] = max |A_b - I|.
__global__ __launch_bounds__(256) void k_stats_batched(int n, const double* const* __restrict__ As, long lda,
                                                       double* __restrict__ out) {
  __shared__ double red[4];
  const double* A = As[blockIdx.x];
  double mx = 0.0;
  for (long e = threadIdx.x; e < (long)n * n; e += blockDim.x) {
    const long i = e / n, j = e - i * n;
    mx = fmax(mx, fabs(A[i * lda + j] - (i == j ? 1.0 : 0.0)));
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < n; i++) t += A[(long)i * lda + i];
    out[2 * blockIdx.x] = t;
    out[2 * blockIdx.x + 1] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  }
}

__global__ __launch_bounds__(256) void k_f32_to_f64_batched(long n, const float* const* __restrict__ src,
                                                            double* const* __restrict__ dst) {
  const float* s = src[blockIdx.y];
  double* d = dst[blockIdx.y];
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    d[e] = (double)s[e];
}

__global__ __launch_bounds__(256) void k_f64_to_f32(long n, const double* __restrict__ src, float* __restrict__ dst) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    dst[e] = (float)src[e];
}

struct Arena {
  char* p;
  size_t used = 0, cap;
  template <typename T>
  T* take(size_t n) {
    const size_t off = (used + 255) & ~(size_t)255;
    used = off + n * sizeof(T);
    return p ? (T*)(p + off) : nullptr;
  }
};

}  // namespace

extern "C" {

size_t exc_batched_local_qr_workspace(int nb, int m, const long* h_cols) {
  if (nb < 1 || m < 1 || !h_cols) return 0;
  Arena a{nullptr, 0, 0};
  size_t xs = 0;
  long kmax = 0;
  for (int b = 0; b < nb; b++) {
    xs += (size_t)m * h_cols[b];
    kmax = std::max(kmax, h_cols[b]);
  }
  const size_t mm = (size_t)m * m;
  a.take<float>(xs);                        // FP32 copies
  a.take<float>(mm * nb);                   // FP32 Grams
  a.take<double>(mm * nb * 4);              // G, L1, L3, scratch (inverse)
  a.take<double>(xs);                       // Z
  a.take<void*>(8 * (size_t)nb);            // pointer tables
  a.take<double>(2 * (size_t)nb + 2 * nb);  // shifts, stats
  a.take<int>(nb);
  a.take<char>(fact::tiles_batched_work_bytes(m, nb));  // tile flags
  std::vector<ExcGemmProblem> probs(nb);
  for (int b = 0; b < nb; b++) {
    probs[b].M = m;
    probs[b].N = m;
    probs[b].K = h_cols[b];
    probs[b].lower_only = 1;
    probs[b].mirror = 1;
  }
  size_t g = exc_gemm_grouped_workspace(8, nb, probs.data(), -1);
  for (int b = 0; b < nb; b++) {
    probs[b].N = h_cols[b];
    probs[b].K = m;
    probs[b].lower_only = probs[b].mirror = 0;
  }
  g = std::max(g, exc_gemm_grouped_workspace(8, nb, probs.data(), -1));
  a.take<char>(g);
  return a.used + 256;
}

// See the file comment. d_x[b]: X_b (m x cols[b], row-major, FP64, not modified); d_rt[b]: R_b^T
// (m x m, lower); d_qt[b] (nullable): Q_b^T (m x cols[b]). h_stats (nullable, 3 nb doubles):
// the pass-1 shift, max |G3 - I| before the last FP64 pass, and the number of FP64 passes per block.
int exc_batched_local_qr(int nb, int m, const long* h_cols, const double* const* h_x, double* const* h_rt,
                         double* const* h_qt, double* h_stats, void* d_work, size_t work_bytes, exc_stream_t stream) {
  EXC_REQUIRE(nb >= 1 && m >= 1 && h_cols && h_x && h_rt && d_work, EXC_ERR_ARG, "exc_batched_local_qr: bad arguments");
  for (int b = 0; b < nb; b++)
    EXC_REQUIRE(h_cols[b] >= m && h_x[b] && h_rt[b], EXC_ERR_ARG, "exc_batched_local_qr: block narrower than m");
  EXC_REQUIRE(work_bytes >= exc_batched_local_qr_workspace(nb, m, h_cols), EXC_ERR_SPACE,
              "exc_batched_local_qr: workspace too small");
  hipStream_t s = exc_detail::stream(stream);
  const size_t mm = (size_t)m * m;
  Arena a{(char*)d_work, 0, work_bytes};
  size_t xs = 0;
  std::vector<size_t> xoff(nb);
  for (int b = 0; b < nb; b++) {
    xoff[b] = xs;
    xs += (size_t)m * h_cols[b];
  }
  float* x32 = a.take<float>(xs);
  float* g32 = a.take<float>(mm * nb);
  double* G = a.take<double>(mm * nb);
  double* L1 = a.take<double>(mm * nb);
  double* L3 = a.take<double>(mm * nb);
  double* S = a.take<double>(mm * nb);
  double* Z = a.take<double>(xs);
  void** tab = a.take<void*>(8 * (size_t)nb);
  double* shift = a.take<double>(2 * (size_t)nb);
  double* stats = a.take<double>(2 * (size_t)nb);
  int* info = a.take<int>(nb);
  const size_t tfb = fact::tiles_batched_work_bytes(m, nb);
  void* tflags = a.take<char>(tfb);
  const size_t gws = work_bytes - a.used - 256;
  void* gwork = a.take<char>(gws);
  // pointer tables: 0 G, 1 L1, 2 L3, 3 S, 4 g32, 5 Z, 6 rt, 7 qt
  std::vector<void*> h(8 * (size_t)nb);
  for (int b = 0; b < nb; b++) {
    h[0 * nb + b] = G + mm * b;
    h[1 * nb + b] = L1 + mm * b;
    h[2 * nb + b] = L3 + mm * b;
    h[3 * nb + b] = S + mm * b;
    h[4 * nb + b] = g32 + mm * b;
    h[5 * nb + b] = Z + xoff[b];
    h[6 * nb + b] = h_rt[b];
    h[7 * nb + b] = h_qt ? h_qt[b] : nullptr;
  }
  EXC_CHECK_HIP(hipMemcpyAsync(tab, h.data(), h.size() * sizeof(void*), hipMemcpyHostToDevice, s));
  double* const* tG = (double* const*)(tab + 0 * nb);
  double* const* tL1 = (double* const*)(tab + 1 * nb);
  double* const* tL3 = (double* const*)(tab + 2 * nb);
  double* const* tS = (double* const*)(tab + 3 * nb);
  const float* const* tg32 = (const float* const*)(tab + 4 * nb);
  const size_t lds = (size_t)m * sizeof(double);
  // Cholesky of every block (shifted when `sh` is given) and the inverses of lower factors: the
  // tile forms (all blocks in one launch) up to their order limit, one workgroup per block above
  auto chol = [&](double* const* tA, const double* sh) -> int {
    if (m <= fact::kTileMaxN) {
      if (sh) {
        hipLaunchKernelGGL(k_add_shift_batched, dim3(exc_detail::ceil_div(m, 256), nb), dim3(256), 0, s, m, tA,
                           (long)m, sh);
        EXC_CHECK_HIP(hipGetLastError());
      }
      return fact::potrf_tiles_batched(m, tA, nb, (long)m, info, tflags, tfb, s);
    }
    hipLaunchKernelGGL(k_potrf_batched, dim3(nb), dim3(256), lds, s, m, tA, (long)m, sh, info);
    EXC_CHECK_HIP(hipGetLastError());
    return EXC_OK;
  };
  auto trinv = [&](double* const* tL, double* const* tX) -> int {
    if (m <= fact::kTileMaxN)
      return fact::trinv_tiles_batched(m, (const double* const*)tL, (long)m, tX, (long)m, nb, tflags, tfb, s);
    hipLaunchKernelGGL(k_trinv_batched, dim3(exc_detail::ceil_div(m, 256), nb), dim3(256), 0, s, m,
                       (const double* const*)tL, (long)m, tX, (long)m);
    EXC_CHECK_HIP(hipGetLastError());
    return EXC_OK;
  };
  // ---- pass 1 (FP32)
  for (int b = 0; b < nb; b++) {
    const long n = (long)m * h_cols[b];
    hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)std::min<long>(4096, (n + 255) / 256)), dim3(256), 0, s, n, h_x[b],
                       x32 + xoff[b]);
  }
  EXC_CHECK_HIP(hipGetLastError());
  std::vector<ExcGemmProblem> P(nb);
  for (int b = 0; b < nb; b++) {
    ExcGemmProblem& q = P[b];
    q = ExcGemmProblem{};
    q.A = x32 + xoff[b];
    q.lda = h_cols[b];
    q.B = x32 + xoff[b];
    q.ldb = h_cols[b];
    q.C = g32 + mm * b;
    q.ldc = m;
    q.M = q.N = m;
    q.K = h_cols[b];
    q.lower_only = q.mirror = 1;
  }
  int rc = exc_gemm_grouped_opts(4, 0, 1, nb, P.data(), 1.0, 0.0, -1, 0, gwork, gws, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_f32_to_f64_batched, dim3((unsigned)std::min<size_t>(1024, (mm + 255) / 256), nb), dim3(256), 0, s,
                     (long)mm, tg32, tG);
  hipLaunchKernelGGL(k_stats_batched, dim3(nb), dim3(256), 0, s, m, (const double* const*)tG, (long)m, stats);
  EXC_CHECK_HIP(hipGetLastError());
  std::vector<double> hs(2 * (size_t)nb), sh(nb);
  EXC_CHECK_HIP(hipMemcpyAsync(hs.data(), stats, hs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  const double u32 = std::ldexp(1.0, -24);
  for (int b = 0; b < nb; b++) sh[b] = u32 * hs[2 * b];
  std::vector<int> hi(nb, 1);
  for (int attempt = 0; attempt < 60; attempt++) {
    EXC_CHECK_HIP(hipMemcpyAsync(L1, G, mm * nb * sizeof(double), hipMemcpyDeviceToDevice, s));
    EXC_CHECK_HIP(hipMemcpyAsync(shift, sh.data(), nb * sizeof(double), hipMemcpyHostToDevice, s));
    rc = chol(tL1, shift);
    if (rc) return rc;
    EXC_CHECK_HIP(hipMemcpyAsync(hi.data(), info, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    EXC_CHECK_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int b = 0; b < nb; b++)
      if (hi[b]) {
        all = false;
        sh[b] *= 2;
      }
    if (all) break;
    EXC_REQUIRE(attempt < 59, EXC_ERR_NUMERIC, "exc_batched_local_qr: no shift made a Gram matrix positive definite");
  }
  if (h_stats)
    for (int b = 0; b < nb; b++) h_stats[3 * b] = sh[b];
  // ---- FP64 passes: Z = Linv X (grouped), G3 = Z Z^T (grouped), L3 = chol(G3)
  rc = trinv(tL1, tS);
  if (rc) return rc;
  for (int b = 0; b < nb; b++) {
    ExcGemmProblem& q = P[b];
    q = ExcGemmProblem{};
    q.A = S + mm * b;
    q.lda = m;
    q.B = h_x[b];
    q.ldb = h_cols[b];
    q.C = Z + xoff[b];
    q.ldc = h_cols[b];
    q.M = m;
    q.N = h_cols[b];
    q.K = m;
    q.tri = EXC_TRI_A_LOWER;
  }
  rc = exc_gemm_grouped_opts(8, 0, 0, nb, P.data(), 1.0, 0.0, -1, 0, gwork, gws, s);
  if (rc) return rc;
  std::vector<int> passes(nb, 1);
  std::vector<double> dev(nb, 0.0);
  // R1 = L1 (kept in L1); every further pass multiplies its factor into it
  for (int pass = 0; pass < 3; pass++) {
    for (int b = 0; b < nb; b++) {
      ExcGemmProblem& q = P[b];
      q = ExcGemmProblem{};
      q.A = Z + xoff[b];
      q.lda = h_cols[b];
      q.B = Z + xoff[b];
      q.ldb = h_cols[b];
      q.C = G + mm * b;
      q.ldc = m;
      q.M = q.N = m;
      q.K = h_cols[b];
      q.lower_only = q.mirror = 1;
    }
    rc = exc_gemm_grouped_opts(8, 0, 1, nb, P.data(), 1.0, 0.0, -1, 0, gwork, gws, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_stats_batched, dim3(nb), dim3(256), 0, s, m, (const double* const*)tG, (long)m, stats);
    EXC_CHECK_HIP(hipMemcpyAsync(L3, G, mm * nb * sizeof(double), hipMemcpyDeviceToDevice, s));
    rc = chol(tL3, nullptr);
    if (rc) return rc;
    EXC_CHECK_HIP(hipMemcpyAsync(hs.data(), stats, hs.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    EXC_CHECK_HIP(hipMemcpyAsync(hi.data(), info, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    EXC_CHECK_HIP(hipStreamSynchronize(s));
    bool again = false;
    for (int b = 0; b < nb; b++) {
      EXC_REQUIRE(hi[b] == 0, EXC_ERR_NUMERIC, "exc_batched_local_qr: an FP64 pass met a non-positive pivot");
      dev[b] = hs[2 * b + 1];
      again = again || dev[b] > 1e-2;
    }
    if (!again || pass == 2) break;
    // another pass for all blocks (cheap for the well-conditioned ones): Z <- L3^-1 Z, L1 <- L1 L3
    rc = trinv(tL3, tS);
    if (rc) return rc;
    for (int b = 0; b < nb; b++) {
      ExcGemmProblem& q = P[b];
      q = ExcGemmProblem{};
      q.A = S + mm * b;
      q.lda = m;
      q.B = Z + xoff[b];
      q.ldb = h_cols[b];
      q.C = h_qt ? h_qt[b] : (void*)(x32 + xoff[b]);  // scratch: the FP32 copies are no longer needed
      q.ldc = h_cols[b];
      q.M = m;
      q.N = h_cols[b];
      q.K = m;
      q.tri = EXC_TRI_A_LOWER;
    }
    // the product may not overwrite its own input: d_qt serves as its FP64 scratch
    EXC_REQUIRE(h_qt != nullptr, EXC_ERR_NUMERIC,
                "exc_batched_local_qr: an ill-conditioned block needs a second FP64 pass; pass d_qt for its scratch");
    rc = exc_gemm_grouped_opts(8, 0, 0, nb, P.data(), 1.0, 0.0, -1, 0, gwork, gws, s);
    if (rc) return rc;
    for (int b = 0; b < nb; b++)
      EXC_CHECK_HIP(hipMemcpyAsync(Z + xoff[b], h_qt[b], (size_t)m * h_cols[b] * sizeof(double),
                                   hipMemcpyDeviceToDevice, s));
    for (int b = 0; b < nb; b++) {
      ExcGemmProblem& q = P[b];
      q = ExcGemmProblem{};
      q.A = L1 + mm * b;
      q.lda = m;
      q.B = L3 + mm * b;
      q.ldb = m;
      q.C = S + mm * b;
      q.ldc = m;
      q.M = q.N = q.K = m;
      q.tri = EXC_TRI_A_LOWER;
    }
    rc = exc_gemm_grouped_opts(8, 0, 0, nb, P.data(), 1.0, 0.0, 0, 0, gwork, gws, s);
    if (rc) return rc;
    EXC_CHECK_HIP(hipMemcpyAsync(L1, S, mm * nb * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (int b = 0; b < nb; b++) passes[b]++;
  }
  // R^T = L1 L3 into the outputs
  for (int b = 0; b < nb; b++) {
    ExcGemmProblem& q = P[b];
    q = ExcGemmProblem{};
    q.A = L1 + mm * b;
    q.lda = m;
    q.B = L3 + mm * b;
    q.ldb = m;
    q.C = h_rt[b];
    q.ldc = m;
    q.M = q.N = q.K = m;
    q.tri = EXC_TRI_A_LOWER;
  }
  rc = exc_gemm_grouped_opts(8, 0, 0, nb, P.data(), 1.0, 0.0, 0, 0, gwork, gws, s);
  if (rc) return rc;
  if (h_qt) {  // Q^T = L3^-1 Z
    rc = trinv(tL3, tS);
    if (rc) return rc;
    for (int b = 0; b < nb; b++) {
      ExcGemmProblem& q = P[b];
      q = ExcGemmProblem{};
      q.A = S + mm * b;
      q.lda = m;
      q.B = Z + xoff[b];
      q.ldb = h_cols[b];
      q.C = h_qt[b];
      q.ldc = h_cols[b];
      q.M = m;
      q.N = h_cols[b];
      q.K = m;
      q.tri = EXC_TRI_A_LOWER;
    }
    rc = exc_gemm_grouped_opts(8, 0, 0, nb, P.data(), 1.0, 0.0, -1, 0, gwork, gws, s);
    if (rc) return rc;
  }
  if (h_stats)
    for (int b = 0; b < nb; b++) {
      h_stats[3 * b + 1] = dev[b];
      h_stats[3 * b + 2] = passes[b];
    }
  return EXC_OK;
}

}  // extern "C"
