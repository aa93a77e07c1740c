// Kernels of the mixed-precision form: precision conversion, diagonal shifts for the shifted
// Cholesky passes, and the FP64 Ogita-Aishima refinement of a square SVD.
//
// The refinement follows Ogita and Aishima, "Iterative refinement for singular value
// decomposition based on matrix multiplication" (J. Comput. Appl. Math. 369, 2020). For an m x m
// matrix L with approximate factors L ~ U diag(sigma) W^T it forms the residual products
//   R = I - U^T U,  S = I - W^T W,  T = U^T L W                     (FP64 GEMMs on the MFMA)
// and the corrections U <- U (I + F), W <- W (I + G) with, for i != j,
//   alpha = T_ij + s_j R_ij,  beta = T_ji + s_j S_ij,
//   F_ij = (s_j alpha + s_i beta) / (s_j^2 - s_i^2),  G_ij = (s_i alpha + s_j beta) / (s_j^2 - s_i^2),
// F_ii = R_ii / 2, G_ii = S_ii / 2 and s_i = T_ii / (1 - (R_ii + S_ii) / 2). Pairs whose values
// agree to a relative tolerance tau form clusters, where that 2 x 2 system is singular: there F and
// G take the symmetric orthogonalization R / 2, S / 2, and the left vectors of cluster members
// (values above floor * max s) are then recomputed as L w_j / s_j, so that within a cluster the
// two sides stay consistent.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "exc_common.hpp"
#include "gemm.hpp"

namespace {

__global__ __launch_bounds__(256) void k_cvt_f64_f32(long rows, long cols, const double* __restrict__ src, long lds,
                                                     float* __restrict__ dst, long ldd) {
  const long total = rows * cols;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e / cols, c = e - r * cols;
    dst[r * ldd + c] = (float)src[r * lds + c];
  }
}

__global__ __launch_bounds__(256) void k_cvt_f32_f64(long rows, long cols, const float* __restrict__ src, long lds,
                                                     double* __restrict__ dst, long ldd) {
  const long total = rows * cols;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e / cols, c = e - r * cols;
    dst[r * ldd + c] = (double)src[r * lds + c];
  }
}

// Contiguous rows: two elements per thread and no index division.
__global__ __launch_bounds__(256) void k_cvt_f64_f32_flat(long n, const double* __restrict__ src,
                                                          float* __restrict__ dst) {
  const long n2 = n / 2;
  const double2* s2 = (const double2*)src;
  float2* d2 = (float2*)dst;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n2; e += (long)gridDim.x * blockDim.x) {
    const double2 v = s2[e];
    d2[e] = make_float2((float)v.x, (float)v.y);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) dst[n - 1] = (float)src[n - 1];
}

// out[0] = trace(A): thread t sums the diagonal entries t, t + 256, ... in order, then the 256
// partial sums are added in a fixed tree order (the same bits on every run and every device).
__global__ __launch_bounds__(256) void k_trace(int n, const double* __restrict__ a, long lda, double* __restrict__ out) {
  __shared__ double part[256];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) t = t + a[(long)i * lda + i];
  part[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] = part[threadIdx.x] + part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

__global__ void k_add_diag(int n, double* __restrict__ a, long lda, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[(long)i * lda + i] = a[(long)i * lda + i] + v;
}

// Per-workgroup partials of max |A - I| and sum (A - I)^2 over rows r = blockIdx.x, blockIdx.x +
// gridDim.x, ...; part[2 b] = max, part[2 b + 1] = sum of squares (fixed order within the block).
constexpr int kDevBlocks = 256;
__global__ __launch_bounds__(256) void k_dev_identity_part(int n, const double* __restrict__ a, long lda,
                                                           double* __restrict__ part) {
  __shared__ double red[4], rsq[4];
  double mx = 0.0, sq = 0.0;
  for (long i = blockIdx.x; i < n; i += gridDim.x)
    for (long j = threadIdx.x; j < n; j += blockDim.x) {
      const double v = a[i * lda + j] - (i == j ? 1.0 : 0.0);
      mx = fmax(mx, fabs(v));
      sq += v * v;
    }
  mx = wave_max(mx);
  sq = wave_sum(sq);
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = mx;
    rsq[threadIdx.x >> 6] = sq;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    part[2 * blockIdx.x + 1] = ((rsq[0] + rsq[1]) + rsq[2]) + rsq[3];
  }
}

// out[0] = max over the partials, out[1] = sqrt of their sums of squares added in block order.
__global__ void k_dev_identity_final(int nparts, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double mx = 0.0, sq = 0.0;
    for (int b = 0; b < nparts; b++) {
      mx = fmax(mx, part[2 * b]);
      sq = sq + part[2 * b + 1];
    }
    out[0] = mx;
    out[1] = sqrt(sq);
  }
}

// Row-packed lower triangle: packed[i (i + 1) / 2 + j] = a[i * lda + j], j <= i. One thread per
// packed entry; the row comes from the index by the inverse of the triangular numbers.
__device__ __forceinline__ void tri_index(long q, long& i, long& j) {
  i = (long)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > q) i--;
  while ((i + 1) * (i + 2) / 2 <= q) i++;
  j = q - i * (i + 1) / 2;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pack_lower(int n, const T* __restrict__ a, long lda, T* __restrict__ packed) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)n * (n + 1) / 2) return;
  long i, j;
  tri_index(q, i, j);
  packed[q] = a[i * lda + j];
}

__global__ __launch_bounds__(256) void k_unpack_lower(int n, const double* __restrict__ packed, double* __restrict__ a,
                                                      long lda) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)n * (n + 1) / 2) return;
  long i, j;
  tri_index(q, i, j);
  const double v = packed[q];
  a[i * lda + j] = v;
  a[j * lda + i] = v;
}

// acc[q] = (first ? 0 : acc[q]) + (double)part[q]: one peer's term of a sum taken in peer order,
// the same FP64 additions in the same order as a host loop `acc = 0; acc = acc + x_j`.
template <typename T>
__global__ __launch_bounds__(256) void k_acc_packed(long n, const T* __restrict__ part, int first,
                                                    double* __restrict__ acc) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double base = first ? 0.0 : acc[q];
  acc[q] = base + (double)part[q];
}

// sigma_i = T_ii / (1 - (R_ii + S_ii) / 2) with R_ii = 1 - (U^T U)_ii, S_ii = 1 - (W^T W)_ii;
// smax[0] = max |sigma_i|. One workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_oa_sigma(int m, const double* __restrict__ T, const double* __restrict__ utu,
                                                   const double* __restrict__ wtw, double* __restrict__ sigma,
                                                   double* __restrict__ smax) {
  __shared__ double red[16];
  double mx = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const long d = (long)i * m + i;
    const double r = 1.0 - utu[d], s = 1.0 - wtw[d];
    const double v = T[d] / (1.0 - (r + s) / 2.0);
    sigma[i] = v;
    mx = fmax(mx, fabs(v));
  }
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t = fmax(t, red[q]);
    smax[0] = t;
  }
}

__device__ __forceinline__ bool oa_cluster(double si, double sj, double tau) {
  return fabs(sj - si) <= tau * fmax(fabs(si), fabs(sj));
}

__global__ __launch_bounds__(256) void k_oa_coeffs(int m, const double* __restrict__ T, const double* __restrict__ utu,
                                                   const double* __restrict__ wtw, const double* __restrict__ sigma,
                                                   double tau, double* __restrict__ F, double* __restrict__ G) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * m) return;
  const int i = (int)(e / m), j = (int)(e - (long)i * m);
  const double r = (i == j ? 1.0 : 0.0) - utu[e];
  const double s = (i == j ? 1.0 : 0.0) - wtw[e];
  const double si = sigma[i], sj = sigma[j];
  double f, g;
  if (i == j || oa_cluster(si, sj, tau)) {
    f = r / 2.0;
    g = s / 2.0;
  } else {
    const double al = T[e] + sj * r;
    const double be = T[(long)j * m + i] + sj * s;
    const double det = sj * sj - si * si;
    f = (sj * al + si * be) / det;
    g = (si * al + sj * be) / det;
  }
  F[e] = f;
  G[e] = g;
}

// flags[j] = 1 when value j has a cluster partner and lies above floor * smax. With STAGED the
// workgroup first copies all m values into LDS (m * 8 bytes of dynamic LDS), so the scan, every
// lane reading the same value at each step, runs on LDS broadcasts instead of one dependent global
// load per step; the comparisons are the same.
template <bool STAGED>
__global__ __launch_bounds__(256) void k_oa_members(int m, const double* __restrict__ sigma,
                                                    const double* __restrict__ smax, double tau, double floor_rel,
                                                    int* __restrict__ flags) {
  extern __shared__ double sg[];
  const double* sv = sigma;
  if (STAGED) {
    for (int i = threadIdx.x; i < m; i += blockDim.x) sg[i] = sigma[i];
    __syncthreads();
    sv = sg;
  }
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const double sj = sv[j];
  int f = 0;
  if (fabs(sj) > floor_rel * smax[0])
    for (int i = 0; i < m && !f; i++)
      if (i != j && oa_cluster(sv[i], sj, tau)) f = 1;
  flags[j] = f;
}

__global__ __launch_bounds__(256) void k_oa_fix(int m, const double* __restrict__ LW, const double* __restrict__ sigma,
                                                const int* __restrict__ flags, double* __restrict__ U, long ldu) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * m) return;
  const int r = (int)(e / m), j = (int)(e - (long)r * m);
  if (flags[j]) U[(long)r * ldu + j] = LW[e] / sigma[j];
}

unsigned grid_for(long n) { return (unsigned)std::min<long>(std::max<long>(1, (n + 255) / 256), 16384); }

}  // namespace

extern "C" {

int exc_convert_f64_f32(long rows, long cols, const double* d_src, long lds, float* d_dst, long ldd,
                        exc_stream_t stream) {
  EXC_REQUIRE(rows >= 0 && cols >= 0 && (rows * cols == 0 || (d_src && d_dst && lds >= cols && ldd >= cols)),
              EXC_ERR_ARG, "exc_convert_f64_f32: bad arguments");
  if (rows * cols == 0) return EXC_OK;
  hipStream_t s = exc_detail::stream(stream);
  if (lds == cols && ldd == cols && ((uintptr_t)d_src % 16 == 0) && ((uintptr_t)d_dst % 8 == 0))
    hipLaunchKernelGGL(k_cvt_f64_f32_flat, dim3(grid_for(rows * cols / 2)), dim3(256), 0, s, rows * cols, d_src,
                       d_dst);
  else
    hipLaunchKernelGGL(k_cvt_f64_f32, dim3(grid_for(rows * cols)), dim3(256), 0, s, rows, cols, d_src, lds, d_dst,
                       ldd);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_convert_f32_f64(long rows, long cols, const float* d_src, long lds, double* d_dst, long ldd,
                        exc_stream_t stream) {
  EXC_REQUIRE(rows >= 0 && cols >= 0 && (rows * cols == 0 || (d_src && d_dst && lds >= cols && ldd >= cols)),
              EXC_ERR_ARG, "exc_convert_f32_f64: bad arguments");
  if (rows * cols == 0) return EXC_OK;
  hipLaunchKernelGGL(k_cvt_f32_f64, dim3(grid_for(rows * cols)), dim3(256), 0, exc_detail::stream(stream), rows, cols,
                     d_src, lds, d_dst, ldd);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_trace(int n, const double* d_a, long lda, double* d_out, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_a && d_out && lda >= n, EXC_ERR_ARG, "exc_trace: bad arguments");
  hipLaunchKernelGGL(k_trace, dim3(1), dim3(256), 0, exc_detail::stream(stream), n, d_a, lda, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_add_diag(int n, double* d_a, long lda, double value, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_a && lda >= n, EXC_ERR_ARG, "exc_add_diag: bad arguments");
  hipLaunchKernelGGL(k_add_diag, dim3(exc_detail::ceil_div(n, 256)), dim3(256), 0, exc_detail::stream(stream), n, d_a,
                     lda, value);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

size_t exc_dev_identity_workspace(int n) {
  (void)n;
  return (size_t)2 * kDevBlocks * sizeof(double);
}

int exc_dev_identity(int n, const double* d_a, long lda, double* d_out, void* d_work, size_t work_bytes,
                     exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_a && d_out && lda >= n && d_work, EXC_ERR_ARG, "exc_dev_identity: bad arguments");
  EXC_REQUIRE(work_bytes >= exc_dev_identity_workspace(n), EXC_ERR_SPACE, "exc_dev_identity: workspace too small");
  const int blocks = std::min(n, kDevBlocks);
  hipStream_t s = exc_detail::stream(stream);
  hipLaunchKernelGGL(k_dev_identity_part, dim3(blocks), dim3(256), 0, s, n, d_a, lda, (double*)d_work);
  hipLaunchKernelGGL(k_dev_identity_final, dim3(1), dim3(64), 0, s, blocks, (const double*)d_work, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_pack_lower_f32(int n, const float* d_a, long lda, float* d_packed, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_a && d_packed && lda >= n, EXC_ERR_ARG, "exc_pack_lower_f32: bad arguments");
  const long t = (long)n * (n + 1) / 2;
  hipLaunchKernelGGL(k_pack_lower<float>, dim3(exc_detail::ceil_div(t, 256)), dim3(256), 0,
                     exc_detail::stream(stream), n, d_a, lda, d_packed);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_pack_lower_f64(int n, const double* d_a, long lda, double* d_packed, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_a && d_packed && lda >= n, EXC_ERR_ARG, "exc_pack_lower_f64: bad arguments");
  const long t = (long)n * (n + 1) / 2;
  hipLaunchKernelGGL(k_pack_lower<double>, dim3(exc_detail::ceil_div(t, 256)), dim3(256), 0,
                     exc_detail::stream(stream), n, d_a, lda, d_packed);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_unpack_lower_f64(int n, const double* d_packed, double* d_a, long lda, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_a && d_packed && lda >= n, EXC_ERR_ARG, "exc_unpack_lower_f64: bad arguments");
  const long t = (long)n * (n + 1) / 2;
  hipLaunchKernelGGL(k_unpack_lower, dim3(exc_detail::ceil_div(t, 256)), dim3(256), 0, exc_detail::stream(stream),
                     n, d_packed, d_a, lda);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_acc_packed(long n, const void* d_part, int part_is_f32, int first, double* d_acc, exc_stream_t stream) {
  EXC_REQUIRE(n >= 0 && (n == 0 || (d_part && d_acc)), EXC_ERR_ARG, "exc_acc_packed: bad arguments");
  if (n == 0) return EXC_OK;
  hipStream_t s = exc_detail::stream(stream);
  if (part_is_f32)
    hipLaunchKernelGGL(k_acc_packed<float>, dim3(exc_detail::ceil_div(n, 256)), dim3(256), 0, s, n,
                       (const float*)d_part, first, d_acc);
  else
    hipLaunchKernelGGL(k_acc_packed<double>, dim3(exc_detail::ceil_div(n, 256)), dim3(256), 0, s, n,
                       (const double*)d_part, first, d_acc);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

// A first step that starts from vectors this close to orthonormal (max |I - U^T U|, |I - W^T W|)
// is the last one: the step's error is about the square of the starting one, below rounding.
constexpr double kOaOneStepDev = 1e-10;

size_t exc_oa_refine_workspace(int m) {
  if (m < 1) return 0;
  const size_t mm = (size_t)m * m * sizeof(double);
  const size_t mat = (mm + 255) / 256 * 256;
  return 8 * mat + 3 * 256 + ((size_t)m * sizeof(double) + 255) / 256 * 256 +
         ((size_t)m * sizeof(int) + 255) / 256 * 256 + exc_gemm_workspace(m, m, m) + 256 +
         (exc_dev_identity_workspace(m) + 255) / 256 * 256 + 256;
}

int exc_oa_refine(int m, const double* d_l, long ldl, int l_lower, double* d_u, long ldu, double* d_w, long ldw,
                  double* d_sigma, int iters, double tau, double floor_rel, void* d_work, size_t work_bytes,
                  exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && d_l && d_u && d_w && d_sigma && ldl >= m && ldu >= m && ldw >= m && iters >= 0 && tau >= 0 &&
                  d_work,
              EXC_ERR_ARG, "exc_oa_refine: bad arguments");
  EXC_REQUIRE(work_bytes >= exc_oa_refine_workspace(m), EXC_ERR_SPACE, "exc_oa_refine: workspace too small");
  hipStream_t s = exc_detail::stream(stream);
  const size_t mm = (size_t)m * m;
  const size_t mat = (mm * sizeof(double) + 255) / 256 * 256;
  char* w = (char*)d_work;
  auto take = [&](size_t bytes) {
    char* p = w;
    w += (bytes + 255) / 256 * 256;
    return p;
  };
  double* LW = (double*)take(mat);
  double* T = (double*)take(mat);
  double* UtU = (double*)take(mat);
  double* WtW = (double*)take(mat);
  double* F = (double*)take(mat);
  double* G = (double*)take(mat);
  double* Un = (double*)take(mat);
  double* Wn = (double*)take(mat);
  double* smax = (double*)take(256);
  int* flags = (int*)take((size_t)m * sizeof(int));
  double* sig = (double*)take((size_t)m * sizeof(double));
  ExcGemmOpts go;
  go.work = take(exc_gemm_workspace(m, m, m));
  go.work_bytes = exc_gemm_workspace(m, m, m);
  void* dev_work = take(exc_dev_identity_workspace(m));
  double* dev = (double*)take(4 * sizeof(double));
  ExcGemmOpts gl = go;
  gl.tri = l_lower ? EXC_TRI_A_LOWER : EXC_TRI_NONE;
  ExcGemmOpts gs = go;
  gs.lower_only = 1;
  gs.mirror = 1;
  const unsigned ge = grid_for((long)mm);
  for (int it = 0; it < iters; it++) {
    int rc = exc_gemm_opts(0, 0, m, m, m, 1.0, d_l, ldl, d_w, ldw, 0.0, LW, m, gl, s);
    if (!rc) rc = exc_gemm_opts(1, 0, m, m, m, 1.0, d_u, ldu, LW, m, 0.0, T, m, go, s);
    if (!rc) rc = exc_gemm_opts(1, 0, m, m, m, 1.0, d_u, ldu, d_u, ldu, 0.0, UtU, m, gs, s);
    if (!rc) rc = exc_gemm_opts(1, 0, m, m, m, 1.0, d_w, ldw, d_w, ldw, 0.0, WtW, m, gs, s);
    if (rc) return rc;
    if (it == 0 && iters > 1) {
      // how far the starting vectors are from orthonormal decides whether later steps run
      rc = exc_dev_identity(m, UtU, m, dev, dev_work, exc_dev_identity_workspace(m), stream);
      if (!rc) rc = exc_dev_identity(m, WtW, m, dev + 2, dev_work, exc_dev_identity_workspace(m), stream);
      if (rc) return rc;
      double h[4];
      EXC_CHECK_HIP(hipMemcpyAsync(h, dev, sizeof h, hipMemcpyDeviceToHost, s));
      EXC_CHECK_HIP(hipStreamSynchronize(s));
      if (std::max(h[0], h[2]) <= kOaOneStepDev) iters = 1;
    }
    hipLaunchKernelGGL(k_oa_sigma, dim3(1), dim3(1024), 0, s, m, T, UtU, WtW, sig, smax);
    hipLaunchKernelGGL(k_oa_coeffs, dim3(ge), dim3(256), 0, s, m, T, UtU, WtW, sig, tau, F, G);
    if ((size_t)m * sizeof(double) <= 64 * 1024)
      hipLaunchKernelGGL(k_oa_members<true>, dim3(exc_detail::ceil_div(m, 256)), dim3(256),
                         (size_t)m * sizeof(double), s, m, sig, smax, tau, floor_rel, flags);
    else
      hipLaunchKernelGGL(k_oa_members<false>, dim3(exc_detail::ceil_div(m, 256)), dim3(256), 0, s, m, sig, smax, tau,
                         floor_rel, flags);
    EXC_CHECK_HIP(hipGetLastError());
    EXC_CHECK_HIP(hipMemcpy2DAsync(Un, m * sizeof(double), d_u, ldu * sizeof(double), m * sizeof(double), m,
                                   hipMemcpyDeviceToDevice, s));
    EXC_CHECK_HIP(hipMemcpy2DAsync(Wn, m * sizeof(double), d_w, ldw * sizeof(double), m * sizeof(double), m,
                                   hipMemcpyDeviceToDevice, s));
    rc = exc_gemm_opts(0, 0, m, m, m, 1.0, d_u, ldu, F, m, 1.0, Un, m, go, s);
    if (!rc) rc = exc_gemm_opts(0, 0, m, m, m, 1.0, d_w, ldw, G, m, 1.0, Wn, m, go, s);
    // left vectors of cluster members from the refined right vectors: L w_j / sigma_j
    if (!rc) rc = exc_gemm_opts(0, 0, m, m, m, 1.0, d_l, ldl, Wn, m, 0.0, LW, m, gl, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_oa_fix, dim3(ge), dim3(256), 0, s, m, LW, sig, flags, Un, (long)m);
    EXC_CHECK_HIP(hipGetLastError());
    EXC_CHECK_HIP(hipMemcpy2DAsync(d_u, ldu * sizeof(double), Un, m * sizeof(double), m * sizeof(double), m,
                                   hipMemcpyDeviceToDevice, s));
    EXC_CHECK_HIP(hipMemcpy2DAsync(d_w, ldw * sizeof(double), Wn, m * sizeof(double), m * sizeof(double), m,
                                   hipMemcpyDeviceToDevice, s));
    EXC_CHECK_HIP(hipMemcpyAsync(d_sigma, sig, (size_t)m * sizeof(double), hipMemcpyDeviceToDevice, s));
  }
  return EXC_OK;
}

}  // extern "C"
