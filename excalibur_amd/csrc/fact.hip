// Blocked dense factorizations on gfx950 (FP64): Cholesky, triangular solves and the
// sign-choosing LU used to recover LAPACK's Householder convention from an orthonormal factor.
// Diagonal blocks (up to 128 x 128) are factored or solved inside one workgroup from LDS; the
// off-diagonal work is MFMA GEMM (gemm.hip). Every triangular solve is a substitution (no
// explicit inverse), so each step is backward stable.
#include "exc_common.hpp"
#include "fact.hpp"
#include "gemm.hpp"

namespace {

using fact::NB;

// Cholesky of one diagonal block (kb <= NB) at A[k0][k0], lower triangle in, L out.
__global__ __launch_bounds__(256) void k_potf2(int kb, double* __restrict__ A, long lda, int k0,
                                              int* __restrict__ info) {
  __shared__ double S[NB][NB + 1];
  const int tid = threadIdx.x;
  double* Ab = A + (long)k0 * lda + k0;
  for (int e = tid; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    S[i][j] = j <= i ? Ab[(long)i * lda + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; j++) {
    if (tid == 0) {
      double d = S[j][j];
      if (!(d > 0.0)) {
        if (info && *info == 0) *info = k0 + j + 1;
        d = 1.0;
      }
      S[j][j] = sqrt(d);
    }
    __syncthreads();
    const double piv = S[j][j];
    for (int i = j + 1 + tid; i < kb; i += 256) S[i][j] = S[i][j] / piv;
    __syncthreads();
    const int m = kb - j - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int ii = e / m, kk = e % m;
      if (kk <= ii) S[j + 1 + ii][j + 1 + kk] -= S[j + 1 + ii][j] * S[j + 1 + kk][j];
    }
    __syncthreads();
  }
  for (int e = tid; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    Ab[(long)i * lda + j] = j <= i ? S[i][j] : 0.0;
  }
}

// LU without pivoting of one diagonal block of A - S, S_jj = -sign(current diagonal).
__global__ __launch_bounds__(256) void k_getf2_sign(int kb, double* __restrict__ A, long lda, int k0,
                                                   double* __restrict__ sgn, int last_plus) {
  __shared__ double S[NB][NB + 1];
  const int tid = threadIdx.x;
  double* Ab = A + (long)k0 * lda + k0;
  for (int e = tid; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    S[i][j] = Ab[(long)i * lda + j];
  }
  __syncthreads();
  for (int j = 0; j < kb; j++) {
    if (tid == 0) {
      const double w = S[j][j];
      double sj = signbit(w) ? 1.0 : -1.0;
      if (last_plus && j == kb - 1) sj = -sj;
      sgn[k0 + j] = sj;
      S[j][j] = w - sj;
    }
    __syncthreads();
    const double piv = S[j][j];
    for (int i = j + 1 + tid; i < kb; i += 256) S[i][j] = S[i][j] / piv;
    __syncthreads();
    const int m = kb - j - 1;
    for (int e = tid; e < m * m; e += 256) {
      const int ii = e / m, kk = e % m;
      S[j + 1 + ii][j + 1 + kk] -= S[j + 1 + ii][j] * S[j][j + 1 + kk];
    }
    __syncthreads();
  }
  for (int e = tid; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    Ab[(long)i * lda + j] = S[i][j];
  }
}

// X_b <- op(L_bb)^-1 X_b for one diagonal block (kb <= NB), one lane per right-hand side: the
// column lives in registers and the triangle is read with wave-uniform addresses (scalar loads),
// so the substitution runs at VALU rate with no LDS traffic.
template <bool UPPER, bool UNIT>
__global__ __launch_bounds__(64) void k_trsm_left_blk(int kb, const double* __restrict__ L, long ldl,
                                                      double* __restrict__ X, long ldx, long nrhs) {
  const long col = (long)blockIdx.x * 64 + threadIdx.x;
  if (col >= nrhs) return;
  double x[NB];
#pragma unroll
  for (int i = 0; i < NB; i++) x[i] = i < kb ? X[(long)i * ldx + col] : 0.0;
  if (!UPPER) {
#pragma unroll
    for (int i = 0; i < NB; i++) {
      if (i < kb) {
        double acc = x[i];
#pragma unroll
        for (int k = 0; k < i; k++) acc -= L[(long)i * ldl + k] * x[k];
        x[i] = UNIT ? acc : acc / L[(long)i * ldl + i];
      }
    }
  } else {
#pragma unroll
    for (int i = NB - 1; i >= 0; i--) {
      if (i < kb) {
        double acc = x[i];
#pragma unroll
        for (int k = i + 1; k < NB; k++)
          if (k < kb) acc -= L[(long)i * ldl + k] * x[k];
        x[i] = UNIT ? acc : acc / L[(long)i * ldl + i];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NB; i++)
    if (i < kb) X[(long)i * ldx + col] = x[i];
}

// X_b <- X_b U_bb^-1 (or X_b L_bb^-T) for one diagonal block, one lane per row (in registers).
template <bool LOWER_T, bool UNIT>
__global__ __launch_bounds__(64) void k_trsm_right_blk(int kb, const double* __restrict__ U, long ldu,
                                                       double* __restrict__ X, long ldx, long rows) {
  const long row = (long)blockIdx.x * 64 + threadIdx.x;
  if (row >= rows) return;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) x[j] = j < kb ? X[row * ldx + j] : 0.0;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    if (j < kb) {
      double acc = x[j];
#pragma unroll
      for (int k = 0; k < j; k++) {
        const double ukj = LOWER_T ? U[(long)j * ldu + k] : U[(long)k * ldu + j];
        acc -= x[k] * ukj;
      }
      x[j] = UNIT ? acc : acc / U[(long)j * ldu + j];
    }
  }
#pragma unroll
  for (int j = 0; j < NB; j++)
    if (j < kb) X[row * ldx + j] = x[j];
}

__global__ __launch_bounds__(256) void k_zero_strict_upper(int n, double* A, long lda) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const long i = e / n, j = e - i * n;
  if (j > i) A[i * lda + j] = 0.0;
}

int launch_left_blk(int upper, int unit, int kb, const double* L, long ldl, double* X, long ldx, long nrhs,
                    hipStream_t s) {
  const dim3 grid(exc_detail::ceil_div(nrhs, 64));
  if (!upper && !unit) hipLaunchKernelGGL((k_trsm_left_blk<false, false>), grid, dim3(64), 0, s, kb, L, ldl, X, ldx, nrhs);
  if (!upper && unit) hipLaunchKernelGGL((k_trsm_left_blk<false, true>), grid, dim3(64), 0, s, kb, L, ldl, X, ldx, nrhs);
  if (upper && !unit) hipLaunchKernelGGL((k_trsm_left_blk<true, false>), grid, dim3(64), 0, s, kb, L, ldl, X, ldx, nrhs);
  if (upper && unit) hipLaunchKernelGGL((k_trsm_left_blk<true, true>), grid, dim3(64), 0, s, kb, L, ldl, X, ldx, nrhs);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int launch_right_blk(int lower_t, int unit, int kb, const double* U, long ldu, double* X, long ldx, long rows,
                     hipStream_t s) {
  const dim3 grid(exc_detail::ceil_div(rows, 64));
  if (!lower_t && !unit) hipLaunchKernelGGL((k_trsm_right_blk<false, false>), grid, dim3(64), 0, s, kb, U, ldu, X, ldx, rows);
  if (!lower_t && unit) hipLaunchKernelGGL((k_trsm_right_blk<false, true>), grid, dim3(64), 0, s, kb, U, ldu, X, ldx, rows);
  if (lower_t && !unit) hipLaunchKernelGGL((k_trsm_right_blk<true, false>), grid, dim3(64), 0, s, kb, U, ldu, X, ldx, rows);
  if (lower_t && unit) hipLaunchKernelGGL((k_trsm_right_blk<true, true>), grid, dim3(64), 0, s, kb, U, ldu, X, ldx, rows);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // namespace

namespace fact {

size_t work_bytes(int n, long wide) {
  (void)wide;
  return (size_t)8 * n * n * sizeof(double);
}

int trsm_left(int upper, int unit, int n, long nrhs, const double* L, long ldl, double* X, long ldx, void* work,
              size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && nrhs >= 0 && L && X, EXC_ERR_ARG, "trsm_left: bad arguments");
  if (n == 0 || nrhs == 0) return EXC_OK;
  const int nblk = (n + NB - 1) / NB;
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  for (int q = 0; q < nblk; q++) {
    const int b = upper ? nblk - 1 - q : q;
    const int r0 = b * NB, kb = std::min(NB, n - r0), r1 = r0 + kb;
    if (!upper && r0 > 0) {
      int rc = exc_gemm_opts(0, 0, kb, nrhs, r0, -1.0, L + (long)r0 * ldl, ldl, X, ldx, 1.0, X + (long)r0 * ldx,
                             ldx, o, s);
      if (rc) return rc;
    }
    if (upper && r1 < n) {
      int rc = exc_gemm_opts(0, 0, kb, nrhs, n - r1, -1.0, L + (long)r0 * ldl + r1, ldl, X + (long)r1 * ldx, ldx,
                             1.0, X + (long)r0 * ldx, ldx, o, s);
      if (rc) return rc;
    }
    int rc = launch_left_blk(upper, unit, kb, L + (long)r0 * ldl + r0, ldl, X + (long)r0 * ldx, ldx, nrhs, s);
    if (rc) return rc;
  }
  return EXC_OK;
}

int trsm_right(int lower_t, int unit, int n, long rows, const double* U, long ldu, double* X, long ldx, void* work,
               size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && rows >= 0 && U && X, EXC_ERR_ARG, "trsm_right: bad arguments");
  if (n == 0 || rows == 0) return EXC_OK;
  const int nblk = (n + NB - 1) / NB;
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  for (int b = 0; b < nblk; b++) {
    const int c0 = b * NB, kb = std::min(NB, n - c0);
    if (c0 > 0) {
      // X[:, c0:c0+kb] -= X[:, 0:c0] op(U)[0:c0, c0:c0+kb], op(U) = L^T for lower_t
      int rc = lower_t ? exc_gemm_opts(0, 1, rows, kb, c0, -1.0, X, ldx, U + (long)c0 * ldu, ldu, 1.0, X + c0, ldx, o, s)
                       : exc_gemm_opts(0, 0, rows, kb, c0, -1.0, X, ldx, U + c0, ldu, 1.0, X + c0, ldx, o, s);
      if (rc) return rc;
    }
    int rc = launch_right_blk(lower_t, unit, kb, U + (long)c0 * ldu + c0, ldu, X + c0, ldx, rows, s);
    if (rc) return rc;
  }
  return EXC_OK;
}

int potrf_lower(int n, double* A, long lda, int* info, void* work, size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && A, EXC_ERR_ARG, "potrf_lower: bad arguments");
  if (n == 0) return EXC_OK;
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  o.lower_only = 1;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    hipLaunchKernelGGL(k_potf2, dim3(1), dim3(256), 0, s, kb, A, lda, k0, info);
    EXC_CHECK_HIP(hipGetLastError());
    const int r0 = k0 + kb;
    if (r0 < n) {
      const long rows = n - r0;
      int rc = launch_right_blk(1, 0, kb, A + (long)k0 * lda + k0, lda, A + (long)r0 * lda + k0, lda, rows, s);
      if (rc) return rc;
      rc = exc_gemm_opts(0, 1, rows, rows, kb, -1.0, A + (long)r0 * lda + k0, lda, A + (long)r0 * lda + k0, lda, 1.0,
                         A + (long)r0 * lda + r0, lda, o, s);
      if (rc) return rc;
    }
  }
  hipLaunchKernelGGL(k_zero_strict_upper, dim3(exc_detail::ceil_div((long)n * n, 256)), dim3(256), 0, s, n, A, lda);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int lu_sign(int n, double* A, long lda, double* sgn, int square_last, void* work, size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && A && sgn, EXC_ERR_ARG, "lu_sign: bad arguments");
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    const int last_plus = (square_last && k0 + kb == n) ? 1 : 0;
    hipLaunchKernelGGL(k_getf2_sign, dim3(1), dim3(256), 0, s, kb, A, lda, k0, sgn, last_plus);
    EXC_CHECK_HIP(hipGetLastError());
    const int r0 = k0 + kb;
    if (r0 < n) {
      const long rest = n - r0;
      double* A11 = A + (long)k0 * lda + k0;
      int rc = launch_left_blk(0, 1, kb, A11, lda, A + (long)k0 * lda + r0, lda, rest, s);
      if (rc) return rc;
      rc = launch_right_blk(0, 0, kb, A11, lda, A + (long)r0 * lda + k0, lda, rest, s);
      if (rc) return rc;
      rc = exc_gemm_opts(0, 0, rest, rest, kb, -1.0, A + (long)r0 * lda + k0, lda, A + (long)k0 * lda + r0, lda, 1.0,
                         A + (long)r0 * lda + r0, lda, o, s);
      if (rc) return rc;
    }
  }
  return EXC_OK;
}

}  // namespace fact

extern "C" {

size_t exc_fact_workspace(int n, long wide) { return fact::work_bytes(n, wide); }

int exc_potrf_lower(int n, double* d_a, long lda, int* d_info, void* d_work, size_t work_bytes, exc_stream_t stream) {
  return fact::potrf_lower(n, d_a, lda, d_info, d_work, work_bytes, exc_detail::stream(stream));
}

int exc_trsm_left(int upper, int unit, int n, long nrhs, const double* d_l, long ldl, double* d_x, long ldx,
                  void* d_work, size_t work_bytes, exc_stream_t stream) {
  return fact::trsm_left(upper, unit, n, nrhs, d_l, ldl, d_x, ldx, d_work, work_bytes, exc_detail::stream(stream));
}

int exc_trsm_right(int lower_t, int unit, int n, long rows, const double* d_u, long ldu, double* d_x, long ldx,
                   void* d_work, size_t work_bytes, exc_stream_t stream) {
  return fact::trsm_right(lower_t, unit, n, rows, d_u, ldu, d_x, ldx, d_work, work_bytes, exc_detail::stream(stream));
}

int exc_lu_sign(int n, double* d_a, long lda, double* d_sgn, int square_last, void* d_work, size_t work_bytes,
                exc_stream_t stream) {
  return fact::lu_sign(n, d_a, lda, d_sgn, square_last, d_work, work_bytes, exc_detail::stream(stream));
}

}  // extern "C"
