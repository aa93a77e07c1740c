// Householder QR in LAPACK's convention for tall-skinny blocks, and the compact-WY
// applications the pipeline needs (dormqr / dorgqr of client.cpp:417-445, 706-794, 1132-1176).
//
// The factor is computed by Cholesky QR (two passes, a shifted pass first when the Gram matrix
// is not numerically positive definite): Gram products, Cholesky and triangular solves are all
// MFMA GEMM plus LDS-resident diagonal blocks. The result is then brought to LAPACK's
// Householder convention (R_jj = -sign(alpha_j) ||x_j||, tau_j in [1, 2]) by Householder
// reconstruction: an LU without pivoting of the leading m x m block of Q minus a sign matrix S
// whose entries are chosen column by column as -sign of the pivot candidate. That choice is
// exactly Householder's cancellation-free sign, so S gives LAPACK's signs of R and Q, L gives the
// Householder vectors and -diag(U) S gives tau (Ballard, Demmel, Grigori, Jacquelin, Knight,
// Nguyen, "Reconstructing Householder vectors from tall-skinny QR", JPDC 2015).
//
// Layout: the pipeline keeps every tall matrix transposed (row-major m x r, a row per column of
// the reference's column-major matrix), so the QR of X^T is computed from X directly and Q^T and
// the reflectors (one per row) come out in the same layout.
#include <cmath>
#include <vector>

#include "exc_common.hpp"
#include "fact.hpp"
#include "gemm.hpp"
#include "qr.hpp"

namespace {

constexpr double kHugeInvTau = 1e300;  // 1/tau for tau == 0: that reflector drops out of T

__global__ __launch_bounds__(256) void k_copy2d(long rows, long cols, const double* __restrict__ src, long lds,
                                                double* __restrict__ dst, long ldd) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const long i = e / cols, j = e - i * cols;
  dst[i * ldd + j] = src[i * lds + j];
}

// dst (cols x rows) = src^T, 64 x 64 LDS tiles
__global__ __launch_bounds__(256) void k_transpose(long rows, long cols, const double* __restrict__ src, long lds,
                                                   double* __restrict__ dst, long ldd) {
  __shared__ double tile[64][65];
  const long i0 = (long)blockIdx.y * 64, j0 = (long)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4)
    if (i0 + r < rows && j0 + tx < cols) tile[r][tx] = src[(i0 + r) * lds + j0 + tx];
  __syncthreads();
  for (int r = ty; r < 64; r += 4)
    if (j0 + r < cols && i0 + tx < rows) dst[(j0 + r) * ldd + i0 + tx] = tile[tx][r];
}

__global__ __launch_bounds__(256) void k_scale_rows(long rows, long cols, double* __restrict__ A, long lda,
                                                    const double* __restrict__ s) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const long i = e / cols, j = e - i * cols;
  A[i * lda + j] *= s[i];
}

// B[i][j] = s[i] A[i][j] (the row scaling of k_scale_rows into another buffer)
__global__ __launch_bounds__(256) void k_scale_rows_copy(long rows, long cols, const double* __restrict__ A, long lda,
                                                         double* __restrict__ B, long ldb,
                                                         const double* __restrict__ s) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const long i = e / cols, j = e - i * cols;
  B[i * ldb + j] = A[i * lda + j] * s[i];
}

__global__ __launch_bounds__(256) void k_scale_cols(long rows, long cols, double* __restrict__ A, long lda,
                                                    const double* __restrict__ s) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const long i = e / cols, j = e - i * cols;
  A[i * lda + j] *= s[j];
}

__global__ void k_add_diag(int n, double* A, long lda, const double* shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[(long)i * lda + i] += *shift;
}

// shift = 11 (m r + m (m + 1)) u trace(G) (Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
// shifted Cholesky QR, SIAM J. Sci. Comput. 2020), with trace(G) = ||X||_F^2 >= ||X||_2^2.
__global__ __launch_bounds__(256) void k_shift_from_trace(int n, const double* G, long ldg, double m, double r,
                                                          double* shift) {
  __shared__ double part[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += G[(long)i * ldg + i];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) *shift = 11.0 * (m * r + m * (m + 1.0)) * 1.1102230246251565e-16 * part[0];
}

// Householder rows of the reflectors: vt[j][c] = 0 (c < j), 1 (c == j), L[c][j] (j < c < m).
__global__ __launch_bounds__(256) void k_vt_top(int m, const double* __restrict__ LU, long ldlu,
                                                double* __restrict__ vt, long ldv) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * m) return;
  const int j = (int)(e / m), c = (int)(e % m);
  vt[(long)j * ldv + c] = c < j ? 0.0 : (c == j ? 1.0 : LU[(long)c * ldlu + j]);
}

// lower (m x m) = upper part of LU, transposed
__global__ __launch_bounds__(256) void k_upper_t(int m, const double* __restrict__ LU, long ldlu,
                                                 double* __restrict__ Lo) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * m) return;
  const int i = (int)(e / m), j = (int)(e % m);
  Lo[e] = j <= i ? LU[(long)j * ldlu + i] : 0.0;
}

__global__ void k_tau_from_lu(int m, const double* __restrict__ LU, long ldlu, const double* __restrict__ sgn,
                              int square_last, double* __restrict__ tau) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  double t = -LU[(long)j * ldlu + j] * sgn[j];
  if (square_last && j == m - 1) t = 0.0;
  tau[j] = t;
}

// Clean reflector matrix from row storage: Vc[j][c] = 0 (c<j), 1 (c==j), vt[j][c] (c>j).
__global__ __launch_bounds__(256) void k_vc_rows(int k, long L, const double* __restrict__ vt, long ldv,
                                                 double* __restrict__ Vc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)k * L) return;
  const long j = e / L, c = e - j * L;
  Vc[e] = c < j ? 0.0 : (c == j ? 1.0 : vt[j * ldv + c]);
}

// Clean reflector matrix from column storage (house, m x m row-major, reflector j in column j):
// Vc[i][j] = 0 (i<j), 1 (i==j), house[i][j] (i>j).
__global__ __launch_bounds__(256) void k_vc_house(int m, const double* __restrict__ house, double* __restrict__ Vc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * m) return;
  const long i = e / m, j = e - i * m;
  Vc[e] = i < j ? 0.0 : (i == j ? 1.0 : house[i * m + j]);
}

// T^-1 = striu(G) + diag(1/tau) (upper), or its transpose (lower) when lower != 0.
__global__ __launch_bounds__(256) void k_tinv(int k, const double* __restrict__ G, const double* __restrict__ tau,
                                              int lower, double* __restrict__ Ti) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)k * k) return;
  const int i = (int)(e / k), j = (int)(e % k);
  double v = 0.0;
  if (i == j) v = tau[i] == 0.0 ? kHugeInvTau : 1.0 / tau[i];
  else if (!lower && i < j) v = G[e];
  else if (lower && i > j) v = G[e];
  Ti[e] = v;
}

__global__ __launch_bounds__(256) void k_identity_prefix(int k, long L, double* __restrict__ Q, long ldq) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)k * L) return;
  const long i = e / L, j = e - i * L;
  Q[i * ldq + j] = i == j ? 1.0 : 0.0;
}

inline unsigned g1(long n) { return exc_detail::ceil_div(n, 256); }

// Bump allocator over the caller's workspace.
struct Ws {
  char* p;
  size_t used = 0, cap;
  Ws(void* b, size_t c) : p((char*)b), cap(c) {}
  double* d(size_t n) {
    size_t off = (used + 255) & ~(size_t)255;
    used = off + n * sizeof(double);
    return p ? (double*)(p + off) : nullptr;
  }
  int* i(size_t n) {
    size_t off = (used + 255) & ~(size_t)255;
    used = off + n * sizeof(int);
    return p ? (int*)(p + off) : nullptr;
  }
  void* rest(size_t* bytes) {
    size_t off = (used + 255) & ~(size_t)255;
    *bytes = cap > off ? cap - off : 0;
    return p ? p + off : nullptr;
  }
};

}  // namespace

namespace qr {

size_t qr_t_workspace(int m, long r, bool want_q) {
  Ws w(nullptr, 0);
  if (!want_q) w.d((size_t)m * r);  // working copy of Q^T
  w.d((size_t)m * m);               // Gram / L
  w.d((size_t)m * m);               // R^T accumulator temp
  w.d((size_t)m * m);               // LU block
  w.d((size_t)m * m);               // U^T
  w.d(m);                           // signs
  w.d(4);                           // shift
  w.i(4);                           // info
  w.d((size_t)m * r);               // the last pass's output (L^-1 Y)
  w.d((size_t)m * m);               // L^-1 of the last pass
  return w.used + fact::work_bytes(m, r) + 4096;
}

int qr_t(int m, long r, const double* X, long ldx, const QrOut& o, int square_last, void* work, size_t wb,
         hipStream_t s) {
  EXC_REQUIRE(m >= 1 && r >= m && X && o.rt, EXC_ERR_ARG, "qr_t: needs r >= m >= 1 and an R output");
  Ws w(work, wb);
  double* Y = o.qt;
  long ldy = o.ldq;
  if (!Y) { Y = w.d((size_t)m * r); ldy = r; }
  double* G = w.d((size_t)m * m);
  double* Tmp = w.d((size_t)m * m);
  double* LU = w.d((size_t)m * m);
  double* Ut = w.d((size_t)m * m);
  double* sgn = w.d(m);
  double* shift = w.d(4);
  int* info = w.i(4);
  double* Yn = w.d((size_t)m * r);
  double* Linv = w.d((size_t)m * m);
  size_t gw_bytes;
  void* gw = w.rest(&gw_bytes);
  EXC_REQUIRE(w.used <= wb, EXC_ERR_SPACE, "qr_t: workspace too small");
  if (Y != X) {
    hipLaunchKernelGGL(k_copy2d, dim3(g1((long)m * r)), dim3(256), 0, s, (long)m, r, X, ldx, Y, ldy);
    EXC_CHECK_HIP(hipGetLastError());
  }
  ExcGemmOpts sym;
  sym.lower_only = 1;
  sym.work = gw;
  sym.work_bytes = gw_bytes;
  ExcGemmOpts plain;
  plain.work = gw;
  plain.work_bytes = gw_bytes;
  // Cholesky QR passes until two clean (unshifted) passes follow the last shifted one.
  bool shifted = false, have_rt = false;
  int clean = 0;
  for (int it = 0; it < 8 && clean < 2; it++) {
    int rc = exc_gemm_opts(0, 1, m, m, r, 1.0, Y, ldy, Y, ldy, 0.0, G, m, sym, s);
    if (rc) return rc;
    if (shifted) {
      hipLaunchKernelGGL(k_shift_from_trace, dim3(1), dim3(256), 0, s, m, G, (long)m, (double)m, (double)r, shift);
      hipLaunchKernelGGL(k_add_diag, dim3(exc_detail::ceil_div(m, 256)), dim3(256), 0, s, m, G, (long)m, shift);
      EXC_CHECK_HIP(hipGetLastError());
    }
    EXC_CHECK_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
    rc = fact::potrf_lower(m, G, m, info, gw, gw_bytes, s);
    if (rc) return rc;
    int hinfo = 0;
    EXC_CHECK_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
    EXC_CHECK_HIP(hipStreamSynchronize(s));
    if (hinfo != 0) {
      EXC_REQUIRE(!shifted, EXC_ERR_NUMERIC, "qr_t: shifted Cholesky QR failed (matrix numerically singular)");
      shifted = true;  // redo this pass with the shift
      continue;
    }
    if (clean == 1 && !shifted) {
      // The last pass: its input is orthonormal to about u cond(X)^2, so L is within that of
      // the identity and multiplying by its explicit inverse (formed by substitution) is as
      // accurate as substitution itself. One triangular-operand product then replaces the wide
      // solve: Yn = L^-1 Y.
      hipLaunchKernelGGL(k_identity_prefix, dim3(g1((long)m * m)), dim3(256), 0, s, m, (long)m, Linv, (long)m);
      EXC_CHECK_HIP(hipGetLastError());
      rc = fact::trsm_left(0, 0, m, m, G, m, Linv, m, gw, gw_bytes, s);
      if (rc) return rc;
      ExcGemmOpts lo = plain;
      lo.tri = EXC_TRI_A_LOWER;
      rc = exc_gemm_opts(0, 0, m, r, m, 1.0, Linv, m, Y, ldy, 0.0, Yn, r, lo, s);
      if (rc) return rc;
      Y = Yn;
      ldy = r;
    } else {
      rc = fact::trsm_left(0, 0, m, r, G, m, Y, ldy, gw, gw_bytes, s);
      if (rc) return rc;
    }
    if (!have_rt) {
      hipLaunchKernelGGL(k_copy2d, dim3(g1((long)m * m)), dim3(256), 0, s, (long)m, (long)m, G, (long)m, o.rt,
                         (long)m);
      have_rt = true;
    } else {
      // R^T <- R^T L (lower x lower)
      ExcGemmOpts tri = plain;
      tri.tri = EXC_TRI_A_LOWER;
      rc = exc_gemm_opts(0, 0, m, m, m, 1.0, o.rt, m, G, m, 0.0, Tmp, m, tri, s);
      if (rc) return rc;
      hipLaunchKernelGGL(k_copy2d, dim3(g1((long)m * m)), dim3(256), 0, s, (long)m, (long)m, Tmp, (long)m, o.rt,
                         (long)m);
    }
    EXC_CHECK_HIP(hipGetLastError());
    if (shifted) { shifted = false; clean = 0; }
    else clean++;
  }
  // Householder reconstruction: LU of Q[0:m, 0:m] - S, Q[0:m, 0:m] = (Y[:, 0:m])^T
  hipLaunchKernelGGL(k_transpose, dim3(exc_detail::ceil_div(m, 64), exc_detail::ceil_div(m, 64)), dim3(256), 0, s,
                     (long)m, (long)m, Y, ldy, LU, (long)m);
  EXC_CHECK_HIP(hipGetLastError());
  int rc = fact::lu_sign(m, LU, m, sgn, (square_last && r == m) ? 1 : 0, gw, gw_bytes, s);
  if (rc) return rc;
  if (o.vt) {
    // reflectors: rows of V^T; leading block from L, the rest U^-T Y[:, m:]
    hipLaunchKernelGGL(k_vt_top, dim3(g1((long)m * m)), dim3(256), 0, s, m, LU, (long)m, o.vt, o.ldv);
    EXC_CHECK_HIP(hipGetLastError());
    if (r > m) {
      hipLaunchKernelGGL(k_copy2d, dim3(g1((long)m * (r - m))), dim3(256), 0, s, (long)m, r - m, Y + m, ldy,
                         o.vt + m, o.ldv);
      hipLaunchKernelGGL(k_upper_t, dim3(g1((long)m * m)), dim3(256), 0, s, m, LU, (long)m, Ut);
      EXC_CHECK_HIP(hipGetLastError());
      rc = fact::trsm_left(0, 0, m, r - m, Ut, m, o.vt + m, o.ldv, gw, gw_bytes, s);
      if (rc) return rc;
    }
  }
  if (o.tau) {
    hipLaunchKernelGGL(k_tau_from_lu, dim3(exc_detail::ceil_div(m, 64)), dim3(64), 0, s, m, LU, (long)m, sgn,
                       square_last && r == m ? 1 : 0, o.tau);
    EXC_CHECK_HIP(hipGetLastError());
  }
  // LAPACK's signs: R <- S R (columns of R^T), Q <- Q S (rows of Q^T)
  hipLaunchKernelGGL(k_scale_cols, dim3(g1((long)m * m)), dim3(256), 0, s, (long)m, (long)m, o.rt, (long)m, sgn);
  if (o.qt) {
    if (Y == o.qt) hipLaunchKernelGGL(k_scale_rows, dim3(g1((long)m * r)), dim3(256), 0, s, (long)m, r, o.qt, o.ldq, sgn);
    else hipLaunchKernelGGL(k_scale_rows_copy, dim3(g1((long)m * r)), dim3(256), 0, s, (long)m, r, Y, ldy, o.qt, o.ldq, sgn);
  }
  EXC_CHECK_HIP(hipGetLastError());
  if (o.sgn) EXC_CHECK_HIP(hipMemcpyAsync(o.sgn, sgn, m * sizeof(double), hipMemcpyDeviceToDevice, s));
  return EXC_OK;
}

size_t wy_workspace(int k, long L, long rows) {
  Ws w(nullptr, 0);
  w.d((size_t)k * L);                        // Vc
  w.d((size_t)k * k);                        // G
  w.d((size_t)k * k);                        // T^-1
  w.d((size_t)std::max(rows, (long)k) * k);  // W
  return w.used + fact::work_bytes(k, L) + 4096;  // + split-K partials of the GEMMs
}

int apply_q_right(long rows, long L, int k, double* Mt, long ldm, const double* vt, long ldv, const double* tau,
                  void* work, size_t wb, hipStream_t s) {
  EXC_REQUIRE(rows >= 0 && L >= k && k >= 0 && Mt && vt && tau, EXC_ERR_ARG, "apply_q_right: bad arguments");
  if (rows == 0 || k == 0) return EXC_OK;
  Ws w(work, wb);
  double* Vc = w.d((size_t)k * L);
  double* G = w.d((size_t)k * k);
  double* Ti = w.d((size_t)k * k);
  double* W = w.d((size_t)rows * k);
  size_t gwb;
  void* gw = w.rest(&gwb);
  EXC_REQUIRE(w.used <= wb, EXC_ERR_SPACE, "apply_q_right: workspace too small");
  ExcGemmOpts o;
  o.work = gw;
  o.work_bytes = gwb;
  hipLaunchKernelGGL(k_vc_rows, dim3(g1((long)k * L)), dim3(256), 0, s, k, L, vt, ldv, Vc);
  EXC_CHECK_HIP(hipGetLastError());
  int rc = exc_gemm_opts(0, 1, k, k, L, 1.0, Vc, L, Vc, L, 0.0, G, k, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_tinv, dim3(g1((long)k * k)), dim3(256), 0, s, k, G, tau, 0, Ti);
  EXC_CHECK_HIP(hipGetLastError());
  // W = Mt Vc^T ; W <- W T (solve with T^-1) ; Mt -= W Vc
  rc = exc_gemm_opts(0, 1, rows, k, L, 1.0, Mt, ldm, Vc, L, 0.0, W, k, o, s);
  if (rc) return rc;
  rc = fact::trsm_right(0, 0, k, rows, Ti, k, W, k, gw, gwb, s);
  if (rc) return rc;
  return exc_gemm_opts(0, 0, rows, L, k, -1.0, W, k, Vc, L, 1.0, Mt, ldm, o, s);
}

int orgqr_t(int k, long L, const double* vt, long ldv, const double* tau, double* qt, long ldq, void* work,
            size_t wb, hipStream_t s) {
  EXC_REQUIRE(k >= 1 && L >= k && vt && tau && qt, EXC_ERR_ARG, "orgqr_t: bad arguments");
  Ws w(work, wb);
  double* Vc = w.d((size_t)k * L);
  double* G = w.d((size_t)k * k);
  double* Ti = w.d((size_t)k * k);
  size_t gwb;
  void* gw = w.rest(&gwb);
  EXC_REQUIRE(w.used <= wb, EXC_ERR_SPACE, "orgqr_t: workspace too small");
  ExcGemmOpts o;
  o.work = gw;
  o.work_bytes = gwb;
  hipLaunchKernelGGL(k_vc_rows, dim3(g1((long)k * L)), dim3(256), 0, s, k, L, vt, ldv, Vc);
  EXC_CHECK_HIP(hipGetLastError());
  int rc = exc_gemm_opts(0, 1, k, k, L, 1.0, Vc, L, Vc, L, 0.0, G, k, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_tinv, dim3(g1((long)k * k)), dim3(256), 0, s, k, G, tau, 1, Ti);
  EXC_CHECK_HIP(hipGetLastError());
  // X = (T^-T)^-1 Vc = T^T Vc (in place in Vc would lose V1; solve into qt's space is wrong too):
  // Q^T = [I 0] - V1^T (T^T Vc). Keep Vc, compute X in G-sized chunks? X is k x L: reuse Vc after
  // copying V1 (k x k) aside into G (G is no longer needed once T^-1 exists).
  hipLaunchKernelGGL(k_copy2d, dim3(g1((long)k * k)), dim3(256), 0, s, (long)k, (long)k, Vc, L, G, (long)k);
  EXC_CHECK_HIP(hipGetLastError());
  rc = fact::trsm_left(0, 0, k, L, Ti, k, Vc, L, gw, gwb, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_identity_prefix, dim3(g1((long)k * L)), dim3(256), 0, s, k, L, qt, ldq);
  EXC_CHECK_HIP(hipGetLastError());
  ExcGemmOpts tri = o;
  tri.tri = EXC_TRI_A_LOWER;  // op(A) = V1^T is unit lower
  return exc_gemm_opts(1, 0, k, L, k, -1.0, G, k, Vc, L, 1.0, qt, ldq, tri, s);
}

int ormqr_house_ln(int m, long n, const double* house, const double* tau, double* C, long ldc, void* work,
                   size_t wb, hipStream_t s) {
  EXC_REQUIRE(m >= 1 && n >= 0 && house && tau && C, EXC_ERR_ARG, "ormqr_house_ln: bad arguments");
  if (n == 0) return EXC_OK;
  Ws w(work, wb);
  double* Vc = w.d((size_t)m * m);
  double* G = w.d((size_t)m * m);
  double* Ti = w.d((size_t)m * m);
  double* W = w.d((size_t)m * n);
  size_t gwb;
  void* gw = w.rest(&gwb);
  EXC_REQUIRE(w.used <= wb, EXC_ERR_SPACE, "ormqr_house_ln: workspace too small");
  ExcGemmOpts o;
  o.work = gw;
  o.work_bytes = gwb;
  hipLaunchKernelGGL(k_vc_house, dim3(g1((long)m * m)), dim3(256), 0, s, m, house, Vc);
  EXC_CHECK_HIP(hipGetLastError());
  int rc = exc_gemm_opts(1, 0, m, m, m, 1.0, Vc, m, Vc, m, 0.0, G, m, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_tinv, dim3(g1((long)m * m)), dim3(256), 0, s, m, G, tau, 0, Ti);
  EXC_CHECK_HIP(hipGetLastError());
  ExcGemmOpts tri = o;
  tri.tri = EXC_TRI_A_UPPER;  // op(A) = Vc^T is unit upper
  rc = exc_gemm_opts(1, 0, m, n, m, 1.0, Vc, m, C, ldc, 0.0, W, n, tri, s);
  if (rc) return rc;
  rc = fact::trsm_left(1, 0, m, n, Ti, m, W, n, gw, gwb, s);
  if (rc) return rc;
  ExcGemmOpts tril = o;
  tril.tri = EXC_TRI_A_LOWER;
  return exc_gemm_opts(0, 0, m, n, m, -1.0, Vc, m, W, n, 1.0, C, ldc, tril, s);
}

size_t ormqr_house_workspace(int m, long n) {
  return (size_t)(3 * (size_t)m * m + (size_t)m * n) * sizeof(double) + fact::work_bytes(m, n) + 8192;
}

void transpose(long rows, long cols, const double* src, long lds, double* dst, long ldd, hipStream_t s) {
  hipLaunchKernelGGL(k_transpose, dim3(exc_detail::ceil_div(cols, 64), exc_detail::ceil_div(rows, 64)), dim3(256),
                     0, s, rows, cols, src, lds, dst, ldd);
}

void copy2d(long rows, long cols, const double* src, long lds, double* dst, long ldd, hipStream_t s) {
  if (rows * cols == 0) return;
  hipLaunchKernelGGL(k_copy2d, dim3(g1(rows * cols)), dim3(256), 0, s, rows, cols, src, lds, dst, ldd);
}

}  // namespace qr

extern "C" {

size_t exc_qr_t_workspace(int m, long r, int want_q) { return qr::qr_t_workspace(m, r, want_q != 0); }

int exc_qr_t(int m, long r, const double* d_x, long ldx, double* d_rt, double* d_qt, long ldq, double* d_vt,
             long ldv, double* d_tau, double* d_sgn, int square_last, void* d_work, size_t work_bytes,
             exc_stream_t stream) {
  QrOut o;
  o.rt = d_rt;
  o.qt = d_qt;
  o.ldq = ldq;
  o.vt = d_vt;
  o.ldv = ldv;
  o.tau = d_tau;
  o.sgn = d_sgn;
  return qr::qr_t(m, r, d_x, ldx, o, square_last, d_work, work_bytes, exc_detail::stream(stream));
}

size_t exc_wy_workspace(int k, long L, long rows) { return qr::wy_workspace(k, L, rows); }

int exc_apply_q_right(long rows, long L, int k, double* d_mt, long ldm, const double* d_vt, long ldv,
                      const double* d_tau, void* d_work, size_t work_bytes, exc_stream_t stream) {
  return qr::apply_q_right(rows, L, k, d_mt, ldm, d_vt, ldv, d_tau, d_work, work_bytes, exc_detail::stream(stream));
}

int exc_orgqr_t(int k, long L, const double* d_vt, long ldv, const double* d_tau, double* d_qt, long ldq,
                void* d_work, size_t work_bytes, exc_stream_t stream) {
  return qr::orgqr_t(k, L, d_vt, ldv, d_tau, d_qt, ldq, d_work, work_bytes, exc_detail::stream(stream));
}

size_t exc_ormqr_house_workspace(int m, long n) { return qr::ormqr_house_workspace(m, n); }

int exc_ormqr_house_ln(int m, long n, const double* d_house, const double* d_tau, double* d_c, long ldc,
                       void* d_work, size_t work_bytes, exc_stream_t stream) {
  return qr::ormqr_house_ln(m, n, d_house, d_tau, d_c, ldc, d_work, work_bytes, exc_detail::stream(stream));
}

}  // extern "C"
