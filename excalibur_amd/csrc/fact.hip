// Dense factorizations on gfx950 (FP64): Cholesky, triangular solves and the sign-choosing LU
// used to recover LAPACK's Householder convention from an orthonormal factor. All four recurse on
// halves (split at a multiple of the 64-wide diagonal block), so the off-diagonal work is a few
// large MFMA GEMMs (gemm.hip) rather than many 64-deep ones; the 64 x 64 diagonal blocks are
// factored in one workgroup from LDS or solved one lane per right-hand side from registers.
// Every triangular solve is a substitution (no explicit inverse), so each step is backward
// stable.
#include "exc_common.hpp"
#include "fact.hpp"
#include "gemm.hpp"
#include "solve.hpp"

namespace {

using fact::NB;
using fact::readlane_f64;

// The 64 x 64 diagonal-block kernels are branch-free: a block smaller than 64 is padded with the
// identity (unit diagonal, zero off-diagonal), which leaves the padded unknowns untouched, so
// every loop has a constant trip count and the compiler can keep many LDS reads in flight. The
// updates are right-looking: once unknown j is final, the later ones take one fused
// multiply-add each with the contiguous row j of the (transposed as needed) triangle in LDS.
// Multiplies by a reciprocal of the pivot and uses fused multiply-adds (these kernels are not
// part of the bit-exact set).

// Cholesky of one diagonal block (kb <= NB) at A[k0][k0], lower triangle in, L out. One
// wavefront; lane i holds row i; the pivot column is exchanged through LDS.
__global__ __launch_bounds__(64) void k_potf2(int kb, double* __restrict__ A, long lda, int k0,
                                             int* __restrict__ info) {
  __shared__ double col[2][NB];
  const int lane = threadIdx.x;
  double* Ab = A + (long)k0 * lda + k0;
  const int rl = lane < kb ? lane : kb - 1;
  double x[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const double v = Ab[(long)rl * lda + (k < kb ? k : kb - 1)];
    x[k] = (lane < kb && k < kb) ? (k <= lane ? v : 0.0) : (k == lane ? 1.0 : 0.0);
  }
  int bad = 0;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const int buf = j & 1;
    // the pivot: x[j] of lane j, published with the column
    col[buf][lane] = x[j];
    __syncthreads();
    double d = col[buf][j];
    if (!(d > 0.0)) {
      bad = bad ? bad : j + 1;
      d = 1.0;
    }
    const double piv = sqrt(d);
    const double rp = 1.0 / piv;
    const double l = lane > j ? x[j] * rp : 0.0;  // l_i below the pivot, zero elsewhere
    x[j] = lane == j ? piv : (lane > j ? l : x[j]);
    __syncthreads();
    col[buf][lane] = l;
    __syncthreads();
#pragma unroll
    for (int k = j + 1; k < NB; k++) x[k] = fma(-l, col[buf][k], x[k]);
  }
  if (bad && lane == 0 && info && *info == 0 && bad <= kb) *info = k0 + bad;
#pragma unroll
  for (int k = 0; k < NB; k++)
    if (lane < kb && k < kb) Ab[(long)lane * lda + k] = k <= lane ? x[k] : 0.0;
}

// LU without pivoting of one diagonal block of A - S, S_jj = -sign(current diagonal). One
// wavefront; lane i holds row i; the pivot row goes through LDS.
__global__ __launch_bounds__(64) void k_getf2_sign(int kb, double* __restrict__ A, long lda, int k0,
                                                  double* __restrict__ sgn, int last_plus) {
  __shared__ double row[NB];
  __shared__ double pv[1];
  const int lane = threadIdx.x;
  double* Ab = A + (long)k0 * lda + k0;
  const int rl = lane < kb ? lane : kb - 1;
  double x[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const double v = Ab[(long)rl * lda + (k < kb ? k : kb - 1)];
    x[k] = (lane < kb && k < kb) ? v : (k == lane ? 1.0 : 0.0);
  }
#pragma unroll
  for (int j = 0; j < NB; j++) {
    double w = 0.0;
    if (lane == j) {
      w = x[j];
      double sj = signbit(w) ? 1.0 : -1.0;
      if (last_plus && j == kb - 1) sj = -sj;
      if (j < kb) sgn[k0 + j] = sj;
      x[j] = w - sj;
      pv[0] = x[j];
#pragma unroll
      for (int k = 0; k < NB; k++) row[k] = x[k];
    }
    __syncthreads();
    const double rp = 1.0 / pv[0];
    const double l = lane > j ? x[j] * rp : 0.0;
    if (lane > j) x[j] = l;
#pragma unroll
    for (int k = j + 1; k < NB; k++) x[k] = fma(-l, row[k], x[k]);
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < NB; k++)
    if (lane < kb && k < kb) Ab[(long)lane * lda + k] = x[k];
}

// X_b <- op(L_bb)^-1 X_b for one diagonal block (kb <= NB), one lane per ROW: lane i holds row i
// of the triangle (strictly below the diagonal for L, strictly above for U, zeros elsewhere) and
// row i of C right-hand sides, both scaled by the pivot's reciprocal. Step j (top down for L,
// bottom up for U) broadcasts lane j's unknown with readlane, and every lane takes one fused
// multiply-add per column; lanes the step does not touch hold a zero coefficient, so there is no
// branch. Unknown j's chain is a readlane and a fused multiply-add (about 50 cycles),
// and the C columns interleave to hide it: a narrow solve (few waves) takes a few microseconds
// instead of a 64-deep chain of dependent dot products and divisions per column. Rows at or
// beyond kb are padding (zero coefficients and right-hand sides). The four wavefronts of a
// workgroup share one staged copy of the triangle and take C columns each; the triangle and the
// right-hand sides pass through LDS so that every global access is coalesced.
template <bool UPPER, bool UNIT, int C>
__global__ __launch_bounds__(256) void k_trsm_left_blk(int kb, const double* __restrict__ L, long ldl,
                                                       double* __restrict__ X, long ldx, long nrhs) {
  // LDS: first the triangle (rows padded to SL), then, once every wave holds its row, each wave's
  // 64 x C block of right-hand sides (rows padded to S; both strides keep 16-byte alignment)
  constexpr int SL = NB + 2, S = C + 2, RPI = 64 / C;
  constexpr int kLds = NB * SL > 4 * NB * S ? NB * SL : 4 * NB * S;
  __shared__ __attribute__((aligned(16))) double lds[kLds];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long c0 = ((long)blockIdx.x * 4 + wave) * C;
  const int row = lane < kb ? lane : kb - 1;
  const bool live = lane < kb;
  const int lc = lane < kb ? lane : kb - 1;
#pragma unroll
  for (int r = wave; r < NB; r += 4) lds[r * SL + lane] = L[(long)(r < kb ? r : kb - 1) * ldl + lc];
  __syncthreads();
  // row i scaled by its pivot's reciprocal, so that each step's unknown is the broadcast value
  const double rinv = UNIT ? 1.0 : 1.0 / lds[row * SL + row];
  double lr[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const double v = lds[row * SL + k];
    lr[k] = (live && k < kb && (UPPER ? k > lane : k < lane)) ? (UNIT ? v : v * rinv) : 0.0;
  }
  __syncthreads();
  // coalesced: each load instruction covers 64 / C whole row segments of C columns
  double* Xs = lds + wave * NB * S;
  const int sr = lane / C, sc = lane % C;
  const bool cok = c0 + sc < nrhs;
  const long gcol = cok ? c0 + sc : nrhs - 1;
#pragma unroll
  for (int t = 0; t < C; t++) {
    const int r = t * RPI + sr;
    const double v = X[(long)(r < kb ? r : kb - 1) * ldx + gcol];
    Xs[r * S + sc] = (r < kb && cok) ? v : 0.0;
  }
  __syncthreads();
  double y[C];
#pragma unroll
  for (int c = 0; c < C; c++) y[c] = UNIT ? Xs[lane * S + c] : Xs[lane * S + c] * rinv;
#pragma unroll
  for (int q = 0; q < NB; q++) {
    const int j = UPPER ? NB - 1 - q : q;
    if (j >= kb) continue;  // uniform
#pragma unroll
    for (int c = 0; c < C; c++) y[c] = fma(-lr[j], readlane_f64(y[c], j), y[c]);
  }
#pragma unroll
  for (int c = 0; c < C; c++) Xs[lane * S + c] = y[c];
  __syncthreads();
#pragma unroll
  for (int t = 0; t < C; t++) {
    const int r = t * RPI + sr;
    if (r < kb && cok) X[(long)r * ldx + c0 + sc] = Xs[r * S + sc];
  }
}

// The same solve for a diagonal block of 64 < kb <= 128 rows, as two 64-row blocks in one launch
// (the update between two leaves would otherwise be a separate small GEMM, latency-bound for
// narrow right-hand sides and memory-bound for wide ones): solve the first block (rows [0, 64) of L, or [64, kb) of U),
// subtract its contribution from the second block's right-hand sides (one fused multiply-add per
// coefficient, the solved unknowns broadcast from LDS), then solve the second block. Each wave
// keeps its 128 x C right-hand sides in LDS for the whole launch.
template <bool UPPER, bool UNIT, int C>
__global__ __launch_bounds__(256) void k_trsm_left_blk2(int kb, const double* __restrict__ L, long ldl,
                                                        double* __restrict__ X, long ldx, long nrhs) {
  constexpr int SL = NB + 2, S = C + 2, RPI = 64 / C;
  __shared__ __attribute__((aligned(16))) double Ls[NB * SL];
  __shared__ __attribute__((aligned(16))) double Xall[4 * 2 * NB * S];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long c0 = ((long)blockIdx.x * 4 + wave) * C;
  double* Xs = Xall + wave * 2 * NB * S;
  const int sr = lane / C, sc = lane % C;
  const bool cok = c0 + sc < nrhs;
  const long gcol = cok ? c0 + sc : nrhs - 1;
#pragma unroll
  for (int t = 0; t < 2 * C; t++) {
    const int r = t * RPI + sr;
    const double v = X[(long)(r < kb ? r : kb - 1) * ldx + gcol];
    Xs[r * S + sc] = (r < kb && cok) ? v : 0.0;
  }
  // rows [b0, b0 + bn) of the first and second blocks
  const int f0 = UPPER ? NB : 0, fn = UPPER ? kb - NB : NB;
  const int s0 = UPPER ? 0 : NB, sn = UPPER ? NB : kb - NB;
  // Ls <- the bn x bn diagonal triangle at (b0, b0); lanes read their scaled row into lr
  auto stage = [&](int r0, int rn, int q0, int qn) {
    const int lc = lane < qn ? lane : qn - 1;
#pragma unroll
    for (int r = wave; r < NB; r += 4)
      Ls[r * SL + lane] = (r < rn && lane < qn) ? L[(long)(r0 + (r < rn ? r : rn - 1)) * ldl + q0 + lc] : 0.0;
  };
  auto solve = [&](int b0, int bn, double (&y)[C]) {
    __syncthreads();  // earlier readers of Ls are done
    stage(b0, bn, b0, bn);
    __syncthreads();
    const bool live = lane < bn;
    const int row = live ? lane : bn - 1;
    const double rinv = UNIT ? 1.0 : 1.0 / Ls[row * SL + row];
    double lr[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
      const double v = Ls[row * SL + k];
      lr[k] = (live && k < bn && (UPPER ? k > lane : k < lane)) ? (UNIT ? v : v * rinv) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < C; c++) y[c] = UNIT ? y[c] : y[c] * rinv;
#pragma unroll
    for (int q = 0; q < NB; q++) {
      const int j = UPPER ? NB - 1 - q : q;
      if (j >= bn) continue;  // uniform
#pragma unroll
      for (int c = 0; c < C; c++) y[c] = fma(-lr[j], readlane_f64(y[c], j), y[c]);
    }
#pragma unroll
    for (int c = 0; c < C; c++) Xs[(b0 + lane) * S + c] = y[c];
  };
  __syncthreads();
  double y[C];
#pragma unroll
  for (int c = 0; c < C; c++) y[c] = Xs[(f0 + lane) * S + c];
  solve(f0, fn, y);
  // second block: y -= O x_first, O = rows [s0, s0 + sn) x columns [f0, f0 + fn) of the triangle
  __syncthreads();
  stage(s0, sn, f0, fn);
  __syncthreads();
  {
    double o[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) o[k] = Ls[lane * SL + k];
#pragma unroll
    for (int c = 0; c < C; c++) y[c] = Xs[(s0 + lane) * S + c];
#pragma unroll
    for (int k = 0; k < NB; k++) {
#pragma unroll
      for (int c = 0; c < C; c++) y[c] = fma(-o[k], Xs[(f0 + k) * S + c], y[c]);
    }
  }
  solve(s0, sn, y);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 2 * C; t++) {
    const int r = t * RPI + sr;
    if (r < kb && cok) X[(long)r * ldx + c0 + sc] = Xs[r * S + sc];
  }
}

// X_b <- X_b U_bb^-1 (or X_b L_bb^-T) for one diagonal block (kb <= NB), one lane per COLUMN j
// of the unknown rows: lane j holds column j of op(U) above the diagonal (U[k][j], or L[j][k]
// for L^T) scaled by the pivot's reciprocal, and entry j of C rows of X (contiguous across the
// lanes, so the accesses to X are coalesced as they are). Step k (ascending) broadcasts lane k's
// unknown with readlane and every lane takes one fused multiply-add per row, the same scheme as
// k_trsm_left_blk. The triangle is staged through LDS with coalesced row loads; the four
// wavefronts of a workgroup share it and take C rows each.
template <bool LOWER_T, bool UNIT, int C>
__global__ __launch_bounds__(256) void k_trsm_right_blk(int kb, const double* __restrict__ U, long ldu,
                                                        double* __restrict__ X, long ldx, long rows) {
  constexpr int SL = NB + 2;
  __shared__ __attribute__((aligned(16))) double lds[NB * SL];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long r0 = ((long)blockIdx.x * 4 + wave) * C;
  const bool live = lane < kb;
  const int lc = lane < kb ? lane : kb - 1;
#pragma unroll
  for (int r = wave; r < NB; r += 4) lds[r * SL + lane] = U[(long)(r < kb ? r : kb - 1) * ldu + lc];
  __syncthreads();
  // op(U)[k][j] = U[k][j], or L[j][k] for L^T
  auto op = [&](int k, int j) { return LOWER_T ? lds[j * SL + k] : lds[k * SL + j]; };
  const double rinv = UNIT ? 1.0 : 1.0 / op(lc, lc);
  double co[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const double v = op(k, lc);
    co[k] = (live && k < lane) ? (UNIT ? v : v * rinv) : 0.0;
  }
  double y[C];
#pragma unroll
  for (int c = 0; c < C; c++) {
    const long rr = r0 + c < rows ? r0 + c : rows - 1;
    const double v = X[rr * ldx + lc];
    y[c] = (live && r0 + c < rows) ? (UNIT ? v : v * rinv) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < NB; k++) {
    if (k >= kb) continue;  // uniform
#pragma unroll
    for (int c = 0; c < C; c++) y[c] = fma(-co[k], readlane_f64(y[c], k), y[c]);
  }
  if (live) {
#pragma unroll
    for (int c = 0; c < C; c++)
      if (r0 + c < rows) X[(r0 + c) * ldx + lane] = y[c];
  }
}

__global__ __launch_bounds__(256) void k_zero_strict_upper(int n, double* A, long lda) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const long i = e / n, j = e - i * n;
  if (j > i) A[i * lda + j] = 0.0;
}

int launch_left_blk(int upper, int unit, int kb, const double* L, long ldl, double* X, long ldx, long nrhs,
                    hipStream_t s) {
  // 4 columns per wavefront up to 4096 right-hand sides (latency-bound: more waves), else 16
  // (throughput-bound: fewer reads of the triangle); four wavefronts per workgroup
  const bool wide = nrhs > 4L * 1024;
  const dim3 grid(exc_detail::ceil_div(nrhs, 4L * (wide ? 16 : 4))), block(256);
#define EXC_LEFT_BLK(U, N)                                                                                      \
  if ((upper != 0) == U && (unit != 0) == N) {                                                                                \
    if (wide) hipLaunchKernelGGL((k_trsm_left_blk<U, N, 16>), grid, block, 0, s, kb, L, ldl, X, ldx, nrhs);    \
    else hipLaunchKernelGGL((k_trsm_left_blk<U, N, 4>), grid, block, 0, s, kb, L, ldl, X, ldx, nrhs);          \
  }
  EXC_LEFT_BLK(false, false) EXC_LEFT_BLK(false, true) EXC_LEFT_BLK(true, false) EXC_LEFT_BLK(true, true)
#undef EXC_LEFT_BLK
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

// Two-block leaves (64 < kb <= 128): 4 columns per wavefront up to 4096 right-hand sides, else 8.
int launch_left_blk2(int upper, int unit, int kb, const double* L, long ldl, double* X, long ldx, long nrhs,
                     hipStream_t s) {
  const bool wide = nrhs > 4L * 1024;
  const dim3 grid(exc_detail::ceil_div(nrhs, 4L * (wide ? 8 : 4))), block(256);
#define EXC_LEFT_BLK2(U, N)                                                                                     \
  if ((upper != 0) == U && (unit != 0) == N) {                                                                  \
    if (wide) hipLaunchKernelGGL((k_trsm_left_blk2<U, N, 8>), grid, block, 0, s, kb, L, ldl, X, ldx, nrhs);     \
    else hipLaunchKernelGGL((k_trsm_left_blk2<U, N, 4>), grid, block, 0, s, kb, L, ldl, X, ldx, nrhs);          \
  }
  EXC_LEFT_BLK2(false, false) EXC_LEFT_BLK2(false, true) EXC_LEFT_BLK2(true, false) EXC_LEFT_BLK2(true, true)
#undef EXC_LEFT_BLK2
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int launch_right_blk(int lower_t, int unit, int kb, const double* U, long ldu, double* X, long ldx, long rows,
                     hipStream_t s) {
  // as launch_left_blk: 4 rows per wavefront up to 4096 rows, else 16; four wavefronts per workgroup
  const bool wide = rows > 4L * 1024;
  const dim3 grid(exc_detail::ceil_div(rows, 4L * (wide ? 16 : 4))), block(256);
#define EXC_RIGHT_BLK(T, N)                                                                                     \
  if ((lower_t != 0) == T && (unit != 0) == N) {                                                                \
    if (wide) hipLaunchKernelGGL((k_trsm_right_blk<T, N, 16>), grid, block, 0, s, kb, U, ldu, X, ldx, rows);    \
    else hipLaunchKernelGGL((k_trsm_right_blk<T, N, 4>), grid, block, 0, s, kb, U, ldu, X, ldx, rows);          \
  }
  EXC_RIGHT_BLK(false, false) EXC_RIGHT_BLK(false, true) EXC_RIGHT_BLK(true, false) EXC_RIGHT_BLK(true, true)
#undef EXC_RIGHT_BLK
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // namespace

namespace fact {

// Room for 16 split-K partials of an n x n product: enough for the GEMM's split rule to fill the
// device with the lower tiles of an n x n Gram matrix over a long K (14 splits at n = 1000).
size_t work_bytes(int n, long wide) {
  (void)wide;
  return (size_t)16 * n * n * sizeof(double);
}

// Split point of a recursive step: a multiple of NB near n / 2, 0 < n1 < n (for n > NB).
static int split(int n) {
  int n1 = ((n / NB + 1) / 2) * NB;
  if (n1 <= 0) n1 = NB;
  if (n1 >= n) n1 = n - NB;
  return n1;
}

int trsm_left(int upper, int unit, int n, long nrhs, const double* L, long ldl, double* X, long ldx, void* work,
              size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && nrhs >= 0 && L && X, EXC_ERR_ARG, "trsm_left: bad arguments");
  if (n == 0 || nrhs == 0) return EXC_OK;
  if (n <= NB) return launch_left_blk(upper, unit, n, L, ldl, X, ldx, nrhs, s);
  if (n <= 2 * NB) return launch_left_blk2(upper, unit, n, L, ldl, X, ldx, nrhs, s);
  const int n1 = split(n), n2 = n - n1;
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  double* X2 = X + (long)n1 * ldx;
  const double* L22 = L + (long)n1 * ldl + n1;
  int rc;
  if (!upper) {
    // [L11 0; L21 L22]: X1 <- L11^-1 X1, X2 <- L22^-1 (X2 - L21 X1)
    if ((rc = trsm_left(0, unit, n1, nrhs, L, ldl, X, ldx, work, wb, s))) return rc;
    if ((rc = exc_gemm_opts(0, 0, n2, nrhs, n1, -1.0, L + (long)n1 * ldl, ldl, X, ldx, 1.0, X2, ldx, o, s)))
      return rc;
    return trsm_left(0, unit, n2, nrhs, L22, ldl, X2, ldx, work, wb, s);
  }
  // [U11 U12; 0 U22]: X2 <- U22^-1 X2, X1 <- U11^-1 (X1 - U12 X2)
  if ((rc = trsm_left(1, unit, n2, nrhs, L22, ldl, X2, ldx, work, wb, s))) return rc;
  if ((rc = exc_gemm_opts(0, 0, n1, nrhs, n2, -1.0, L + n1, ldl, X2, ldx, 1.0, X, ldx, o, s))) return rc;
  return trsm_left(1, unit, n1, nrhs, L, ldl, X, ldx, work, wb, s);
}

int trsm_right(int lower_t, int unit, int n, long rows, const double* U, long ldu, double* X, long ldx, void* work,
               size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && rows >= 0 && U && X, EXC_ERR_ARG, "trsm_right: bad arguments");
  if (n == 0 || rows == 0) return EXC_OK;
  if (n <= NB) return launch_right_blk(lower_t, unit, n, U, ldu, X, ldx, rows, s);
  const int n1 = split(n), n2 = n - n1;
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  // op(U) = [V11 V12; 0 V22] (V = U, or L^T): X1 <- X1 V11^-1, X2 <- (X2 - X1 V12) V22^-1
  int rc;
  if ((rc = trsm_right(lower_t, unit, n1, rows, U, ldu, X, ldx, work, wb, s))) return rc;
  rc = lower_t ? exc_gemm_opts(0, 1, rows, n2, n1, -1.0, X, ldx, U + (long)n1 * ldu, ldu, 1.0, X + n1, ldx, o, s)
               : exc_gemm_opts(0, 0, rows, n2, n1, -1.0, X, ldx, U + n1, ldu, 1.0, X + n1, ldx, o, s);
  if (rc) return rc;
  return trsm_right(lower_t, unit, n2, rows, U + (long)n1 * ldu + n1, ldu, X + n1, ldx, work, wb, s);
}

static int potrf_rec(int n, double* A, long lda, int k0, int* info, void* work, size_t wb, hipStream_t s) {
  if (n <= NB) {
    hipLaunchKernelGGL(k_potf2, dim3(1), dim3(64), 0, s, n, A, lda, k0, info);
    EXC_CHECK_HIP(hipGetLastError());
    return EXC_OK;
  }
  const int n1 = split(n), n2 = n - n1;
  double* A11 = A + (long)k0 * lda + k0;
  double* A21 = A11 + (long)n1 * lda;
  int rc;
  if ((rc = potrf_rec(n1, A, lda, k0, info, work, wb, s))) return rc;
  if ((rc = trsm_right(1, 0, n1, n2, A11, lda, A21, lda, work, wb, s))) return rc;  // A21 <- A21 L11^-T
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  o.lower_only = 1;
  if ((rc = exc_gemm_opts(0, 1, n2, n2, n1, -1.0, A21, lda, A21, lda, 1.0, A21 + n1, lda, o, s))) return rc;
  return potrf_rec(n2, A, lda, k0 + n1, info, work, wb, s);
}

int potrf_lower(int n, double* A, long lda, int* info, void* work, size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && A, EXC_ERR_ARG, "potrf_lower: bad arguments");
  // *info is 0 unless this call meets a non-positive pivot (the kernels only write failures)
  if (info) EXC_CHECK_HIP(hipMemsetAsync(info, 0, sizeof(int), s));
  if (n == 0) return EXC_OK;
  if (n > NB && n <= kTileMaxN) return potrf_tiles(n, A, lda, info, work, wb, s);
  int rc = potrf_rec(n, A, lda, 0, info, work, wb, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_strict_upper, dim3(exc_detail::ceil_div((long)n * n, 256)), dim3(256), 0, s, n, A, lda);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

static int lu_rec(int n, double* A, long lda, int k0, int total, double* sgn, int square_last, void* work, size_t wb,
                  hipStream_t s) {
  double* A11 = A + (long)k0 * lda + k0;
  if (n <= NB) {
    const int last_plus = (square_last && k0 + n == total) ? 1 : 0;
    hipLaunchKernelGGL(k_getf2_sign, dim3(1), dim3(64), 0, s, n, A, lda, k0, sgn, last_plus);
    EXC_CHECK_HIP(hipGetLastError());
    return EXC_OK;
  }
  const int n1 = split(n), n2 = n - n1;
  double* A12 = A11 + n1;
  double* A21 = A11 + (long)n1 * lda;
  int rc;
  if ((rc = lu_rec(n1, A, lda, k0, total, sgn, square_last, work, wb, s))) return rc;
  if ((rc = trsm_left(0, 1, n1, n2, A11, lda, A12, lda, work, wb, s))) return rc;   // A12 <- L11^-1 A12
  if ((rc = trsm_right(0, 0, n1, n2, A11, lda, A21, lda, work, wb, s))) return rc;  // A21 <- A21 U11^-1
  ExcGemmOpts o;
  o.work = work;
  o.work_bytes = wb;
  if ((rc = exc_gemm_opts(0, 0, n2, n2, n1, -1.0, A21, lda, A12, lda, 1.0, A21 + n1, lda, o, s))) return rc;
  return lu_rec(n2, A, lda, k0 + n1, total, sgn, square_last, work, wb, s);
}

int lu_sign(int n, double* A, long lda, double* sgn, int square_last, void* work, size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 0 && A && sgn, EXC_ERR_ARG, "lu_sign: bad arguments");
  if (n == 0) return EXC_OK;
  if (n > NB && n <= kTileMaxN) return lu_tiles(n, A, lda, sgn, square_last, work, wb, s);
  return lu_rec(n, A, lda, 0, n, sgn, square_last, work, wb, s);
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
