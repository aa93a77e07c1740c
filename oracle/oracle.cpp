// CPU FP64 restatement of Excalibur's per-peer matrix pipeline (reference: src/bbtsvd.cpp,
// src/client.cpp, src/utils.cpp, include/utils.hpp).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the GPU implementation in
// excalibur_amd/. It is built into oracle/_build/liboracle.so by oracle/build.sh and may be
// loaded only by tests/, by __graft_entry__.smoke() (as the checker) and by bench.py's
// cpu_baseline leg. Nothing in excalibur_amd/ links, loads or calls it.
//
// Numerics. Built with -O2 -ffp-contract=off and without fast-math, so every expression is
// evaluated in the literal IEEE order of the reference source (separately rounded multiply and
// add, correctly rounded division). pow(x, 2) is written x*x (exact same result); pow(x, 0.5)
// stays a call to the C library's pow, as in the reference when built without fast-math.
// The BLAS/LAPACK pieces the reference takes from MKL (dnrm2, idamax/idamin, drotg, drot,
// dlarfg, dgeqrf, dorgqr, dormqr, dgemm, dgemv, dger) are restated from the published
// reference-BLAS/LAPACK algorithms (dnrm2 as sqrt of the sequential sum of squares, dgeqrf as
// the unblocked dgeqr2, dorgqr as dorg2r, dormqr as dorm2r). MKL's exact rounding is not
// public, so those agree with the reference to rounding, not bit for bit; integer outputs
// (Sturm counts, twist indices, Givens index sequences) and everything computed by the
// reference's own loops follow the source exactly.
//
// Pinning: tests/test_oracle_*.py check every function here against closed forms, brute force
// (eigenvalues of B^T B), LAPACK through scipy, numpy's SVD and orthogonality invariants.

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <vector>

namespace {

using vec = std::vector<double>;

// ---------------------------------------------------------------------------------------------
// BLAS level-1 restatements (reference BLAS semantics)
// ---------------------------------------------------------------------------------------------
double nrm2(long n, const double* x, long inc) {
  double s = 0.0;
  for (long i = 0; i < n; i++) {
    double v = x[i * inc];
    s = s + v * v;
  }
  return std::sqrt(s);
}

long idamax(long n, const double* x, long inc) {  // first index of the largest |x_i|
  long p = 0;
  double best = -1.0;
  for (long i = 0; i < n; i++) {
    double a = std::fabs(x[i * inc]);
    if (a > best) { best = a; p = i; }
  }
  return p;
}

long idamin(long n, const double* x, long inc) {  // first index of the smallest |x_i|
  long p = 0;
  double best = INFINITY;
  bool any = false;
  for (long i = 0; i < n; i++) {
    double a = std::fabs(x[i * inc]);
    if (!any || a < best) { best = a; p = i; any = true; }
  }
  return p;
}

double ddot(long n, const double* x, long incx, const double* y, long incy) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s = s + x[i * incx] * y[i * incy];
  return s;
}

void daxpy(long n, double a, const double* x, long incx, double* y, long incy) {
  for (long i = 0; i < n; i++) y[i * incy] = y[i * incy] + a * x[i * incx];
}

void dscal(long n, double a, double* x, long inc) {
  for (long i = 0; i < n; i++) x[i * inc] = a * x[i * inc];
}

// Reference BLAS drotg (classic form, as in reference BLAS before 3.10).
void drotg(double* a, double* b, double* c, double* s) {
  double roe = *b;
  if (std::fabs(*a) > std::fabs(*b)) roe = *a;
  double scale = std::fabs(*a) + std::fabs(*b);
  double r, z;
  if (scale == 0.0) {
    *c = 1.0; *s = 0.0; r = 0.0; z = 0.0;
  } else {
    double as = *a / scale, bs = *b / scale;
    r = scale * std::sqrt(as * as + bs * bs);
    r = std::copysign(1.0, roe) * r;
    *c = *a / r;
    *s = *b / r;
    z = 1.0;
    if (std::fabs(*a) > std::fabs(*b)) z = *s;
    if (std::fabs(*b) >= std::fabs(*a) && *c != 0.0) z = 1.0 / *c;
  }
  *a = r;
  *b = z;
}

// Reference BLAS drot: x' = c x + s y, y' = c y - s x.
void drot(long n, double* x, long incx, double* y, long incy, double c, double s) {
  for (long i = 0; i < n; i++) {
    double dx = x[i * incx], dy = y[i * incy];
    double t = c * dx + s * dy;
    y[i * incy] = c * dy - s * dx;
    x[i * incx] = t;
  }
}

// ---------------------------------------------------------------------------------------------
// LAPACK restatements (column-major, leading dimension lda), reference-LAPACK algorithms
// ---------------------------------------------------------------------------------------------
double dlapy2(double x, double y) {
  double xa = std::fabs(x), ya = std::fabs(y);
  double w = std::max(xa, ya), z = std::min(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  double q = z / w;
  return w * std::sqrt(1.0 + q * q);
}

void dlarfg(long n, double* alpha, double* x, long incx, double* tau) {
  if (n <= 1) { *tau = 0.0; return; }
  double xnorm = nrm2(n - 1, x, incx);
  if (xnorm == 0.0) { *tau = 0.0; return; }
  double beta = -std::copysign(dlapy2(*alpha, xnorm), *alpha);
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  int knt = 0;
  if (std::fabs(beta) < safmin) {
    const double rsafmn = 1.0 / safmin;
    do {
      knt++;
      dscal(n - 1, rsafmn, x, incx);
      beta = beta * rsafmn;
      *alpha = *alpha * rsafmn;
    } while (std::fabs(beta) < safmin && knt < 20);
    xnorm = nrm2(n - 1, x, incx);
    beta = -std::copysign(dlapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  dscal(n - 1, 1.0 / (*alpha - beta), x, incx);
  for (int j = 0; j < knt; j++) beta = beta * safmin;
  *alpha = beta;
}

// Apply H = I - tau v v^T from the left to C (rows x cols, column-major ldc); v has unit stride
// and v[0] is taken as stored (callers set it to 1).
// Columns are independent, so they are split over threads; each element keeps the sequential
// order above, and the results do not depend on the thread count.
void dlarf_left(long rows, long cols, const double* v, double tau, double* C, long ldc) {
  if (tau == 0.0) return;
#pragma omp parallel for schedule(static) if (rows * cols > (1L << 20))
  for (long j = 0; j < cols; j++) {
    double t = 0.0;  // w_j = (C^T v)_j (reference dgemv 'T')
    for (long i = 0; i < rows; i++) t = t + C[i + j * ldc] * v[i];
    t = -tau * t;  // C_j = C_j - tau v w_j (reference dger)
    for (long i = 0; i < rows; i++) C[i + j * ldc] = C[i + j * ldc] + v[i] * t;
  }
}

void dgeqr2(long m, long n, double* A, long lda, double* tau) {
  long k = std::min(m, n);
  for (long i = 0; i < k; i++) {
    dlarfg(m - i, &A[i + i * lda], &A[std::min(i + 1, m - 1) + i * lda], 1, &tau[i]);
    if (i < n - 1) {
      double aii = A[i + i * lda];
      A[i + i * lda] = 1.0;
      dlarf_left(m - i, n - i - 1, &A[i + i * lda], tau[i], &A[i + (i + 1) * lda], lda);
      A[i + i * lda] = aii;
    }
  }
}

void dorg2r(long m, long n, long k, double* A, long lda, const double* tau) {
  for (long j = k; j < n; j++) {
    for (long l = 0; l < m; l++) A[l + j * lda] = 0.0;
    A[j + j * lda] = 1.0;
  }
  for (long i = k - 1; i >= 0; i--) {
    if (i < n - 1) {
      A[i + i * lda] = 1.0;
      dlarf_left(m - i, n - i - 1, &A[i + i * lda], tau[i], &A[i + (i + 1) * lda], lda);
    }
    if (i < m - 1) dscal(m - i - 1, -tau[i], &A[i + 1 + i * lda], 1);
    A[i + i * lda] = 1.0 - tau[i];
    for (long l = 0; l < i; l++) A[l + i * lda] = 0.0;
  }
}

// dorm2r, side 'L' only: C (m x n, column-major ldc) <- op(Q) C, Q = H_0 ... H_{k-1}, reflector i
// is column i of A from row i down (unit diagonal implied).
void dorm2r_left(bool trans, long m, long n, long k, const double* A, long lda, const double* tau,
                 double* C, long ldc) {
  vec v(m);
  auto apply = [&](long i) {
    long len = m - i;
    v[0] = 1.0;
    for (long l = 1; l < len; l++) v[l] = A[i + l + i * lda];
    dlarf_left(len, n, v.data(), tau[i], &C[i], ldc);
  };
  if (trans) for (long i = 0; i < k; i++) apply(i);
  else for (long i = k - 1; i >= 0; i--) apply(i);
}

// Row-major GEMM C (M x N) = op(A) op(B): every element is the sequential sum over l = 0..K-1
// starting from 0. Rows of C are split over threads; for a non-transposed B the l loop runs
// outside the j loop so that B is read along its rows, which adds the same terms in the same
// order to each element.
void gemm_rm(bool ta, bool tb, long M, long N, long K, const double* A, long lda, const double* B,
             long ldb, double* C, long ldc) {
#pragma omp parallel for schedule(static) if ((double)M * N * K > 1e6)
  for (long i = 0; i < M; i++) {
    double* c = C + i * ldc;
    if (!tb) {
      for (long j = 0; j < N; j++) c[j] = 0.0;
      for (long l = 0; l < K; l++) {
        const double a = ta ? A[l * lda + i] : A[i * lda + l];
        const double* b = B + l * ldb;
        for (long j = 0; j < N; j++) c[j] = c[j] + a * b[j];
      }
    } else {
      for (long j = 0; j < N; j++) {
        double s = 0.0;
        for (long l = 0; l < K; l++) s = s + (ta ? A[l * lda + i] : A[i * lda + l]) * B[j * ldb + l];
        c[j] = s;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Sturm count and bisection (bbtsvd.cpp:7-52)
// ---------------------------------------------------------------------------------------------
int negative_count(int n, const double* a, const double* b, double mu) {
  int count = 0;
  double d = 1.0, t = 0.0;
  for (int i = 0; i < n; i++) {
    if (i > 0) t = t * ((b[i - 1] * b[i - 1]) / d) - mu * mu;
    else t = -(mu * mu);
    d = a[i] * a[i] + t;
    if (d < 0) count++;
  }
  return count;
}

void bisection(int n, const double* a, const double* b, double lo, double hi, double tol, int top_k,
               double* sigma, int* count, double* leaf_lo, double* leaf_hi) {
  int nc_a = negative_count(n, a, b, lo);
  int nc_b = negative_count(n, a, b, hi);
  *count = 0;
  if (!(nc_a <= nc_b && top_k > 0)) return;
  struct Job { double low; int nlow; double high; int nhigh; };
  std::vector<Job> st;
  st.push_back({lo, nc_a, hi, nc_b});
  while (!st.empty()) {
    Job j = st.back();
    st.pop_back();
    double mid = (j.low + j.high) / 2;
    if ((j.high - j.low) <= tol || mid == j.low || mid == j.high) {
      if (leaf_lo) leaf_lo[*count] = j.low;
      if (leaf_hi) leaf_hi[*count] = j.high;
      sigma[(*count)++] = mid;
      if (*count >= top_k) break;
    } else {
      int nm = negative_count(n, a, b, mid);
      if (nm > j.nlow) st.push_back({j.low, j.nlow, mid, nm});
      if (j.nhigh > nm) st.push_back({mid, nm, j.high, j.nhigh});
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Twisted factorization (bbtsvd.cpp:54-147)
// ---------------------------------------------------------------------------------------------
void ldl(int n, const double* a, const double* b, double* l, double* d) {
  d[0] = a[0] * a[0];
  for (int i = 1; i < n; i++) {
    double ab = a[i - 1] * b[i - 1];
    l[i - 1] = ab / d[i - 1];
    d[i] = a[i] * a[i] + b[i - 1] * b[i - 1] - (ab * ab) / d[i - 1];
  }
}

void inverse_iteration(int n, int k, const double* l, const double* u, double* v, long ldv) {
  v[(long)k * ldv] = 1.0;
  for (int i = k - 1; i >= 0; i--) v[i * ldv] = -l[i] * v[(i + 1) * ldv];
  for (int i = k + 1; i < n; i++) v[i * ldv] = -u[i - 1] * v[(i - 1) * ldv];
  double inv = 1.0 / nrm2(n, v, ldv);
  dscal(n, inv, v, ldv);
}

void dstqds(int n, const double* l, const double* d, double tau, double* lp, double* dp, double* s) {
  s[0] = -tau;
  for (int i = 0; i < n - 1; i++) {
    dp[i] = d[i] + s[i];
    lp[i] = d[i] * l[i] / dp[i];
    s[i + 1] = lp[i] * l[i] * s[i] - tau;
  }
  dp[n - 1] = d[n - 1] + s[n - 1];
}

void dpqds(int n, const double* l, const double* d, double tau, double* up, double* rp, double* p) {
  p[n - 1] = d[n - 1] - tau;
  for (int i = n - 2; i >= 0; i--) {
    rp[i + 1] = d[i] * (l[i] * l[i]) + p[i + 1];
    up[i] = l[i] * d[i] / rp[i + 1];
    p[i] = p[i + 1] * d[i] / rp[i + 1] - tau;
  }
  rp[0] = p[0];
}

void one_pair(int n, long ldu, double omega, const double* l, const double* d, const double* a,
              const double* b, double* v, double* u, int* twist_v, int* twist_u) {
  double w2 = omega * omega;
  size_t n1 = n > 1 ? (size_t)(n - 1) : 1;
  vec lp(n1), dp(n), s(n + 1), up(n1), rp(n + 1), p(n + 1), g(n);
  dstqds(n, l, d, w2, lp.data(), dp.data(), s.data());
  dpqds(n, l, d, w2, up.data(), rp.data(), p.data());
  for (int i = 0; i < n; i++) g[i] = s[i] + p[i] + w2;
  int k = (int)idamin(n, g.data(), 1);
  if (twist_v) *twist_v = k;
  inverse_iteration(n, k, lp.data(), up.data(), v, 1);
  // coupling to the twisted factorization of B B^T (bbtsvd.cpp:113-124)
  s[n] = -w2; p[n] = 1.0; rp[n] = 1.0;
  for (int i = 0; i < n; i++) {
    dp[i] = s[i + 1] / s[i] * dp[i];
    g[i] = -w2 * g[i] * rp[i + 1] / (s[i] * p[i + 1]);
  }
  k = (int)idamin(n, g.data(), 1);
  if (twist_u) *twist_u = k;
  for (int i = 0; i < n; i++) rp[i] = p[i] / p[i + 1] * rp[i + 1];
  for (int i = 0; i < n - 1; i++) {
    lp[i] = a[i + 1] / dp[i] * b[i];
    up[i] = a[i + 1] / rp[i + 1] * b[i];
  }
  inverse_iteration(n, k, lp.data(), up.data(), u, ldu);
}

void two_side(int n, int kcount, const double* omega, const double* a, const double* b, double* u,
              double* v, int* twists) {
  vec l(n > 1 ? n - 1 : 1), d(n);
  ldl(n, a, b, l.data(), d.data());
  for (int i = 0; i < kcount; i++)
    one_pair(n, kcount, omega[i], l.data(), d.data(), a, b, v + (long)i * n, u + i,
             twists ? &twists[2 * i] : nullptr, twists ? &twists[2 * i + 1] : nullptr);
}

// 7-argument bidiagonal_bisection_twisted_svd (bbtsvd.cpp:932-968). Returns the bisection's
// sigma count, or -1 when it is below top_k (the reference asserts there).
int bidiag_svd(int k, const double* alpha, const double* beta, double* u, double* v, double* sigma,
               int top_k, int* twists) {
  int cnt = 0;
  double na = nrm2(k, alpha, 1), nb = nrm2(k - 1, beta, 1);
  double ub = na * na + nb * nb;
  bisection(k, alpha, beta, 0, ub, 1e-20, top_k, sigma, &cnt, nullptr, nullptr);
  if (cnt < top_k) return -1;
  two_side(k, top_k, sigma, alpha, beta, u, v, twists);
  for (int i = 0; i < top_k; i++) {  // sign of v from (B^T u)_p / sigma
    long p = idamax(k, v + (long)i * k, 1);
    double signer;
    if (p == 0) signer = u[p * top_k + i] * alpha[p] / sigma[i];
    else signer = (u[(p - 1) * top_k + i] * beta[p - 1] + u[p * top_k + i] * alpha[p]) / sigma[i];
    if ((signer * v[(long)i * k + p]) < 0) dscal(k, -1.0, v + (long)i * k, 1);
  }
  for (int i = 0; i < top_k; i++) {  // "check None" (bbtsvd.cpp:957-967)
    if (std::fabs(sigma[i] - alpha[i]) < 1e-15) {
      u[i + (long)i * top_k] = 1.0;
      for (int j = i + 1; j < top_k; j++) u[(long)i * top_k + j] = 0.0;
      for (int j = i + 1; j < k; j++) u[(long)j * top_k + i] = 0.0;
      for (int j = 0; j < i; j++) v[(long)i * k + j] = 0.0;
      v[(long)i * k + i] = 1.0;
      for (int j = i + 1; j < k; j++) v[(long)i * k + j] = 0.0;
    }
  }
  return cnt;
}

// ---------------------------------------------------------------------------------------------
// Givens masking (bbtsvd.cpp:727-876, utils.hpp:31-33)
// ---------------------------------------------------------------------------------------------
inline double random_value(double lo, double hi) {
  return (rand() / (double)RAND_MAX) * (hi - lo) + lo;
}

struct GivensPlan {
  size_t K = 0, passes = 0, half = 0;
  std::vector<std::vector<size_t>> idx;
  std::vector<vec> c, s;
};

// Index sequences and rotation values exactly as bbtsvd.cpp:729-779 draws them.
void givens_plan(size_t K, size_t passes, bool mask, unsigned seed, GivensPlan& P) {
  srand(seed);
  auto dre = std::default_random_engine(seed);
  P.K = K; P.passes = passes;
  P.idx.clear(); P.c.clear(); P.s.clear();
  for (size_t i = 0; i < passes; i++) {
    std::vector<size_t> t;
    for (size_t j = 0; j < K; j++) t.push_back(j);
    std::shuffle(t.begin(), t.end(), dre);
    if (K % 2 == 1) {
      t.push_back(rand() % K);
      while (t[K] == t[K - 1]) t[K] = rand() % K;
    }
    if (!mask) std::reverse(t.begin(), t.end());
    for (size_t j = 0; j < ((K % 2) == 0 ? K : (K + 1)); j += 2)
      if (t[j] > t[j + 1]) std::swap(t[j], t[j + 1]);
    P.idx.push_back(t);
  }
  if (!mask) std::reverse(P.idx.begin(), P.idx.end());
  P.half = ((K % 2) == 0 ? K : (K + 1)) / 2;
  for (size_t i = 0; i < passes; i++) {
    vec ci(P.half, 0), si(P.half, 0);
    for (size_t j = 0; j < P.half; j++) {
      ci[j] = random_value(-1, 1);
      si[j] = random_value(-1, 1);
      double root = std::pow(ci[j] * ci[j] + si[j] * si[j], 0.5);
      ci[j] /= root;
      si[j] /= root;
    }
    if (!mask) { std::reverse(ci.begin(), ci.end()); std::reverse(si.begin(), si.end()); }
    P.c.push_back(ci);
    P.s.push_back(si);
  }
  if (!mask) { std::reverse(P.c.begin(), P.c.end()); std::reverse(P.s.begin(), P.s.end()); }
}

// x (m x n row-major) in place. Per element the rotations run in (pass, pair) order with
// x1' = c x1 + s' x2 and x2' = s'' x1 + c x2, s' = (mask ? s : -s), s'' = -s'.
void givens_mask(size_t m, size_t n, size_t passes, double* x, bool mask, bool is_left, unsigned seed) {
  size_t K = is_left ? m : n;
  if (K <= 1) return;
  GivensPlan P;
  givens_plan(K, passes, mask, seed, P);
  for (size_t i = 0; i < passes; i++)
    for (size_t j = 0; j < P.half; j++) {
      size_t r1 = P.idx[i][2 * j], r2 = P.idx[i][2 * j + 1];
      double cc = P.c[i][j];
      double ss = (mask ? 1 : -1) * P.s[i][j];
      double _ss = (mask ? -1 : 1) * P.s[i][j];
      if (is_left) {
        double* p1 = x + r1 * n;
        double* p2 = x + r2 * n;
        for (size_t col = 0; col < n; col++) {
          double a = p1[col], b = p2[col];
          p1[col] = a * cc + b * ss;
          p2[col] = a * _ss + b * cc;
        }
      } else {
        for (size_t row = 0; row < m; row++) {
          double* pr = x + row * n;
          double a = pr[r1], b = pr[r2];
          pr[r1] = a * cc + b * ss;
          pr[r2] = a * _ss + b * cc;
        }
      }
    }
}

struct SquareMaskConvert { size_t k, m, n, remain; };

// bbtsvd.cpp:166-188 (the recursion written as a loop over `remain`).
void find_best_convert(SquareMaskConvert* smc) {
  const size_t threshold = 10000;
  while (true) {
    size_t k = smc->k - smc->remain;
    size_t tm = k, tn = 1;
    for (size_t i = threshold; i > 1; i--)
      if (k % i == 0) {
        size_t kd = k / i;
        if (i > kd && kd < threshold && kd > 1 && (i - kd) < (tm - tn) &&
            ((double)i / (double)kd) < 2.0) {
          tm = i;
          tn = k / i;
        }
      }
    if (tm < threshold) { smc->m = tm; smc->n = tn; return; }
    smc->remain += 1;
  }
}

// block_shift_mask (bbtsvd.cpp:533-669).
void block_shift_mask(size_t m, size_t n, size_t passes, double* x, bool mask, bool is_left, unsigned seed) {
  srand(seed);
  size_t k = is_left ? m : n;
  if (k <= 10000) { givens_mask(m, n, passes, x, mask, is_left, seed); return; }
  SquareMaskConvert smc = {k, 0, 0, 0};
  find_best_convert(&smc);
  size_t vm = smc.m, vn = smc.n, rem = smc.remain;
  bool remain = rem > 0;
  size_t kk = k - rem;
  vec blk(kk), sh(kk), rb(remain ? rem + vm : 1);
  size_t lines = is_left ? n : m;
  long stride = is_left ? (long)n : 1;            // step between the K entries of one line
  auto line = [&](size_t i) { return x + i * (is_left ? 1 : n); };
  for (size_t i = 0; i < lines; i++) {
    double* L = line(i);
    if (mask) {
      for (size_t q = 0; q < kk; q++) blk[q] = L[q * stride];
      if (remain) {
        for (size_t q = 0; q < vm; q++) rb[q] = blk[q * vn];
        for (size_t q = 0; q < rem; q++) rb[vm + q] = L[(kk + q) * stride];
        givens_mask(1, rem + vm, passes, rb.data(), true, false, seed + 1000);
        for (size_t q = 0; q < vm; q++) blk[q * vn] = rb[q];
      }
      givens_mask(vm, vn, passes, blk.data(), true, true, seed + 2000);
      for (size_t j = 0; j < vm; j++)
        for (size_t q = 0; q < vn; q++) {
          int shift = ((int)q - (int)j) % ((int)vn);
          if (shift < 0) shift += (int)vn;
          sh[j * vn + q] = blk[j * vn + shift];
        }
      givens_mask(vm, vn, passes, sh.data(), true, true, seed + 3000);
      givens_mask(vm, vn, passes, sh.data(), true, false, seed + 4000);
      if (remain) {
        for (size_t q = 0; q < vm; q++) rb[q] = sh[q * vn];
        givens_mask(1, rem + vm, passes, rb.data(), true, false, seed + 5000);
        for (size_t q = 0; q < vm; q++) sh[q * vn] = rb[q];
        for (size_t q = 0; q < rem; q++) L[(kk + q) * stride] = rb[vm + q];
      }
      for (size_t q = 0; q < kk; q++) L[q * stride] = sh[q];
    } else {
      for (size_t q = 0; q < kk; q++) sh[q] = L[q * stride];
      if (remain) {
        for (size_t q = 0; q < rem; q++) rb[vm + q] = L[(kk + q) * stride];
        for (size_t q = 0; q < vm; q++) rb[q] = sh[q * vn];
        givens_mask(1, rem + vm, passes, rb.data(), false, false, seed + 5000);
        for (size_t q = 0; q < vm; q++) sh[q * vn] = rb[q];
      }
      givens_mask(vm, vn, passes, sh.data(), false, true, seed + 3000);
      givens_mask(vm, vn, passes, sh.data(), false, false, seed + 4000);
      for (size_t j = 0; j < vm; j++)
        for (size_t q = 0; q < vn; q++) blk[j * vn + q] = sh[j * vn + (q + j) % ((int)vn)];
      givens_mask(vm, vn, passes, blk.data(), false, true, seed + 2000);
      if (remain) {
        for (size_t q = 0; q < vm; q++) rb[q] = blk[q * vn];
        givens_mask(1, rem + vm, passes, rb.data(), false, false, seed + 1000);
        for (size_t q = 0; q < vm; q++) blk[q * vn] = rb[q];
        for (size_t q = 0; q < rem; q++) L[(kk + q) * stride] = rb[vm + q];
      }
      for (size_t q = 0; q < kk; q++) L[q * stride] = blk[q];
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Pipeline stages of DecentralizedClient::run() (client.cpp:417-1199), short-and-wide branch
// ---------------------------------------------------------------------------------------------

// local_qr_reduction(local_n, m, ...) (client.cpp:417-445). x: m x r row-major (= X_i, i.e. the
// column-major r x m matrix the reference factors). rt: m x m row-major R^T. qt (nullable):
// m x r row-major Q^T (= the reference's local_v).
void local_qr(long r, long m, const double* x, double* rt, double* qt) {
  vec A(x, x + r * m), tau(m, 0.0);
  dgeqr2(r, m, A.data(), r, tau.data());
  if (qt) {
    std::fill(qt, qt + r * m, 0.0);
    for (long i = 0; i < m; i++)
      for (long l = i; l < r; l++) qt[i * r + l] = A[i * r + l];
    dorg2r(r, m, m, qt, r, tau.data());
  }
  std::fill(rt, rt + m * m, 0.0);
  for (long i = 0; i < m; i++)
    for (long j = 0; j <= i; j++) rt[i * m + j] = A[i * r + j];
}

// Global QR over P simulated peers (client.cpp:706-794). panels[c]: m_part[c] x R row-major
// (= column-major R x m_part[c] with ld R), updated in place. qg (nullable): R x m column-major
// Q (the reference's mask_shared_v after dorgqr). shrink[c]: m x m_part[c] row-major R[:, own].
void global_qr(int P, long m, long R, const long* m_part, double** panels, double* qg,
               double** shrink, double* tau_out) {
  std::vector<long> pos(P + 1, 0);
  for (int i = 0; i < P; i++) pos[i + 1] = pos[i] + m_part[i];
  vec V((size_t)R * m, 0.0), tau(m, 0.0);
  for (int i = 0; i < P; i++) {
    double* A = panels[i];
    dgeqr2(R - pos[i], m_part[i], A + pos[i], R, tau.data() + pos[i]);
    for (long j = pos[i]; j < pos[i + 1]; j++) {
      long lj = j - pos[i];
      for (long l = j + 1; l < R; l++) {
        V[l + j * R] = A[l + lj * R];
        A[l + lj * R] = 0.0;
      }
    }
    for (int c = i + 1; c < P; c++)
      dorm2r_left(true, R - pos[i], m_part[c], m_part[i], V.data() + pos[i] * R + pos[i], R,
                  tau.data() + pos[i], panels[c] + pos[i], R);
  }
  if (qg) {
    std::copy(V.begin(), V.end(), qg);
    dorg2r(R, m, m, qg, R, tau.data());
  }
  if (tau_out) std::copy(tau.begin(), tau.end(), tau_out);
  for (int c = 0; c < P; c++)
    for (long j = 0; j < m; j++)
      for (long l = 0; l < m_part[c]; l++) shrink[c][j * m_part[c] + l] = panels[c][l * R + j];
}

// Per-peer all-reduce of the small path (client.cpp:313-323): own buffer, then the others'
// buffers added in increasing peer id.
void allreduce_small(int P, long len, std::vector<vec>& msg) {
  std::vector<vec> out(P);
  for (int p = 0; p < P; p++) {
    out[p] = vec(msg[p].begin(), msg[p].begin() + len);
    for (int q = 0; q < P; q++) {
      if (q == p) continue;
      for (long i = 0; i < len; i++) out[p][i] = out[p][i] + msg[q][i];
    }
  }
  for (int p = 0; p < P; p++) std::copy(out[p].begin(), out[p].end(), msg[p].begin());
}

// Ralha one-sided bidiagonalization, overlapped variant (client.cpp:829-910), over P simulated
// peers. F[p]: m x c[p] row-major (destroyed). Per peer: alpha (m), beta (m-1) lower
// bidiagonal, house (m x m row-major, reflector j in column j from row j), tau (m), U (m x c[p]).
void ralha(int P, long m, const long* c, double** F, double** alpha, double** beta, double** house,
           double** tau, double** U) {
  std::vector<vec> msg(P, vec(m + 3, 0.0));
  for (int p = 0; p < P; p++) {
    std::fill(house[p], house[p] + m * m, 0.0);
    std::fill(tau[p], tau[p] + m, 0.0);
    std::fill(U[p], U[p] + m * c[p], 0.0);
    std::fill(alpha[p], alpha[p] + m, 0.0);
    std::fill(beta[p], beta[p] + (m - 1), 0.0);
    house[p][0] = 1.0;
  }
  // alpha_0 = ||f_0||, u_0 = f_0 / alpha_0
  for (int p = 0; p < P; p++) { double t = nrm2(c[p], F[p], 1); msg[p][0] = t * t; }
  allreduce_small(P, 1, msg);
  for (int p = 0; p < P; p++) {
    alpha[p][0] = std::pow(msg[p][0], 0.5);
    daxpy(c[p], 1.0 / alpha[p][0], F[p], 1, U[p], 1);
  }
  for (int p = 0; p < P; p++) std::fill(msg[p].begin(), msg[p].end(), 0.0);
  for (long i = 0; i < m - 2; i++) {
    for (int p = 0; p < P; p++) {
      long cp = c[p];
      const double* fi = F[p] + i * cp;
      for (long r = 0; r < m - 1 - i; r++) msg[p][r] = ddot(cp, F[p] + (i + 1 + r) * cp, 1, fi, 1);
      if (i > 0) {
        const double* ui = U[p] + (i - 1) * cp;
        msg[p][m - i - 1] = ddot(cp, fi, 1, ui, 1);
        msg[p][m - i] = ddot(cp, ui, 1, ui, 1);
        msg[p][m - i + 1] = ddot(cp, fi, 1, fi, 1);
        msg[p][m - i + 2] = ddot(cp, ui, 1, fi, 1);
      }
    }
    allreduce_small(P, m + 3 - i, msg);
    for (int p = 0; p < P; p++) {
      long cp = c[p];
      double* H = house[p];
      const vec& g = msg[p];
      for (long r = 0; r < m - 1 - i; r++) H[(i + 1 + r) * m + (i + 1)] = g[r];
      if (i > 0) beta[p][i - 1] = g[m - i - 1];
      dlarfg(m - 1 - i, &H[(i + 1) * (m + 1)], &H[(i + 1) * (m + 1) + m], m, &tau[p][i + 1]);
      H[(i + 1) * (m + 1)] = 1.0;
      // vt = F[i+1:]^T h (reference dgemv 'N' on the column-major view), F -= tau h vt^T (dger)
      vec vt(cp, 0.0);
      for (long r = 0; r < m - 1 - i; r++) {
        double t = H[(i + 1 + r) * m + (i + 1)];
        const double* fr = F[p] + (i + 1 + r) * cp;
        for (long col = 0; col < cp; col++) vt[col] = vt[col] + t * fr[col];
      }
      for (long r = 0; r < m - 1 - i; r++) {
        double t = -tau[p][i + 1] * H[(i + 1 + r) * m + (i + 1)];
        double* fr = F[p] + (i + 1 + r) * cp;
        for (long col = 0; col < cp; col++) fr[col] = fr[col] + vt[col] * t;
      }
      if (i > 0) {
        double b = beta[p][i - 1];
        daxpy(cp, -b, U[p] + (i - 1) * cp, 1, F[p] + i * cp, 1);
        alpha[p][i] = std::pow(g[m - i] * (b * b) + g[m - i + 1] - 2 * b * g[m - i + 2], 0.5);
        daxpy(cp, 1.0 / alpha[p][i], F[p] + i * cp, 1, U[p] + i * cp, 1);
      }
    }
  }
  for (long i = (m >= 2 ? m - 2 : 1); i < m; i++) {
    for (int p = 0; p < P; p++) msg[p][0] = ddot(c[p], F[p] + i * c[p], 1, U[p] + (i - 1) * c[p], 1);
    allreduce_small(P, 1, msg);
    for (int p = 0; p < P; p++) {
      beta[p][i - 1] = msg[p][0];
      daxpy(c[p], -beta[p][i - 1], U[p] + (i - 1) * c[p], 1, F[p] + i * c[p], 1);
      double t = nrm2(c[p], F[p] + i * c[p], 1);
      msg[p][0] = t * t;
    }
    allreduce_small(P, 1, msg);
    for (int p = 0; p < P; p++) {
      alpha[p][i] = std::pow(msg[p][0], 0.5);
      daxpy(c[p], 1.0 / alpha[p][i], F[p] + i * c[p], 1, U[p] + i * c[p], 1);
    }
  }
}

// Lower to upper bidiagonal by drotg (client.cpp:947-954). cs: 2(m-1) values (c_i, s_i).
void lower_to_upper(long m, double* alpha, double* beta, double* cs) {
  for (long i = 0; i < m - 1; i++) {
    double bu = beta[i];
    drotg(alpha + i, &bu, cs + 2 * i, cs + 2 * i + 1);
    beta[i] = cs[2 * i + 1] * alpha[i + 1];
    alpha[i + 1] = cs[2 * i] * alpha[i + 1];
  }
}

// Peer-consistency sign rule (client.cpp:987-995): the largest-magnitude entry of each column
// of bu (m x cnt) becomes positive; the matching row of bv (cnt x m) flips with it.
void sign_normalize(long m, long cnt, double* bu, double* bv) {
  for (long i = 0; i < cnt; i++) {
    long p = idamax(m, bu + i, cnt);
    if (bu[i + p * cnt] < 0) {
      dscal(m, -1.0, bu + i, cnt);
      dscal(m, -1.0, bv + i * m, 1);
    }
  }
}

// dormqr('L','N') on a row-major C (m x n) with reflectors stored as columns of row-major
// house (m x m): C <- H_0 H_1 ... H_{m-1} C.
// Per reflector, w_q = sum over l of C[j+l][q] v_l (sequential in l from 0), then
// C[j+l][q] += v_l (-tau_j w_q). The loops run l outside q so that C is read along its rows, and
// column blocks go to threads; each element sees the same operations in the same order.
void ormqr_house_ln(long m, long n, const double* house, const double* tau, double* C) {
  vec v(m), w(n);
  const long QB = 64;
  for (long j = m - 1; j >= 0; j--) {
    if (tau[j] == 0.0) continue;
    long len = m - j;
    v[0] = 1.0;
    for (long l = 1; l < len; l++) v[l] = house[(j + l) * m + j];
#pragma omp parallel for schedule(static) if (len * n > (1L << 18))
    for (long q0 = 0; q0 < n; q0 += QB) {
      const long q1 = std::min(n, q0 + QB);
      for (long q = q0; q < q1; q++) w[q] = 0.0;
      for (long l = 0; l < len; l++) {
        const double* c = C + (j + l) * n;
        for (long q = q0; q < q1; q++) w[q] = w[q] + c[q] * v[l];
      }
      for (long q = q0; q < q1; q++) w[q] = -tau[j] * w[q];
      for (long l = 0; l < len; l++) {
        double* c = C + (j + l) * n;
        for (long q = q0; q < q1; q++) c[q] = c[q] + v[l] * w[q];
      }
    }
  }
}

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

// One simulated run() of P peers. Results and intermediates are kept by name per peer.
struct RunState {
  int P = 0;
  long m = 0;
  int count = 0;
  std::map<std::string, std::vector<vec>> arrays;  // name -> per-peer array
  std::map<std::string, double> timers;
  vec& get(const std::string& name, int p) {
    auto& a = arrays[name];
    if ((int)a.size() < P) a.resize(P);
    return a[p];
  }
};

}  // namespace

extern "C" {

// ----- elementary restatements, exported for the oracle's own pinning tests -----
double orc_nrm2(long n, const double* x, long inc) { return nrm2(n, x, inc); }
long orc_idamax(long n, const double* x, long inc) { return idamax(n, x, inc); }
long orc_idamin(long n, const double* x, long inc) { return idamin(n, x, inc); }
void orc_drotg(double* a, double* b, double* c, double* s) { drotg(a, b, c, s); }
void orc_dlarfg(long n, double* alpha, double* x, long incx, double* tau) { dlarfg(n, alpha, x, incx, tau); }
void orc_dgeqrf(long m, long n, double* A, long lda, double* tau) { dgeqr2(m, n, A, lda, tau); }
void orc_dorgqr(long m, long n, long k, double* A, long lda, const double* tau) { dorg2r(m, n, k, A, lda, tau); }
void orc_dormqr_left(int trans, long m, long n, long k, const double* A, long lda, const double* tau,
                     double* C, long ldc) {
  dorm2r_left(trans != 0, m, n, k, A, lda, tau, C, ldc);
}

// ----- bbtsvd.hpp functions -----
int orc_negative_count(int n, const double* a, const double* b, double mu) { return negative_count(n, a, b, mu); }
void orc_negative_count_batch(int n, const double* a, const double* b, const double* mu, int nmu, int* out) {
  for (int i = 0; i < nmu; i++) out[i] = negative_count(n, a, b, mu[i]);
}
void orc_bisection(int n, const double* a, const double* b, double lo, double hi, double tol, int top_k,
                   double* sigma, int* count, double* leaf_lo, double* leaf_hi) {
  bisection(n, a, b, lo, hi, tol, top_k, sigma, count, leaf_lo, leaf_hi);
}
void orc_ldl(int n, const double* a, const double* b, double* l, double* d) { ldl(n, a, b, l, d); }
void orc_inverse_iteration(int n, int k, const double* l, const double* u, double* v, long ldv) {
  inverse_iteration(n, k, l, u, v, ldv);
}
void orc_dstqds(int n, const double* l, const double* d, double tau, double* lp, double* dp, double* s) {
  dstqds(n, l, d, tau, lp, dp, s);
}
void orc_dpqds(int n, const double* l, const double* d, double tau, double* up, double* rp, double* p) {
  dpqds(n, l, d, tau, up, rp, p);
}
void orc_one_pair(int n, long ldu, double omega, const double* l, const double* d, const double* a,
                  const double* b, double* v, double* u, int* twist_v, int* twist_u) {
  one_pair(n, ldu, omega, l, d, a, b, v, u, twist_v, twist_u);
}
void orc_two_side(int n, int k, const double* omega, const double* a, const double* b, double* u, double* v,
                  int* twists) {
  two_side(n, k, omega, a, b, u, v, twists);
}
int orc_bidiag_svd(int k, const double* alpha, const double* beta, double* u, double* v, double* sigma,
                   int top_k, int* twists) {
  return bidiag_svd(k, alpha, beta, u, v, sigma, top_k, twists);
}

// Givens plan: pairs (passes x 2*half, as indices), c and s (passes x half).
long orc_givens_half(long K) { return (K % 2 == 0 ? K : K + 1) / 2; }
void orc_givens_plan(long K, long passes, int mask, unsigned seed, long* pairs, double* c, double* s) {
  GivensPlan P;
  givens_plan((size_t)K, (size_t)passes, mask != 0, seed, P);
  for (size_t i = 0; i < P.passes; i++)
    for (size_t j = 0; j < P.half; j++) {
      pairs[(i * P.half + j) * 2] = (long)P.idx[i][2 * j];
      pairs[(i * P.half + j) * 2 + 1] = (long)P.idx[i][2 * j + 1];
      c[i * P.half + j] = P.c[i][j];
      s[i * P.half + j] = P.s[i][j];
    }
}
void orc_givens_mask(long m, long n, long passes, double* x, int mask, int is_left, unsigned seed) {
  givens_mask((size_t)m, (size_t)n, (size_t)passes, x, mask != 0, is_left != 0, seed);
}
void orc_block_shift_mask(long m, long n, long passes, double* x, int mask, int is_left, unsigned seed) {
  block_shift_mask((size_t)m, (size_t)n, (size_t)passes, x, mask != 0, is_left != 0, seed);
}
void orc_find_best_convert(long k, long* m, long* n, long* remain) {
  SquareMaskConvert smc = {(size_t)k, 0, 0, 0};
  find_best_convert(&smc);
  *m = (long)smc.m; *n = (long)smc.n; *remain = (long)smc.remain;
}

// ----- pipeline stages -----
void orc_local_qr(long r, long m, const double* x, double* rt, double* qt) { local_qr(r, m, x, rt, qt); }
void orc_global_qr(int P, long m, long R, const long* m_part, double** panels, double* qg, double** shrink,
                   double* tau) {
  global_qr(P, m, R, m_part, panels, qg, shrink, tau);
}
void orc_ralha(int P, long m, const long* c, double** F, double** alpha, double** beta, double** house,
               double** tau, double** U) {
  ralha(P, m, c, F, alpha, beta, house, tau, U);
}
void orc_lower_to_upper(long m, double* alpha, double* beta, double* cs) { lower_to_upper(m, alpha, beta, cs); }
void orc_sign_normalize(long m, long cnt, double* bu, double* bv) { sign_normalize(m, cnt, bu, bv); }
void orc_ormqr_house_ln(long m, long n, const double* house, const double* tau, double* C) {
  ormqr_house_ln(m, n, house, tau, C);
}
void orc_drot(long n, double* x, long incx, double* y, long incy, double c, double s) { drot(n, x, incx, y, incy, c, s); }
void orc_gemm(int ta, int tb, long M, long N, long K, const double* A, long lda, const double* B, long ldb,
              double* C, long ldc) {
  gemm_rm(ta != 0, tb != 0, M, N, K, A, lda, B, ldb, C, ldc);
}

// Bidiagonal SVD and sign rule of every simulated peer on its own copy of (alpha, beta) of order
// n (client.cpp:963-996). Returns the count kept, or -1 when the bisection fails.
static int bsvd_all(RunState* S, int P, long n, std::vector<double*>& al, std::vector<double*>& be, int top_k) {
  int cnt = 0;
  for (int p = 0; p < P; p++) {
    double* a = al[p];
    double* b = be[p];
    double na = nrm2(n, a, 1), nb = nrm2(n - 1, b, 1);
    double ub = na * na + nb * nb;
    int sc = negative_count(n, a, b, ub) - negative_count(n, a, b, 0);
    if (top_k > 0) sc = std::min(top_k, sc);
    vec& bu = S->get("bidiag_u", p);
    vec& bv = S->get("bidiag_v", p);
    vec& sg = S->get("sigma", p);
    bu.assign((size_t)n * sc, 0.0);
    bv.assign((size_t)n * sc, 0.0);
    sg.assign(sc > 0 ? sc : 1, 0.0);
    vec tw(2 * (size_t)std::max(sc, 1), 0.0);
    std::vector<int> twi(2 * (size_t)std::max(sc, 1), 0);
    int rc = bidiag_svd((int)n, a, b, bu.data(), bv.data(), sg.data(), sc, twi.data());
    if (rc < 0) return -1;
    for (size_t q = 0; q < twi.size(); q++) tw[q] = twi[q];
    S->get("twists", p) = tw;
    sign_normalize(n, sc, bu.data(), bv.data());
    cnt = sc;
  }
  return cnt;
}

// The right vectors of the tall-and-skinny combine (client.cpp:1113-1154, taken when
// m >= global_n): bidiag_u rotated back and multiplied by the reflectors of order n, then this
// peer's rows npos .. npos + ln - 1 transposed into FinalVT (cnt x ln) and unmasked on the right
// with the private plan.
static void tall_final_v(RunState* S, int p, long n, int cnt, long npos, long ln, unsigned priv) {
  vec bu = S->get("bidiag_u", p);
  const vec& cs = S->get("cs", p);
  for (long i = n - 2; i >= 0; i--)
    drot(cnt, bu.data() + i * cnt, 1, bu.data() + (i + 1) * cnt, 1, cs[2 * i], -cs[2 * i + 1]);
  ormqr_house_ln(n, cnt, S->get("house", p).data(), S->get("tau", p).data(), bu.data());
  vec& fv = S->get("FinalVT", p);
  fv.assign((size_t)cnt * ln, 0.0);
  for (long a = 0; a < cnt; a++)
    for (long l = 0; l < ln; l++) fv[a * ln + l] = bu[(npos + l) * cnt + a];
  block_shift_mask(cnt, ln, 32, fv.data(), false, false, priv);
}

// ----- the tall-and-skinny branch (m > global_n) of run(): client.cpp:543-569 (masks), 627-678
// (re-partition, local QR of m_part x global_n), 800-996 with the bidiagonal of order global_n,
// 998-1154 (combine). x[p] (m * local_n[p] doubles, the Client{p}.mat bytes) is read as a
// local_n[p] x m row-major block, as the reference reads it in this branch. FinalU is cnt x m
// row-major, FinalVT cnt x local_n[p]. -----
static void* run_tall(RunState* S, int P, long m, const long* local_n, const double* const* x, unsigned pub,
                      const std::vector<unsigned>& priv, long gn, int top_k, Clock::time_point t0) {
  std::vector<long> npos(P + 1, 0);
  for (int p = 0; p < P; p++) npos[p + 1] = npos[p] + local_n[p];
  S->timers["time_init"] = secs(t0, Clock::now());
  // masks (client.cpp:543-569): local_k = local_n; private plan on the left, public on the right
  auto ta0 = Clock::now();
  std::vector<vec> masked(P);
  for (int p = 0; p < P; p++) {
    masked[p].assign(x[p], x[p] + local_n[p] * m);
    block_shift_mask(local_n[p], m, 32, masked[p].data(), true, true, priv[p]);
    block_shift_mask(local_n[p], m, 32, masked[p].data(), true, false, pub);
    S->get("masked_x", p) = masked[p];
  }
  S->timers["time_apply_mask"] = secs(ta0, Clock::now());
  // re-partition (client.cpp:627-659): peer c gets columns mpos[c] .. of every block
  std::vector<long> mp(P, m / P), mpos(P + 1, 0), bl(P);
  for (int i = 0; i < (int)(m % P); i++) mp[i] += 1;
  for (int i = 0; i < P; i++) mpos[i + 1] = mpos[i] + mp[i];
  std::vector<vec> F(P), ured(P);
  auto tq0 = Clock::now();
  for (int c = 0; c < P; c++) {
    vec rep((size_t)gn * mp[c], 0.0);
    for (int i = 0; i < P; i++)
      for (long a = 0; a < local_n[i]; a++)
        for (long b = 0; b < mp[c]; b++) rep[(npos[i] + a) * mp[c] + b] = masked[i][a * m + mpos[c] + b];
    S->get("repart", c) = rep;
    bl[c] = std::min(mp[c], gn);
    if (mp[c] > gn) {  // local QR of the m_part x global_n column-major block (client.cpp:661-676)
      F[c].assign((size_t)gn * gn, 0.0);
      ured[c].assign((size_t)gn * mp[c], 0.0);
      local_qr(mp[c], gn, rep.data(), F[c].data(), ured[c].data());
      S->get("u_reduction", c) = ured[c];
    } else {
      F[c] = rep;
    }
    S->get("ralha_in", c) = F[c];
  }
  S->timers["time_local_qr"] = secs(tq0, Clock::now());
  // Ralha of order global_n (client.cpp:800-910) and lower->upper (947-954)
  auto tb0 = Clock::now();
  std::vector<double*> Fp(P), al(P), be(P), ho(P), tau(P), Up(P);
  for (int p = 0; p < P; p++) {
    Fp[p] = F[p].data();
    S->get("alpha", p).assign(gn, 0.0); al[p] = S->get("alpha", p).data();
    S->get("beta", p).assign(gn - 1, 0.0); be[p] = S->get("beta", p).data();
    S->get("house", p).assign((size_t)gn * gn, 0.0); ho[p] = S->get("house", p).data();
    S->get("tau", p).assign(gn, 0.0); tau[p] = S->get("tau", p).data();
    S->get("ralha_u", p).assign((size_t)gn * bl[p], 0.0); Up[p] = S->get("ralha_u", p).data();
  }
  ralha(P, gn, bl.data(), Fp.data(), al.data(), be.data(), ho.data(), tau.data(), Up.data());
  for (int p = 0; p < P; p++) {
    S->get("cs", p).assign(2 * (gn - 1), 0.0);
    lower_to_upper(gn, al[p], be[p], S->get("cs", p).data());
  }
  S->timers["time_bi_diag"] = secs(tb0, Clock::now());
  auto ts0 = Clock::now();
  const int cnt = bsvd_all(S, P, gn, al, be, top_k);
  if (cnt < 0) { delete S; return nullptr; }
  S->count = cnt;
  S->timers["time_bsvd"] = secs(ts0, Clock::now());
  // combine (client.cpp:998-1154): FinalU cnt x m, the column blocks m_part of every peer
  auto tc0 = Clock::now();
  vec finalU((size_t)cnt * m, 0.0);
  for (int c = 0; c < P; c++) {
    vec tmp((size_t)cnt * bl[c], 0.0);
    gemm_rm(false, false, cnt, bl[c], gn, S->get("bidiag_v", c).data(), gn, Up[c], bl[c], tmp.data(), bl[c]);
    vec part;
    if (mp[c] > gn) {
      part.assign((size_t)cnt * mp[c], 0.0);
      gemm_rm(false, false, cnt, mp[c], gn, tmp.data(), gn, ured[c].data(), mp[c], part.data(), mp[c]);
    } else {
      part = tmp;
    }
    for (long i = 0; i < cnt; i++)
      for (long j = 0; j < mp[c]; j++) finalU[i * m + mpos[c] + j] = part[i * mp[c] + j];
  }
  block_shift_mask(cnt, m, 32, finalU.data(), false, false, pub);
  for (int p = 0; p < P; p++) S->get("FinalU", p) = finalU;
  for (int p = 0; p < P; p++) tall_final_v(S, p, gn, cnt, npos[p], local_n[p], priv[p]);
  S->timers["time_combine_svd_remove_mask"] = secs(tc0, Clock::now());
  S->timers["total"] = secs(t0, Clock::now());
  S->arrays["seeds"].assign(P, vec());
  for (int p = 0; p < P; p++) S->arrays["seeds"][p] = vec{(double)pub, (double)priv[p]};
  return S;
}

// ----- whole run() of P peers, mode 1 (two-sided, bisection + twisted), opt 3 (Givens masks,
// overlapped Ralha). x[p]: the Client{p}.mat bytes, m x local_n[p] row-major in the
// short-and-wide branch (m <= global_n; m == global_n takes the tall-and-skinny combine of the
// right vectors, as the reference does), local_n[p] x m row-major in the tall-and-skinny branch
// (m > global_n, run_tall). seeds[p]: each peer's CLI seed. Returns a handle, or null when the
// bisection fails. -----
void* orc_run(int P, long m, const long* local_n, const double* const* x, const unsigned* seeds, int top_k) {
  long global_n = 0;
  for (int p = 0; p < P; p++) global_n += local_n[p];
  if (m < 3 || P < 1 || global_n < 3) return nullptr;
  RunState* S = new RunState();
  S->P = P; S->m = m;
  auto t0 = Clock::now();
  // seeds (client.cpp:196, 497-531)
  std::vector<unsigned> priv(P);
  unsigned pub = 0;
  for (int p = 0; p < P; p++) {
    srand(seeds[p]);
    if (p == 0) pub = (unsigned)rand();
    priv[p] = (unsigned)rand();
  }
  if (m > global_n) return run_tall(S, P, m, local_n, x, pub, priv, global_n, top_k, t0);
  const double thr = 5.0;
  std::vector<long> rl(P), rnp(P + 1, 0);
  std::vector<bool> did_lqr(P);
  for (int p = 0; p < P; p++) {
    did_lqr[p] = ((size_t)local_n[p] / (size_t)m) > thr;
    rl[p] = did_lqr[p] ? std::min(local_n[p], m) : local_n[p];
  }
  for (int p = 0; p < P; p++) rnp[p + 1] = rnp[p] + rl[p];
  long R = rnp[P];
  S->timers["time_init"] = secs(t0, Clock::now());

  // local QR (client.cpp:572-589) and masks (client.cpp:592-605)
  std::vector<vec> masked(P);
  double tq = 0, tm = 0;
  for (int p = 0; p < P; p++) {
    auto a = Clock::now();
    if (did_lqr[p] && m < local_n[p]) {
      masked[p].assign((size_t)m * m, 0.0);
      vec& qt = S->get("local_v", p);
      qt.assign((size_t)m * local_n[p], 0.0);
      local_qr(local_n[p], m, x[p], masked[p].data(), qt.data());
      S->get("local_rt", p) = masked[p];
    } else {
      masked[p].assign(x[p], x[p] + m * local_n[p]);
    }
    auto b = Clock::now();
    block_shift_mask(m, rl[p], 32, masked[p].data(), true, true, pub);
    block_shift_mask(m, rl[p], 32, masked[p].data(), true, false, priv[p]);
    S->get("masked_x", p) = masked[p];
    auto c = Clock::now();
    tq += secs(a, b);
    tm += secs(b, c);
  }
  S->timers["time_local_qr"] = tq;
  S->timers["time_apply_mask"] = tm;

  // re-partition (client.cpp:681-703)
  std::vector<long> mp(P, m / P), mpos(P + 1, 0);
  for (int i = 0; i < (int)(m % P); i++) mp[i] += 1;
  for (int i = 0; i < P; i++) mpos[i + 1] = mpos[i] + mp[i];
  std::vector<vec> rep(P);
  for (int c = 0; c < P; c++) {
    rep[c].assign((size_t)mp[c] * R, 0.0);
    for (int i = 0; i < P; i++)
      for (long r = 0; r < mp[c]; r++)
        for (long q = 0; q < rl[i]; q++)
          rep[c][r * R + rnp[i] + q] = masked[i][(mpos[c] + r) * rl[i] + q];
  }
  // global QR (client.cpp:706-798)
  auto tg0 = Clock::now();
  std::vector<vec> F(P);
  vec qg;
  if (R > m) {
    qg.assign((size_t)R * m, 0.0);
    std::vector<double*> pan(P), shr(P);
    for (int c = 0; c < P; c++) {
      F[c].assign((size_t)m * mp[c], 0.0);
      pan[c] = rep[c].data();
      shr[c] = F[c].data();
    }
    vec tau(m);
    global_qr(P, m, R, mp.data(), pan.data(), qg.data(), shr.data(), tau.data());
    for (int c = 0; c < P; c++) S->get("qg", c) = qg;
  } else {
    for (int c = 0; c < P; c++) {  // in-place transpose (client.cpp:797)
      F[c].assign((size_t)R * mp[c], 0.0);
      for (long r = 0; r < mp[c]; r++)
        for (long q = 0; q < R; q++) F[c][q * mp[c] + r] = rep[c][r * R + q];
    }
  }
  for (int c = 0; c < P; c++) S->get("ralha_in", c) = F[c];
  S->timers["time_global_qr"] = secs(tg0, Clock::now());

  // Ralha (client.cpp:800-910) and lower->upper (947-954)
  auto tb0 = Clock::now();
  std::vector<double*> Fp(P), al(P), be(P), ho(P), ta(P), Up(P);
  for (int p = 0; p < P; p++) {
    Fp[p] = F[p].data();
    S->get("alpha", p).assign(m, 0.0); al[p] = S->get("alpha", p).data();
    S->get("beta", p).assign(m - 1, 0.0); be[p] = S->get("beta", p).data();
    S->get("house", p).assign((size_t)m * m, 0.0); ho[p] = S->get("house", p).data();
    S->get("tau", p).assign(m, 0.0); ta[p] = S->get("tau", p).data();
    S->get("ralha_u", p).assign((size_t)m * mp[p], 0.0); Up[p] = S->get("ralha_u", p).data();
  }
  ralha(P, m, mp.data(), Fp.data(), al.data(), be.data(), ho.data(), ta.data(), Up.data());
  for (int p = 0; p < P; p++) {
    S->get("alpha_lower", p) = S->get("alpha", p);
    S->get("beta_lower", p) = S->get("beta", p);
    S->get("cs", p).assign(2 * (m - 1), 0.0);
    lower_to_upper(m, al[p], be[p], S->get("cs", p).data());
  }
  S->timers["time_bi_diag"] = secs(tb0, Clock::now());

  // bidiagonal SVD + sign rule (client.cpp:963-996), every peer on its own copy
  auto ts0 = Clock::now();
  std::vector<double*> alv(al.begin(), al.end()), bev(be.begin(), be.end());
  const int cnt = bsvd_all(S, P, m, alv, bev, top_k);
  if (cnt < 0) { delete S; return nullptr; }
  S->count = cnt;
  S->timers["time_bsvd"] = secs(ts0, Clock::now());

  // combine + unmask (client.cpp:998-1199)
  auto tc0 = Clock::now();
  vec finalU((size_t)m * cnt, 0.0);
  for (int p = 0; p < P; p++) {
    vec part((size_t)cnt * mp[p], 0.0);
    gemm_rm(false, false, cnt, mp[p], m, S->get("bidiag_v", p).data(), m, Up[p], mp[p], part.data(), mp[p]);
    for (long i = 0; i < cnt; i++)
      for (long j = 0; j < mp[p]; j++) finalU[(mpos[p] + j) * cnt + i] = part[i * mp[p] + j];
  }
  block_shift_mask(m, cnt, 32, finalU.data(), false, true, pub);
  for (int p = 0; p < P; p++) S->get("FinalU", p) = finalU;
  if (m == global_n) {  // m >= global_n: the tall-and-skinny combine of the right vectors
    for (int p = 0; p < P; p++) tall_final_v(S, p, m, cnt, rnp[p], local_n[p], priv[p]);
    S->timers["time_combine_svd_remove_mask"] = secs(tc0, Clock::now());
    S->timers["total"] = secs(t0, Clock::now());
    S->arrays["seeds"].assign(P, vec());
    for (int p = 0; p < P; p++) S->arrays["seeds"][p] = vec{(double)pub, (double)priv[p]};
    return S;
  }
  for (int p = 0; p < P; p++) {
    vec bu = S->get("bidiag_u", p);
    const vec& cs = S->get("cs", p);
    for (long i = m - 2; i >= 0; i--)
      drot(cnt, bu.data() + i * cnt, 1, bu.data() + (i + 1) * cnt, 1, cs[2 * i], -cs[2 * i + 1]);
    ormqr_house_ln(m, cnt, ho[p], ta[p], bu.data());
    vec tmp((size_t)cnt * rl[p], 0.0);
    if (R > m) {
      // tmp = bu^T (cnt x m) . Qg[rnp:rnp+rl, :]^T  (client.cpp:1179)
#pragma omp parallel for schedule(static)
      for (long a = 0; a < cnt; a++)
        for (long l = 0; l < rl[p]; l++) {
          double s = 0.0;
          for (long j = 0; j < m; j++) s = s + bu[j * cnt + a] * qg[(rnp[p] + l) + j * R];
          tmp[a * rl[p] + l] = s;
        }
    }
    block_shift_mask(cnt, rl[p], 32, tmp.data(), false, false, priv[p]);
    vec& fv = S->get("FinalVT", p);
    if (did_lqr[p] && m < local_n[p]) {
      fv.assign((size_t)cnt * local_n[p], 0.0);
      gemm_rm(false, false, cnt, local_n[p], rl[p], tmp.data(), rl[p], S->get("local_v", p).data(), local_n[p],
              fv.data(), local_n[p]);
    } else {
      fv = tmp;
    }
  }
  S->timers["time_combine_svd_remove_mask"] = secs(tc0, Clock::now());
  S->timers["total"] = secs(t0, Clock::now());
  S->arrays["seeds"].assign(P, vec());
  for (int p = 0; p < P; p++) S->arrays["seeds"][p] = vec{(double)pub, (double)priv[p]};
  return S;
}

int orc_run_count(void* h) { return h ? ((RunState*)h)->count : -1; }

// Copies array `name` of peer p into out (capacity cap). Returns its length, or -1.
long orc_run_get(void* h, const char* name, int p, double* out, long cap) {
  if (!h) return -1;
  RunState* S = (RunState*)h;
  auto it = S->arrays.find(name);
  if (it == S->arrays.end() || p < 0 || p >= (int)it->second.size()) return -1;
  const vec& a = it->second[p];
  if (out) std::copy(a.begin(), a.begin() + std::min<long>(cap, (long)a.size()), out);
  return (long)a.size();
}

double orc_run_timer(void* h, const char* name) {
  if (!h) return -1;
  RunState* S = (RunState*)h;
  auto it = S->timers.find(name);
  return it == S->timers.end() ? -1.0 : it->second;
}

void orc_run_free(void* h) { delete (RunState*)h; }

}  // extern "C"
