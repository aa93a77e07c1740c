// Host-pointer entry points with the semantics of the reference's bbtsvd.hpp functions
// (include/bbtsvd.hpp:22-47 of the reference): arguments are host arrays, results are written
// back to host arrays before returning. Each call copies its inputs to the GPU, runs the same
// kernels as the device-pointer API on the default stream, and copies the outputs back.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "excalibur_bbtsvd.h"
#include "exc_common.hpp"

namespace {

// dstqds (bbtsvd.cpp:73-81): L D L^T - tau I = L+ D+ L+^T with the auxiliary s; one thread.
__global__ void k_dstqds1(int n, const double* __restrict__ l, const double* __restrict__ d, double tau,
                          double* __restrict__ lp, double* __restrict__ dp, double* __restrict__ s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double sc = -tau;
  s[0] = sc;
  for (int i = 0; i + 1 < n; i++) {
    const double dpi = d[i] + sc;
    const double lpi = d[i] * l[i] / dpi;
    dp[i] = dpi;
    lp[i] = lpi;
    sc = lpi * l[i] * sc - tau;
    s[i + 1] = sc;
  }
  dp[n - 1] = d[n - 1] + sc;
}

// dpqds (bbtsvd.cpp:83-91): L D L^T - tau I = U+ R+ U+^T with the auxiliary p; one thread.
__global__ void k_dpqds1(int n, const double* __restrict__ l, const double* __restrict__ d, double tau,
                         double* __restrict__ up, double* __restrict__ rp, double* __restrict__ p) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double pc = d[n - 1] - tau;
  p[n - 1] = pc;
  for (int i = n - 2; i >= 0; i--) {
    const double r = d[i] * (l[i] * l[i]) + pc;
    rp[i + 1] = r;
    up[i] = l[i] * d[i] / r;
    pc = pc * d[i] / r - tau;
    p[i] = pc;
  }
  rp[0] = pc;
}

// inverse_iteration (bbtsvd.cpp:65-71) on a strided vector: z_k = 1, recurrences outwards,
// scaled by 1 / sqrt(sum of squares in index order).
__device__ void inverse_iteration_dev(int n, int k, const double* l, const double* u, double* v, long ldv) {
  v[(long)k * ldv] = 1.0;
  for (int i = k - 1; i >= 0; i--) v[(long)i * ldv] = -l[i] * v[(long)(i + 1) * ldv];
  for (int i = k + 1; i < n; i++) v[(long)i * ldv] = -u[i - 1] * v[(long)(i - 1) * ldv];
  double ss = 0.0;
  for (int i = 0; i < n; i++) ss = ss + v[(long)i * ldv] * v[(long)i * ldv];
  const double inv = 1.0 / sqrt(ss);
  for (int i = 0; i < n; i++) v[(long)i * ldv] = inv * v[(long)i * ldv];
}

__global__ void k_inverse_iteration1(int n, int k, const double* l, const double* u, double* v, long ldv) {
  if (threadIdx.x == 0 && blockIdx.x == 0) inverse_iteration_dev(n, k, l, u, v, ldv);
}

// First index of the smallest |g_i| (cblas_idamin).
__device__ int first_min_abs(int n, const double* g) {
  int best = 0;
  double bv = fabs(g[0]);
  for (int i = 1; i < n; i++)
    if (fabs(g[i]) < bv) {
      bv = fabs(g[i]);
      best = i;
    }
  return best;
}

// one_pair_singular_vector_twisted_factorization (bbtsvd.cpp:93-135) for one omega with the
// caller's L, D: v is contiguous, u has stride ldu. Scratch w holds 7n + 3 doubles.
__global__ void k_one_pair1(int n, long ldu, double omega, const double* __restrict__ l, const double* __restrict__ d,
                            const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ v,
                            double* __restrict__ u, double* __restrict__ w) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* lp = w;            // n - 1 (+1 spare)
  double* dp = lp + n;       // n
  double* s = dp + n;        // n + 1
  double* up = s + n + 1;    // n - 1 (+1 spare)
  double* rp = up + n;       // n + 1
  double* p = rp + n + 1;    // n + 1
  double* g = p + n + 1;     // n
  const double w2 = omega * omega;
  // L D L^T - w2 I, both ways
  double sc = -w2;
  s[0] = sc;
  for (int i = 0; i + 1 < n; i++) {
    dp[i] = d[i] + sc;
    lp[i] = d[i] * l[i] / dp[i];
    sc = lp[i] * l[i] * sc - w2;
    s[i + 1] = sc;
  }
  dp[n - 1] = d[n - 1] + sc;
  double pc = d[n - 1] - w2;
  p[n - 1] = pc;
  for (int i = n - 2; i >= 0; i--) {
    rp[i + 1] = d[i] * (l[i] * l[i]) + pc;
    up[i] = l[i] * d[i] / rp[i + 1];
    pc = pc * d[i] / rp[i + 1] - w2;
    p[i] = pc;
  }
  rp[0] = pc;
  for (int i = 0; i < n; i++) g[i] = s[i] + p[i] + w2;
  inverse_iteration_dev(n, first_min_abs(n, g), lp, up, v, 1);
  // coupling to B B^T (Grosser-Lang): shifted D+, gamma and R+ for the other Gram matrix
  s[n] = -w2;
  p[n] = 1.0;
  rp[n] = 1.0;
  for (int i = 0; i < n; i++) {
    dp[i] = s[i + 1] / s[i] * dp[i];
    g[i] = -w2 * g[i] * rp[i + 1] / (s[i] * p[i + 1]);
  }
  const int k1 = first_min_abs(n, g);
  for (int i = 0; i < n; i++) rp[i] = p[i] / p[i + 1] * rp[i + 1];
  for (int i = 0; i + 1 < n; i++) {
    lp[i] = a[i + 1] / dp[i] * b[i];
    up[i] = a[i + 1] / rp[i + 1] * b[i];
  }
  inverse_iteration_dev(n, k1, lp, up, u, ldu);
}

// Device buffers of one host-pointer call, freed when the call returns.
struct Scope {
  std::vector<void*> ptrs;
  ~Scope() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, (n ? n : 1) * sizeof(T)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return (T*)p;
  }
  template <typename T>
  T* upload(const T* h, size_t n) {
    T* d = alloc<T>(n);
    if (d && n && hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return d;
  }
};

#define EXC_DEV(ptr)                                                                  \
  do {                                                                                \
    if (!(ptr)) {                                                                     \
      exc_detail::set_last_message("device allocation or copy failed", __FILE__, __LINE__); \
      return EXC_ERR_HIP;                                                             \
    }                                                                                 \
  } while (0)

int download(void* h, const void* d, size_t bytes) {
  if (bytes == 0) return EXC_OK;
  EXC_CHECK_HIP(hipDeviceSynchronize());
  EXC_CHECK_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  return EXC_OK;
}

}  // namespace

extern "C" {

int exc_bb_negative_count(int n, const double* x_diagonal, const double* x_upper_diagonal, double mu, int* count) {
  EXC_REQUIRE(n >= 1 && x_diagonal && (n == 1 || x_upper_diagonal) && count, EXC_ERR_ARG,
              "exc_bb_negative_count: bad arguments");
  Scope S;
  double* a = S.upload(x_diagonal, n);
  double* b = S.upload(n > 1 ? x_upper_diagonal : x_diagonal, n > 1 ? n - 1 : 1);
  double* dm = S.upload(&mu, 1);
  int* dc = S.alloc<int>(1);
  EXC_DEV(a && b && dm && dc);
  int rc = exc_sturm_count(n, a, b, dm, 1, dc, nullptr);
  if (rc) return rc;
  return download(count, dc, sizeof(int));
}

int exc_bb_bisection_singular_values(int n, const double* x_diagonal, const double* x_upper_diagonal, double a,
                                     double b, double tol, int top_k, double* sigma, int* sigma_count) {
  EXC_REQUIRE(n >= 1 && x_diagonal && (n == 1 || x_upper_diagonal) && sigma_count && (top_k <= 0 || sigma),
              EXC_ERR_ARG, "exc_bb_bisection_singular_values: bad arguments");
  *sigma_count = 0;
  if (top_k <= 0) return EXC_OK;
  Scope S;
  double* da = S.upload(x_diagonal, n);
  double* db = S.upload(n > 1 ? x_upper_diagonal : x_diagonal, n > 1 ? n - 1 : 1);
  const double ends[2] = {a, b};
  double* dends = S.upload(ends, 2);
  int* dc = S.alloc<int>(2);
  EXC_DEV(da && db && dends && dc);
  int rc = exc_sturm_count(n, da, db, dends, 2, dc, nullptr);
  if (rc) return rc;
  int nc[2];
  rc = download(nc, dc, sizeof nc);
  if (rc) return rc;
  if (nc[0] > nc[1]) return EXC_OK;  // empty interval: no singular value recorded
  const int k = top_k < n ? top_k : n;
  const size_t wb = exc_bisection_workspace(n, k);
  void* dw = S.alloc<char>(wb);
  double* ds = S.alloc<double>(k);
  EXC_DEV(dw && ds);
  rc = exc_bisection(n, da, db, a, b, tol, k, ds, dc, nullptr, nullptr, dw, wb, nullptr);
  if (rc) return rc;
  rc = download(sigma_count, dc, sizeof(int));
  if (rc) return rc;
  return download(sigma, ds, (size_t)*sigma_count * sizeof(double));
}

int exc_bb_ldl_decomposition(int n, const double* x_diagonal, const double* x_upper_diagonal, double* l, double* d) {
  EXC_REQUIRE(n >= 1 && x_diagonal && d && (n == 1 || (x_upper_diagonal && l)), EXC_ERR_ARG,
              "exc_bb_ldl_decomposition: bad arguments");
  Scope S;
  double* da = S.upload(x_diagonal, n);
  double* db = S.upload(n > 1 ? x_upper_diagonal : x_diagonal, n > 1 ? n - 1 : 1);
  double* dl = S.alloc<double>(n);
  double* dd = S.alloc<double>(n);
  EXC_DEV(da && db && dl && dd);
  int rc = exc_ldl(n, da, db, dl, dd, nullptr);
  if (rc) return rc;
  if (n > 1 && (rc = download(l, dl, (n - 1) * sizeof(double)))) return rc;
  return download(d, dd, n * sizeof(double));
}

int exc_bb_inverse_iteration(int n, int k, const double* l, const double* u, double* v, int ldv) {
  EXC_REQUIRE(n >= 1 && k >= 0 && k < n && ldv >= 1 && v && (n == 1 || (l && u)), EXC_ERR_ARG,
              "exc_bb_inverse_iteration: bad arguments");
  Scope S;
  const double zero = 0.0;
  double* dl = n > 1 ? S.upload(l, n - 1) : S.upload(&zero, 1);
  double* du = n > 1 ? S.upload(u, n - 1) : S.upload(&zero, 1);
  const size_t span = (size_t)(n - 1) * ldv + 1;
  double* dv = S.upload(v, span);  // entries between the strided ones are kept
  EXC_DEV(dl && du && dv);
  hipLaunchKernelGGL(k_inverse_iteration1, dim3(1), dim3(64), 0, nullptr, n, k, dl, du, dv, (long)ldv);
  EXC_CHECK_HIP(hipGetLastError());
  return download(v, dv, span * sizeof(double));
}

int exc_bb_dstqds(int n, const double* l, const double* d, double tau, double* l_plus, double* d_plus, double* s) {
  EXC_REQUIRE(n >= 1 && d && d_plus && s && (n == 1 || (l && l_plus)), EXC_ERR_ARG, "exc_bb_dstqds: bad arguments");
  Scope S;
  const double zero = 0.0;
  double* dl = n > 1 ? S.upload(l, n - 1) : S.upload(&zero, 1);
  double* dd = S.upload(d, n);
  double* dlp = S.alloc<double>(n);
  double* ddp = S.alloc<double>(n);
  double* ds = S.alloc<double>(n);
  EXC_DEV(dl && dd && dlp && ddp && ds);
  hipLaunchKernelGGL(k_dstqds1, dim3(1), dim3(64), 0, nullptr, n, dl, dd, tau, dlp, ddp, ds);
  EXC_CHECK_HIP(hipGetLastError());
  int rc;
  if (n > 1 && (rc = download(l_plus, dlp, (n - 1) * sizeof(double)))) return rc;
  if ((rc = download(d_plus, ddp, n * sizeof(double)))) return rc;
  return download(s, ds, n * sizeof(double));
}

int exc_bb_dpqds(int n, const double* l, const double* d, double tau, double* u_plus, double* r_plus, double* p) {
  EXC_REQUIRE(n >= 1 && d && r_plus && p && (n == 1 || (l && u_plus)), EXC_ERR_ARG, "exc_bb_dpqds: bad arguments");
  Scope S;
  const double zero = 0.0;
  double* dl = n > 1 ? S.upload(l, n - 1) : S.upload(&zero, 1);
  double* dd = S.upload(d, n);
  double* dup = S.alloc<double>(n);
  double* drp = S.alloc<double>(n);
  double* dp = S.alloc<double>(n);
  EXC_DEV(dl && dd && dup && drp && dp);
  hipLaunchKernelGGL(k_dpqds1, dim3(1), dim3(64), 0, nullptr, n, dl, dd, tau, dup, drp, dp);
  EXC_CHECK_HIP(hipGetLastError());
  int rc;
  if (n > 1 && (rc = download(u_plus, dup, (n - 1) * sizeof(double)))) return rc;
  if ((rc = download(r_plus, drp, n * sizeof(double)))) return rc;
  return download(p, dp, n * sizeof(double));
}

int exc_bb_one_pair_twisted(int n, int omega_count, double omega, const double* l, const double* d,
                            const double* x_diagonal, const double* x_upper_diagonal, double* v, double* u) {
  EXC_REQUIRE(n >= 1 && omega_count >= 1 && d && x_diagonal && v && u &&
                  (n == 1 || (l && x_upper_diagonal)),
              EXC_ERR_ARG, "exc_bb_one_pair_twisted: bad arguments");
  Scope S;
  const double zero = 0.0;
  double* dl = n > 1 ? S.upload(l, n - 1) : S.upload(&zero, 1);
  double* dd = S.upload(d, n);
  double* da = S.upload(x_diagonal, n);
  double* db = n > 1 ? S.upload(x_upper_diagonal, n - 1) : S.upload(&zero, 1);
  const size_t span = (size_t)(n - 1) * omega_count + 1;
  double* dv = S.alloc<double>(n);
  double* du = S.upload(u, span);  // entries between the strided ones are kept
  double* dw = S.alloc<double>(7 * (size_t)n + 3);
  EXC_DEV(dl && dd && da && db && dv && du && dw);
  hipLaunchKernelGGL(k_one_pair1, dim3(1), dim3(64), 0, nullptr, n, (long)omega_count, omega, dl, dd, da, db, dv, du,
                     dw);
  EXC_CHECK_HIP(hipGetLastError());
  int rc = download(v, dv, n * sizeof(double));
  if (rc) return rc;
  return download(u, du, span * sizeof(double));
}

int exc_bb_two_side_twisted(int n, int omega_count, const double* omega, const double* x_diagonal,
                            const double* x_upper_diagonal, double* u, double* v) {
  EXC_REQUIRE(n >= 1 && omega_count >= 0 && x_diagonal && (n == 1 || x_upper_diagonal) &&
                  (omega_count == 0 || (omega && u && v)),
              EXC_ERR_ARG, "exc_bb_two_side_twisted: bad arguments");
  if (omega_count == 0) return EXC_OK;
  Scope S;
  const double zero = 0.0;
  double* dw_ = S.upload(omega, omega_count);
  double* da = S.upload(x_diagonal, n);
  double* db = n > 1 ? S.upload(x_upper_diagonal, n - 1) : S.upload(&zero, 1);
  double* du = S.alloc<double>((size_t)n * omega_count);
  double* dv = S.alloc<double>((size_t)n * omega_count);
  const size_t wb = exc_twisted_workspace(n, omega_count);
  void* dwork = S.alloc<char>(wb);
  EXC_DEV(dw_ && da && db && du && dv && dwork);
  int rc = exc_twisted_vectors(n, omega_count, dw_, da, db, du, dv, nullptr, dwork, wb, nullptr);
  if (rc) return rc;
  if ((rc = download(u, du, (size_t)n * omega_count * sizeof(double)))) return rc;
  return download(v, dv, (size_t)n * omega_count * sizeof(double));
}

int exc_bb_bidiagonal_svd(int k, const double* alpha, const double* beta, double* bidiagonal_u,
                          double* bidiagonal_v, double* sigma, int top_k) {
  EXC_REQUIRE(k >= 1 && top_k >= 0 && top_k <= k && alpha && (k == 1 || beta) && sigma &&
                  (top_k == 0 || (bidiagonal_u && bidiagonal_v)),
              EXC_ERR_ARG, "exc_bb_bidiagonal_svd: bad arguments");
  Scope S;
  const double zero = 0.0;
  double* da = S.upload(alpha, k);
  double* db = k > 1 ? S.upload(beta, k - 1) : S.upload(&zero, 1);
  double* du = S.alloc<double>((size_t)k * (top_k ? top_k : 1));
  double* dv = S.alloc<double>((size_t)k * (top_k ? top_k : 1));
  double* ds = S.alloc<double>(top_k ? top_k : 1);
  const size_t wb = exc_bidiag_svd_workspace(k, top_k);
  void* dwork = S.alloc<char>(wb);
  EXC_DEV(da && db && du && dv && ds && dwork);
  int rc = exc_bidiag_svd(k, da, db, du, dv, ds, top_k, nullptr, dwork, wb, nullptr);
  if (rc) return rc;
  if ((rc = download(sigma, ds, (size_t)top_k * sizeof(double)))) return rc;
  if ((rc = download(bidiagonal_u, du, (size_t)k * top_k * sizeof(double)))) return rc;
  return download(bidiagonal_v, dv, (size_t)k * top_k * sizeof(double));
}

int exc_bb_givens_mask(size_t m, size_t n, size_t mask_times, double* x, int mask, int is_left, unsigned int seed) {
  EXC_REQUIRE(x || m == 0 || n == 0, EXC_ERR_ARG, "exc_bb_givens_mask: null matrix");
  const size_t K = is_left ? m : n;
  srand(seed);
  if (K <= 1 || m == 0 || n == 0) return EXC_OK;  // nothing to rotate (the reference prints and returns)
  const long half = exc_givens_half((long)K);
  std::vector<int32_t> pairs((size_t)mask_times * half * 2);
  std::vector<double> cs((size_t)mask_times * half * 2);
  int rc = exc_givens_plan((long)K, (long)mask_times, mask, seed, pairs.data(), cs.data());
  if (rc) return rc;
  Scope S;
  double* dx = S.upload(x, m * n);
  int32_t* dp = S.upload(pairs.data(), pairs.size());
  double* dcs = S.upload(cs.data(), cs.size());
  EXC_DEV(dx && dp && dcs);
  rc = exc_givens_apply(dx, (long)m, (long)n, dp, dcs, (long)mask_times, mask, is_left, nullptr);
  if (rc) return rc;
  return download(x, dx, m * n * sizeof(double));
}

int exc_bb_block_shift_mask(size_t m, size_t n, size_t mask_times, double* x, int mask, int is_left,
                            unsigned int seed) {
  const size_t K = is_left ? m : n;
  if (K <= 10000) return exc_bb_givens_mask(m, n, mask_times, x, mask, is_left, seed);
  // the virtual-matrix mask (bbtsvd.cpp:548-668) on the device
  EXC_REQUIRE(x, EXC_ERR_ARG, "exc_bb_block_shift_mask: null matrix");
  srand(seed);
  Scope S;
  double* dx = S.upload(x, m * n);
  const size_t wb = exc_block_shift_mask_workspace((long)m, (long)n, (long)mask_times, is_left);
  void* dw = S.alloc<char>(wb);
  EXC_DEV(dx && dw);
  int rc = exc_block_shift_mask(dx, (long)m, (long)n, (long)mask_times, mask, is_left, seed, dw, wb, nullptr);
  if (rc) return rc;
  return download(x, dx, m * n * sizeof(double));
}

int exc_bb_find_best_convert(size_t k, size_t* out_m, size_t* out_n, size_t* out_remain) {
  // bbtsvd.cpp:166-188 with OPT_THRESHOLD = 10000 and RATIO_THRESHOLD = 2 (bbtsvd.cpp:3-4): the
  // most square factorisation k - remain = a * b, 1 < b < a < 10000, a / b < 2, with the
  // smallest remain.
  EXC_REQUIRE(out_m && out_n && out_remain && k <= (size_t)10000 * 10000, EXC_ERR_ARG,
              "exc_bb_find_best_convert: bad arguments");
  const size_t T = 10000;
  for (size_t remain = 0; remain <= k; remain++) {
    const size_t kk = k - remain;
    size_t bm = kk, bn = 1;
    for (size_t i = T; i > 1; i--) {
      if (kk % i) continue;
      const size_t q = kk / i;
      if (i > q && q < T && q > 1 && (i - q) < (bm - bn) && ((double)i / (double)q) < 2.0) {
        bm = i;
        bn = q;
      }
    }
    if (bm < T) {
      *out_m = bm;
      *out_n = bn;
      *out_remain = remain;
      return EXC_OK;
    }
  }
  EXC_REQUIRE(false, EXC_ERR_ARG, "exc_bb_find_best_convert: no factorisation");
}

}  // extern "C"
