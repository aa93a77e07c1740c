// Bidiagonal SVD on gfx950 in FP64: Sturm counts, tree-consistent multisection bisection and
// the twisted-factorization singular vectors of the reference's bbtsvd.cpp:7-147 and 932-968.
//
// Every recurrence below keeps the reference's operation order with separately rounded
// multiplies and adds (this translation unit is built with -ffp-contract=off and the pragma
// below), correctly rounded division and square root, so counts, twist indices, singular values
// and vectors are bit-identical to a sequential evaluation of the same formulas.
#include "exc_common.hpp"
#include "sturm.hpp"

#include <mutex>

#pragma clang fp contract(off)

namespace {

struct BisectHdr {
  double lo, hi, root_mid;
  int nclo, nchi;
  int root_leaf;
  int status;  // 1: more targets needed than were evaluated
  int count;
  int flags;  // bit 1: a non-monotone Sturm count was met on some target's path
};


__global__ __launch_bounds__(256) void k_sturm_count(int n, const double* __restrict__ a,
                                                     const double* __restrict__ b, const double* __restrict__ mu,
                                                     int nmu, int* __restrict__ out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nmu) out[j] = sturm_count(n, a, b, mu[j]);
}

// ||alpha||^2 + ||beta||^2 with the sequential sum of squares of the reference's nrm2 restated
// (sqrt, then squared, as bbtsvd.cpp:936 writes pow(dnrm2(.), 2)). One workgroup stages chunks
// in LDS; thread 0 accumulates in index order.
__global__ __launch_bounds__(256) void k_frob_bound(int n, const double* __restrict__ a,
                                                    const double* __restrict__ b, BisectHdr* hdr) {
  __shared__ double buf[2048];
  double sums[2] = {0.0, 0.0};
  for (int which = 0; which < 2; which++) {
    const double* x = which == 0 ? a : b;
    const int len = which == 0 ? n : n - 1;
    double s = 0.0;
    for (int base = 0; base < len; base += 2048) {
      const int cnt = min(2048, len - base);
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) buf[i] = x[base + i];
      __syncthreads();
      if (threadIdx.x == 0)
        for (int i = 0; i < cnt; i++) {
          const double v = buf[i];
          s = s + v * v;
        }
      __syncthreads();
    }
    sums[which] = s;
  }
  if (threadIdx.x == 0) {
    const double na = sqrt(sums[0]), nb = sqrt(sums[1]);
    hdr->lo = 0.0;
    hdr->hi = na * na + nb * nb;
  }
}

__global__ void k_set_interval(BisectHdr* hdr, double lo, double hi) {
  if (threadIdx.x == 0) { hdr->lo = lo; hdr->hi = hi; }
}

// Counts at the root's ends and whether the root itself is a leaf (bbtsvd.cpp:21-36).
// ldl_decomposition (bbtsvd.cpp:54-63): d = diagonal, l = subdiagonal of B^T B = L D L^T.
__device__ __forceinline__ void ldl_seq(int n, const double* __restrict__ a, const double* __restrict__ b,
                                        double* __restrict__ l, double* __restrict__ d) {
  double dp = a[0] * a[0];
  d[0] = dp;
  for (int i = 1; i < n; i++) {
    const double ab = a[i - 1] * b[i - 1];
    l[i - 1] = ab / dp;
    dp = a[i] * a[i] + b[i - 1] * b[i - 1] - (ab * ab) / dp;
    d[i] = dp;
  }
}

// The root's Sturm counts (lanes 0 and 1 of wavefront 0, side by side: one call with the lane's
// end point) and, when l and d are given, the LDL factorization the twisted factorization needs
// later, in wavefront 1 at the same time. 128 threads.
__global__ void k_bisect_init(int n, const double* __restrict__ a, const double* __restrict__ b, double tol,
                              BisectHdr* hdr, double* __restrict__ l, double* __restrict__ d) {
  const double lo = hdr->lo, hi = hdr->hi;
  if (threadIdx.x < 2) {
    const int c = sturm_count(n, a, b, threadIdx.x == 0 ? lo : hi);
    if (threadIdx.x == 0) hdr->nclo = c;
    else hdr->nchi = c;
  }
  if (threadIdx.x == 64 && l) ldl_seq(n, a, b, l, d);
  if (threadIdx.x == 2) {
    const double mid = (lo + hi) / 2;
    hdr->root_mid = mid;
    hdr->root_leaf = ((hi - lo) <= tol || mid == lo || mid == hi) ? 1 : 0;
    hdr->status = 0;
    hdr->flags = 0;
  }
}

// One group of W lanes per singular value (target). Target q (descending order) is the
// ascending index t = nchi - 1 - q. Per round the W-1 lanes evaluate the Sturm counts of the
// W-1 interior nodes of the depth-log2(W) subtree below the current node (node intervals
// replayed from the current one with the same (lo+hi)/2 midpoints), then the group descends up
// to log2(W) levels: left when t < count(mid) (the reference pushes that child, because it
// holds singular values nc_low..nc_mid-1), right otherwise, stopping at the first leaf
// ((hi-lo) <= tol, or mid equal to an end). These are exactly the nodes the reference's
// depth-first search visits on the way to t, so the leaf is the same node.
template <int W>
__global__ __launch_bounds__(256) void k_bisect_targets(int n, const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        const BisectHdr* __restrict__ hdr, double tol, int cap,
                                                        double* __restrict__ llo, double* __restrict__ lhi,
                                                        double* __restrict__ lmid, int* __restrict__ flags) {
  constexpr int D = (W == 1) ? 1 : (W == 2 ? 1 : W == 4 ? 2 : W == 8 ? 3 : W == 16 ? 4 : W == 32 ? 5 : 6);
  const int lane = threadIdx.x;
  const int gid = (blockIdx.x * blockDim.x + lane) / W;
  const int j = lane % W;
  const int nclo = hdr->nclo, nchi = hdr->nchi;
  int T = (nclo <= nchi) ? min(cap, nchi - nclo) : 0;
  if (hdr->root_leaf) T = 0;
  if (gid >= T) return;  // uniform within the group
  const int t = nchi - 1 - gid;
  double lo = hdr->lo, hi = hdr->hi;
  int clo = nclo, chi = nchi;
  bool monotone = true;
  const int h = j + 1;
  const int lvl = 31 - __clz(h);
  const int wlane = lane & (kWave - 1);
  const int gbase = wlane - j;
  while (true) {
    const double mid = (lo + hi) / 2;
    if ((hi - lo) <= tol || mid == lo || mid == hi) break;
    int c = 0;
    if (W == 1) {
      c = sturm_count(n, a, b, mid);
    } else if (j < W - 1) {
      double l = lo, r = hi;
      for (int bit = lvl - 1; bit >= 0; --bit) {
        const double mm = (l + r) / 2;
        if ((h >> bit) & 1) l = mm;
        else r = mm;
      }
      c = sturm_count(n, a, b, (l + r) / 2);
    }
    int hc = 1;
    for (int lev = 0; lev < D; ++lev) {
      const double md = (lo + hi) / 2;
      if (lev > 0 && ((hi - lo) <= tol || md == lo || md == hi)) break;
      const int cnt = (W == 1) ? c : __shfl(c, gbase + hc - 1, kWave);
      monotone = monotone && clo <= cnt && cnt <= chi;
      if (t < cnt) { hi = md; chi = cnt; hc = 2 * hc; }
      else { lo = md; clo = cnt; hc = 2 * hc + 1; }
    }
  }
  if (j == 0) {
    llo[gid] = lo;
    lhi[gid] = hi;
    lmid[gid] = (lo + hi) / 2;
    if (!monotone) atomicOr(&flags[0], 2);
  }
}

// Leaves of consecutive targets that are the same node are one output of the reference's
// search (a leaf holding several singular values records one value, bbtsvd.cpp:34-37). One
// workgroup: flags, block scan, scatter of the first top_k distinct leaves.
__global__ __launch_bounds__(1024) void k_bisect_compact(BisectHdr* hdr, int top_k, int cap,
                                                         const double* __restrict__ llo,
                                                         const double* __restrict__ lhi,
                                                         const double* __restrict__ lmid, double* __restrict__ sigma,
                                                         int* __restrict__ count, double* __restrict__ olo,
                                                         double* __restrict__ ohi) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int nclo = hdr->nclo, nchi = hdr->nchi;
  if (!(nclo <= nchi && top_k > 0)) {
    if (tid == 0) { *count = 0; hdr->count = 0; hdr->status = 0; }
    return;
  }
  if (hdr->root_leaf) {
    if (tid == 0) {
      sigma[0] = hdr->root_mid;
      if (olo) olo[0] = hdr->lo;
      if (ohi) ohi[0] = hdr->hi;
      *count = 1; hdr->count = 1; hdr->status = 0;
    }
    return;
  }
  const int T = min(cap, nchi - nclo);
  const int per = (T + 1023) / 1024;
  const int s0 = min(T, tid * per), s1 = min(T, s0 + per);
  auto flag = [&](int q) { return q == 0 || llo[q] != llo[q - 1] || lhi[q] != lhi[q - 1]; };
  int local = 0;
  for (int q = s0; q < s1; q++) local += flag(q) ? 1 : 0;
  part[tid] = local;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
    int v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int pos = part[tid] - local;
  for (int q = s0; q < s1; q++)
    if (flag(q)) {
      if (pos < top_k) {
        sigma[pos] = lmid[q];
        if (olo) olo[pos] = llo[q];
        if (ohi) ohi[pos] = lhi[q];
      }
      pos++;
    }
  if (tid == 1023) {
    const int total = part[1023];
    const int c = min(total, top_k);
    *count = c;
    hdr->count = c;
    hdr->status = (total < top_k && T < nchi - nclo) ? 1 : 0;
  }
}

// ldl_decomposition (bbtsvd.cpp:54-63), one thread.
__global__ void k_ldl(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ l,
                      double* __restrict__ d) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ldl_seq(n, a, b, l, d);
}

// One lane per singular value: dstqds, dpqds, twist, inverse iteration for v, coupling, twist
// and inverse iteration for u (bbtsvd.cpp:65-135). The per-value sequences s, p, d+, r+ live in
// HBM interleaved by value (element i of value j at i*k + j), so a wavefront's accesses are
// contiguous. l+ and u+ are recomputed from them with the reference's expressions instead of
// being stored. v is written interleaved (vs) and transposed afterwards; u is written in place
// (column j of the n x k row-major u is interleaved by construction).
__global__ __launch_bounds__(256) void k_twisted(int n, int k, const double* __restrict__ a,
                                                 const double* __restrict__ b, const double* __restrict__ l,
                                                 const double* __restrict__ d, const double* __restrict__ omega,
                                                 double* __restrict__ u, double* __restrict__ vs,
                                                 double* __restrict__ S, double* __restrict__ Pp,
                                                 double* __restrict__ DP, double* __restrict__ RP,
                                                 int* __restrict__ twists) {
  // four wavefronts per 64 values: wavefront 0 runs the recurrences; then wavefronts 0 and 1
  // find the two twist indices, the four inverse-iteration sweeps (v and u, each backward and
  // forward from its twist) run one per wavefront, and wavefronts 0 and 1 normalize v and u
  const int wave = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const bool active = j < k;
  const size_t K = (size_t)k;
  auto at = [&](int i) -> size_t { return (size_t)i * K + j; };
  const double w = active ? omega[j] : 0.0;  // lanes past k read nothing (they only join the barriers)
  const double w2 = w * w;

  // dstqds (bbtsvd.cpp:73-81, forward) and dpqds (bbtsvd.cpp:83-91, backward) are independent
  // recurrences: one loop runs both, so the two chains of dependent divisions overlap. Each
  // chain evaluates the reference's expressions in the reference's order.
  if (wave == 0 && active) {
  double sc = -w2;
  S[at(0)] = sc;
  double pc = d[n - 1] - w2;
  Pp[at(n - 1)] = pc;
  #pragma unroll 2
  for (int t = 0; t < n - 1; t++) {
    const int i = t;          // dstqds step
    const int q = n - 2 - t;  // dpqds step
    const double dp = d[i] + sc;
    const double rp = d[q] * (l[q] * l[q]) + pc;
    const double lp = d[i] * l[i] / dp;
    const double pq = pc * d[q] / rp;
    DP[at(i)] = dp;
    RP[at(q + 1)] = rp;
    sc = lp * l[i] * sc - w2;
    pc = pq - w2;
    S[at(i + 1)] = sc;
    Pp[at(q)] = pc;
  }
  DP[at(n - 1)] = d[n - 1] + sc;
  RP[at(0)] = pc;
  }
  __syncthreads();  // the sequences are visible to every wavefront of the workgroup
  const int lane = threadIdx.x & 63;
  __shared__ int twist_at[2][64];  // k0 (v side) and k1 (u side) of each value
  // coupling (bbtsvd.cpp:113-124): s[n] = -omega^2, p[n] = 1, r+[n] = 1
  auto Sv = [&](int i) -> double { return i == n ? -w2 : S[at(i)]; };
  auto Pv = [&](int i) -> double { return i == n ? 1.0 : Pp[at(i)]; };
  auto Rv = [&](int i) -> double { return i == n ? 1.0 : RP[at(i)]; };
  if (active && wave == 0) {
    // twist: first index of min |s + p + omega^2|
    int k0 = 0;
    double best = 0.0;
    #pragma unroll 16
    for (int i = 0; i < n; i++) {
      const double g = S[at(i)] + Pp[at(i)] + w2;
      const double ag = fabs(g);
      if (i == 0 || ag < best) { best = ag; k0 = i; }
    }
    twist_at[0][lane] = k0;
    vs[at(k0)] = 1.0;
    if (twists) twists[2 * j] = k0;
  }
  if (active && wave == 1) {
    int k1 = 0;
    double best = 0.0;
    #pragma unroll 16
    for (int i = 0; i < n; i++) {
      const double si = Sv(i), pi = Pv(i);
      const double g = si + pi + w2;
      const double gn = -w2 * g * Rv(i + 1) / (si * Pv(i + 1));
      const double ag = fabs(gn);
      if (i == 0 || ag < best) { best = ag; k1 = i; }
    }
    twist_at[1][lane] = k1;
    u[at(k1)] = 1.0;
    if (twists) twists[2 * j + 1] = k1;
  }
  __syncthreads();
  // the four inverse-iteration sweeps, one per wavefront: v (bbtsvd.cpp:65-71) with
  // l+ = d l / d+, u+ = l d / r+; u with l+_i = a_{i+1} / d+'_i b_i, u+_i = a_{i+1} / r+'_{i+1} b_i,
  // d+'_i = s_{i+1} / s_i d+_i, r+'_i = p_i / p_{i+1} r+_{i+1}
  if (active) {
    double z = 1.0;
    if (wave == 0) {
      const int k0 = twist_at[0][lane];
      #pragma unroll 16
      for (int i = k0 - 1; i >= 0; i--) {
        const double lp = d[i] * l[i] / DP[at(i)];
        z = -lp * z;
        vs[at(i)] = z;
      }
    } else if (wave == 1) {
      const int k0 = twist_at[0][lane];
      #pragma unroll 16
      for (int i = k0 + 1; i < n; i++) {
        const double up = l[i - 1] * d[i - 1] / RP[at(i)];
        z = -up * z;
        vs[at(i)] = z;
      }
    } else if (wave == 2) {
      const int k1 = twist_at[1][lane];
      #pragma unroll 16
      for (int i = k1 - 1; i >= 0; i--) {
        const double dpn = Sv(i + 1) / Sv(i) * DP[at(i)];
        const double lpn = a[i + 1] / dpn * b[i];
        z = -lpn * z;
        u[at(i)] = z;
      }
    } else {
      const int k1 = twist_at[1][lane];
      #pragma unroll 16
      for (int i = k1 + 1; i < n; i++) {
        const int q = i - 1;
        const double rpn = Pv(q + 1) / Pv(q + 2) * Rv(q + 2);
        const double upn = a[q + 1] / rpn * b[q];
        z = -upn * z;
        u[at(i)] = z;
      }
    }
  }
  __syncthreads();
  // normalization: the sum of squares in index order, then the scaling (wave 0: v, wave 1: u)
  if (active && wave < 2) {
    double* vec = wave == 0 ? vs : u;
    double ss = 0.0;
    #pragma unroll 16
    for (int i = 0; i < n; i++) {
      const double v = vec[at(i)];
      ss = ss + v * v;
    }
    const double inv = 1.0 / sqrt(ss);
    #pragma unroll 16
    for (int i = 0; i < n; i++) vec[at(i)] = inv * vec[at(i)];
  }
}

// Sign of v (bbtsvd.cpp:943-955): p = first index of max |v_j|, signer = (B^T u_j)_p / sigma_j.
__global__ __launch_bounds__(256) void k_vsign(int n, int k, const double* __restrict__ a,
                                               const double* __restrict__ b, const double* __restrict__ sigma,
                                               const double* __restrict__ u, const double* __restrict__ vs,
                                               double* __restrict__ sgn) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  const size_t K = (size_t)k;
  int p = 0;
  double best = -1.0;
#pragma unroll 8
  for (int i = 0; i < n; i++) {
    const double av = fabs(vs[(size_t)i * K + j]);
    if (av > best) { best = av; p = i; }
  }
  double signer;
  if (p == 0) signer = u[(size_t)p * K + j] * a[p] / sigma[j];
  else signer = (u[(size_t)(p - 1) * K + j] * b[p - 1] + u[(size_t)p * K + j] * a[p]) / sigma[j];
  sgn[j] = ((signer * vs[(size_t)p * K + j]) < 0) ? -1.0 : 1.0;
}

// v (k x n) = sign_j * vs^T, through a 64 x 64 LDS tile.
__global__ __launch_bounds__(256) void k_transpose_sign(int n, int k, const double* __restrict__ vs,
                                                        const double* __restrict__ sgn, double* __restrict__ v) {
  __shared__ double tile[64][65];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < 64; r += 4) {
    const int i = i0 + r, j = j0 + tx;
    if (i < n && j < k) tile[r][tx] = vs[(size_t)i * k + j];
  }
  __syncthreads();
  for (int r = ty; r < 64; r += 4) {
    const int j = j0 + r, i = i0 + tx;
    if (i < n && j < k) v[(size_t)j * n + i] = sgn ? sgn[j] * tile[tx][r] : tile[tx][r];
  }
}

// "check None" (bbtsvd.cpp:957-967). The writes of different j never touch the same entry.
__global__ __launch_bounds__(256) void k_check_none(int n, int k, const double* __restrict__ a,
                                                    const double* __restrict__ sigma, double* __restrict__ u,
                                                    double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  if (fabs(sigma[i] - a[i]) < 1e-15) {
    const size_t K = (size_t)k;
    u[i + (size_t)i * K] = 1.0;
    for (int j = i + 1; j < k; j++) u[(size_t)i * K + j] = 0.0;
    for (int j = i + 1; j < n; j++) u[(size_t)j * K + i] = 0.0;
    for (int j = 0; j < i; j++) v[(size_t)i * n + j] = 0.0;
    v[(size_t)i * n + i] = 1.0;
    for (int j = i + 1; j < n; j++) v[(size_t)i * n + j] = 0.0;
  }
}

// Bump allocator over a caller workspace (256-byte aligned pieces).
struct Bump {
  char* p;
  size_t used = 0, cap;
  Bump(void* base, size_t c) : p((char*)base), cap(c) {}
  template <typename T>
  T* take(size_t n) {
    size_t off = (used + 255) & ~(size_t)255;
    used = off + n * sizeof(T);
    return (T*)(p ? p + off : nullptr);
  }
  bool ok() const { return used <= cap; }
};

int pick_w(int targets) {
  long lanes = 262144;  // about four wavefronts per SIMD across the chip
  int w = 64;
  while (w > 1 && (long)targets * w > lanes) w >>= 1;
  return w;
}

int launch_targets(int n, const double* a, const double* b, const BisectHdr* hdr, double tol, int cap, double* llo,
                   double* lhi, double* lmid, int* flags, hipStream_t s) {
  if (cap <= 0) return EXC_OK;
  const int W = pick_w(cap);
  const unsigned grid = exc_detail::ceil_div((long)cap * W, 256);
#define EXC_TGT(WW)                                                                                  \
  case WW:                                                                                           \
    hipLaunchKernelGGL(k_bisect_targets<WW>, dim3(grid), dim3(256), 0, s, n, a, b, hdr, tol, cap,   \
                       llo, lhi, lmid, flags);                                                       \
    break;
  switch (W) {
    EXC_TGT(1) EXC_TGT(2) EXC_TGT(4) EXC_TGT(8) EXC_TGT(16) EXC_TGT(32) EXC_TGT(64)
  }
#undef EXC_TGT
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

struct BisectWs {
  BisectHdr* hdr;
  double *llo, *lhi, *lmid;
  // exact breadth-first replay of the depth-first search (used when counts are not monotone)
  int cap;  // frontier / leaf capacity
  double *f_lo[2], *f_hi[2];
  int *f_clo[2], *f_chi[2];
  int* ctr;  // [0] frontier size (next), [1] leaf count, [2] overflow
  double *L_lo, *L_hi;
};

size_t bisect_ws_bytes(int n, void* base, BisectWs* ws, Bump* bump) {
  Bump& B = *bump;
  BisectWs w;
  w.hdr = B.take<BisectHdr>(1);
  w.llo = B.take<double>(n + 1);
  w.lhi = B.take<double>(n + 1);
  w.lmid = B.take<double>(n + 1);
  w.cap = 4 * n + 1024;
  for (int q = 0; q < 2; q++) {
    w.f_lo[q] = B.take<double>(w.cap);
    w.f_hi[q] = B.take<double>(w.cap);
    w.f_clo[q] = B.take<int>(w.cap);
    w.f_chi[q] = B.take<int>(w.cap);
  }
  w.ctr = B.take<int>(4);
  w.L_lo = B.take<double>(w.cap);
  w.L_hi = B.take<double>(w.cap);
  if (ws) *ws = w;
  (void)base;
  return B.used;
}

__global__ void k_bfs_seed(const BisectHdr* hdr, double* flo, double* fhi, int* fclo, int* fchi, int* ctr) {
  if (threadIdx.x == 0) {
    flo[0] = hdr->lo; fhi[0] = hdr->hi; fclo[0] = hdr->nclo; fchi[0] = hdr->nchi;
    ctr[0] = 0; ctr[1] = 0; ctr[2] = 0; ctr[3] = 1;  // ctr[3]: current frontier size
  }
}

// One level of the reference's search in breadth-first order: every frontier node is popped
// (leaf test, else one Sturm count at its midpoint) and its children are pushed under the
// reference's rules (left if nc_mid > nc_low, right if nc_high > nc_mid; bbtsvd.cpp:38-48).
__global__ __launch_bounds__(64) void k_bfs_level(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                  double tol, int cap, int fn, const double* __restrict__ flo,
                                                  const double* __restrict__ fhi, const int* __restrict__ fclo,
                                                  const int* __restrict__ fchi, double* __restrict__ nlo,
                                                  double* __restrict__ nhi, int* __restrict__ nclo,
                                                  int* __restrict__ nchi, double* __restrict__ Llo,
                                                  double* __restrict__ Lhi, int* __restrict__ ctr) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= fn) return;
  const double lo = flo[f], hi = fhi[f];
  const double mid = (lo + hi) / 2;
  if ((hi - lo) <= tol || mid == lo || mid == hi) {
    const int q = atomicAdd(&ctr[1], 1);
    if (q < cap) { Llo[q] = lo; Lhi[q] = hi; }
    else atomicOr(&ctr[2], 1);
    return;
  }
  const int c = sturm_count(n, a, b, mid);
  if (c > fclo[f]) {
    const int q = atomicAdd(&ctr[0], 1);
    if (q < cap) { nlo[q] = lo; nhi[q] = mid; nclo[q] = fclo[f]; nchi[q] = c; }
    else atomicOr(&ctr[2], 1);
  }
  if (fchi[f] > c) {
    const int q = atomicAdd(&ctr[0], 1);
    if (q < cap) { nlo[q] = mid; nhi[q] = hi; nclo[q] = c; nchi[q] = fchi[f]; }
    else atomicOr(&ctr[2], 1);
  }
}

// Leaves in the order the depth-first search records them: descending (lo, hi).
__global__ __launch_bounds__(256) void k_leaf_rank(int L, int top_k, const double* __restrict__ Llo,
                                                   const double* __restrict__ Lhi, double* __restrict__ sigma,
                                                   double* __restrict__ olo, double* __restrict__ ohi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  const double lo = Llo[i], hi = Lhi[i];
  int rank = 0;
  for (int j = 0; j < L; j++) {
    const double l2 = Llo[j], h2 = Lhi[j];
    rank += (l2 > lo || (l2 == lo && (h2 > hi || (h2 == hi && j < i)))) ? 1 : 0;
  }
  if (rank < top_k) {
    sigma[rank] = (lo + hi) / 2;
    if (olo) olo[rank] = lo;
    if (ohi) ohi[rank] = hi;
  }
}

__global__ void k_set_count(int* count, BisectHdr* hdr, const int* ctr, int top_k) {
  if (threadIdx.x == 0) {
    const int c = min(ctr[1], top_k);
    *count = c;
    hdr->count = c;
  }
}

int run_bfs(int n, const double* a, const double* b, double tol, int top_k, double* sigma, int* count, double* olo,
            double* ohi, const BisectWs& w, hipStream_t s) {
  hipLaunchKernelGGL(k_bfs_seed, dim3(1), dim3(64), 0, s, w.hdr, w.f_lo[0], w.f_hi[0], w.f_clo[0], w.f_chi[0],
                     w.ctr);
  EXC_CHECK_HIP(hipGetLastError());
  int fn = 1, cur = 0;
  int h[4];
  for (int level = 0; fn > 0; level++) {
    EXC_REQUIRE(level < 4096, EXC_ERR_NUMERIC, "bisection: search deeper than 4096 levels");
    const int nxt = cur ^ 1;
    EXC_CHECK_HIP(hipMemsetAsync(w.ctr, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_bfs_level, dim3(exc_detail::ceil_div(fn, 64)), dim3(64), 0, s, n, a, b, tol, w.cap, fn,
                       w.f_lo[cur], w.f_hi[cur], w.f_clo[cur], w.f_chi[cur], w.f_lo[nxt], w.f_hi[nxt],
                       w.f_clo[nxt], w.f_chi[nxt], w.L_lo, w.L_hi, w.ctr);
    EXC_CHECK_HIP(hipGetLastError());
    EXC_CHECK_HIP(hipMemcpyAsync(h, w.ctr, sizeof(h), hipMemcpyDeviceToHost, s));
    EXC_CHECK_HIP(hipStreamSynchronize(s));
    EXC_REQUIRE(h[2] == 0, EXC_ERR_NUMERIC, "bisection: search frontier exceeded its capacity");
    fn = h[0];
    cur = nxt;
  }
  const int L = h[1];
  if (L > 0)
    hipLaunchKernelGGL(k_leaf_rank, dim3(exc_detail::ceil_div(L, 256)), dim3(256), 0, s, L, top_k, w.L_lo, w.L_hi,
                       sigma, olo, ohi);
  hipLaunchKernelGGL(k_set_count, dim3(1), dim3(64), 0, s, count, w.hdr, w.ctr, top_k);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

// Shared by exc_bisection and exc_bidiag_svd: the interval is already in hdr.
int run_bisection(int n, const double* a, const double* b, double tol, int top_k, double* sigma, int* count,
                  double* olo, double* ohi, const BisectWs& w, hipStream_t s, double* ldl_l = nullptr,
                  double* ldl_d = nullptr) {
  hipLaunchKernelGGL(k_bisect_init, dim3(1), dim3(128), 0, s, n, a, b, tol, w.hdr, ldl_l, ldl_d);
  EXC_CHECK_HIP(hipGetLastError());
  int cap = min(n, top_k + 64);
  for (int attempt = 0; attempt < 2; attempt++) {
    int rc = launch_targets(n, a, b, w.hdr, tol, cap, w.llo, w.lhi, w.lmid, &w.hdr->flags, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bisect_compact, dim3(1), dim3(1024), 0, s, w.hdr, top_k, cap, w.llo, w.lhi, w.lmid,
                       sigma, count, olo, ohi);
    EXC_CHECK_HIP(hipGetLastError());
    BisectHdr h;
    EXC_CHECK_HIP(hipMemcpyAsync(&h, w.hdr, sizeof(h), hipMemcpyDeviceToHost, s));
    EXC_CHECK_HIP(hipStreamSynchronize(s));
    if (h.flags & 2) {
      // A count left its bracket on some path: the search can reach leaves no singular value
      // index leads to. Replay the search node by node instead.
      return run_bfs(n, a, b, tol, top_k, sigma, count, olo, ohi, w, s);
    }
    if (!h.status || cap >= n) break;
    cap = n;  // more merged leaves than the margin: evaluate every target
  }
  return EXC_OK;
}

}  // namespace

extern "C" {

int exc_sturm_count(int n, const double* d_alpha, const double* d_beta, const double* d_mu, int nmu, int* d_count,
                    exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && nmu >= 0 && d_alpha && d_mu && d_count && (n == 1 || d_beta), EXC_ERR_ARG,
              "exc_sturm_count: bad arguments");
  if (nmu == 0) return EXC_OK;
  hipLaunchKernelGGL(k_sturm_count, dim3(exc_detail::ceil_div(nmu, 256)), dim3(256), 0, exc_detail::stream(stream),
                     n, d_alpha, d_beta, d_mu, nmu, d_count);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

size_t exc_bisection_workspace(int n, int top_k) {
  (void)top_k;
  Bump B(nullptr, 0);
  return bisect_ws_bytes(n, nullptr, nullptr, &B) + 256;
}

int exc_bisection(int n, const double* d_alpha, const double* d_beta, double lo, double hi, double tol, int top_k,
                  double* d_sigma, int* d_count, double* d_leaf_lo, double* d_leaf_hi, void* d_work,
                  size_t work_bytes, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_alpha && (n == 1 || d_beta) && d_count && (top_k <= 0 || d_sigma) && d_work, EXC_ERR_ARG,
              "exc_bisection: bad arguments");
  Bump B(d_work, work_bytes);
  BisectWs w;
  bisect_ws_bytes(n, d_work, &w, &B);
  EXC_REQUIRE(B.ok(), EXC_ERR_SPACE, "exc_bisection: workspace too small");
  hipStream_t s = exc_detail::stream(stream);
  hipLaunchKernelGGL(k_set_interval, dim3(1), dim3(64), 0, s, w.hdr, lo, hi);
  EXC_CHECK_HIP(hipGetLastError());
  if (top_k > n) top_k = n;
  return run_bisection(n, d_alpha, d_beta, tol, top_k, d_sigma, d_count, d_leaf_lo, d_leaf_hi, w, s);
}

int exc_ldl(int n, const double* d_alpha, const double* d_beta, double* d_l, double* d_d, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_alpha && d_d && (n == 1 || (d_beta && d_l)), EXC_ERR_ARG, "exc_ldl: bad arguments");
  hipLaunchKernelGGL(k_ldl, dim3(1), dim3(64), 0, exc_detail::stream(stream), n, d_alpha, d_beta, d_l, d_d);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

namespace {
struct TwistWs {
  double *l, *d, *S, *P, *DP, *RP, *vs, *sgn;
};
void twist_ws(int n, int k, Bump& B, TwistWs* w) {
  w->l = B.take<double>(n);
  w->d = B.take<double>(n);
  w->S = B.take<double>((size_t)(n + 1) * k);
  w->P = B.take<double>((size_t)(n + 1) * k);
  w->DP = B.take<double>((size_t)(n + 1) * k);
  w->RP = B.take<double>((size_t)(n + 1) * k);
  w->vs = B.take<double>((size_t)n * k);
  w->sgn = B.take<double>(k);
}

int run_twisted(int n, int k, const double* omega, const double* a, const double* b, double* u, double* v,
                int* twists, const TwistWs& w, bool apply_vsign, const double* sigma_for_sign, hipStream_t s,
                bool ldl_done = false) {
  if (!ldl_done) {
    hipLaunchKernelGGL(k_ldl, dim3(1), dim3(64), 0, s, n, a, b, w.l, w.d);
    EXC_CHECK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_twisted, dim3(exc_detail::ceil_div(k, 64)), dim3(256), 0, s, n, k, a, b, w.l, w.d, omega, u,
                     w.vs, w.S, w.P, w.DP, w.RP, twists);
  EXC_CHECK_HIP(hipGetLastError());
  if (apply_vsign) {
    hipLaunchKernelGGL(k_vsign, dim3(exc_detail::ceil_div(k, 64)), dim3(64), 0, s, n, k, a, b, sigma_for_sign, u,
                       w.vs, w.sgn);
    EXC_CHECK_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_transpose_sign, dim3(exc_detail::ceil_div(k, 64), exc_detail::ceil_div(n, 64)), dim3(256), 0, s,
                     n, k, w.vs, apply_vsign ? w.sgn : nullptr, v);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}
}  // namespace

size_t exc_twisted_workspace(int n, int k) {
  Bump B(nullptr, 0);
  TwistWs w;
  twist_ws(n, k, B, &w);
  return B.used + 256;
}

int exc_twisted_vectors(int n, int k, const double* d_omega, const double* d_alpha, const double* d_beta,
                        double* d_u, double* d_v, int* d_twists, void* d_work, size_t work_bytes,
                        exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && k >= 0 && d_omega && d_alpha && (n == 1 || d_beta) && d_u && d_v && d_work, EXC_ERR_ARG,
              "exc_twisted_vectors: bad arguments");
  if (k == 0) return EXC_OK;
  Bump B(d_work, work_bytes);
  TwistWs w;
  twist_ws(n, k, B, &w);
  EXC_REQUIRE(B.ok(), EXC_ERR_SPACE, "exc_twisted_vectors: workspace too small");
  return run_twisted(n, k, d_omega, d_alpha, d_beta, d_u, d_v, d_twists, w, false, nullptr,
                     exc_detail::stream(stream));
}

size_t exc_bidiag_svd_workspace(int k, int top_k) {
  Bump B(nullptr, 0);
  bisect_ws_bytes(k, nullptr, nullptr, &B);
  B.take<int>(1);
  TwistWs w;
  twist_ws(k, top_k, B, &w);
  return B.used + 256;
}

int exc_bidiag_svd(int k, const double* d_alpha, const double* d_beta, double* d_u, double* d_v, double* d_sigma,
                   int top_k, int* d_twists, void* d_work, size_t work_bytes, exc_stream_t stream) {
  EXC_REQUIRE(k >= 1 && top_k >= 0 && top_k <= k && d_alpha && (k == 1 || d_beta) && d_sigma && d_work &&
                  (top_k == 0 || (d_u && d_v)),
              EXC_ERR_ARG, "exc_bidiag_svd: bad arguments");
  Bump B(d_work, work_bytes);
  BisectWs bw;
  bisect_ws_bytes(k, d_work, &bw, &B);
  int* dcount = B.take<int>(1);
  TwistWs tw;
  twist_ws(k, top_k, B, &tw);
  EXC_REQUIRE(B.ok(), EXC_ERR_SPACE, "exc_bidiag_svd: workspace too small");
  hipStream_t s = exc_detail::stream(stream);
  hipLaunchKernelGGL(k_frob_bound, dim3(1), dim3(256), 0, s, k, d_alpha, d_beta, bw.hdr);
  EXC_CHECK_HIP(hipGetLastError());
  int rc = run_bisection(k, d_alpha, d_beta, 1e-20, top_k, d_sigma, dcount, nullptr, nullptr, bw, s,
                         top_k > 0 ? tw.l : nullptr, top_k > 0 ? tw.d : nullptr);
  if (rc) return rc;
  int cnt = 0;
  EXC_CHECK_HIP(hipMemcpyAsync(&cnt, dcount, sizeof(int), hipMemcpyDeviceToHost, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  EXC_REQUIRE(cnt >= top_k, EXC_ERR_COUNT, "exc_bidiag_svd: bisection found fewer singular values than top_k");
  if (top_k == 0) return EXC_OK;
  rc = run_twisted(k, top_k, d_sigma, d_alpha, d_beta, d_u, d_v, d_twists, tw, true, d_sigma, s, true);
  if (rc) return rc;
  hipLaunchKernelGGL(k_check_none, dim3(exc_detail::ceil_div(top_k, 256)), dim3(256), 0, s, k, top_k, d_alpha,
                     d_sigma, d_u, d_v);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free variant
// Scratch of exc_bidiag_svd_noalloc: one copy per device, alive until the process ends. Host code
// never passes `scratch` itself to a kernel (in host code the name is only a stand-in): it takes
// the device address once per device with hipGetSymbolAddress and uses that.
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)48 << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];
void* g_scratch_ptr[kNoAllocMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

int scratch_for_current_device(void** out) {
  int dev = 0;
  EXC_CHECK_HIP(hipGetDevice(&dev));
  EXC_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, EXC_ERR_ARG, "exc_bidiag_svd_noalloc: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    EXC_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    EXC_REQUIRE(p != nullptr, EXC_ERR_HIP, "exc_bidiag_svd_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return EXC_OK;
}
}  // namespace

extern "C" {

size_t exc_bidiag_svd_noalloc_capacity(void) { return kNoAllocScratchBytes; }

int exc_bidiag_svd_noalloc(int k, const double* d_alpha, const double* d_beta, double* d_u, double* d_v,
                           double* d_sigma, int top_k, int* d_twists, exc_stream_t stream) {
  EXC_REQUIRE(k >= 1 && top_k >= 0 && top_k <= k, EXC_ERR_ARG, "exc_bidiag_svd_noalloc: bad arguments");
  const size_t need = exc_bidiag_svd_workspace(k, top_k);
  EXC_REQUIRE(need <= kNoAllocScratchBytes, EXC_ERR_SPACE,
              "exc_bidiag_svd_noalloc: k and top_k need more scratch than the fixed 48 MiB per device "
              "(k = top_k = 1024 fits); use exc_bidiag_svd with a workspace");
  void* p = nullptr;
  const int rc = scratch_for_current_device(&p);
  if (rc) return rc;
  return exc_bidiag_svd(k, d_alpha, d_beta, d_u, d_v, d_sigma, top_k, d_twists, p, kNoAllocScratchBytes, stream);
}

}  // extern "C"
