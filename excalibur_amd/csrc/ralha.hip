// One-sided (Ralha) bidiagonalization step kernels (client.cpp:829-910) and the small pipeline
// kernels around the bidiagonal SVD (client.cpp:947-954, 987-995, 1127-1128).
//
// This peer's share of the rows f_j (the reference's row-major m x c F) is kept column-major,
// Fc[col * m + j], so every column is contiguous; U likewise. One kernel per column step applies
// step i column-locally (a workgroup owns four columns, one wavefront each, and recomputes the
// reflector from the reduced message in a fixed order, so all workgroups see the same h and tau)
// and then writes its partial sums of step i+1's message; a second kernel adds the partials in
// workgroup order. The host all-reduces that message over the peers between steps.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cstring>

#include "exc_common.hpp"
#include "solve.hpp"

namespace {

using fact::readlane_f64;

constexpr int kColsPerWg = 4;  // columns per step workgroup (one wavefront each)

__device__ inline double dlapy2_dev(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// Block-wide sum with a fixed order (same result in every workgroup for the same input): the
// wavefront totals, then every thread adds them in wavefront order. The leading barrier keeps a
// second call from overwriting red before every thread has read the first call's totals.
__device__ inline double block_sum(double v, double* red) {
  v = wave_total(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q];
  return t;
}

// The reflector of step i from the reduced message g (LAPACK dlarfg on n = m-1-i entries, the
// same fixed-order reduction in every workgroup): hbuf[0..n) = h (h_0 = 1), returns tau; also
// beta_{i-1}, alpha_i and 1 / alpha_i from the four reduced dots (client.cpp:858-884). Workgroup 0
// stores column i+1 of house, tau[i+1], beta[i-1] and alpha[i]. Ends with a barrier.
//
// Every dependent FP64 divide or square root is a chain of about ten operations, so the chain is
// kept short: alpha_i and its reciprocal are formed before the reflector's branch (they are
// independent of it), and beta = -sign(alpha) sqrt(alpha^2 + ||x||^2) is taken directly while
// both terms are far from overflow and underflow; outside that range dlapy2's scaled form and
// dlarfg's rescaling are used as in LAPACK.
__device__ double ralha_reflector(int i, int m, const double* __restrict__ g, double* hbuf, double* red,
                                  double* __restrict__ house, double* __restrict__ tau, double* __restrict__ alpha,
                                  double* __restrict__ beta, double* out_b, double* out_a, double* out_ia) {
  const int tid = threadIdx.x;
  const int n = m - 1 - i;
  double ss = 0.0;
  for (int r = 1 + tid; r < n; r += blockDim.x) ss += g[r] * g[r];
  const double x2 = block_sum(ss, red);
  const double b = i > 0 ? g[m - i - 1] : 0.0;
  const double a = i > 0 ? sqrt(g[m - i] * (b * b) + g[m - i + 1] - 2 * b * g[m - i + 2]) : 0.0;
  *out_ia = i > 0 ? 1.0 / a : 0.0;
  double alph = g[0], t = 0.0, scale = 0.0;
  if (n > 1 && x2 != 0.0) {
    const double a2 = alph * alph;
    const double big = fmax(a2, x2);
    double bet = (big > 1e-280 && big < 1e280) ? -copysign(sqrt(a2 + x2), alph)
                                               : -copysign(dlapy2_dev(alph, sqrt(x2)), alph);
    double xs = 1.0;
    const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    if (fabs(bet) < safmin) {
      const double rs = 1.0 / safmin;
      int knt = 0;
      do {
        knt++;
        xs *= rs;
        bet *= rs;
        alph *= rs;
      } while (fabs(bet) < safmin && knt < 20);
      double s2 = 0.0;
      for (int r = 1 + tid; r < n; r += blockDim.x) s2 += (g[r] * xs) * (g[r] * xs);
      const double xnorm = sqrt(block_sum(s2, red));
      bet = -copysign(dlapy2_dev(alph, xnorm), alph);
    }
    t = (bet - alph) / bet;
    scale = xs * (1.0 / (alph - bet));
  }
  for (int r = tid; r < n; r += blockDim.x) hbuf[r] = r == 0 ? 1.0 : g[r] * scale;
  if (blockIdx.x == 0) {
    for (int r = tid; r < n; r += blockDim.x) house[(long)(i + 1 + r) * m + (i + 1)] = r == 0 ? 1.0 : g[r] * scale;
    if (tid == 0) {
      tau[i + 1] = t;
      if (i > 0) {
        beta[i - 1] = b;
        alpha[i] = a;
      }
    }
  }
  __syncthreads();
  *out_b = b;
  *out_a = a;
  return t;
}

// Sum over the workgroups' partial rows of entry e (partial[v * ld + e], v < nwg) in a fixed
// order: sixteen groups of 16 threads (group g = threadIdx.x / 16, entry lane threadIdx.x % 16)
// each add the partials v = g, g + 16, g + 32, ... in turn, up to eight loads in flight at
// once; then the sixteen group sums are added in group order. A 256-thread workgroup handles 16 entries per
// call; the result is returned to threads 0..15 (entry lane threadIdx.x). With COHERENT the
// partials are read with device-scope atomics (written by other workgroups of a running grid).
template <bool COHERENT>
__device__ __forceinline__ double partials_sum(const double* __restrict__ partial, long ld, int nwg, int e,
                                               bool active, double (*acc)[16]) {
  const int g = threadIdx.x >> 4, l = threadIdx.x & 15;
  double s = 0.0;
  if (active) {
    auto ld1 = [&](int v) {
      const double* p = partial + (long)v * ld + e;
      return COHERENT ? __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *p;
    };
    for (int v0 = g; v0 < nwg; v0 += 128) {
      double x[8];
#pragma unroll
      for (int q = 0; q < 8; q++) x[q] = v0 + 16 * q < nwg ? ld1(v0 + 16 * q) : 0.0;
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (v0 + 16 * q < nwg) s += x[q];
    }
  }
  acc[g][l] = s;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < 16) {
    t = acc[0][l];
#pragma unroll
    for (int q = 1; q < 16; q++) t += acc[q][l];
  }
  __syncthreads();
  return t;
}

// Register-resident columns for m <= 64 * KM: each wavefront holds CPW whole columns of F (row
// lane + 64q in slot q), with u_{i-1} and u_i of each column.
template <int KM, int CPW>
struct RegCols {
  double fr[CPW][KM];
  double um1[CPW];  // u_{i-1} (i > 0)
  double ui[CPW];   // u_i (i >= 0), the u_{j-1} of the next message; the same on every lane
  int cols[CPW];
};

// Loads rows >= max(i, 0) of this wavefront's columns (zeros elsewhere) and u_{i-1}.
template <int KM, int CPW>
__device__ __forceinline__ void reg_load(RegCols<KM, CPW>& R, int i, int m, int c, const double* __restrict__ Fc,
                                         const double* __restrict__ Uc) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int cc = 0; cc < CPW; cc++) {
    const int col = (blockIdx.x * 4 + wave) * CPW + cc;
    R.cols[cc] = col;
    R.um1[cc] = 0.0;
    R.ui[cc] = 0.0;
#pragma unroll
    for (int q = 0; q < KM; q++) {
      const int row = lane + 64 * q;
      R.fr[cc][q] = (col < c && row < m && row >= (i < 0 ? 0 : i)) ? Fc[(long)col * m + row] : 0.0;
    }
    if (col < c && i > 0) R.um1[cc] = Uc[(long)col * m + i - 1];
  }
}

// Step i >= 0 on the registers (client.cpp:858-891): the reflector from the reduced message g,
// F[i+1:, col] -= tau h (h . F[i+1:, col]), then f_i -= beta u_{i-1} and u_i = f_i / alpha_i
// (stored to U); for i = 0, u_0 comes from the tail kernels. Leaves u_i in R.ui on every lane.
// The CPW columns' dot products run side by side (four fused partial sums each, then a DPP wave
// total per column), so their latencies overlap.
template <int KM, int CPW>
__device__ __forceinline__ void reg_apply(RegCols<KM, CPW>& R, int i, int m, int c, const double* __restrict__ g,
                                          double* hbuf, double* red, double* __restrict__ Uc,
                                          double* __restrict__ house, double* __restrict__ tau,
                                          double* __restrict__ alpha, double* __restrict__ beta) {
  const int lane = threadIdx.x & 63;
  double b, a;
  double ia;
  const double t = ralha_reflector(i, m, g, hbuf, red, house, tau, alpha, beta, &b, &a, &ia);
  if (t != 0.0) {
    // h from LDS (zero outside rows i+1 .. m-1); read twice rather than held in KM registers
    auto h = [&](int q) {
      const int row = lane + 64 * q;
      return (row > i && row < m) ? hbuf[row - i - 1] : 0.0;
    };
    double s[CPW];
#pragma unroll
    for (int cc = 0; cc < CPW; cc++) {
      double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
#pragma unroll
      for (int q = 0; q < KM; q += 4) {
        p0 = fma(R.fr[cc][q], h(q), p0);
        if (q + 1 < KM) p1 = fma(R.fr[cc][q + 1], h(q + 1), p1);
        if (q + 2 < KM) p2 = fma(R.fr[cc][q + 2], h(q + 2), p2);
        if (q + 3 < KM) p3 = fma(R.fr[cc][q + 3], h(q + 3), p3);
      }
      s[cc] = (p0 + p1) + (p2 + p3);
    }
#pragma unroll
    for (int cc = 0; cc < CPW; cc++) s[cc] = wave_total(s[cc]);
#pragma unroll
    for (int q = 0; q < KM; q++) {
      const int row = lane + 64 * q;
      const double th = -t * h(q);
#pragma unroll
      for (int cc = 0; cc < CPW; cc++)
        if (R.cols[cc] < c && row > i && row < m) R.fr[cc][q] = fma(s[cc], th, R.fr[cc][q]);
    }
  }
  const int qi = i >> 6, li = i & 63;
  if (i > 0) {
    const double inv_a = ia;
#pragma unroll
    for (int q = 0; q < KM; q++) {
      if (q == qi) {  // uniform; row i sits in lane i % 64 of slot i / 64
#pragma unroll
        for (int cc = 0; cc < CPW; cc++) {
          const double fi = R.fr[cc][q] + (-b) * R.um1[cc];
          const double ui = inv_a * fi;
          if (lane == li && R.cols[cc] < c) {
            R.fr[cc][q] = fi;
            Uc[(long)R.cols[cc] * m + i] = ui;
          }
          const double u = readlane_f64(ui, li);
          R.ui[cc] = R.cols[cc] < c ? u : 0.0;
        }
      }
    }
  } else {
#pragma unroll
    for (int cc = 0; cc < CPW; cc++) R.ui[cc] = R.cols[cc] < c ? Uc[(long)R.cols[cc] * m] : 0.0;
  }
}

// Stores rows >= rowmin of this wavefront's columns.
template <int KM, int CPW>
__device__ __forceinline__ void reg_store(const RegCols<KM, CPW>& R, int rowmin, int m, int c,
                                          double* __restrict__ Fc) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int cc = 0; cc < CPW; cc++) {
    if (R.cols[cc] >= c) continue;
#pragma unroll
    for (int q = 0; q < KM; q++) {
      const int row = lane + 64 * q;
      if (row >= rowmin && row < m) Fc[(long)R.cols[cc] * m + row] = R.fr[cc][q];
    }
  }
}

// This workgroup's share of message j (client.cpp:842-855): f_{j+1+r} . f_j summed over its
// columns, then the four dots with u_{j-1} (j > 0); the four wavefronts' shares are added in
// wavefront order through mp (4 x (m + 4) doubles of LDS) into out[0, m - 1 - j + 4); with
// COHERENT the row is stored with device-scope atomics for other workgroups to read.
template <int KM, int CPW, bool COHERENT = false>
__device__ __forceinline__ void reg_share(const RegCols<KM, CPW>& R, int j, int m, int c, double* mp,
                                          double* __restrict__ out) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ldm = m + 4;
  const int nrows = m - 1 - j;
  double pr[KM];
#pragma unroll
  for (int q = 0; q < KM; q++) pr[q] = 0.0;
  double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
#pragma unroll
  for (int cc = 0; cc < CPW; cc++) {
    if (R.cols[cc] >= c) continue;
    double fj = 0.0;
#pragma unroll
    for (int q = 0; q < KM; q++)
      if (lane + 64 * q == j) fj = R.fr[cc][q];
    fj = readlane_f64(fj, j & 63);
#pragma unroll
    for (int q = 0; q < KM; q++) pr[q] = fma(R.fr[cc][q], fj, pr[q]);
    if (j > 0) {
      const double u = R.ui[cc];
      d0 += fj * u;
      d1 += u * u;
      d2 += fj * fj;
      d3 += u * fj;
    }
  }
  double* mw = mp + wave * ldm;
#pragma unroll
  for (int q = 0; q < KM; q++) {
    const int row = lane + 64 * q;
    if (row > j && row < m) mw[row - j - 1] = pr[q];
  }
  if (lane == 0) {
    mw[nrows] = d0;
    mw[nrows + 1] = d1;
    mw[nrows + 2] = d2;
    mw[nrows + 3] = d3;
  }
  __syncthreads();
  const int cnt = nrows + (j > 0 ? 4 : 0);
  for (int r = tid; r < cnt; r += blockDim.x) {
    const double v = ((mp[r] + mp[ldm + r]) + mp[2 * ldm + r]) + mp[3 * ldm + r];
    if (COHERENT)
      __hip_atomic_store(&out[r], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // read by other workgroups
    else
      out[r] = v;
  }
}

// One launch per step: the register form of k_ralha_step below. A column is read once,
// updated, stored once, and its share of the next message is formed from the registers.
template <int KM, int CPW>
__global__ __launch_bounds__(256) void k_ralha_step_reg(int i, int m, int c, double* __restrict__ Fc,
                                                        double* __restrict__ Uc, const double* __restrict__ g,
                                                        double* __restrict__ house, double* __restrict__ tau,
                                                        double* __restrict__ alpha, double* __restrict__ beta,
                                                        double* __restrict__ partial) {
  extern __shared__ double lds[];  // hbuf: m, then mp: 4 x (m + 4)
  double* hbuf = lds;
  double* mp = lds + m;
  __shared__ double red[8];
  RegCols<KM, CPW> R;
  reg_load(R, i, m, c, Fc, Uc);
  if (i >= 0) {
    reg_apply(R, i, m, c, g, hbuf, red, Uc, house, tau, alpha, beta);
    reg_store(R, i, m, c, Fc);
  }
  const int j = i + 1;
  if (j > m - 3) return;  // uniform: no message after the last full step
  reg_share(R, j, m, c, mp, partial + (long)blockIdx.x * (m + 3));
}

// ---- The whole loop in one launch -------------------------------------------------------------
//
// k_ralha_loop runs steps -1 .. m-3 in one launch. Every workgroup keeps its columns in registers
// for the whole loop, and the host only moves the messages over the network:
//
//   GPU: message j published (each workgroup reduces a slice of the partials into link.msg, then
//        sets link.done[wg] = j + 1)
//   host: waits for every done[wg] > j, all-reduces msg in place, sets link.in_seq = j + 1
//   GPU: workgroup 0 copies msg to device memory and sets dsync[kGin]; every workgroup applies
//        step j and forms its share of message j + 1.
//
// Memory ordering without cache maintenance: on gfx950 a device-scope fence writes back and
// invalidates the whole L2 of the XCD, and with dozens of workgroups fencing several times per
// step those operations stall every memory access on the chip. So every value that crosses
// workgroups is stored and loaded with relaxed device-scope atomics (sc1: coherent across the
// XCDs' L2s) or system-scope ones for the host link (sc0 sc1), and a wavefront waits for its own
// stores to complete (s_waitcnt vmcnt(0)) before the workgroup barrier that precedes a flag
// store. A flag is read with the same relaxed loads, and the data after it with them too, so no
// cache can hold a stale copy.
//
// The grid barrier before the slice reduction is a flag per workgroup (its step number), polled
// by one wavefront with one load per 64 flags: no atomic on a shared counter, whose updates from
// every workgroup would be serialized at one address.
//
// One owner (boxes with P = 1, the bidiagonalization of a matrix a single peer holds) needs no
// flags at all: its partial rows and message slots start filled with a sentinel (a signaling NaN
// bit pattern that no arithmetic produces), a reader polls the values themselves until none is the
// sentinel, and the reader of a partial or the owner of a message slice puts the sentinel back
// once every workgroup is known to have read it. That removes the barrier and the wait for each
// flag's data to be acknowledged before the flag is raised. Ordering of the resets: a workgroup
// waits for its own stores (vmcnt(0)) before each block of its slice is stored, and again at the
// end of each receive, before it stores its next partial row; every later write to a reset
// location, and every read that must see the reset, happens only after some workgroup has read a
// store issued after that wait. (The first wait alone suffices: a workgroup whose slice holds an
// entry that is written again two steps later still owns a slice at the next step, as slices
// shrink with the message; the second is a margin.)
//
// Every wait is bounded by a wall-clock limit and also ends on dsync[kStop] (set by any
// workgroup that gave up) or link.abort (set by the host), so the grid always drains. The grid is
// small (one workgroup per 4 x CPW columns) and the host launches it only when twice that many
// workgroups fit on the device at once, so all workgroups are resident together.

constexpr int kLoopMaxWg = 256;
constexpr int kLoopMaxM = 3000;       // LDS: hbuf, the message copy and the four wavefronts' shares
constexpr int kLinkMsgOffset = 1280;  // bytes: 64-byte header, done[256], then the message

struct LoopLink {  // pinned, coherent host memory
  int in_seq;      // host: reduced messages delivered so far
  int abort;       // host: leave at the next wait
  int error;       // GPU: why the first workgroup gave up: 1 = a wait ran out of time, 2 = the abort flag
  int nwg, m;      // set by the host at launch
  int pad[11];
  int done[kLoopMaxWg];  // GPU: messages published by workgroup w's slice
};
static_assert(sizeof(LoopLink) <= kLinkMsgOffset, "link header");

// Peers that share the device can exchange the messages through device memory instead of the
// host: every peer owns a box (a counter, then two message slots used by step parity), opened by
// the other peers' processes through IPC handles. Peer p's workgroups publish their slices of
// message j into slot j % 2 of p's box and add 1 to its counter; message j of peer q is complete
// when q's counter reaches nwg[q] (j + 1). The reduced message is then, on every peer, its own
// message plus the others' in peer order, the order of the host all-reduce
// (client.cpp:313-327), so both exchanges give the same bits. Two slots suffice: a peer writes
// message j + 2 only after every peer has published message j + 1, which each does only after
// reading all of message j.
constexpr int kMaxPeers = 16;
constexpr int kBoxMsgOffset = 256;
struct PeerBoxes {  // kernel argument; P = 0 selects the host link
  int P, me;
  int nwg[kMaxPeers];
  char* box[kMaxPeers];
};

enum { kGin = 0, kStop = 2 };  // device words; the barrier flags follow at kFlagsOffset bytes
constexpr int kFlagsOffset = 256;      // workgroup w's barrier flag: ((int*)(dsync + 64))[w]
constexpr int kSyncBytes = 256 + 4 * kLoopMaxWg;
constexpr unsigned long long kSentinel = 0x7FF0DEAD0000BEEFull;  // a signaling NaN: no arithmetic yields it

__device__ __forceinline__ double ld_dev(const double* p) { return coh_ld(p); }
__device__ __forceinline__ void st_dev(double* p, double v) { coh_st(p, v); }
__device__ __forceinline__ double ld_sys(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void st_sys(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
// This wavefront's stores have completed (vmcnt 0; expcnt and lgkmcnt left alone).
__device__ __forceinline__ void stores_done() { coh_stores_done(); }

__device__ __forceinline__ void loop_fail(int* dsync, LoopLink* link, int why) {
  __hip_atomic_store(&dsync[kStop], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&link->error, why, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Thread 0 waits until *word >= target (device memory, relaxed device-scope loads); the verdict
// reaches every thread through LDS. Returns false when the wait gave up.
__device__ __forceinline__ bool loop_wait_dev(int* word, int target, int* dsync, LoopLink* link, long long limit,
                                              int* sh) {
  if (threadIdx.x == 0) {
    int ok = 1;
    const long long t0 = wall_clock64();
    for (unsigned k = 0; __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; k++) {
      if ((k & 15) != 15) continue;  // the stop flag and the clock every 16 polls: one load in flight per poll
      if (__hip_atomic_load(&dsync[kStop], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        ok = 0;
        break;
      }
      if (wall_clock64() - t0 > limit) {
        loop_fail(dsync, link, 1);
        ok = 0;
        break;
      }
    }
    sh[0] = ok;
  }
  __syncthreads();
  const bool ok = sh[0] != 0;
  __syncthreads();
  return ok;
}

__device__ __forceinline__ int* barrier_flags(int* dsync) { return (int*)((char*)dsync + kFlagsOffset); }

// Wavefront 0 waits until flags[0 .. G) (G <= 256) all reach target, one load per 64 flags per
// poll, all in flight; the verdict reaches every thread through LDS.
__device__ __forceinline__ bool wait_all(const int* flags, int G, int target, int* dsync, LoopLink* link,
                                         long long limit, int* sh) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int ok = 1;
    const long long t0 = wall_clock64();
    for (unsigned k = 0;; k++) {
      int f[4];
#pragma unroll
      for (int q = 0; q < 4; q++)
        f[q] = lane + 64 * q < G ? __hip_atomic_load(&flags[lane + 64 * q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                 : target;
      if (__all(f[0] >= target && f[1] >= target && f[2] >= target && f[3] >= target)) break;
      if ((k & 15) == 15) {
        const int stop = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(&dsync[kStop], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (stop != 0) {
          ok = 0;
          break;
        }
        if (wall_clock64() - t0 > limit) {
          if (lane == 0) loop_fail(dsync, link, 1);
          ok = 0;
          break;
        }
      }
    }
    if (lane == 0) sh[0] = ok;
  }
  __syncthreads();
  const bool ok = sh[0] != 0;
  __syncthreads();
  return ok;
}

// Workgroup 0's wait for the host (link.in_seq >= target, relaxed system-scope loads).
__device__ __forceinline__ bool loop_wait_host(LoopLink* link, int target, int* dsync, long long limit, int* sh) {
  if (threadIdx.x == 0) {
    int ok = 1;
    const long long t0 = wall_clock64();
    for (unsigned k = 0;; k++) {
      if (__hip_atomic_load(&link->in_seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) >= target) break;
      if ((k & 15) == 15) {
        if (__hip_atomic_load(&dsync[kStop], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;  // another workgroup gave up and recorded why
          break;
        }
        if (__hip_atomic_load(&link->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) {
          loop_fail(dsync, link, 2);
          ok = 0;
          break;
        }
        if (wall_clock64() - t0 > limit) {
          loop_fail(dsync, link, 1);
          ok = 0;
          break;
        }
      }
    }
    sh[0] = ok;
  }
  __syncthreads();
  const bool ok = sh[0] != 0;
  __syncthreads();
  return ok;
}

// Publishes message j. This workgroup's partial row is in `partial` (stored with st_dev). After a
// grid barrier each workgroup adds the partials of its slice of the message (partials_sum, the
// order of k_ralha_reduce) into link.msg, zeros after the valid entries, and sets its done word.
__device__ __forceinline__ bool loop_publish(int j, int m, const double* __restrict__ partial, LoopLink* link,
                                             double* __restrict__ hmsg, int* dsync, long long limit, int* sh,
                                             double (*acc)[16]) {
  const int G = gridDim.x, w = blockIdx.x;
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(&barrier_flags(dsync)[w], j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!wait_all(barrier_flags(dsync), G, j + 1, dsync, link, limit, sh)) return false;
  const int len = m + 3 - j;
  const int valid = (m - 1 - j) + (j > 0 ? 4 : 0);
  const int per = (len + G - 1) / G;
  const int r0 = w * per, r1 = min(len, r0 + per);
  for (int base = r0; base < r1; base += 16) {
    const int e = base + (threadIdx.x & 15);
    const double v = partials_sum<true>(partial, m + 3, G, e, e < r1 && e < valid, acc);
    if (threadIdx.x < 16 && e < r1) st_sys(&hmsg[e], v);
  }
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(&link->done[w], j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  return true;
}

// Brings the reduced message i into gl (LDS, m + 3 doubles): workgroup 0 copies it from the host
// into gdev and sets dsync[kGin]; every workgroup then loads gdev.
template <int KM>
__device__ __forceinline__ bool loop_receive(int i, int m, LoopLink* link, const double* hmsg,
                                             double* __restrict__ gdev, double* gl, int* dsync, long long limit,
                                             int* sh) {
  const int len = m + 3 - i;
  if (blockIdx.x == 0) {
    if (!loop_wait_host(link, i + 1, dsync, limit, sh)) return false;
    constexpr int kPer = (64 * KM + 3 + 255) / 256;  // m <= 64 KM
    double v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int r = threadIdx.x + 256 * q;
      v[q] = r < len ? ld_sys(&hmsg[r]) : 0.0;  // all loads in flight before the first store
    }
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      const int r = threadIdx.x + 256 * q;
      if (r < len) {
        st_dev(&gdev[r], v[q]);
        gl[r] = v[q];
      }
    }
    stores_done();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(&dsync[kGin], i + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
  }
  if (!loop_wait_dev(&dsync[kGin], i + 1, dsync, link, limit, sh)) return false;
  constexpr int kPer = (64 * KM + 3 + 255) / 256;
  double v[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int r = threadIdx.x + 256 * q;
    v[q] = r < len ? ld_dev(&gdev[r]) : 0.0;
  }
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int r = threadIdx.x + 256 * q;
    if (r < len) gl[r] = v[q];
  }
  __syncthreads();
  return true;
}

// Message j through the boxes: the slices go to this peer's box, then its counter is raised.
__device__ __forceinline__ bool box_publish(int j, int m, const double* __restrict__ partial, const PeerBoxes& B,
                                            LoopLink* link, int* dsync, long long limit, int* sh,
                                            double (*acc)[16]) {
  const int G = gridDim.x, w = blockIdx.x;
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(&barrier_flags(dsync)[w], j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!wait_all(barrier_flags(dsync), G, j + 1, dsync, link, limit, sh)) return false;
  const int len = m + 3 - j;
  const int valid = (m - 1 - j) + (j > 0 ? 4 : 0);
  const int per = (len + G - 1) / G;
  const int r0 = w * per, r1 = min(len, r0 + per);
  double* slot = (double*)(B.box[B.me] + kBoxMsgOffset) + (long)(j & 1) * (m + 3);
  for (int base = r0; base < r1; base += 16) {
    const int e = base + (threadIdx.x & 15);
    const double v = partials_sum<true>(partial, m + 3, G, e, e < r1 && e < valid, acc);
    if (threadIdx.x < 16 && e < r1) st_dev(&slot[e], v);
  }
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_fetch_add((int*)B.box[B.me], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// Waits until every peer has published message i, then forms the reduced message in gl.
template <int KM>
__device__ __forceinline__ bool box_receive(int i, int m, const PeerBoxes& B, LoopLink* link, double* gl, int* dsync,
                                            long long limit, int* sh) {
  if (threadIdx.x == 0) {
    int ok = 1;
    const long long t0 = wall_clock64();
    for (int q = 0; q < B.P && ok; q++) {
      const int* cnt = (const int*)B.box[q];
      const int target = B.nwg[q] * (i + 1);
      for (unsigned k = 0; __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target; k++) {
        if ((k & 15) == 15) {
          if (__hip_atomic_load(&dsync[kStop], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            ok = 0;
            break;
          }
          if (__hip_atomic_load(&link->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0) {
            loop_fail(dsync, link, 2);
            ok = 0;
            break;
          }
          if (wall_clock64() - t0 > limit) {
            loop_fail(dsync, link, 1);
            ok = 0;
            break;
          }
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    sh[0] = ok;
  }
  __syncthreads();
  const bool ok = sh[0] != 0;
  __syncthreads();
  if (!ok) return false;
  const int len = m + 3 - i;
  const long off = (long)(i & 1) * (m + 3);
  constexpr int kPer = (64 * KM + 3 + 255) / 256;
  double v[kPer];
#pragma unroll
  for (int qq = 0; qq < kPer; qq++) {
    const int r = threadIdx.x + 256 * qq;
    v[qq] = r < len ? ld_dev((const double*)(B.box[B.me] + kBoxMsgOffset) + off + r) : 0.0;
  }
  for (int q = 0; q < B.P; q++) {
    if (q == B.me) continue;
    const double* x = (const double*)(B.box[q] + kBoxMsgOffset) + off;
    double y[kPer];
#pragma unroll
    for (int qq = 0; qq < kPer; qq++) {
      const int r = threadIdx.x + 256 * qq;
      y[qq] = r < len ? ld_dev(x + r) : 0.0;
    }
#pragma unroll
    for (int qq = 0; qq < kPer; qq++) v[qq] = v[qq] + y[qq];
  }
#pragma unroll
  for (int qq = 0; qq < kPer; qq++) {
    const int r = threadIdx.x + 256 * qq;
    if (r < len) gl[r] = v[qq];
  }
  __syncthreads();
  return true;
}

// ---- One owner: partial rows and message slots polled for the sentinel ----
//
// Workspace (exc_ralha_loop_workspace): solo[(2 G + 3) L], L = loop_ld(m): partial row v of step
// j at solo + ((j & 1) G + v) L, message slot s (= j % 3) at solo + (2 G + s) L.
__host__ __device__ inline int loop_ld(int m) { return (m + 3 + 31) / 32 * 32; }

__device__ __forceinline__ bool is_sentinel(double v) {
  return (unsigned long long)__double_as_longlong(v) == kSentinel;
}

// Every 16th poll: false when another workgroup gave up or the wall-clock limit passed.
__device__ __forceinline__ bool poll_ok(unsigned k, long long t0, long long limit, int* dsync, LoopLink* link) {
  if ((k & 15) != 15) return true;
  if (__hip_atomic_load(&dsync[kStop], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return false;
  if (wall_clock64() - t0 > limit) {
    loop_fail(dsync, link, 1);
    return false;
  }
  return true;
}

__global__ void k_solo_fill(double* __restrict__ p, long count) {
  const double sv = __longlong_as_double((long long)kSentinel);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x) p[e] = sv;
}

// Message j of one owner: this workgroup's slice of the G partial rows of step j, polled until no
// entry is the sentinel and summed in the order of partials_sum, goes to slot j % 3. The partial
// entries read go back to the sentinel, and so does this slice of message j - 1's slot (all G
// partials of step j have arrived, so every workgroup has read message j - 1); the last slice
// also resets entry len, which message j - 1 had and message j lacks.
__device__ __forceinline__ bool solo_publish(int j, int m, double* __restrict__ solo, int* dsync, LoopLink* link,
                                             long long limit, double (*acc)[16]) {
  const int G = gridDim.x, w = blockIdx.x, L = loop_ld(m);
  const double sv = __longlong_as_double((long long)kSentinel);
  const int len = m + 3 - j;
  const int valid = (m - 1 - j) + (j > 0 ? 4 : 0);
  const int per = (len + G - 1) / G;
  const int r0 = w * per, r1 = min(len, r0 + per);
  double* P = solo + (long)(j & 1) * G * L;
  double* M = solo + (2L * G + j % 3) * L;
  const int gq = threadIdx.x >> 4, l = threadIdx.x & 15;
  int ok = 1;
  for (int base = r0; base < r1; base += 16) {
    const int e = base + l;
    const bool active = e < r1 && e < valid;
    double s = 0.0;
    for (int v0 = gq; v0 < G; v0 += 128) {
      double x[8];
#pragma unroll
      for (int q = 0; q < 8; q++) x[q] = (active && v0 + 16 * q < G) ? ld_dev(&P[(long)(v0 + 16 * q) * L + e]) : 0.0;
      const long long t0 = wall_clock64();
      for (unsigned k = 0; ok; k++) {
        bool pend = false;
#pragma unroll
        for (int q = 0; q < 8; q++) pend |= is_sentinel(x[q]);
        if (!pend) break;
#pragma unroll
        for (int q = 0; q < 8; q++)
          if (is_sentinel(x[q])) x[q] = ld_dev(&P[(long)(v0 + 16 * q) * L + e]);
        if (!poll_ok(k, t0, limit, dsync, link)) ok = 0;
      }
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (active && v0 + 16 * q < G) s += x[q];
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (active && v0 + 16 * q < G) st_dev(&P[(long)(v0 + 16 * q) * L + e], sv);
    }
    acc[gq][l] = s;
    __syncthreads();
    if (threadIdx.x < 16 && e < r1) {
      double t = acc[0][l];
#pragma unroll
      for (int q = 1; q < 16; q++) t += acc[q][l];
      st_dev(&M[e], t);
    }
    __syncthreads();
  }
  if (j >= 1 && r0 < r1) {
    double* Mp = solo + (2L * G + (j - 1) % 3) * L;
    const int p1 = r1 == len ? len + 1 : r1;
    for (int e = r0 + threadIdx.x; e < p1; e += blockDim.x) st_dev(&Mp[e], sv);
  }
  return __syncthreads_and(ok) != 0;
}

// The whole message i (m + 3 - i doubles) from slot i % 3 into gl, polled until complete; then
// this workgroup's stores (the resets) are waited for, before its next partial row.
template <int KM>
__device__ __forceinline__ bool solo_receive(int i, int m, const double* __restrict__ solo, double* gl, int* dsync,
                                             LoopLink* link, long long limit) {
  const int G = gridDim.x, L = loop_ld(m);
  const int len = m + 3 - i;
  const double* M = solo + (2L * G + i % 3) * L;
  constexpr int kPer = (64 * KM + 3 + 255) / 256;
  double v[kPer];
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int r = threadIdx.x + 256 * q;
    v[q] = r < len ? ld_dev(&M[r]) : 0.0;
  }
  int ok = 1;
  const long long t0 = wall_clock64();
  for (unsigned k = 0; ok; k++) {
    bool pend = false;
#pragma unroll
    for (int q = 0; q < kPer; q++) pend |= is_sentinel(v[q]);
    if (!pend) break;
#pragma unroll
    for (int q = 0; q < kPer; q++)
      if (is_sentinel(v[q])) v[q] = ld_dev(&M[threadIdx.x + 256 * q]);
    if (!poll_ok(k, t0, limit, dsync, link)) ok = 0;
  }
#pragma unroll
  for (int q = 0; q < kPer; q++) {
    const int r = threadIdx.x + 256 * q;
    if (r < len) gl[r] = v[q];
  }
  stores_done();
  return __syncthreads_and(ok) != 0;
}

template <int KM, int CPW>
__global__ __launch_bounds__(256, KM * CPW <= 32 ? 2 : 1) void k_ralha_loop(int m, int c, double* __restrict__ Fc, double* __restrict__ Uc,
                                                    double* __restrict__ house, double* __restrict__ tau,
                                                    double* __restrict__ alpha, double* __restrict__ beta,
                                                    char* link_bytes, int* dsync, double* __restrict__ gdev,
                                                    double* __restrict__ partial, double* __restrict__ solo,
                                                    long long limit, PeerBoxes B) {
  extern __shared__ double lds[];  // hbuf: m, mp: 4 x (m + 4), gl: m + 3
  double* hbuf = lds;
  double* mp = lds + m;
  double* gl = mp + 4 * (m + 4);
  __shared__ double red[8];
  __shared__ double acc[16][16];
  __shared__ int sh[2];
  LoopLink* link = (LoopLink*)link_bytes;
  double* hmsg = (double*)(link_bytes + kLinkMsgOffset);
  double* mine = partial + (long)blockIdx.x * (m + 3);
  RegCols<KM, CPW> R;
  reg_load(R, -1, m, c, Fc, Uc);
  const bool boxes = B.P > 0;   // uniform
  const bool owner = B.P == 1;  // one owner: the sentinel-polled buffers
  // this workgroup's partial row of message j
  auto row = [&](int j) { return owner ? solo + ((long)(j & 1) * gridDim.x + blockIdx.x) * loop_ld(m) : mine; };
  auto publish = [&](int j) {
    return owner ? solo_publish(j, m, solo, dsync, link, limit, acc)
                 : boxes ? box_publish(j, m, partial, B, link, dsync, limit, sh, acc)
                         : loop_publish(j, m, partial, link, hmsg, dsync, limit, sh, acc);
  };
  reg_share<KM, CPW, true>(R, 0, m, c, mp, row(0));
  if (!publish(0)) return;
  for (int i = 0; i <= m - 3; i++) {
    if (!(owner ? solo_receive<KM>(i, m, solo, gl, dsync, link, limit)
                : boxes ? box_receive<KM>(i, m, B, link, gl, dsync, limit, sh)
                        : loop_receive<KM>(i, m, link, hmsg, gdev, gl, dsync, limit, sh)))
      return;
    reg_apply(R, i, m, c, gl, hbuf, red, Uc, house, tau, alpha, beta);
#pragma unroll
    for (int cc = 0; cc < CPW; cc++) R.um1[cc] = R.ui[cc];
    if (i + 1 <= m - 3) {
      reg_share<KM, CPW, true>(R, i + 1, m, c, mp, row(i + 1));
      if (!publish(i + 1)) return;
    }
  }
  reg_store(R, 0, m, c, Fc);
}

// The device words and every workgroup's barrier flag to zero (256 threads).
__global__ void k_loop_clear(int* dsync) {
  if (threadIdx.x < 64) dsync[threadIdx.x] = 0;
  barrier_flags(dsync)[threadIdx.x] = 0;
}

// One column step of the overlapped Ralha loop on column-major F and U (Fc[col*m + row]):
//  apply (i >= 0): the reduced message g of step i -> LAPACK dlarfg (n = m-1-i) into column i+1
//    of house and tau[i+1]; F[i+1:, col] -= tau h (h . F[i+1:, col]) for this workgroup's columns
//    (one wavefront per column); f_i -= beta_{i-1} u_{i-1}, alpha_i, u_i = f_i / alpha_i
//    (client.cpp:858-891);
//  message (j = i+1 <= m-3): this workgroup's share of step j's message (f_{j+1+r} . f_j and
//    the four dots with u_{j-1}, client.cpp:842-855) into its row of `partial`.
__global__ __launch_bounds__(256) void k_ralha_step(int i, int m, int c, double* __restrict__ Fc,
                                                    double* __restrict__ Uc, const double* __restrict__ g,
                                                    double* __restrict__ house, double* __restrict__ tau,
                                                    double* __restrict__ alpha, double* __restrict__ beta,
                                                    double* __restrict__ partial) {
  extern __shared__ double hbuf[];  // m doubles
  __shared__ double red[8];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int c0 = blockIdx.x * kColsPerWg;
  if (i >= 0) {
    const int n = m - 1 - i;
    double b, a;
    double ia;
  const double t = ralha_reflector(i, m, g, hbuf, red, house, tau, alpha, beta, &b, &a, &ia);
    for (int q = wave; q < kColsPerWg; q += (int)(blockDim.x >> 6)) {
      const int col = c0 + q;
      if (col >= c) break;
      double* f = Fc + (long)col * m;
      if (t != 0.0) {
        double s = 0.0;
        for (int r = lane; r < n; r += 64) s += f[i + 1 + r] * hbuf[r];
        s = wave_sum(s);
        for (int r = lane; r < n; r += 64) f[i + 1 + r] += s * (-t * hbuf[r]);
      }
      if (i > 0 && lane == 0) {
        const double fi = f[i] + (-b) * Uc[(long)col * m + i - 1];
        f[i] = fi;
        Uc[(long)col * m + i] = (1.0 / a) * fi;
      }
    }
    __syncthreads();
  }
  const int j = i + 1;
  if (j > m - 3) return;
  const int nrows = m - 1 - j;
  const int c1 = min(c, c0 + kColsPerWg);
  double* out = partial + (long)blockIdx.x * (m + 3);
  for (int r = tid; r < nrows; r += blockDim.x) {
    double s = 0.0;
    for (int col = c0; col < c1; col++) s += Fc[(long)col * m + j + 1 + r] * Fc[(long)col * m + j];
    out[r] = s;
  }
  if (j > 0 && tid < 4) {
    double s = 0.0;
    for (int col = c0; col < c1; col++) {
      const double f = Fc[(long)col * m + j], u = Uc[(long)col * m + j - 1];
      s += tid == 0 ? f * u : (tid == 1 ? u * u : (tid == 2 ? f * f : u * f));
    }
    out[nrows + tid] = s;
  }
}

// msg[r] = sum over the step workgroups of partial[wg][r]; zeros after the valid entries. A
// workgroup owns 16 message entries (partials_sum gives the order).
__global__ __launch_bounds__(256) void k_ralha_reduce(int j, int m, int nwg, const double* __restrict__ partial,
                                                      double* __restrict__ msg) {
  __shared__ double acc[16][16];
  const int e = blockIdx.x * 16 + (threadIdx.x & 15);
  const int len = m + 3 - j;
  const int valid = (m - 1 - j) + (j > 0 ? 4 : 0);
  const double v = partials_sum<false>(partial, m + 3, nwg, e, e < valid, acc);
  if (threadIdx.x < 16 && e < len) msg[e] = v;
}

// dot = f_i . u_{i-1} over the columns (column-major F and U)
__global__ __launch_bounds__(256) void k_ralha_dot(int i, int m, int c, const double* __restrict__ Fc,
                                                   const double* __restrict__ Uc, double* __restrict__ out) {
  __shared__ double red[8];
  double s = 0.0;
  for (int k = threadIdx.x; k < c; k += blockDim.x) s += Fc[(long)k * m + i] * Uc[(long)k * m + i - 1];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = s;
}

// f_i -= beta u_{i-1} (beta = reduced value), then out = ||f_i||^2 (i = -1: ||f_0||^2 only)
__global__ __launch_bounds__(256) void k_ralha_tail_update(int i, int m, int c, double* __restrict__ Fc,
                                                           const double* __restrict__ Uc,
                                                           const double* __restrict__ breduced,
                                                           double* __restrict__ beta, double* __restrict__ out) {
  __shared__ double red[8];
  const int row = i < 0 ? 0 : i;
  double s = 0.0;
  const double b = i > 0 ? breduced[0] : 0.0;
  for (int k = threadIdx.x; k < c; k += blockDim.x) {
    double f = Fc[(long)k * m + row];
    if (i > 0) {
      f = f + (-b) * Uc[(long)k * m + i - 1];
      Fc[(long)k * m + row] = f;
    }
    s += f * f;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    out[0] = s;
    if (i > 0) beta[i - 1] = b;
  }
}

// alpha_i = sqrt(reduced ||f_i||^2); u_i = f_i / alpha_i
__global__ __launch_bounds__(256) void k_ralha_tail_scale(int i, int m, int c, const double* __restrict__ Fc,
                                                          double* __restrict__ Uc, const double* __restrict__ n2,
                                                          double* __restrict__ alpha) {
  const double a = sqrt(n2[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) alpha[i] = a;
  const double inv = 1.0 / a;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x)
    Uc[(long)k * m + i] = inv * Fc[(long)k * m + i];
}

// Lower to upper bidiagonal with the reference BLAS drotg (client.cpp:947-954), one thread. The
// updated alpha_{i+1} is carried in a register, so the loads of beta and alpha ahead can be
// issued before the dependent chain needs them.
__global__ void k_lower_to_upper(int m, double* __restrict__ alpha, double* __restrict__ beta,
                                 double* __restrict__ cs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = alpha[0];
#pragma unroll 8
  for (int i = 0; i < m - 1; i++) {
    const double b = beta[i], an = alpha[i + 1];
    double roe = fabs(a) > fabs(b) ? a : b;
    double scl = fabs(a) + fabs(b);
    double c, s, r;
    if (scl == 0.0) {
      c = 1.0; s = 0.0; r = 0.0;
    } else {
      const double as = a / scl, bs = b / scl;
      r = scl * sqrt(as * as + bs * bs);
      r = copysign(1.0, roe) * r;
      c = a / r;
      s = b / r;
    }
    alpha[i] = r;
    cs[2 * i] = c;
    cs[2 * i + 1] = s;
    beta[i] = s * an;
    a = c * an;
  }
  alpha[m - 1] = a;
}

// Sign rule of client.cpp:987-995: flip[j] = -1 when the first largest-magnitude entry of
// column j of bu (m x k) is negative. A workgroup takes 32 columns; its 8 row groups scan every
// eighth row (coalesced across the columns), then the candidates are merged keeping the larger
// magnitude and, on ties, the smaller row (cblas_idamax's first maximum).
__global__ __launch_bounds__(256) void k_sign_flags(int m, int k, const double* __restrict__ bu,
                                                    double* __restrict__ flip) {
  __shared__ double bv[8][32];
  __shared__ int bi[8][32];
  const int cl = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int j = blockIdx.x * 32 + cl;
  int p = -1;
  double best = -1.0;
  if (j < k) {
#pragma unroll 8
    for (int i = g; i < m; i += 8) {
      const double v = fabs(bu[(long)i * k + j]);
      if (v > best) { best = v; p = i; }
    }
  }
  bv[g][cl] = best;
  bi[g][cl] = p;
  __syncthreads();
  if (g == 0 && j < k) {
    for (int q = 1; q < 8; q++) {
      const double v = bv[q][cl];
      const int iq = bi[q][cl];
      if (iq >= 0 && (v > best || (v == best && iq < p))) { best = v; p = iq; }
    }
    flip[j] = (p >= 0 && bu[(long)p * k + j] < 0) ? -1.0 : 1.0;
  }
}

__global__ __launch_bounds__(256) void k_apply_flags(int m, int k, double* __restrict__ bu, double* __restrict__ bv,
                                                     const double* __restrict__ flip) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * k) return;
  const long i = e / k, j = e - i * k;
  bu[e] *= flip[j];            // column j of bu (m x k)
  bv[j * m + i] *= flip[j];    // row j of bv (k x m)
}

// Back-rotation of client.cpp:1127-1128: for i = m-2..0, rows (i, i+1) of bu (m x k) rotated by
// drot with (c_i, -s_i). One lane per column carries row i+1 down the recurrence.
__global__ __launch_bounds__(64) void k_rot_back(int m, int k, double* __restrict__ bu,
                                                 const double* __restrict__ cs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k || m < 2) return;
  // Rows are read 16 at a time into registers before the 16 rotations that use them: the
  // stores of a block go to rows at or above its first row, which were all read already, so
  // the loads of a block are issued together instead of each waiting behind the previous
  // step's store. Every rotation evaluates the same expressions in the same order.
  constexpr int RB = 16;
  double y = bu[(long)(m - 1) * k + j];
  int i = m - 2;
  for (; i >= RB - 1; i -= RB) {
    double x[RB];
#pragma unroll
    for (int q = 0; q < RB; q++) x[q] = bu[(long)(i - q) * k + j];
#pragma unroll
    for (int q = 0; q < RB; q++) {
      const int r = i - q;
      const double c = cs[2 * r], s = -cs[2 * r + 1];
      const double xn = c * x[q] + s * y;
      const double yn = c * y - s * x[q];
      bu[(long)(r + 1) * k + j] = yn;
      y = xn;
    }
  }
  for (; i >= 0; i--) {
    const double c = cs[2 * i], s = -cs[2 * i + 1];
    const double x = bu[(long)i * k + j];
    const double xn = c * x + s * y;
    const double yn = c * y - s * x;
    bu[(long)(i + 1) * k + j] = yn;
    y = xn;
  }
  bu[j] = y;
}

}  // namespace

extern "C" {

size_t exc_ralha_workspace(int m, int c) {
  return (size_t)exc_detail::ceil_div(c, kColsPerWg) * (m + 3) * sizeof(double) + 256;
}

}  // extern "C"

namespace {

template <int KM, int CPW>
int launch_step_reg(int i, int m, int c, double* fc, double* uc, const double* g, double* house, double* tau,
                    double* alpha, double* beta, double* partial, hipStream_t s, unsigned* nwg_out) {
  const unsigned nwg = exc_detail::ceil_div(c, 4 * CPW);
  const size_t lds = (size_t)(5 * m + 16) * sizeof(double);
  if (lds > 64 * 1024)
    EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_ralha_step_reg<KM, CPW>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_ralha_step_reg<KM, CPW>), dim3(nwg), dim3(256), lds, s, i, m, c, fc, uc, g, house, tau, alpha,
                     beta, partial);
  EXC_CHECK_HIP(hipGetLastError());
  *nwg_out = nwg;
  return EXC_OK;
}

int ralha_step(bool generic, int i, int m, int c, double* d_fc, double* d_uc, const double* d_g, double* d_house,
               double* d_tau, double* d_alpha, double* d_beta, double* d_msg, void* d_work, size_t work_bytes,
               exc_stream_t stream) {
  EXC_REQUIRE(m >= 3 && i >= -1 && i <= m - 3 && c >= 1 && d_fc && d_uc && (i < 0 || d_g) && d_house && d_tau &&
                  d_alpha && d_beta && d_msg && d_work,
              EXC_ERR_ARG, "exc_ralha_step: bad arguments");
  EXC_REQUIRE(work_bytes >= (size_t)exc_detail::ceil_div(c, kColsPerWg) * (m + 3) * sizeof(double), EXC_ERR_SPACE,
              "exc_ralha_step: workspace too small");
  hipStream_t s = exc_detail::stream(stream);
  double* partial = (double*)d_work;
  unsigned nwg = 0;
  int rc = EXC_OK;
  if (!generic && m <= 256) rc = launch_step_reg<4, 2>(i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, partial, s, &nwg);
  else if (!generic && m <= 512) rc = launch_step_reg<8, 2>(i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, partial, s, &nwg);
  else if (!generic && m <= 1024) rc = launch_step_reg<16, 2>(i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, partial, s, &nwg);
  else if (!generic && m <= 2048) rc = launch_step_reg<32, 1>(i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, partial, s, &nwg);
  else if (!generic && m <= 4000) rc = launch_step_reg<64, 1>(i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, partial, s, &nwg);
  else {
    nwg = exc_detail::ceil_div(c, kColsPerWg);
    const size_t lds = (size_t)m * sizeof(double);
    EXC_REQUIRE(lds <= 160 * 1024, EXC_ERR_ARG, "exc_ralha_step: m above 20480 does not fit the reflector in LDS");
    if (lds > 64 * 1024)
      EXC_CHECK_HIP(
          hipFuncSetAttribute((const void*)k_ralha_step, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_ralha_step, dim3(nwg), dim3(256), lds, s, i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha,
                       d_beta, partial);
    EXC_CHECK_HIP(hipGetLastError());
  }
  if (rc) return rc;
  if (i + 1 <= m - 3) {
    hipLaunchKernelGGL(k_ralha_reduce, dim3(exc_detail::ceil_div(m + 3, 16)), dim3(256), 0, s, i + 1, m, (int)nwg,
                       partial, d_msg);
    EXC_CHECK_HIP(hipGetLastError());
  }
  return EXC_OK;
}

}  // namespace

extern "C" {

int exc_ralha_step(int i, int m, int c, double* d_fc, double* d_uc, const double* d_g, double* d_house,
                   double* d_tau, double* d_alpha, double* d_beta, double* d_msg, void* d_work, size_t work_bytes,
                   exc_stream_t stream) {
  return ralha_step(false, i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, d_msg, d_work, work_bytes,
                    stream);
}

int exc_ralha_step_generic(int i, int m, int c, double* d_fc, double* d_uc, const double* d_g, double* d_house,
                           double* d_tau, double* d_alpha, double* d_beta, double* d_msg, void* d_work,
                           size_t work_bytes, exc_stream_t stream) {
  return ralha_step(true, i, m, c, d_fc, d_uc, d_g, d_house, d_tau, d_alpha, d_beta, d_msg, d_work, work_bytes,
                    stream);
}

// [device words and barrier flags][gdev: m + 3][partial rows: G x (m + 3)][one owner's sentinel
// buffers: (2 G + 3) x loop_ld(m)], G the largest grid (one column per wavefront).
size_t exc_ralha_loop_workspace(int m, int c) {
  if (m < 3 || c < 1) return 0;
  const size_t g = ((size_t)(m + 3) * sizeof(double) + 255) / 256 * 256;
  const size_t G = exc_detail::ceil_div(c, 4);
  const size_t part = (G * (m + 3) * sizeof(double) + 255) / 256 * 256;
  return kSyncBytes + g + part + (2 * G + 3) * (size_t)loop_ld(m) * sizeof(double);
}

int exc_ralha_link_alloc(int m, void** h_link) {
  EXC_REQUIRE(m >= 3 && h_link, EXC_ERR_ARG, "exc_ralha_link_alloc: bad arguments");
  EXC_CHECK_HIP(hipHostMalloc(h_link, kLinkMsgOffset + (size_t)(m + 3) * sizeof(double),
                              hipHostMallocCoherent | hipHostMallocMapped));
  memset(*h_link, 0, kLinkMsgOffset + (size_t)(m + 3) * sizeof(double));
  ((LoopLink*)*h_link)->m = m;
  return EXC_OK;
}

int exc_ralha_link_free(void* h_link) {
  if (h_link) EXC_CHECK_HIP(hipHostFree(h_link));
  return EXC_OK;
}

}  // extern "C"

namespace {

struct LoopShape {
  int km = 0, cpw = 0, nwg = 0;
  size_t lds = 0;
};

int loop_km(int m) { return m <= 256 ? 4 : m <= 512 ? 8 : m <= 1024 ? 16 : m <= 2048 ? 32 : m <= 4000 ? 64 : 0; }

// Columns per wavefront: the fewest that keep the grid at 64 workgroups or below, at most 2 and
// within KM * CPW <= 64 registers of columns (a hint of 1 or 2 overrides, within the same bound).
// Four columns per wavefront is not offered: the 32-workgroup grid of k_ralha_loop<16, 4>
// (256 VGPRs + 132 AGPRs, one workgroup per CU) ended a two-process run on one device with a
// GPU memory fault whose cause has not been found, so that form is not built.
LoopShape loop_shape(int m, int c, int hint) {
  LoopShape sh;
  sh.km = m <= kLoopMaxM ? loop_km(m) : 0;
  if (!sh.km) return sh;
  const int maxc = sh.km >= 64 ? 1 : 2;
  int cpw = 1;
  if (hint != 0) {
    if (hint != 1 && hint != 2) return LoopShape{};
    cpw = hint;
  } else {
    while (cpw < maxc && (long)exc_detail::ceil_div(c, 4 * cpw) > 64) cpw *= 2;
  }
  if (cpw > maxc) return LoopShape{};
  sh.cpw = cpw;
  sh.nwg = (int)exc_detail::ceil_div(c, 4 * cpw);
  sh.lds = (size_t)(6 * m + 19) * sizeof(double);
  return sh;
}

template <int KM, int CPW>
const void* loop_kernel() {
  return (const void*)k_ralha_loop<KM, CPW>;
}

const void* loop_kernel_for(const LoopShape& sh) {
#define EXC_LOOP_CASE(K, C) \
  if (sh.km == K && sh.cpw == C) return loop_kernel<K, C>();
  EXC_LOOP_CASE(4, 1) EXC_LOOP_CASE(4, 2)
  EXC_LOOP_CASE(8, 1) EXC_LOOP_CASE(8, 2)
  EXC_LOOP_CASE(16, 1) EXC_LOOP_CASE(16, 2)
  EXC_LOOP_CASE(32, 1) EXC_LOOP_CASE(32, 2)
  EXC_LOOP_CASE(64, 1)
#undef EXC_LOOP_CASE
  return nullptr;
}

// True when `need` copies of the grid fit on the device at once: two for peers that share the
// device (this peer's loop and another's), one for a loop that waits on no other grid.
int loop_resident(const LoopShape& sh, const void* fn, bool* fits, int need = 2) {
  *fits = false;
  if (!fn || sh.nwg < 1 || sh.nwg > kLoopMaxWg) return EXC_OK;
  int dev = 0, cus = 0, per_cu = 0;
  EXC_CHECK_HIP(hipGetDevice(&dev));
  EXC_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  if (sh.lds > 64 * 1024)
    EXC_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds));
  EXC_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, sh.lds));
  *fits = (long)per_cu * cus >= (long)need * sh.nwg;
  return EXC_OK;
}

double wall_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int exc_ralha_loop_fits(int m, int c, int cols_per_wave) {
  if (m < 3 || c < 1) return 0;
  const LoopShape sh = loop_shape(m, c, cols_per_wave);
  bool fits = false;
  if (loop_resident(sh, loop_kernel_for(sh), &fits) != EXC_OK) return 0;
  return fits ? 1 : 0;
}

int exc_ralha_loop_shape(int m, int c, int cols_per_wave, int* nwg, size_t* lds_bytes) {
  if (nwg) *nwg = 0;
  if (lds_bytes) *lds_bytes = 0;
  if (m < 3 || c < 1) return EXC_OK;
  const LoopShape sh = loop_shape(m, c, cols_per_wave);
  if (!sh.km || sh.nwg > kLoopMaxWg) return EXC_OK;
  if (nwg) *nwg = sh.nwg;
  if (lds_bytes) *lds_bytes = sh.lds;
  return EXC_OK;
}

int exc_ralha_loop_slots(int m, int c, int cols_per_wave) {
  if (m < 3 || c < 1) return 0;
  const LoopShape sh = loop_shape(m, c, cols_per_wave);
  const void* fn = loop_kernel_for(sh);
  if (!fn || sh.nwg < 1 || sh.nwg > kLoopMaxWg) return 0;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (sh.lds > 64 * 1024 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh.lds) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, sh.lds) != hipSuccess) return 0;
  return per_cu * cus;
}

int exc_ralha_loops_colocate(int P, const int* nwg, const int* slots, int same_device) {
  if (P < 1 || !nwg || !slots) return 0;
  long total = 0, fewest = -1;
  for (int q = 0; q < P; q++) {
    if (nwg[q] < 1 || slots[q] < nwg[q]) return 0;  // a loop that does not fit even alone
    total += nwg[q];
    fewest = fewest < 0 ? slots[q] : std::min<long>(fewest, slots[q]);
  }
  // On one device every peer's loop waits on every other's message, so all grids must be
  // resident together; on separate devices each grid only needs its own device.
  return (!same_device || total <= fewest) ? 1 : 0;
}

}  // extern "C"

namespace {

__global__ void k_box_clear(int* box) {
  if (threadIdx.x == 0) box[0] = 0;
}

int loop_start(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau, double* d_alpha,
               double* d_beta, void* h_link, void* d_work, size_t work_bytes, int cols_per_wave, double timeout_s,
               const PeerBoxes& B, exc_stream_t stream) {
  EXC_REQUIRE(m >= 3 && c >= 1 && d_fc && d_uc && d_house && d_tau && d_alpha && d_beta && h_link && d_work &&
                  timeout_s > 0,
              EXC_ERR_ARG, "exc_ralha_loop_start: bad arguments");
  EXC_REQUIRE(work_bytes >= exc_ralha_loop_workspace(m, c), EXC_ERR_SPACE, "exc_ralha_loop_start: workspace too small");
  LoopLink* link = (LoopLink*)h_link;
  EXC_REQUIRE(link->m == m, EXC_ERR_ARG, "exc_ralha_loop_start: link allocated for another m");
  const LoopShape sh = loop_shape(m, c, cols_per_wave);
  const void* fn = loop_kernel_for(sh);
  bool fits = false;
  const int rc = loop_resident(sh, fn, &fits, B.P == 1 ? 1 : 2);
  if (rc) return rc;
  EXC_REQUIRE(fits, EXC_ERR_ARG, "exc_ralha_loop_start: the grid would not be resident (see exc_ralha_loop_fits)");
  if (B.P > 0)
    EXC_REQUIRE(B.nwg[B.me] == sh.nwg, EXC_ERR_ARG, "exc_ralha_loop_start_boxes: nwg[me] is not this grid's size");
  hipStream_t s = exc_detail::stream(stream);
  EXC_CHECK_HIP(hipStreamSynchronize(s));  // no earlier loop still uses the link
  memset(link, 0, kLinkMsgOffset);
  link->nwg = sh.nwg;
  link->m = m;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  void* d_link = nullptr;
  EXC_CHECK_HIP(hipHostGetDevicePointer(&d_link, h_link, 0));
  int dev = 0, khz = 0;
  EXC_CHECK_HIP(hipGetDevice(&dev));
  EXC_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
  EXC_REQUIRE(khz > 0, EXC_ERR_HIP, "exc_ralha_loop_start: no wall clock rate");
  const long long limit = (long long)(timeout_s * 1e3 * khz);
  char* w = (char*)d_work;
  int* dsync = (int*)w;
  const size_t g = ((size_t)(m + 3) * sizeof(double) + 255) / 256 * 256;
  const size_t part = ((size_t)exc_detail::ceil_div(c, 4) * (m + 3) * sizeof(double) + 255) / 256 * 256;
  double* gdev = (double*)(w + kSyncBytes);
  double* partial = (double*)(w + kSyncBytes + g);
  double* solo = (double*)(w + kSyncBytes + g + part);
  hipLaunchKernelGGL(k_loop_clear, dim3(1), dim3(kLoopMaxWg), 0, s, dsync);
  if (B.P == 1) {
    const long count = (2L * sh.nwg + 3) * loop_ld(m);
    hipLaunchKernelGGL(k_solo_fill, dim3(std::min<long>(1024, (count + 255) / 256)), dim3(256), 0, s, solo, count);
  }
  PeerBoxes b = B;
  void* args[] = {&m, &c, &d_fc, &d_uc, &d_house, &d_tau, &d_alpha, &d_beta, &d_link, &dsync, &gdev, &partial,
                  &solo, (void*)&limit, &b};
  EXC_CHECK_HIP(hipLaunchKernel(fn, dim3(sh.nwg), dim3(256), args, sh.lds, s));
  return EXC_OK;
}

}  // namespace

extern "C" {

int exc_ralha_loop_start(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau, double* d_alpha,
                         double* d_beta, void* h_link, void* d_work, size_t work_bytes, int cols_per_wave,
                         double timeout_s, exc_stream_t stream) {
  PeerBoxes none{};
  return loop_start(m, c, d_fc, d_uc, d_house, d_tau, d_alpha, d_beta, h_link, d_work, work_bytes, cols_per_wave,
                    timeout_s, none, stream);
}

int exc_ralha_loop_grid(int m, int c, int cols_per_wave) {
  return exc_ralha_loop_fits(m, c, cols_per_wave) ? loop_shape(m, c, cols_per_wave).nwg : 0;
}

size_t exc_ralha_box_bytes(int m) { return m < 3 ? 0 : kBoxMsgOffset + 2 * (size_t)(m + 3) * sizeof(double); }

int exc_ralha_box_clear(void* d_box, exc_stream_t stream) {
  EXC_REQUIRE(d_box, EXC_ERR_ARG, "exc_ralha_box_clear: null box");
  hipLaunchKernelGGL(k_box_clear, dim3(1), dim3(64), 0, exc_detail::stream(stream), (int*)d_box);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_loop_start_boxes(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau,
                               double* d_alpha, double* d_beta, void* h_link, void* d_work, size_t work_bytes,
                               int cols_per_wave, double timeout_s, int npeers, int me, void* const* d_boxes,
                               const int* nwg, exc_stream_t stream) {
  EXC_REQUIRE(npeers >= 1 && npeers <= kMaxPeers && me >= 0 && me < npeers && d_boxes && nwg, EXC_ERR_ARG,
              "exc_ralha_loop_start_boxes: bad peer arguments (1 <= npeers <= 16)");
  PeerBoxes B{};
  B.P = npeers;
  B.me = me;
  for (int q = 0; q < npeers; q++) {
    EXC_REQUIRE(d_boxes[q] && nwg[q] >= 1 && nwg[q] <= kLoopMaxWg, EXC_ERR_ARG,
                "exc_ralha_loop_start_boxes: bad box or grid size");
    B.box[q] = (char*)d_boxes[q];
    B.nwg[q] = nwg[q];
  }
  return loop_start(m, c, d_fc, d_uc, d_house, d_tau, d_alpha, d_beta, h_link, d_work, work_bytes, cols_per_wave,
                    timeout_s, B, stream);
}

int exc_ralha_loop_running(exc_stream_t stream) {
  const hipError_t e = hipStreamQuery(exc_detail::stream(stream));
  if (e == hipErrorNotReady) return 1;
  EXC_CHECK_HIP(e);
  return 0;
}

int exc_ipc_handle(void* d_ptr, void* handle) {
  EXC_REQUIRE(d_ptr && handle, EXC_ERR_ARG, "exc_ipc_handle: bad arguments");
  static_assert(sizeof(hipIpcMemHandle_t) == EXC_IPC_HANDLE_BYTES, "IPC handle size");
  hipIpcMemHandle_t h;
  EXC_CHECK_HIP(hipIpcGetMemHandle(&h, d_ptr));
  memcpy(handle, &h, sizeof h);
  return EXC_OK;
}

int exc_ipc_open(const void* handle, void** d_ptr) {
  EXC_REQUIRE(handle && d_ptr, EXC_ERR_ARG, "exc_ipc_open: bad arguments");
  hipIpcMemHandle_t h;
  memcpy(&h, handle, sizeof h);
  EXC_CHECK_HIP(hipIpcOpenMemHandle(d_ptr, h, hipIpcMemLazyEnablePeerAccess));
  return EXC_OK;
}

int exc_ipc_close(void* d_ptr) {
  if (d_ptr) EXC_CHECK_HIP(hipIpcCloseMemHandle(d_ptr));
  return EXC_OK;
}

int exc_ralha_loop_wait(void* h_link, int j, double timeout_s, double** h_msg) {
  EXC_REQUIRE(h_link && h_msg && j >= 0, EXC_ERR_ARG, "exc_ralha_loop_wait: bad arguments");
  LoopLink* link = (LoopLink*)h_link;
  const int G = link->nwg;
  const double t0 = wall_seconds();
  for (int w = 0; w < G; w++) {
    unsigned k = 0;
    while (__atomic_load_n(&link->done[w], __ATOMIC_ACQUIRE) < j + 1) {
      if ((++k & 255) == 0) {
        const int e = __atomic_load_n(&link->error, __ATOMIC_RELAXED);
        EXC_REQUIRE(e == 0, EXC_ERR_NUMERIC,
                    e == 1 ? "exc_ralha_loop_wait: a wait of the loop kernel ran out of time"
                           : "exc_ralha_loop_wait: the loop kernel stopped");
        if (wall_seconds() - t0 > timeout_s) {
          __atomic_store_n(&link->abort, 1, __ATOMIC_RELEASE);
          EXC_REQUIRE(false, EXC_ERR_NUMERIC, "exc_ralha_loop_wait: timed out waiting for the loop kernel");
        }
      }
    }
  }
  *h_msg = (double*)((char*)h_link + kLinkMsgOffset);
  return EXC_OK;
}

int exc_ralha_loop_post(void* h_link, int j) {
  EXC_REQUIRE(h_link && j >= 0, EXC_ERR_ARG, "exc_ralha_loop_post: bad arguments");
  __atomic_store_n(&((LoopLink*)h_link)->in_seq, j + 1, __ATOMIC_RELEASE);
  return EXC_OK;
}

int exc_ralha_loop_abort(void* h_link) {
  EXC_REQUIRE(h_link, EXC_ERR_ARG, "exc_ralha_loop_abort: bad arguments");
  __atomic_store_n(&((LoopLink*)h_link)->abort, 1, __ATOMIC_RELEASE);
  return EXC_OK;
}

int exc_ralha_loop_finish(void* h_link, exc_stream_t stream) {
  EXC_REQUIRE(h_link, EXC_ERR_ARG, "exc_ralha_loop_finish: bad arguments");
  EXC_CHECK_HIP(hipStreamSynchronize(exc_detail::stream(stream)));
  const int e = __atomic_load_n(&((LoopLink*)h_link)->error, __ATOMIC_ACQUIRE);
  EXC_REQUIRE(e == 0, EXC_ERR_NUMERIC,
              e == 1 ? "exc_ralha_loop_finish: a wait of the loop kernel ran out of time"
                     : "exc_ralha_loop_finish: the loop kernel stopped");
  return EXC_OK;
}

int exc_ralha_dot(int i, int m, int c, const double* d_fc, const double* d_uc, double* d_out, exc_stream_t stream) {
  EXC_REQUIRE(i >= 1 && i < m && c >= 1, EXC_ERR_ARG, "exc_ralha_dot: bad arguments");
  hipLaunchKernelGGL(k_ralha_dot, dim3(1), dim3(256), 0, exc_detail::stream(stream), i, m, c, d_fc, d_uc, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_tail_update(int i, int m, int c, double* d_fc, const double* d_uc, const double* d_breduced,
                          double* d_beta, double* d_out, exc_stream_t stream) {
  EXC_REQUIRE(i >= -1 && i < m && c >= 1, EXC_ERR_ARG, "exc_ralha_tail_update: bad arguments");
  hipLaunchKernelGGL(k_ralha_tail_update, dim3(1), dim3(256), 0, exc_detail::stream(stream), i, m, c, d_fc, d_uc,
                     d_breduced, d_beta, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_tail_scale(int i, int m, int c, const double* d_fc, double* d_uc, const double* d_n2, double* d_alpha,
                         exc_stream_t stream) {
  EXC_REQUIRE(i >= 0 && i < m && c >= 1, EXC_ERR_ARG, "exc_ralha_tail_scale: bad arguments");
  hipLaunchKernelGGL(k_ralha_tail_scale, dim3(exc_detail::ceil_div(c, 256)), dim3(256), 0,
                     exc_detail::stream(stream), i, m, c, d_fc, d_uc, d_n2, d_alpha);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_lower_to_upper(int m, double* d_alpha, double* d_beta, double* d_cs, exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && d_alpha && d_beta && d_cs, EXC_ERR_ARG, "exc_lower_to_upper: bad arguments");
  hipLaunchKernelGGL(k_lower_to_upper, dim3(1), dim3(64), 0, exc_detail::stream(stream), m, d_alpha, d_beta, d_cs);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_sign_normalize(int m, int k, double* d_bu, double* d_bv, double* d_flip, exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && k >= 0 && d_bu && d_bv && d_flip, EXC_ERR_ARG, "exc_sign_normalize: bad arguments");
  if (k == 0) return EXC_OK;
  hipStream_t s = exc_detail::stream(stream);
  hipLaunchKernelGGL(k_sign_flags, dim3(exc_detail::ceil_div(k, 32)), dim3(256), 0, s, m, k, d_bu, d_flip);
  hipLaunchKernelGGL(k_apply_flags, dim3(exc_detail::ceil_div((long)m * k, 256)), dim3(256), 0, s, m, k, d_bu, d_bv,
                     d_flip);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_rot_back(int m, int k, double* d_bu, const double* d_cs, exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && k >= 0 && d_bu && d_cs, EXC_ERR_ARG, "exc_rot_back: bad arguments");
  if (k == 0 || m < 2) return EXC_OK;
  hipLaunchKernelGGL(k_rot_back, dim3(exc_detail::ceil_div(k, 64)), dim3(64), 0, exc_detail::stream(stream), m, k,
                     d_bu, d_cs);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // extern "C"

namespace {
// Workspace layout of exc_bidiag_local: [scalars 256 B][box][loop or step workspace][2 messages].
struct LocalWs {
  double* scal;
  void* box;
  char* work;
  size_t work_bytes;
  double* msg[2];
};
size_t local_ws(int m, int c, void* base, LocalWs* w) {
  const size_t box = (exc_ralha_box_bytes(m) + 255) / 256 * 256;
  const size_t lw = std::max(exc_ralha_loop_workspace(m, c), exc_ralha_workspace(m, c));
  const size_t lwp = (lw + 255) / 256 * 256;
  const size_t msg = ((size_t)(m + 3) * sizeof(double) + 255) / 256 * 256;
  if (w) {
    char* p = (char*)base;
    w->scal = (double*)p;
    w->box = p + 256;
    w->work = p + 256 + box;
    w->work_bytes = lwp;
    w->msg[0] = (double*)(p + 256 + box + lwp);
    w->msg[1] = (double*)(p + 256 + box + lwp + msg);
  }
  return 256 + box + lwp + 2 * msg;
}
}  // namespace

extern "C" {

size_t exc_bidiag_local_workspace(int m, int c) { return m < 3 || c < 1 ? 0 : local_ws(m, c, nullptr, nullptr); }

int exc_bidiag_local(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau, double* d_alpha,
                     double* d_beta, double* d_cs, void* h_link, int colocated, void* d_work, size_t work_bytes,
                     double timeout_s, exc_stream_t stream) {
  EXC_REQUIRE(m >= 3 && c >= 1 && d_fc && d_uc && d_house && d_tau && d_alpha && d_beta && d_cs && d_work &&
                  timeout_s > 0 && colocated >= 1,
              EXC_ERR_ARG, "exc_bidiag_local: bad arguments");
  EXC_REQUIRE(work_bytes >= exc_bidiag_local_workspace(m, c), EXC_ERR_SPACE, "exc_bidiag_local: workspace too small");
  LocalWs w;
  local_ws(m, c, d_work, &w);
  hipStream_t s = exc_detail::stream(stream);
  // the first column (client.cpp:830-833): the all-reduce of one owner is the value itself
  int rc = exc_ralha_tail_update(-1, m, c, d_fc, d_uc, nullptr, d_beta, w.scal, stream);
  if (!rc) rc = exc_ralha_tail_scale(0, m, c, d_fc, d_uc, w.scal, d_alpha, stream);
  if (rc) return rc;
  const LoopShape sh = loop_shape(m, c, 0);
  const void* fn = loop_kernel_for(sh);
  bool fits = false;
  if (h_link && fn) {
    // every colocated caller's grid must be resident at once: a grid that is only partly resident
    // waits at its barriers for workgroups another caller's waiting grid keeps off the device
    rc = loop_resident(sh, fn, &fits, colocated);
    if (rc) return rc;
  }
  if (fits) {
    rc = exc_ralha_box_clear(w.box, stream);
    if (rc) return rc;
    void* boxes[1] = {w.box};
    int nwg[1] = {sh.nwg};
    rc = exc_ralha_loop_start_boxes(m, c, d_fc, d_uc, d_house, d_tau, d_alpha, d_beta, h_link, w.work, w.work_bytes,
                                    0, timeout_s, 1, 0, boxes, nwg, stream);
    if (!rc) rc = exc_ralha_loop_finish(h_link, stream);
    if (rc) return rc;
  } else {
    // one launch per step, back to back on the stream: step i reads message i and writes i + 1
    rc = exc_ralha_step(-1, m, c, d_fc, d_uc, nullptr, d_house, d_tau, d_alpha, d_beta, w.msg[0], w.work,
                        w.work_bytes, stream);
    for (int i = 0; !rc && i + 2 < m; i++)
      rc = exc_ralha_step(i, m, c, d_fc, d_uc, w.msg[i & 1], d_house, d_tau, d_alpha, d_beta, w.msg[(i + 1) & 1],
                          w.work, w.work_bytes, stream);
    if (rc) return rc;
  }
  // the last two columns (client.cpp:897-905)
  for (int i = m - 2; i < m; i++) {
    rc = exc_ralha_dot(i, m, c, d_fc, d_uc, w.scal, stream);
    if (!rc) rc = exc_ralha_tail_update(i, m, c, d_fc, d_uc, w.scal, d_beta, w.scal + 1, stream);
    if (!rc) rc = exc_ralha_tail_scale(i, m, c, d_fc, d_uc, w.scal + 1, d_alpha, stream);
    if (rc) return rc;
  }
  return exc_lower_to_upper(m, d_alpha, d_beta, d_cs, stream);
}

}  // extern "C"
