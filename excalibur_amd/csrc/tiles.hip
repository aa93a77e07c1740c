// Tile-parallel Cholesky and sign-choosing LU (the factorizations of the Cholesky QR and of the
// Householder reconstruction) in one launch each: one workgroup per 64 x 64 tile.
//
// A tile accumulates its updates with MFMA (v_mfma_f64_16x16x4f64) as soon as the tiles it needs
// are final, then factors itself (diagonal tile) or solves against the diagonal tile of its
// column or row, and publishes a per-tile flag. The recursive launch chain of the blocked
// algorithm (about sixty dependent launches per factorization of order 1000) becomes one
// launch whose critical path is the chain of diagonal tiles.
//
// Forward progress without co-residency: workgroups are numbered so that every tile depends only
// on tiles with smaller numbers, and a compute unit's workgroups are dispatched in increasing
// order, so the lowest unfinished workgroup is always resident and never waits. Every wait is
// also bounded: after a long spin the kernel records a timeout and all workgroups exit.
#include <algorithm>
#include <vector>

#include "exc_common.hpp"
#include "fact.hpp"
#include "solve.hpp"

namespace {

using fact::readlane_f64;

constexpr int TB = 64;    // tile edge
constexpr int TNT = 64;   // threads per workgroup: one wavefront per tile
constexpr int LD = TB + 1;
constexpr long kSpinLimit = 1L << 22;

typedef double d4 __attribute__((ext_vector_type(4)));

// Waits until *f != 0: thread 0 spins with relaxed device-scope loads, and the finished tile is
// then read with coherent loads (coh_ld), so no cache needs invalidating. Returns false when the
// wait timed out or another workgroup reported an error.
__device__ __forceinline__ bool wait_flag(int* f, int* error, int* sh) {
  if (threadIdx.x == 0) {
    int ok = 1;
    long spins = 0;
    while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) { ok = 0; break; }
      if (++spins > kSpinLimit) {
        __hip_atomic_store(error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    sh[0] = ok;
  }
  __syncthreads();
  const bool ok = sh[0] != 0;
  __syncthreads();  // every thread has read sh before the next wait can overwrite it
  return ok;
}

// Publishes a tile: every wavefront's coherent stores of it have completed before the flag.
__device__ __forceinline__ void signal_flag(int* f) {
  coh_stores_done();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(f, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Accumulator layout: the wavefront holds the tile as 4 x 4 MFMA blocks of 16 x 16; element v of
// lane l of block (a, b) is row 16a + (l / 16) + 4v, column 16b + l % 16.
struct Quad {
  int li, lk;
  __device__ Quad() {
    li = threadIdx.x & 15;
    lk = threadIdx.x >> 4;
  }
  __device__ int row(int a, int v) const { return 16 * a + lk + 4 * v; }
  __device__ int col(int b) const { return 16 * b + li; }
};

// acc <- the nr x nc tile of A at (r0, c0), zero outside.
__device__ __forceinline__ void acc_load(d4 (&acc)[4][4], const Quad& q, const double* A, long lda, long r0, long c0, int nr,
                         int nc) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = q.row(a, v), c = q.col(b);
        acc[a][b][v] = (r < nr && c < nc) ? A[(r0 + r) * lda + c0 + c] : 0.0;
      }
}

// S (LDS, 64 x 65) <- acc
__device__ __forceinline__ void acc_to_lds(const d4 (&acc)[4][4], const Quad& q, double* S) {
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) S[q.row(a, v) * LD + q.col(b)] = acc[a][b][v];
}

// T (LDS, 64 x 65) <- the nr x nc tile of A at (r0, c0), zero outside (a finished tile of
// another workgroup: coherent loads).
// Lane c reads column c of 16 rows at a time, all 16 loads in flight before the first LDS store.
__device__ __forceinline__ void tile_to_lds(double* T, const double* A, long lda, long r0, long c0, int nr, int nc) {
  const int c = threadIdx.x;
#pragma unroll
  for (int rb = 0; rb < TB; rb += 16) {
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      const int r = rb + q;
      v[q] = (r < nr && c < nc) ? coh_ld(&A[(r0 + r) * lda + c0 + c]) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; q++) T[(rb + q) * LD + c] = v[q];
  }
}

// T (LDS) <- the transpose of the nr x nc tile of A at (r0, c0), zero outside.
__device__ __forceinline__ void tile_to_lds_t(double* T, const double* A, long lda, long r0, long c0, int nr, int nc) {
  for (int e = threadIdx.x; e < TB * TB; e += TNT) {
    const int r = e >> 6, c = e & 63;
    T[c * LD + r] = (r < nr && c < nc) ? A[(r0 + r) * lda + c0 + c] : 0.0;
  }
}

// The nr x nc tile of A at (r0, c0) <- S (LDS); lower_zero writes zeros above the diagonal.
// Coherent stores: other workgroups read the finished tile.
__device__ __forceinline__ void lds_to_tile(const double* S, double* A, long lda, long r0, long c0, int nr, int nc, bool lower_zero) {
  for (int e = threadIdx.x; e < TB * TB; e += TNT) {
    const int r = e >> 6, c = e & 63;
    if (r < nr && c < nc) coh_st(&A[(r0 + r) * lda + c0 + c], (lower_zero && c > r) ? 0.0 : S[r * LD + c]);
  }
}

// acc -= X Y^T (TRANS_Y) or X Y, X and Y 64 x 64 in LDS.
template <bool TRANS_Y>
__device__ __forceinline__ void acc_update(d4 (&acc)[4][4], const Quad& q, const double* X, const double* Y) {
#pragma unroll 2
  for (int kk = 0; kk < TB; kk += 4) {
    double xa[4], yb[4];
#pragma unroll
    for (int a = 0; a < 4; a++) xa[a] = -X[(16 * a + q.li) * LD + kk + q.lk];
#pragma unroll
    for (int b = 0; b < 4; b++)
      yb[b] = TRANS_Y ? Y[(16 * b + q.li) * LD + kk + q.lk] : Y[(kk + q.lk) * LD + 16 * b + q.li];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[a], yb[b], acc[a][b], 0, 0, 0);
  }
}

// 1 / sqrt(d) and sqrt(d) for d > 0: hardware reciprocal square root refined by two Newton steps.
__device__ inline void rsqrt_pair(double d, double* r, double* s) {
  double y = __builtin_amdgcn_rsq(d);
  y = y * (1.5 - 0.5 * d * y * y);
  y = y * (1.5 - 0.5 * d * y * y);
  *r = y;
  *s = d * y;
}

// In-tile factorizations and solves run with one row (or column) of the tile per lane in
// registers; the pivot column, pivot row or reciprocals go through LDS between steps.
// The loops are unrolled so every register index is a constant; wave_sync orders the LDS traffic
// of the one wavefront (and keeps the compiler from moving LDS reads across steps).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


// Cholesky of the diagonal tile in S (padded rows/cols identity), lane i = row i. Returns the
// first non-positive pivot (1-based) or 0; such a pivot is replaced by 1 like the blocked path.
// Every element receives the same fused multiply-adds in the same order as a column-by-column
// right-looking factorization.
//
// The routine runs once per diagonal tile on one wavefront, so every instruction is fetched cold:
// a fully unrolled form (tens of KB of straight-line code) was bound by instruction fetch, not
// arithmetic. So only a panel of PW columns is unrolled (inside it the pivot and the multipliers
// L[k][j] reach the other lanes by v_readlane, no LDS, no barrier); the panel loop is not, and the
// panel's rank-PW update of the later columns runs as a loop over blocks of 8 columns with eight
// independent chains, the panel's multipliers read back from LDS (xc).
constexpr int PW = 8;
__device__ __forceinline__ int potf2_loop(double* S, double* xc, int kb) {
  const int lane = threadIdx.x;
  __asm__ volatile("" ::: "memory");
  int bad = 0;
#pragma unroll 1
  for (int j0 = 0; j0 < TB; j0 += PW) {
    double x[PW], lown[PW];
#pragma unroll
    for (int jj = 0; jj < PW; jj++) x[jj] = S[lane * LD + j0 + jj];
#pragma unroll
    for (int jj = 0; jj < PW; jj++) {
      const int j = j0 + jj;
      double d = readlane_f64(x[jj], j);
      if (!(d > 0.0)) {
        if (!bad && j < kb) bad = j + 1;
        d = 1.0;
      }
      double rp, piv;
      rsqrt_pair(d, &rp, &piv);
      const double l = lane > j ? x[jj] * rp : 0.0;
      x[jj] = lane == j ? piv : (lane > j ? l : x[jj]);
      lown[jj] = l;
#pragma unroll
      for (int kk = jj + 1; kk < PW; kk++) x[kk] = fma(-l, readlane_f64(l, j0 + kk), x[kk]);
    }
#pragma unroll
    for (int jj = 0; jj < PW; jj++) {
      S[lane * LD + j0 + jj] = x[jj];
      xc[lane * (PW + 1) + jj] = lown[jj];
    }
    wave_sync();
#pragma unroll 1
    for (int k0 = j0 + PW; k0 < TB; k0 += 8) {
      double a[8];
#pragma unroll
      for (int kk = 0; kk < 8; kk++) a[kk] = S[lane * LD + k0 + kk];
#pragma unroll
      for (int jj = 0; jj < PW; jj++)
#pragma unroll
        for (int kk = 0; kk < 8; kk++) a[kk] = fma(-lown[jj], xc[(k0 + kk) * (PW + 1) + jj], a[kk]);
#pragma unroll
      for (int kk = 0; kk < 8; kk++) S[lane * LD + k0 + kk] = a[kk];
    }
    wave_sync();
  }
#pragma unroll 1
  for (int k = lane + 1; k < TB; k++) S[lane * LD + k] = 0.0;
  return bad;
}

// LU without pivoting of S - diag(s) (sign rule of the Householder reconstruction), lane i = row
// i; row j is published through LDS at step j. sgn receives s for the kb valid pivots.
__device__ __forceinline__ void getf2_reg(double* S, double* xr, int kb, double* sgn, bool last_plus) {
  const int lane = threadIdx.x;
  __asm__ volatile("" ::: "memory");  // keep this routine's loads inside its branch
  double x[TB];
#pragma unroll
  for (int k = 0; k < TB; k++) x[k] = S[lane * LD + k];
#pragma unroll
  for (int j = 0; j < TB; j++) {
    double* r = xr + (j & 1) * (TB + 1);  // row j (r[k], k > j) and the pivot (r[TB])
    if (lane == j) {
      const double w = x[j];
      double sj = signbit(w) ? 1.0 : -1.0;
      if (last_plus && j == kb - 1) sj = -sj;
      x[j] = w - sj;
      r[TB] = x[j];
      if (j < kb) sgn[j] = sj;
#pragma unroll
      for (int k = j + 1; k < TB; k++) r[k] = x[k];
    }
    wave_sync();
    const double rp = 1.0 / r[TB];
    const double l = lane > j ? x[j] * rp : 0.0;
    if (lane > j) x[j] = l;
#pragma unroll
    for (int k = j + 1; k < TB; k++) x[k] = fma(-l, r[k], x[k]);
  }
#pragma unroll
  for (int k = 0; k < TB; k++) S[lane * LD + k] = x[k];
}

// Off-diagonal tile solves against the final diagonal tile, read from HBM with wave-uniform
// addresses (the flag's acquire made it visible): ROWS: lane = row r of S, S[r] <- S[r] op(D)^-1
// (LOWER_T: op(D) = L^T, Cholesky; else U, LU); else lane = column c, S[:, c] <- L^-1 S[:, c]
// with L the unit lower part of D (LU) or (MODE 3) the lower triangle of the tile the caller
// staged in T (triangular inverse).
template <int MODE>  // 0: rows, L^T; 1: rows, U; 2: columns, unit L; 3: columns, L
__device__ __forceinline__ void tile_solve(double* S, double* T, const double* Dg, long lda, int kd) {
  // The diagonal tile was finished by another workgroup. The LU modes copy it into LDS (T) with
  // one batched coherent copy and substitute from there; the Cholesky mode, whose substitution
  // walks rows of L^T, is faster reading it through the caches after one acquire fence.
  if (MODE == 1 || MODE == 2) {
    tile_to_lds(T, Dg, lda, 0, 0, kd, kd);
    __syncthreads();
  } else if (MODE == 0) {
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }  // MODE 3: the caller staged the diagonal tile in T
  __asm__ volatile("" ::: "memory");
  const int lane = threadIdx.x;
  double x[TB];
#pragma unroll
  for (int k = 0; k < TB; k++) x[k] = MODE >= 2 ? S[k * LD + lane] : S[lane * LD + k];
  if (MODE == 3) {
    if (kd == TB) fact::left_solve<false, false, true>(kd, T, LD, x);
    else fact::left_solve<false, false, false>(kd, T, LD, x);
  } else if (MODE == 0) {
    if (kd == TB) fact::right_solve<true, false, true>(kd, Dg, lda, x);
    else fact::right_solve<true, false, false>(kd, Dg, lda, x);
  } else if (MODE == 1) {
    if (kd == TB) fact::right_solve<false, false, true>(kd, T, LD, x);
    else fact::right_solve<false, false, false>(kd, T, LD, x);
  } else {
    if (kd == TB) fact::left_solve<false, true, true>(kd, T, LD, x);
    else fact::left_solve<false, true, false>(kd, T, LD, x);
  }
#pragma unroll
  for (int k = 0; k < TB; k++) {
    if (MODE >= 2) S[k * LD + lane] = x[k];
    else S[lane * LD + k] = x[k];
  }
}

// Identity padding of a diagonal tile with kb < 64 valid rows/cols.
__device__ __forceinline__ void pad_identity(double* S, int kb) {
  for (int e = threadIdx.x; e < TB * TB; e += TNT) {
    const int r = e >> 6, c = e & 63;
    if (r >= kb || c >= kb) S[r * LD + c] = (r == c) ? 1.0 : 0.0;
  }
}

// Tile numbering of the lower triangle, column by column: t -> (i, j), i >= j.
__device__ void chol_tile(int t, int nb, int* i, int* j) {
  int col = 0, rem = t;
  while (rem >= nb - col) {
    rem -= nb - col;
    col++;
  }
  *j = col;
  *i = col + rem;
}
__host__ __device__ inline int chol_index(int i, int j, int nb) { return j * nb - j * (j - 1) / 2 + (i - j); }


// ---- the same Cholesky with four wavefronts per tile
// The one-wavefront tile spends most of a diagonal step outside the factorization itself: the last
// 64 x 64 x 64 update (256 MFMAs), staging the tiles it needs and storing the finished tile. Here
// wavefront w holds rows 16w .. 16w + 15 of the tile (four 16 x 16 MFMA blocks), so the updates,
// the staging and the stores split four ways; the in-register factorization and the solve, chains
// of dependent operations on 64 rows, stay on wavefront 0. Every block receives the same MFMAs in
// the same order as in the one-wavefront form, and the factorization and solve are the same code,
// so the factor is bit-identical.
constexpr int TNT4 = 256;

// element v of lane l of block b of wavefront w: row 16w + (l / 16) + 4v, column 16b + l % 16
struct Quad4 {
  int li, lk, w;
  __device__ Quad4() {
    li = threadIdx.x & 15;
    lk = (threadIdx.x & 63) >> 4;
    w = threadIdx.x >> 6;
  }
  __device__ int row(int v) const { return 16 * w + lk + 4 * v; }
  __device__ int col(int b) const { return 16 * b + li; }
};

__device__ __forceinline__ void acc_load4(d4 (&acc)[4], const Quad4& q, const double* A, long lda, long r0, long c0,
                                          int nr, int nc) {
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int r = q.row(v), c = q.col(b);
      acc[b][v] = (r < nr && c < nc) ? A[(r0 + r) * lda + c0 + c] : 0.0;
    }
}

__device__ __forceinline__ void acc_to_lds4(const d4 (&acc)[4], const Quad4& q, double* S) {
#pragma unroll
  for (int b = 0; b < 4; b++)
#pragma unroll
    for (int v = 0; v < 4; v++) S[q.row(v) * LD + q.col(b)] = acc[b][v];
}

// T (LDS) <- the nr x nc tile of A at (r0, c0), zero outside (coherent loads): thread t reads
// column t % 64 of rows 16 (t / 64) .. 16 (t / 64) + 15, all 16 loads in flight.
__device__ __forceinline__ void tile_to_lds4(double* T, const double* A, long lda, long r0, long c0, int nr, int nc) {
  const int c = threadIdx.x & 63, rb = (threadIdx.x >> 6) * 16;
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int r = rb + q;
    v[q] = (r < nr && c < nc) ? coh_ld(&A[(r0 + r) * lda + c0 + c]) : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 16; q++) T[(rb + q) * LD + c] = v[q];
}

__device__ __forceinline__ void lds_to_tile4(const double* S, double* A, long lda, long r0, long c0, int nr, int nc,
                                             bool lower_zero) {
  for (int e = threadIdx.x; e < TB * TB; e += TNT4) {
    const int r = e >> 6, c = e & 63;
    if (r < nr && c < nc) coh_st(&A[(r0 + r) * lda + c0 + c], (lower_zero && c > r) ? 0.0 : S[r * LD + c]);
  }
}

__device__ __forceinline__ void pad_identity4(double* S, int kb) {
  for (int e = threadIdx.x; e < TB * TB; e += TNT4) {
    const int r = e >> 6, c = e & 63;
    if (r >= kb || c >= kb) S[r * LD + c] = (r == c) ? 1.0 : 0.0;
  }
}

// acc -= X Y^T for this wavefront's row block (X, Y 64 x 64 in LDS)
__device__ __forceinline__ void acc_update4(d4 (&acc)[4], const Quad4& q, const double* X, const double* Y) {
#pragma unroll 2
  for (int kk = 0; kk < TB; kk += 4) {
    const double xa = -X[(16 * q.w + q.li) * LD + kk + q.lk];
    double yb[4];
#pragma unroll
    for (int b = 0; b < 4; b++) yb[b] = Y[(16 * b + q.li) * LD + kk + q.lk];
#pragma unroll
    for (int b = 0; b < 4; b++) acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa, yb[b], acc[b], 0, 0, 0);
  }
}

// Tile t of the lower triangle of A's Cholesky factor (t in column-major tile order; every tile
// waits only on tiles of lower index, so in-order dispatch cannot deadlock).
__device__ __forceinline__ void potrf_tile_run(int n, double* __restrict__ A, long lda, int nb, int* __restrict__ done,
                                               int* __restrict__ error, int t) {
  extern __shared__ double smem[];  // S, X, Y: three 64 x 65 tiles
  double* S = smem;
  double* X = smem + TB * LD;
  double* Y = smem + 2 * TB * LD;
  __shared__ int sh[1];
  int ti, tj;
  chol_tile(t, nb, &ti, &tj);
  const long r0 = (long)ti * TB, c0 = (long)tj * TB;
  const int nr = min(TB, n - (int)r0), nc = min(TB, n - (int)c0);
  const Quad4 q;
  d4 acc[4];
  acc_load4(acc, q, A, lda, r0, c0, nr, nc);
  for (int k = 0; k < tj; k++) {
    if (!wait_flag(&done[chol_index(ti, k, nb)], error, sh)) return;
    if (ti != tj && !wait_flag(&done[chol_index(tj, k, nb)], error, sh)) return;
    const int kc = min(TB, n - k * TB);
    tile_to_lds4(X, A, lda, r0, (long)k * TB, nr, kc);
    if (ti != tj) tile_to_lds4(Y, A, lda, c0, (long)k * TB, nc, kc);
    __syncthreads();
    acc_update4(acc, q, X, ti != tj ? Y : X);
    __syncthreads();
  }
  acc_to_lds4(acc, q, S);
  __syncthreads();
  if (ti == tj) {
    pad_identity4(S, nr);
    __syncthreads();
    if (threadIdx.x < TB) {
      const int bad = potf2_loop(S, X, nr);
      if (threadIdx.x == 0 && bad) atomicCAS(&error[1], 0, (int)r0 + bad);
    }
    __syncthreads();
    lds_to_tile4(S, A, lda, r0, c0, nr, nc, true);
  } else {
    if (!wait_flag(&done[chol_index(tj, tj, nb)], error, sh)) return;
    if (threadIdx.x < TB) tile_solve<0>(S, X, A + c0 * lda + c0, lda, nc);  // L(i, j) = A(i, j) L(j, j)^-T
    __syncthreads();
    lds_to_tile4(S, A, lda, r0, c0, nr, nc, false);
  }
  signal_flag(&done[t]);
}

__global__ __launch_bounds__(TNT4) void k_potrf_tiles(int n, double* __restrict__ A, long lda, int nb,
                                                      int* __restrict__ done, int* __restrict__ error) {
  potrf_tile_run(n, A, lda, nb, done, error, blockIdx.x);
}

// The same over a batch of matrices of one order (blockIdx.y picks the matrix).
__global__ __launch_bounds__(TNT4) void k_potrf_tiles_batched(int n, double* const* __restrict__ As, long lda, int nb,
                                                              int* __restrict__ done, int* __restrict__ error) {
  const int b = blockIdx.y;
  potrf_tile_run(n, As[b], lda, nb, done + (long)b * nb * nb, error + 2 * b, blockIdx.x);
}


// X = L^-1 by tiles (both lower, n x n): tile (i, j), i >= j, in the column-major order of the
// Cholesky tiles, is X(i, j) = L(i, i)^-1 (delta_ij I - sum_{k=j}^{i-1} L(i, k) X(k, j)); it waits
// only on the tiles X(k, j) above it in its column. The upper triangle of X is not written.
__global__ __launch_bounds__(TNT) void k_trinv_tiles_batched(int n, const double* const* __restrict__ Ls, long ldl,
                                                             double* const* __restrict__ Xs, long ldx, int nb,
                                                             int* __restrict__ done_all, int* __restrict__ error_all) {
  extern __shared__ double smem[];
  double* S = smem;
  double* P = smem + TB * LD;
  double* Q = smem + 2 * TB * LD;
  __shared__ int sh[1];
  const int b = blockIdx.y;
  const double* L = Ls[b];
  double* Xo = Xs[b];
  int* done = done_all + (long)b * nb * nb;
  int* error = error_all + 2 * b;
  int ti, tj;
  chol_tile(blockIdx.x, nb, &ti, &tj);
  const long r0 = (long)ti * TB, c0 = (long)tj * TB;
  const int nr = min(TB, n - (int)r0), nc = min(TB, n - (int)c0);
  const Quad q;
  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int bb = 0; bb < 4; bb++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[a][bb][v] = (ti == tj && q.row(a, v) == q.col(bb)) ? 1.0 : 0.0;
  for (int k = tj; k < ti; k++) {
    if (!wait_flag(&done[chol_index(k, tj, nb)], error, sh)) return;
    const int kc = min(TB, n - k * TB);
    tile_to_lds(P, L, ldl, r0, (long)k * TB, nr, kc);   // L(i, k)
    tile_to_lds(Q, Xo, ldx, (long)k * TB, c0, kc, nc);  // X(k, j), finished by another workgroup
    __syncthreads();
    acc_update<false>(acc, q, P, Q);
    __syncthreads();
  }
  acc_to_lds(acc, q, S);
  tile_to_lds(P, L, ldl, r0, r0, nr, nr);  // L(i, i)
  __syncthreads();
  if (nr < TB) pad_identity(P, nr);
  __syncthreads();
  tile_solve<3>(S, P, nullptr, 0, TB);
  __syncthreads();
  lds_to_tile(S, Xo, ldx, r0, c0, nr, nc, ti == tj);
  signal_flag(&done[blockIdx.x]);
}

// LU tiles in order of d = min(i, j): (d, d), then (d, d+1), (d+1, d), (d, d+2), (d+2, d), ...
__device__ void lu_tile(int t, int nb, int* i, int* j) {
  int d = 0, rem = t;
  while (rem >= 2 * (nb - d) - 1) {
    rem -= 2 * (nb - d) - 1;
    d++;
  }
  if (rem == 0) { *i = d; *j = d; return; }
  const int off = (rem + 1) / 2;
  if (rem & 1) { *i = d; *j = d + off; }  // U tile in row d
  else { *i = d + off; *j = d; }          // L tile in column d
}
__host__ __device__ inline int lu_index(int i, int j, int nb) {
  const int d = i < j ? i : j;
  const int before = d * (2 * nb - d);  // sum over e < d of 2 (nb - e) - 1
  if (i == j) return before;
  const int off = (i < j) ? (j - d) : (i - d);
  return before + 2 * off - ((i < j) ? 1 : 0);
}

__global__ __launch_bounds__(TNT) void k_lu_tiles(int n, double* __restrict__ A, long lda, int nb,
                                                  double* __restrict__ sgn, int square_last,
                                                  int* __restrict__ done, int* __restrict__ error) {
  extern __shared__ double smem[];  // S, X, Y: three 64 x 65 tiles
  double* S = smem;
  double* X = smem + TB * LD;
  double* Y = smem + 2 * TB * LD;
  __shared__ int sh[1];
  int ti, tj;
  lu_tile(blockIdx.x, nb, &ti, &tj);
  const long r0 = (long)ti * TB, c0 = (long)tj * TB;
  const int nr = min(TB, n - (int)r0), nc = min(TB, n - (int)c0);
  const int d = min(ti, tj);
  const Quad q;
  d4 acc[4][4];
  acc_load(acc, q, A, lda, r0, c0, nr, nc);
  for (int k = 0; k < d; k++) {
    if (!wait_flag(&done[lu_index(ti, k, nb)], error, sh)) return;
    if (!wait_flag(&done[lu_index(k, tj, nb)], error, sh)) return;
    const int kc = min(TB, n - k * TB);
    tile_to_lds(X, A, lda, r0, (long)k * TB, nr, kc);  // L(i, k)
    tile_to_lds(Y, A, lda, (long)k * TB, c0, kc, nc);  // U(k, j)
    __syncthreads();
    acc_update<false>(acc, q, X, Y);
    __syncthreads();
  }
  acc_to_lds(acc, q, S);
  __syncthreads();
  if (ti == tj) {
    pad_identity(S, nr);
    __syncthreads();
    getf2_reg(S, X, nr, sgn + r0, square_last && ti == nb - 1);
    __syncthreads();
    lds_to_tile(S, A, lda, r0, c0, nr, nc, false);
  } else {
    if (!wait_flag(&done[lu_index(d, d, nb)], error, sh)) return;
    const long dd = (long)d * TB;
    const int kd = min(TB, n - (int)dd);
    if (ti > tj) tile_solve<1>(S, X, A + dd * lda + dd, lda, kd);  // L(i, j) = A(i, j) U(j, j)^-1
    else tile_solve<2>(S, X, A + dd * lda + dd, lda, kd);          // U(i, j) = L(i, i)^-1 A(i, j)
    __syncthreads();
    lds_to_tile(S, A, lda, r0, c0, nr, nc, false);
  }
  signal_flag(&done[blockIdx.x]);
}

__global__ void k_zero_ints(int* p, int count) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += gridDim.x * blockDim.x) p[e] = 0;
}

__global__ __launch_bounds__(256) void k_zero_upper_blocks(int n, double* A, long lda) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const long i = e / n, j = e - i * n;
  if (j > i) A[i * lda + j] = 0.0;
}

__global__ __launch_bounds__(256) void k_zero_upper_batched(int n, double* const* As, long lda) {
  double* A = As[blockIdx.y];
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const long i = e / n, j = e - i * n;
    if (j > i) A[i * lda + j] = 0.0;
  }
}

// info[b] = error[2 b + 1] (the first failing pivot of matrix b, 0 when none)
__global__ void k_tile_infos(int count, const int* __restrict__ error, int* __restrict__ info) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < count) info[b] = error[2 * b + 1];
}

constexpr size_t kTileLds = 3 * TB * LD * sizeof(double);

}  // namespace

namespace fact {

size_t tiles_work_bytes(int n) {
  const int nb = (n + TB - 1) / TB;
  return (size_t)(nb * nb + 8) * sizeof(int);
}

int potrf_tiles(int n, double* A, long lda, int* info, void* work, size_t wb, hipStream_t s) {
  const int nb = (n + TB - 1) / TB;
  const int T = nb * (nb + 1) / 2;
  EXC_REQUIRE(wb >= tiles_work_bytes(n), EXC_ERR_SPACE, "potrf: workspace too small");
  int* done = (int*)work;
  int* err = done + nb * nb;
  static bool attr = false;
  if (!attr) {
    EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_potrf_tiles, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kTileLds));
    attr = true;
  }
  hipLaunchKernelGGL(k_zero_ints, dim3(exc_detail::ceil_div(nb * nb + 8, 256)), dim3(256), 0, s, done, nb * nb + 8);
  hipLaunchKernelGGL(k_potrf_tiles, dim3(T), dim3(TNT4), kTileLds, s, n, A, lda, nb, done, err);
  hipLaunchKernelGGL(k_zero_upper_blocks, dim3(exc_detail::ceil_div((long)n * n, 256)), dim3(256), 0, s, n, A, lda);
  EXC_CHECK_HIP(hipGetLastError());
  int herr[2] = {0, 0};
  EXC_CHECK_HIP(hipMemcpyAsync(herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  EXC_REQUIRE(herr[0] == 0, EXC_ERR_NUMERIC, "potrf: a tile wait timed out");
  if (info && herr[1]) EXC_CHECK_HIP(hipMemcpyAsync(info, &herr[1], sizeof(int), hipMemcpyHostToDevice, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  return EXC_OK;
}

size_t tiles_batched_work_bytes(int n, int count) {
  const int nb = (n + TB - 1) / TB;
  return ((size_t)count * nb * nb + 2 * (size_t)count + 8) * sizeof(int);
}

// The Cholesky factors (lower, in place, upper triangle zeroed) of `count` matrices whose device
// addresses d_As (a device array) lists, in one launch; d_info[b] = 0 or 1 + the row of the first
// non-positive pivot of matrix b. Synchronizes the stream to check the bounded waits.
int potrf_tiles_batched(int n, double* const* d_As, int count, long lda, int* d_info, void* work, size_t wb,
                        hipStream_t s) {
  EXC_REQUIRE(n >= 1 && count >= 1 && d_As && d_info, EXC_ERR_ARG, "potrf_tiles_batched: bad arguments");
  EXC_REQUIRE(n <= kTileMaxN, EXC_ERR_ARG, "potrf_tiles_batched: order above the tile limit");
  EXC_REQUIRE(wb >= tiles_batched_work_bytes(n, count), EXC_ERR_SPACE, "potrf_tiles_batched: workspace too small");
  const int nb = (n + TB - 1) / TB;
  const int T = nb * (nb + 1) / 2;
  int* done = (int*)work;
  int* err = done + (long)count * nb * nb;
  static bool attr = false;
  if (!attr) {
    EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_potrf_tiles_batched, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kTileLds));
    attr = true;
  }
  const long flags = (long)count * nb * nb + 2L * count;
  hipLaunchKernelGGL(k_zero_ints, dim3(exc_detail::ceil_div(flags, 256)), dim3(256), 0, s, done, (int)flags);
  hipLaunchKernelGGL(k_potrf_tiles_batched, dim3(T, count), dim3(TNT4), kTileLds, s, n, d_As, lda, nb, done, err);
  hipLaunchKernelGGL(k_zero_upper_batched, dim3(std::min<unsigned>(1024, exc_detail::ceil_div((long)n * n, 256)), count),
                     dim3(256), 0, s, n, d_As, lda);
  hipLaunchKernelGGL(k_tile_infos, dim3(exc_detail::ceil_div(count, 256)), dim3(256), 0, s, count, err, d_info);
  EXC_CHECK_HIP(hipGetLastError());
  std::vector<int> herr(2 * (size_t)count);
  EXC_CHECK_HIP(hipMemcpyAsync(herr.data(), err, herr.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < count; b++) EXC_REQUIRE(herr[2 * b] == 0, EXC_ERR_NUMERIC, "potrf: a tile wait timed out");
  return EXC_OK;
}

// X_b = L_b^-1 (lower; the upper triangle of X_b zeroed) for `count` matrices, in one launch.
// Synchronizes the stream to check the bounded waits.
int trinv_tiles_batched(int n, const double* const* d_Ls, long ldl, double* const* d_Xs, long ldx, int count, void* work,
                        size_t wb, hipStream_t s) {
  EXC_REQUIRE(n >= 1 && count >= 1 && d_Ls && d_Xs, EXC_ERR_ARG, "trinv_tiles_batched: bad arguments");
  EXC_REQUIRE(n <= kTileMaxN, EXC_ERR_ARG, "trinv_tiles_batched: order above the tile limit");
  EXC_REQUIRE(wb >= tiles_batched_work_bytes(n, count), EXC_ERR_SPACE, "trinv_tiles_batched: workspace too small");
  const int nb = (n + TB - 1) / TB;
  const int T = nb * (nb + 1) / 2;
  int* done = (int*)work;
  int* err = done + (long)count * nb * nb;
  static bool attr = false;
  if (!attr) {
    EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_trinv_tiles_batched, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kTileLds));
    attr = true;
  }
  const long flags = (long)count * nb * nb + 2L * count;
  hipLaunchKernelGGL(k_zero_ints, dim3(exc_detail::ceil_div(flags, 256)), dim3(256), 0, s, done, (int)flags);
  hipLaunchKernelGGL(k_trinv_tiles_batched, dim3(T, count), dim3(TNT), kTileLds, s, n, d_Ls, ldl, d_Xs, ldx, nb, done,
                     err);
  hipLaunchKernelGGL(k_zero_upper_batched, dim3(std::min<unsigned>(1024, exc_detail::ceil_div((long)n * n, 256)), count),
                     dim3(256), 0, s, n, d_Xs, ldx);
  EXC_CHECK_HIP(hipGetLastError());
  std::vector<int> herr(2 * (size_t)count);
  EXC_CHECK_HIP(hipMemcpyAsync(herr.data(), err, herr.size() * sizeof(int), hipMemcpyDeviceToHost, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < count; b++)
    EXC_REQUIRE(herr[2 * b] == 0, EXC_ERR_NUMERIC, "trinv: a tile wait timed out");
  return EXC_OK;
}

int lu_tiles(int n, double* A, long lda, double* sgn, int square_last, void* work, size_t wb, hipStream_t s) {
  const int nb = (n + TB - 1) / TB;
  const int T = nb * nb;
  EXC_REQUIRE(wb >= tiles_work_bytes(n), EXC_ERR_SPACE, "lu_sign: workspace too small");
  int* done = (int*)work;
  int* err = done + nb * nb;
  static bool attr = false;
  if (!attr) {
    EXC_CHECK_HIP(hipFuncSetAttribute((const void*)k_lu_tiles, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)kTileLds));
    attr = true;
  }
  hipLaunchKernelGGL(k_zero_ints, dim3(exc_detail::ceil_div(nb * nb + 8, 256)), dim3(256), 0, s, done, nb * nb + 8);
  hipLaunchKernelGGL(k_lu_tiles, dim3(T), dim3(TNT), kTileLds, s, n, A, lda, nb, sgn, square_last, done, err);
  EXC_CHECK_HIP(hipGetLastError());
  int herr = 0;
  EXC_CHECK_HIP(hipMemcpyAsync(&herr, err, sizeof herr, hipMemcpyDeviceToHost, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  EXC_REQUIRE(herr == 0, EXC_ERR_NUMERIC, "lu_sign: a tile wait timed out");
  return EXC_OK;
}

}  // namespace fact
