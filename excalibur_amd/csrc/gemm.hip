// Dense GEMM on the gfx950 matrix cores: C = alpha * op(A) op(B) + beta * C, row-major, FP64
// (v_mfma_f64_16x16x4_f64) and FP32 (v_mfma_f32_16x16x4_f32).
//
// Workgroup tile 128 x 128, four wavefronts in a 2 x 2 arrangement, each wavefront a 64 x 64
// block of 4 x 4 MFMA tiles; K is staged through LDS in chunks of 16 with register
// double-buffering of the next chunk. Options used by the factorizations: split-K (partial tiles
// summed in a fixed order by a second kernel, so results do not depend on scheduling), skipping
// the zero K range of a triangular operand, and computing only the lower tiles of a symmetric
// product (mirrored into the upper half by the reduction).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "exc_common.hpp"
#include "gemm.hpp"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef d4 acc_t;
  __device__ static inline acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // result element v of lane l sits in row (l/16) + 4v, column l%16
  __device__ static inline int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef f4 acc_t;
  __device__ static inline acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // result element v of lane l sits in row 4(l/16) + v, column l%16
  __device__ static inline int row(int lk, int v) { return 4 * lk + v; }
};

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;

struct GemmArgs {
  int ta, tb;
  long M, N, K;
  const void* A;
  long lda;
  const void* B;
  long ldb;
  void* C;
  long ldc;
  double alpha, beta;
  int tri;        // EXC_TRI_* restriction of the K range per tile
  int lower_only; // compute only tiles with tile_row >= tile_col
  long kchunk;    // split-K chunk (multiple of BK); 0 = no split
  void* part;     // split-K partials: splits x M x N (type T)
  long tm;        // tile rows
  long ntile;     // tiles per split (tm x tn, or the lower tiles tm (tm + 1) / 2)
  long total;     // ntile x splits = the 1-D grid
};

constexpr int kXcds = 8;  // dies of the device; the dispatcher deals workgroups round-robin over them

// Tile of workgroup blockIdx.x. The dispatcher sends linear id L to die L % 8, so consecutive tiles
// of the natural order would land on eight dies with eight separate L2s. The remap gives die x a
// contiguous range of the tile order instead, and that order runs the row tiles of one column
// (m fastest) back to back, so their shared B column tile is fetched from HBM once per die and
// the row tiles read it from that die's L2. Splits of K are the slowest index.
__device__ inline void tile_of(const GemmArgs& g, long& tile_m, long& tile_n, long& split) {
  const long L = blockIdx.x;
  const long q = g.total / kXcds, r = g.total % kXcds;
  const long x = L % kXcds, j = L / kXcds;
  const long t = x * q + min(x, r) + j;
  split = t / g.ntile;
  long u = t - split * g.ntile;
  if (g.lower_only) {
    // column n holds the tm - n tiles m = n .. tm - 1
    long n = 0;
    while (u >= g.tm - n) {
      u -= g.tm - n;
      n++;
    }
    tile_n = n;
    tile_m = n + u;
  } else {
    tile_n = u / g.tm;
    tile_m = u - tile_n * g.tm;
  }
}

// K range [k0, k1) that can contribute to output rows [i0, i0+BM) / columns [j0, j0+BN).
__device__ inline void k_range(const GemmArgs& g, long i0, long j0, long& k0, long& k1) {
  k0 = 0;
  k1 = g.K;
  switch (g.tri) {
    case EXC_TRI_A_LOWER: k1 = min(k1, i0 + BM); break;  // op(A)[i][k] = 0 for k > i
    case EXC_TRI_A_UPPER: k0 = max(k0, i0); break;       // op(A)[i][k] = 0 for k < i
    case EXC_TRI_B_LOWER: k0 = max(k0, j0); break;       // op(B)[k][j] = 0 for k < j
    case EXC_TRI_B_UPPER: k1 = min(k1, j0 + BN); break;  // op(B)[k][j] = 0 for k > j
    default: break;
  }
}

// One k-chunk of A and B (BM x BK and BK x BN) staged through registers into LDS. The per-thread
// element coordinates do not change between chunks: for a non-transposed A every thread owns one
// k column (tid % 16) of 8 rows 16 apart; for a transposed A one row (tid % 128) of 8 k steps 2
// apart (and likewise for B), so each chunk is 8 loads at a fixed stride from one base pointer.
template <typename T, int TA, int TB>
struct Stager {
  const T* pa;  // base of this thread's A elements at k = 0
  const T* pb;
  long sa, sb;  // stride between the thread's 8 elements
  long ka, kb;  // stride of one k step
  int ai, ak, bj, bk;  // LDS coordinates of element q = 0 (q advances ai or ak by 16 / 2)
  __device__ Stager(const GemmArgs& g, long i0, long j0, int tid) {
    if (!TA) { ai = tid >> 4; ak = tid & 15; pa = (const T*)g.A + (i0 + ai) * g.lda + ak; sa = 16 * g.lda; ka = 1; }
    else { ak = tid >> 7; ai = tid & 127; pa = (const T*)g.A + (long)ak * g.lda + i0 + ai; sa = 2 * g.lda; ka = g.lda; }
    if (!TB) { bk = tid >> 7; bj = tid & 127; pb = (const T*)g.B + (long)bk * g.ldb + j0 + bj; sb = 2 * g.ldb; kb = g.ldb; }
    else { bj = tid >> 4; bk = tid & 15; pb = (const T*)g.B + (j0 + bj) * g.ldb + bk; sb = 16 * g.ldb; kb = 1; }
  }
  __device__ inline int a_i(int q) const { return TA ? ai : ai + 16 * q; }
  __device__ inline int a_k(int q) const { return TA ? ak + 2 * q : ak; }
  __device__ inline int b_j(int q) const { return TB ? bj + 16 * q : bj; }
  __device__ inline int b_k(int q) const { return TB ? bk : bk + 2 * q; }
};

// One output tile [i0, i0+BM) x [j0, j0+BN) over the K range [k0, k1). The result element (gi, gj)
// goes to Cb[(gi - ci0) * ldc + (gj - cj0)]: scaled by alpha (plus beta times the old value) when
// direct, else stored as it is (a split-K partial).
template <typename T, int TA, int TB>
__device__ __forceinline__ void gemm_tile(const GemmArgs& g, long i0, long j0, long k0, long k1, T* Cb, long ldc,
                                          long ci0, long cj0, bool direct, T (*As)[BK][BM + 1],
                                          T (*Bs)[BK][BN + 1]) {
  using acc_t = typename Mfma<T>::acc_t;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const Stager<T, TA, TB> st(g, i0, j0, tid);
  // interior tile: no row or column bounds, chunks aligned to BK. A whole chunk of it then loads
  // without checks unless it crosses the diagonal of a triangular operand (the k range already
  // excludes the chunks wholly on the zero side), so only such chunks and a last partial chunk
  // check their elements.
  const bool interior = i0 + BM <= g.M && j0 + BN <= g.N && (k0 % BK) == 0;
  auto fast = [&](long kb) {
    if (!interior || kb + BK > k1) return false;
    switch (g.tri) {
      case EXC_TRI_A_LOWER: return kb + BK <= i0 + 1;       // every k <= i0 <= row
      case EXC_TRI_A_UPPER: return kb >= i0 + BM - 1;       // every k >= last row
      case EXC_TRI_B_LOWER: return kb >= j0 + BN - 1;       // every k >= last column
      case EXC_TRI_B_UPPER: return kb + BK <= j0 + 1;       // every k <= j0 <= column
      default: return true;
    }
  };
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  acc_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = acc_t{0, 0, 0, 0};
  T ra[8], rb[8];
  auto load_fast = [&](long kbase) {
    const T* pa = st.pa + kbase * st.ka;
    const T* pb = st.pb + kbase * st.kb;
#pragma unroll
    for (int q = 0; q < 8; q++) ra[q] = pa[q * st.sa];
#pragma unroll
    for (int q = 0; q < 8; q++) rb[q] = pb[q * st.sb];
  };
  auto load_checked = [&](long kbase) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const long gi = i0 + st.a_i(q), gk = kbase + st.a_k(q);
      T v = T(0);
      if (gi < g.M && gk >= k0 && gk < k1) v = st.pa[kbase * st.ka + q * st.sa];
      if ((g.tri == EXC_TRI_A_LOWER && gk > gi) || (g.tri == EXC_TRI_A_UPPER && gk < gi)) v = T(0);
      ra[q] = v;
      const long gj = j0 + st.b_j(q), gk2 = kbase + st.b_k(q);
      T w = T(0);
      if (gj < g.N && gk2 >= k0 && gk2 < k1) w = st.pb[kbase * st.kb + q * st.sb];
      if ((g.tri == EXC_TRI_B_LOWER && gk2 < gj) || (g.tri == EXC_TRI_B_UPPER && gk2 > gj)) w = T(0);
      rb[q] = w;
    }
  };
  auto store_chunk = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; q++) As[buf][st.a_k(q)][st.a_i(q)] = ra[q];
#pragma unroll
    for (int q = 0; q < 8; q++) Bs[buf][st.b_k(q)][st.b_j(q)] = rb[q];
  };
  const long kstart = (k0 / BK) * BK;
  int buf = 0;
  if (kstart < k1) {
    if (fast(kstart)) load_fast(kstart);
    else load_checked(kstart);
    store_chunk(0);
  }
  __syncthreads();
  const int li = lane & 15, lk = lane >> 4;
  for (long kb = kstart; kb < k1; kb += BK) {
    const bool more = kb + BK < k1;
    if (more) {
      if (fast(kb + BK)) load_fast(kb + BK);
      else load_checked(kb + BK);
    }
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      T af[4], bf[4];
#pragma unroll
      for (int a = 0; a < 4; a++) af[a] = As[buf][ks + lk][wm + a * 16 + li];
#pragma unroll
      for (int b = 0; b < 4; b++) bf[b] = Bs[buf][ks + lk][wn + b * 16 + li];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = Mfma<T>::step(af[a], bf[b], acc[a][b]);
    }
    if (more) {
      store_chunk(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
  }
  // epilogue: operands were supplied as A[lane%16][lane/16], B[lane/16][lane%16]
  T* C = Cb;
  const T alpha = (T)g.alpha, beta = (T)g.beta;
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const long gi = i0 + wm + a * 16 + Mfma<T>::row(lk, v);
        const long gj = j0 + wn + b * 16 + li;
        if (gi < g.M && gj < g.N) {
          const T val = acc[a][b][v];
          T* cp = C + (gi - ci0) * ldc + (gj - cj0);
          if (direct) {
            T out = alpha * val;
            if (g.beta != 0.0) out += beta * *cp;
            *cp = out;
          } else {
            *cp = val;
          }
        }
      }
}

template <typename T, int TA, int TB>
__global__ __launch_bounds__(NT, 2) void k_gemm(GemmArgs g) {
  __shared__ T As[2][BK][BM + 1];
  __shared__ T Bs[2][BK][BN + 1];
  long tile_m, tile_n, split;
  tile_of(g, tile_m, tile_n, split);
  const long i0 = tile_m * BM, j0 = tile_n * BN;
  long k0, k1;
  k_range(g, i0, j0, k0, k1);
  if (g.kchunk > 0) {
    const long ks = split * g.kchunk;
    k0 = max(k0, ks);
    k1 = min(k1, ks + g.kchunk);
  }
  T* Cb = g.kchunk > 0 ? (T*)g.part + (size_t)split * g.M * g.N : (T*)g.C;
  const long ldc = g.kchunk > 0 ? g.N : g.ldc;
  gemm_tile<T, TA, TB>(g, i0, j0, k0, k1, Cb, ldc, 0, 0, g.kchunk == 0, As, Bs);
}

// Grouped products: many independent problems of one transposition pattern in one launch. The
// host lays out work items (problem, tile, K range, partial slot); a persistent grid strides over
// them, each item running the same tile code as k_gemm. Items of a tile split over K (or of a
// mirrored symmetric product) store partial tiles; k_grouped_reduce then sums each tile's
// partials in slot order (deterministic), applies alpha and beta, and mirrors lower tiles.
struct GroupProb {
  const void* A;
  const void* B;
  void* C;
  long lda, ldb, ldc, M, N, K;
  int tri, lower_only, mirror, pad;
};
struct GroupItem {
  int p, tile_m, tile_n, slot;  // slot < 0: the item owns its whole tile and stores it directly
  long k0, k1;
};
struct GroupRed {
  int p, tile_m, tile_n, first_slot, nslots, pad;
};

template <typename T, int TA, int TB>
__global__ __launch_bounds__(NT, 2) void k_gemm_grouped(const GroupProb* __restrict__ probs,
                                                        const GroupItem* __restrict__ items, long nitems,
                                                        T* __restrict__ part, double alpha, double beta) {
  __shared__ T As[2][BK][BM + 1];
  __shared__ T Bs[2][BK][BN + 1];
  for (long it = blockIdx.x; it < nitems; it += gridDim.x) {
    const GroupItem w = items[it];
    const GroupProb pr = probs[w.p];
    GemmArgs g{};
    g.A = pr.A; g.lda = pr.lda; g.B = pr.B; g.ldb = pr.ldb; g.M = pr.M; g.N = pr.N; g.K = pr.K;
    g.tri = pr.tri; g.alpha = alpha; g.beta = beta;
    const long i0 = (long)w.tile_m * BM, j0 = (long)w.tile_n * BN;
    long k0, k1;
    k_range(g, i0, j0, k0, k1);
    k0 = max(k0, w.k0);
    k1 = min(k1, w.k1);
    if (w.slot < 0) gemm_tile<T, TA, TB>(g, i0, j0, k0, k1, (T*)pr.C, pr.ldc, 0, 0, true, As, Bs);
    else gemm_tile<T, TA, TB>(g, i0, j0, k0, k1, part + (size_t)w.slot * BM * BN, BN, i0, j0, false, As, Bs);
    __syncthreads();  // the next item restages the LDS tiles
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_grouped_reduce(const GroupProb* __restrict__ probs,
                                                        const GroupRed* __restrict__ reds, const T* __restrict__ part,
                                                        double alpha, double beta) {
  const GroupRed r = reds[blockIdx.x];
  const GroupProb pr = probs[r.p];
  T* C = (T*)pr.C;
  for (int e = threadIdx.x; e < BM * BN; e += blockDim.x) {
    const int li = e / BN, lj = e - li * BN;
    const long gi = (long)r.tile_m * BM + li, gj = (long)r.tile_n * BN + lj;
    if (gi >= pr.M || gj >= pr.N) continue;
    if (pr.lower_only && gj > gi) continue;  // the upper part of a diagonal tile comes by mirroring
    T s = T(0);
    for (int q = 0; q < r.nslots; q++) s += part[(size_t)(r.first_slot + q) * BM * BN + e];
    T out = (T)alpha * s;
    if (beta != 0.0) out += (T)beta * C[gi * pr.ldc + gj];
    C[gi * pr.ldc + gj] = out;
    if (pr.mirror && gj < gi) C[gj * pr.ldc + gi] = out;
  }
}

// Sums split-K partials in split order, applies alpha/beta, and mirrors lower tiles when asked.
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_reduce(GemmArgs g, int splits) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = g.M * g.N;
  if (idx >= total) return;
  const long i = idx / g.N, j = idx - i * g.N;
  if (g.lower_only && (j / BN) > (i / BM)) return;
  const T* P = (const T*)g.part;
  T s = T(0);
  for (int z = 0; z < splits; z++) s += P[(size_t)z * total + idx];
  T* C = (T*)g.C;
  T out = (T)g.alpha * s;
  if (g.beta != 0.0) out += (T)g.beta * C[i * g.ldc + j];
  C[i * g.ldc + j] = out;
}

// Copies the lower triangle onto the upper one (square C).
template <typename T>
__global__ __launch_bounds__(256) void k_mirror_lower(long n, T* C, long ldc) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const long i = idx / n, j = idx - i * n;
  if (j > i) C[i * ldc + j] = C[j * ldc + i];
}

// Workgroups of k_gemm resident at once on the current device: per_cu per CU (two for the FP64
// kernel's 256 VGPRs and 64 KB of LDS, three for FP32), times the CU count (cached per device).
long resident_slots(int per_cu) {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256L * per_cu;
  if (cus[dev] <= 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return (long)cus[dev] * per_cu;
}

template <typename T>
int gemm_impl(int ta, int tb, long M, long N, long K, double alpha, const T* A, long lda, const T* B, long ldb,
              double beta, T* C, long ldc, const ExcGemmOpts& o, hipStream_t s) {
  EXC_REQUIRE(M >= 0 && N >= 0 && K >= 0 && C, EXC_ERR_ARG, "exc_gemm: bad shape");
  if (M == 0 || N == 0) return EXC_OK;
  GemmArgs g{};
  g.ta = ta; g.tb = tb; g.M = M; g.N = N; g.K = K;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.alpha = alpha; g.beta = beta; g.tri = o.tri; g.lower_only = o.lower_only;
  const long tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  long tiles = o.lower_only ? tm * (tm + 1) / 2 : tm * tn;
  int splits = 1;
  const long slots = resident_slots(sizeof(T) == 8 ? 2 : 3);
  if (o.work && K > 4 * BK && tiles < slots) {
    // Split K (only while the tiles alone leave resident slots idle) into the number of
    // chunks that minimizes an estimate of the time: (rounds of resident workgroups) x (chunk
    // depth) x ~3 us per 16-deep chunk (a workgroup alone on its CU waits on its loads), plus,
    // when split, the fixed-order reduction: ~5 us and a read of every partial at ~3 TB/s. A grid
    // just above the resident slots would run a second, nearly empty round; a large output
    // would spend more summing partials than the split saves. Ties go to fewer splits.
    const long maxs = std::min<long>(K / BK, 4 * ((slots + tiles - 1) / tiles));
    long best = 1;
    double best_cost = -1.0;
    for (long sp = 1; sp <= std::max(1L, maxs); sp++) {
      if (sp > 1 && (size_t)sp * M * N * sizeof(T) > o.work_bytes) break;
      const long rounds = (tiles * sp + slots - 1) / slots;
      const long depth = ((K + sp - 1) / sp + BK - 1) / BK;
      double cost = 3.0 * (double)(rounds * depth);
      if (sp > 1) cost += 5.0 + (double)sp * M * N * sizeof(T) / 3.0e6;
      if (best_cost < 0.0 || cost < best_cost) { best = sp; best_cost = cost; }
    }
    splits = (int)best;
  }
  if (K == 0) splits = 1;
  if (splits > 1) {
    long chunk = (K + splits - 1) / splits;
    chunk = ((chunk + BK - 1) / BK) * BK;
    splits = (int)((K + chunk - 1) / chunk);
    g.kchunk = chunk;
    g.part = o.work;
  }
  g.tm = tm;
  g.ntile = tiles;
  g.total = tiles * splits;
  EXC_REQUIRE(g.total < (1L << 31), EXC_ERR_ARG, "exc_gemm: too many tiles");
  dim3 grid((unsigned)g.total);
  hipEvent_t e0 = exc_prof::mark(s);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<T, 0, 0>), grid, dim3(NT), 0, s, g);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<T, 0, 1>), grid, dim3(NT), 0, s, g);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<T, 1, 0>), grid, dim3(NT), 0, s, g);
  else hipLaunchKernelGGL((k_gemm<T, 1, 1>), grid, dim3(NT), 0, s, g);
  EXC_CHECK_HIP(hipGetLastError());
  if (e0) {
    // useful flops: a triangular operand halves the products, a symmetric (lower-only) result too
    double flops = 2.0 * (double)M * (double)N * (double)K;
    if (o.tri != EXC_TRI_NONE) flops *= 0.5;
    if (o.lower_only) flops *= 0.5;
    const double bytes = (double)sizeof(T) * ((double)M * K + (double)K * N + (double)M * N * (beta != 0.0 ? 2 : 1));
    exc_prof::book(sizeof(T) == 8 ? "gemm_f64" : "gemm_f32", e0, exc_prof::mark(s), flops, bytes);
  }
  if (splits > 1) {
    hipLaunchKernelGGL(k_splitk_reduce<T>, dim3(exc_detail::ceil_div(M * N, 256)), dim3(256), 0, s, g, splits);
    EXC_CHECK_HIP(hipGetLastError());
  }
  if (o.lower_only && o.mirror) {
    hipLaunchKernelGGL(k_mirror_lower<T>, dim3(exc_detail::ceil_div(M * M, 256)), dim3(256), 0, s, M, C, ldc);
    EXC_CHECK_HIP(hipGetLastError());
  }
  return EXC_OK;
}

// ---- grouped products: tables, launch and the autotuner
struct GroupPlan {
  std::vector<GroupProb> probs;
  std::vector<GroupItem> items;
  std::vector<GroupRed> reds;
  long slots = 0;
  double flops = 0.0;
};

GroupPlan group_plan(int nprob, const ExcGemmProblem* P, long kchunk) {
  GroupPlan g;
  if (kchunk > 0) kchunk = ((kchunk + BK - 1) / BK) * BK;
  for (int p = 0; p < nprob; p++) {
    const ExcGemmProblem& q = P[p];
    GroupProb gp{q.A, q.B, q.C, q.lda, q.ldb, q.ldc, q.M, q.N, q.K, q.tri, q.lower_only, q.mirror, 0};
    g.probs.push_back(gp);
    if (q.M <= 0 || q.N <= 0) continue;
    double f = 2.0 * (double)q.M * q.N * q.K;
    if (q.tri != EXC_TRI_NONE) f *= 0.5;
    if (q.lower_only) f *= 0.5;
    g.flops += f;
    const long tm = (q.M + BM - 1) / BM, tn = (q.N + BN - 1) / BN;
    const long splits = (kchunk > 0 && q.K > kchunk) ? (q.K + kchunk - 1) / kchunk : 1;
    for (long b = 0; b < tn; b++)
      for (long a = 0; a < tm; a++) {
        if (q.lower_only && a < b) continue;
        if (splits == 1 && !q.mirror) {
          g.items.push_back(GroupItem{p, (int)a, (int)b, -1, 0, q.K});
          continue;
        }
        const long first = g.slots;
        for (long sp = 0; sp < splits; sp++) {
          const long k0 = splits == 1 ? 0 : sp * kchunk;
          const long k1 = splits == 1 ? q.K : std::min(q.K, (sp + 1) * kchunk);
          g.items.push_back(GroupItem{p, (int)a, (int)b, (int)(g.slots++), k0, k1});
        }
        g.reds.push_back(GroupRed{p, (int)a, (int)b, (int)first, (int)splits, 0});
      }
  }
  return g;
}

size_t group_bytes(const GroupPlan& g, size_t tsize) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  return up(g.probs.size() * sizeof(GroupProb)) + up(g.items.size() * sizeof(GroupItem)) +
         up(std::max<size_t>(1, g.reds.size()) * sizeof(GroupRed)) + (size_t)g.slots * BM * BN * tsize;
}

template <typename T>
int group_run(int ta, int tb, const GroupPlan& g, double alpha, double beta, int grid, void* work, size_t wb,
              hipStream_t s) {
  EXC_REQUIRE(group_bytes(g, sizeof(T)) <= wb, EXC_ERR_SPACE, "exc_gemm_grouped: workspace too small");
  if (g.items.empty()) return EXC_OK;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  char* w = (char*)work;
  GroupProb* dp = (GroupProb*)w;
  w += up(g.probs.size() * sizeof(GroupProb));
  GroupItem* di = (GroupItem*)w;
  w += up(g.items.size() * sizeof(GroupItem));
  GroupRed* dr = (GroupRed*)w;
  w += up(std::max<size_t>(1, g.reds.size()) * sizeof(GroupRed));
  T* part = (T*)w;
  // the tables are small; copy them synchronously so no host buffer must outlive the call
  EXC_CHECK_HIP(hipMemcpyAsync(dp, g.probs.data(), g.probs.size() * sizeof(GroupProb), hipMemcpyHostToDevice, s));
  EXC_CHECK_HIP(hipMemcpyAsync(di, g.items.data(), g.items.size() * sizeof(GroupItem), hipMemcpyHostToDevice, s));
  if (!g.reds.empty())
    EXC_CHECK_HIP(hipMemcpyAsync(dr, g.reds.data(), g.reds.size() * sizeof(GroupRed), hipMemcpyHostToDevice, s));
  EXC_CHECK_HIP(hipStreamSynchronize(s));
  const long nitems = (long)g.items.size();
  if (grid <= 0) grid = (int)std::min<long>(nitems, resident_slots(sizeof(T) == 8 ? 2 : 3));
  grid = (int)std::min<long>(grid, nitems);
  hipEvent_t e0 = exc_prof::mark(s);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm_grouped<T, 0, 0>), dim3(grid), dim3(NT), 0, s, dp, di, nitems, part, alpha, beta);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm_grouped<T, 0, 1>), dim3(grid), dim3(NT), 0, s, dp, di, nitems, part, alpha, beta);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm_grouped<T, 1, 0>), dim3(grid), dim3(NT), 0, s, dp, di, nitems, part, alpha, beta);
  else hipLaunchKernelGGL((k_gemm_grouped<T, 1, 1>), dim3(grid), dim3(NT), 0, s, dp, di, nitems, part, alpha, beta);
  EXC_CHECK_HIP(hipGetLastError());
  if (!g.reds.empty()) {
    hipLaunchKernelGGL(k_grouped_reduce<T>, dim3((unsigned)g.reds.size()), dim3(256), 0, s, dp, dr, part, alpha,
                       beta);
    EXC_CHECK_HIP(hipGetLastError());
  }
  if (e0) exc_prof::book(sizeof(T) == 8 ? "gemm_grouped_f64" : "gemm_grouped_f32", e0, exc_prof::mark(s), g.flops, 0.0);
  return EXC_OK;
}

// Tuning results per problem set: the K chunk and the persistent grid that ran fastest.
struct TuneEntry {
  long key[8];
  long kchunk;
  int grid;
  double ms;
};
std::vector<TuneEntry> g_tune;
std::mutex g_tune_mu;

template <typename T>
int grouped_impl(int ta, int tb, int nprob, const ExcGemmProblem* P, double alpha, double beta, long kchunk, int grid,
                 void* work, size_t wb, hipStream_t s) {
  EXC_REQUIRE(nprob >= 0 && (nprob == 0 || P), EXC_ERR_ARG, "exc_gemm_grouped: bad arguments");
  if (kchunk >= 0) return group_run<T>(ta, tb, group_plan(nprob, P, kchunk), alpha, beta, grid, work, wb, s);
  // autotune: one timed run per candidate on the caller's data (each run overwrites the outputs
  // completely, so the last one leaves the result), the fastest kept per problem-set key
  long key[8] = {(long)sizeof(T), ta, tb, nprob, 0, 0, 0, 0};
  for (int p = 0; p < nprob; p++) {
    key[4] += P[p].K;
    key[5] = std::max(key[5], P[p].K);
    key[6] = std::max(key[6], P[p].M * 100003 + P[p].N);
    key[7] = key[7] * 31 + P[p].lower_only * 2 + P[p].tri * 7 + (P[p].beta_zero_hint ? 1 : 0);
  }
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const TuneEntry& e : g_tune)
      if (std::equal(e.key, e.key + 8, key))
        return group_run<T>(ta, tb, group_plan(nprob, P, e.kchunk), alpha, beta, e.grid, work, wb, s);
  }
  EXC_REQUIRE(beta == 0.0, EXC_ERR_ARG, "exc_gemm_grouped: tuning runs need beta = 0 (each run overwrites C)");
  const long slots = resident_slots(sizeof(T) == 8 ? 2 : 3);
  const long chunks[] = {0, 16384, 8192, 4096, 2048};
  const int mults[] = {1, 2};
  hipEvent_t a, b;
  EXC_CHECK_HIP(hipEventCreate(&a));
  EXC_CHECK_HIP(hipEventCreate(&b));
  TuneEntry best{};
  best.ms = -1.0;
  for (long kc : chunks) {
    const GroupPlan g = group_plan(nprob, P, kc);
    if (group_bytes(g, sizeof(T)) > wb) continue;
    for (int mu : mults) {
      const int gr = (int)std::min<long>((long)g.items.size(), slots * mu);
      int rc = group_run<T>(ta, tb, g, alpha, beta, gr, work, wb, s);  // warm (and tables uploaded)
      if (rc) return rc;
      EXC_CHECK_HIP(hipEventRecord(a, s));
      rc = group_run<T>(ta, tb, g, alpha, beta, gr, work, wb, s);
      if (rc) return rc;
      EXC_CHECK_HIP(hipEventRecord(b, s));
      EXC_CHECK_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      EXC_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
      if (best.ms < 0 || ms < best.ms) {
        best.ms = ms;
        best.kchunk = kc;
        best.grid = gr;
      }
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  EXC_REQUIRE(best.ms >= 0, EXC_ERR_SPACE, "exc_gemm_grouped: workspace too small for any tuning candidate");
  std::copy(key, key + 8, best.key);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.push_back(best);
  }
  return group_run<T>(ta, tb, group_plan(nprob, P, best.kchunk), alpha, beta, best.grid, work, wb, s);
}

}  // namespace

int exc_gemm_grouped_opts(int dtype, int ta, int tb, int nprob, const ExcGemmProblem* probs, double alpha, double beta,
                          long kchunk, int grid, void* work, size_t work_bytes, hipStream_t s) {
  EXC_REQUIRE(dtype == 4 || dtype == 8, EXC_ERR_ARG, "exc_gemm_grouped: dtype is 4 (float) or 8 (double)");
  if (dtype == 8) return grouped_impl<double>(ta, tb, nprob, probs, alpha, beta, kchunk, grid, work, work_bytes, s);
  return grouped_impl<float>(ta, tb, nprob, probs, alpha, beta, kchunk, grid, work, work_bytes, s);
}

size_t exc_gemm_grouped_workspace(int dtype, int nprob, const ExcGemmProblem* probs, long kchunk) {
  // the autotuner's smallest chunk needs the most partial slots
  const long kc = kchunk < 0 ? 2048 : kchunk;
  return group_bytes(group_plan(nprob, probs, kc), dtype == 8 ? 8 : 4) + 256;
}

extern "C" int exc_gemm_grouped_tuning(char* buf, size_t cap) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  std::string out = "[";
  for (size_t i = 0; i < g_tune.size(); i++) {
    char line[256];
    snprintf(line, sizeof line, "%s{\"dtype\": %ld, \"problems\": %ld, \"sum_k\": %ld, \"kchunk\": %ld, \"grid\": %d, \"ms\": %.4f}",
             i ? ", " : "", g_tune[i].key[0], g_tune[i].key[3], g_tune[i].key[4], g_tune[i].kchunk, g_tune[i].grid,
             g_tune[i].ms);
    out += line;
  }
  out += "]";
  if (buf && cap) {
    strncpy(buf, out.c_str(), cap - 1);
    buf[cap - 1] = 0;
  }
  return EXC_OK;
}

int exc_gemm_opts(int ta, int tb, long M, long N, long K, double alpha, const double* A, long lda, const double* B,
                  long ldb, double beta, double* C, long ldc, const ExcGemmOpts& o, hipStream_t s) {
  return gemm_impl<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, o, s);
}

int exc_sgemm_opts(int ta, int tb, long M, long N, long K, double alpha, const float* A, long lda, const float* B,
                   long ldb, double beta, float* C, long ldc, const ExcGemmOpts& o, hipStream_t s) {
  return gemm_impl<float>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, o, s);
}

extern "C" {

size_t exc_gemm_workspace(long M, long N, long K) {
  (void)K;
  return (size_t)8 * M * N * sizeof(double);
}

int exc_dgemm(int trans_a, int trans_b, long M, long N, long K, double alpha, const double* d_a, long lda,
              const double* d_b, long ldb, double beta, double* d_c, long ldc, void* d_work, size_t work_bytes,
              exc_stream_t stream) {
  ExcGemmOpts o;
  o.work = d_work;
  o.work_bytes = work_bytes;
  return gemm_impl<double>(trans_a, trans_b, M, N, K, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, o,
                           exc_detail::stream(stream));
}

int exc_sgemm(int trans_a, int trans_b, long M, long N, long K, double alpha, const float* d_a, long lda,
              const float* d_b, long ldb, double beta, float* d_c, long ldc, void* d_work, size_t work_bytes,
              exc_stream_t stream) {
  ExcGemmOpts o;
  o.work = d_work;
  o.work_bytes = work_bytes;
  return gemm_impl<float>(trans_a, trans_b, M, N, K, alpha, d_a, lda, d_b, ldb, beta, d_c, ldc, o,
                          exc_detail::stream(stream));
}

}  // extern "C"
