// block_shift_mask for more than 10000 rotated lines (bbtsvd.cpp:548-668): every line of K values
// is viewed as a vm x vn matrix (K - rem = vm vn, find_best_convert) plus a remainder of rem
// values. The mask is a left Givens sweep of that matrix, a cyclic shift of its rows, a left and a
// right sweep, and two sweeps over the vector [first column, remainder]. All lines share the
// plans (seeds + 1000 .. + 5000), so each sweep runs once over all lines through a reshaped view:
//   Z (lines*vm x vn): line i's matrix in rows i*vm .. i*vm+vm-1 (its natural layout) takes the
//     right sweeps;
//   Y (vm x lines*vn): Y[j][i*vn + q] = line i's [j][q] takes the left sweeps;
//   RB (lines x (vm + rem)): the first column and the remainder of every line takes the vector
//     sweeps.
// The sweeps are exc_givens_apply's, so every rotation is the reference's to the bit; the views
// are exact permutations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "exc_common.hpp"
#include "excalibur_bbtsvd.h"

namespace {

// lines (L x K, row-major; or the transposed layout when lines are columns: element q of line i at
// x[q * ld + i]) <-> Z (L*vm x vn) and RB[:, vm:] (the remainder).
__global__ __launch_bounds__(256) void k_lines_to_z(long L, long K, long kk, const double* __restrict__ x, long ld,
                                                    int col_lines, double* __restrict__ Z, double* __restrict__ rb,
                                                    long ldrb, long vm) {
  const long total = L * K;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e / K, q = e - i * K;
    const double v = col_lines ? x[q * ld + i] : x[i * ld + q];
    if (q < kk) Z[i * kk + q] = v;
    else rb[i * ldrb + vm + (q - kk)] = v;
  }
}

__global__ __launch_bounds__(256) void k_z_to_lines(long L, long K, long kk, const double* __restrict__ Z,
                                                    const double* __restrict__ rb, long ldrb, long vm,
                                                    double* __restrict__ x, long ld, int col_lines) {
  const long total = L * K;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e / K, q = e - i * K;
    const double v = q < kk ? Z[i * kk + q] : rb[i * ldrb + vm + (q - kk)];
    if (col_lines) x[q * ld + i] = v;
    else x[i * ld + q] = v;
  }
}

// RB[i][j] <- Z[i*vm + j][0] (to_rb) or back.
__global__ __launch_bounds__(256) void k_first_col(long L, long vm, long vn, double* __restrict__ Z,
                                                   double* __restrict__ rb, long ldrb, int to_rb) {
  const long total = L * vm;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e / vm, j = e - i * vm;
    if (to_rb) rb[i * ldrb + j] = Z[(i * vm + j) * vn];
    else Z[(i * vm + j) * vn] = rb[i * ldrb + j];
  }
}

// Y[j][i*vn + q] = Z[i*vm + j][q] (to_y) or back.
__global__ __launch_bounds__(256) void k_z_y(long L, long vm, long vn, double* __restrict__ Z, double* __restrict__ Y,
                                             int to_y) {
  const long total = L * vm * vn;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long i = e / (vm * vn), r = e - i * vm * vn, j = r / vn, q = r - j * vn;
    if (to_y) Y[j * (L * vn) + i * vn + q] = Z[e];
    else Z[e] = Y[j * (L * vn) + i * vn + q];
  }
}

// Cyclic shift of the rows of every line's matrix: forward dst[j][q] = src[j][(q - j) mod vn]
// (bbtsvd.cpp:582-588), backward dst[j][q] = src[j][(q + j) mod vn] (bbtsvd.cpp:641-644).
__global__ __launch_bounds__(256) void k_shift(long L, long vm, long vn, const double* __restrict__ src,
                                               double* __restrict__ dst, int forward) {
  const long total = L * vm * vn;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e % (vm * vn), j = r / vn, q = r - j * vn;
    long s = forward ? (q - j) % vn : (q + j) % vn;
    if (s < 0) s += vn;
    dst[e] = src[e - q + s];
  }
}

unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>(16384, (n + 255) / 256)); }

void vm_shape(long K, long* vm, long* vn, long* rem) {
  size_t a = 0, b = 0, r = 0;
  exc_bb_find_best_convert((size_t)K, &a, &b, &r);
  *vm = (long)a;
  *vn = (long)b;
  *rem = (long)r;
}

size_t plan_doubles(long K, long passes) {
  const long half = exc_givens_half(K);
  return (size_t)(passes * half * 2 * sizeof(int32_t) + 7) / 8 + (size_t)passes * half * 2;
}

}  // namespace

extern "C" {

size_t exc_block_shift_mask_workspace(long m, long n, long passes, int is_left) {
  const long K = is_left ? m : n, L = is_left ? n : m;
  if (K <= 1) return 256;
  if (K <= 10000) return plan_doubles(K, passes) * 8 + 512;
  long vm, vn, rem;
  vm_shape(K, &vm, &vn, &rem);
  const size_t kk = (size_t)vm * vn;
  size_t d = 3 * kk * (size_t)L + (size_t)L * (vm + rem) + 16;
  d += 2 * plan_doubles(vm + std::max<long>(rem, 1), passes) + 2 * plan_doubles(vm, passes) + plan_doubles(vn, passes) + 64;
  return d * sizeof(double) + 10 * 256;
}

// d_x: m x n row-major, in place; lines are the columns when is_left (K = m), else the rows.
int exc_block_shift_mask(double* d_x, long m, long n, long passes, int mask, int is_left, unsigned seed, void* d_work,
                         size_t work_bytes, exc_stream_t stream) {
  const long K = is_left ? m : n, L = is_left ? n : m;
  EXC_REQUIRE(d_x && m >= 1 && n >= 1 && passes >= 0, EXC_ERR_ARG, "exc_block_shift_mask: bad arguments");
  hipStream_t s = exc_detail::stream(stream);
  if (K <= 10000) {
    // bbtsvd.cpp:541-545: the plain sweep
    const long half = exc_givens_half(K);
    if (K <= 1) return EXC_OK;
    std::vector<int32_t> pairs((size_t)passes * half * 2);
    std::vector<double> cs((size_t)passes * half * 2);
    int rc = exc_givens_plan(K, passes, mask, seed, pairs.data(), cs.data());
    if (rc) return rc;
    EXC_REQUIRE(work_bytes >= plan_doubles(K, passes) * 8 + 256 && d_work, EXC_ERR_SPACE,
                "exc_block_shift_mask: workspace too small for the plan");
    int32_t* dp = (int32_t*)d_work;
    double* dc = (double*)d_work + (pairs.size() * sizeof(int32_t) + 7) / 8;
    EXC_CHECK_HIP(hipMemcpyAsync(dp, pairs.data(), pairs.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    EXC_CHECK_HIP(hipMemcpyAsync(dc, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice, s));
    rc = exc_givens_apply(d_x, m, n, dp, dc, passes, mask, is_left, stream);
    EXC_CHECK_HIP(hipStreamSynchronize(s));  // the host plan buffers go out of scope
    return rc;
  }
  EXC_REQUIRE(d_work && work_bytes >= exc_block_shift_mask_workspace(m, n, passes, is_left), EXC_ERR_SPACE,
              "exc_block_shift_mask: workspace too small");
  long vm, vn, rem;
  vm_shape(K, &vm, &vn, &rem);
  const long kk = vm * vn, wrb = vm + rem;
  const bool remain = rem > 0;
  char* w = (char*)d_work;
  auto take = [&](size_t doubles) {
    double* p = (double*)w;
    w += (doubles * sizeof(double) + 255) / 256 * 256;
    return p;
  };
  double* Z = take((size_t)kk * L);
  double* Z2 = take((size_t)kk * L);
  double* Y = take((size_t)kk * L);
  double* rb = take((size_t)L * wrb + 1);
  // plans: 0 remain (seed+1000), 1 block 1 (vm, +2000), 2 block 2 left (vm, +3000), 3 block 2 right
  // (vn, +4000), 4 remain again (seed+5000)
  const long Ks[5] = {wrb, vm, vm, vn, wrb};
  const unsigned sd[5] = {seed + 1000, seed + 2000, seed + 3000, seed + 4000, seed + 5000};
  int32_t* dp[5] = {nullptr};
  double* dc[5] = {nullptr};
  std::vector<std::vector<int32_t>> hp(5);
  std::vector<std::vector<double>> hc(5);
  for (int q = 0; q < 5; q++) {
    if ((q == 0 || q == 4) && !remain) continue;
    const long half = exc_givens_half(Ks[q]);
    hp[q].resize((size_t)passes * half * 2);
    hc[q].resize((size_t)passes * half * 2);
    const int rc = exc_givens_plan(Ks[q], passes, mask, sd[q], hp[q].data(), hc[q].data());
    if (rc) return rc;
    dp[q] = (int32_t*)take((hp[q].size() * sizeof(int32_t) + 7) / 8);
    dc[q] = take(hc[q].size());
    EXC_CHECK_HIP(hipMemcpyAsync(dp[q], hp[q].data(), hp[q].size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    EXC_CHECK_HIP(hipMemcpyAsync(dc[q], hc[q].data(), hc[q].size() * sizeof(double), hipMemcpyHostToDevice, s));
  }
  const long ld = n;
  const int col_lines = is_left ? 1 : 0;
  const unsigned gz = grid_for((long)L * kk);
  auto sweep = [&](double* v, long rows, long cols, int q, int left) {
    return exc_givens_apply(v, rows, cols, dp[q], dc[q], passes, mask, left, stream);
  };
  auto left_sweep = [&](double* Zm, int q) {  // left sweep (K = vm) of every line's matrix in Zm
    hipLaunchKernelGGL(k_z_y, dim3(gz), dim3(256), 0, s, (long)L, vm, vn, Zm, Y, 1);
    int rc = sweep(Y, vm, (long)L * vn, q, 1);
    hipLaunchKernelGGL(k_z_y, dim3(gz), dim3(256), 0, s, (long)L, vm, vn, Zm, Y, 0);
    return rc;
  };
  hipLaunchKernelGGL(k_lines_to_z, dim3(grid_for(L * K)), dim3(256), 0, s, (long)L, K, kk, d_x, ld, col_lines,
                     mask ? Z : Z2, rb, wrb, vm);
  int rc = EXC_OK;
  if (mask) {
    if (remain) {
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z, rb, wrb, 1);
      rc = sweep(rb, L, wrb, 0, 0);
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z, rb, wrb, 0);
    }
    if (!rc) rc = left_sweep(Z, 1);
    hipLaunchKernelGGL(k_shift, dim3(gz), dim3(256), 0, s, (long)L, vm, vn, Z, Z2, 1);
    if (!rc) rc = left_sweep(Z2, 2);
    if (!rc) rc = sweep(Z2, (long)L * vm, vn, 3, 0);
    if (!rc && remain) {
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z2, rb, wrb, 1);
      rc = sweep(rb, L, wrb, 4, 0);
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z2, rb, wrb, 0);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_z_to_lines, dim3(grid_for(L * K)), dim3(256), 0, s, (long)L, K, kk, Z2, rb, wrb, vm, d_x, ld,
                       col_lines);
  } else {
    // the masked data sits in Z2 (the shifted block) and rb's remainder
    if (remain) {
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z2, rb, wrb, 1);
      rc = sweep(rb, L, wrb, 4, 0);
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z2, rb, wrb, 0);
    }
    if (!rc) rc = left_sweep(Z2, 2);
    if (!rc) rc = sweep(Z2, (long)L * vm, vn, 3, 0);
    hipLaunchKernelGGL(k_shift, dim3(gz), dim3(256), 0, s, (long)L, vm, vn, Z2, Z, 0);
    if (!rc) rc = left_sweep(Z, 1);
    if (!rc && remain) {
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z, rb, wrb, 1);
      rc = sweep(rb, L, wrb, 0, 0);
      hipLaunchKernelGGL(k_first_col, dim3(grid_for(L * vm)), dim3(256), 0, s, (long)L, vm, vn, Z, rb, wrb, 0);
    }
    if (rc) return rc;
    hipLaunchKernelGGL(k_z_to_lines, dim3(grid_for(L * K)), dim3(256), 0, s, (long)L, K, kk, Z, rb, wrb, vm, d_x, ld,
                       col_lines);
  }
  EXC_CHECK_HIP(hipGetLastError());
  EXC_CHECK_HIP(hipStreamSynchronize(s));  // the host plan buffers go out of scope
  return EXC_OK;
}

}  // extern "C"
