// Substitution on one lane's register vector against a triangular block of at most 64 x 64
// read with wave-uniform addresses (shared by the triangular-solve leaves and the tile
// factorizations). FULL: the block is 64 wide, so every bound is a constant.
#pragma once

#include "exc_common.hpp"
#include "fact.hpp"

namespace fact {

// Lane src's value of v, as a wave-uniform (scalar) value.
__device__ __forceinline__ double readlane_f64(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// A triangle element: a plain load, or (COH) a coherent one for a block another workgroup of the
// running grid has written.
template <bool COH>
__device__ __forceinline__ double tri_rd(const double* p) {
  return COH ? coh_ld(p) : *p;
}

// Substitution on one lane's column x (FULL: kb == NB, so every bound is a constant).
template <bool UPPER, bool UNIT, bool FULL, bool COH = false>
__device__ __forceinline__ void left_solve(int kb, const double* __restrict__ L, long ldl, double (&x)[NB]) {
#pragma unroll
  for (int q = 0; q < NB; q++) {
    const int i = UPPER ? NB - 1 - q : q;
    if (FULL || i < kb) {
      const double* Li = L + (long)i * ldl;
      double a0 = x[i], a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
      for (int k = UPPER ? i + 1 : 0; k < (UPPER ? NB : i); k++) {
        if (FULL || k < kb) {
          if ((k & 3) == 0) a0 = fma(-tri_rd<COH>(Li + (k)), x[k], a0);
          else if ((k & 3) == 1) a1 = fma(-tri_rd<COH>(Li + (k)), x[k], a1);
          else if ((k & 3) == 2) a2 = fma(-tri_rd<COH>(Li + (k)), x[k], a2);
          else a3 = fma(-tri_rd<COH>(Li + (k)), x[k], a3);
        }
      }
      const double acc = (a0 + a1) + (a2 + a3);
      x[i] = UNIT ? acc : acc / tri_rd<COH>(Li + (i));
    }
  }
}

// Substitution from the right on one lane's row x: x op(U) = b.
template <bool LOWER_T, bool UNIT, bool FULL, bool COH = false>
__device__ __forceinline__ void right_solve(int kb, const double* __restrict__ U, long ldu, double (&x)[NB]) {
  if (LOWER_T) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if (FULL || j < kb) {
        const double* Lj = U + (long)j * ldu;  // op(U)[k][j] = L[j][k]: row j of L
        double a0 = x[j], a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll
        for (int k = 0; k < j; k++) {
          if ((k & 3) == 0) a0 = fma(-x[k], tri_rd<COH>(Lj + (k)), a0);
          else if ((k & 3) == 1) a1 = fma(-x[k], tri_rd<COH>(Lj + (k)), a1);
          else if ((k & 3) == 2) a2 = fma(-x[k], tri_rd<COH>(Lj + (k)), a2);
          else a3 = fma(-x[k], tri_rd<COH>(Lj + (k)), a3);
        }
        const double acc = (a0 + a1) + (a2 + a3);
        x[j] = UNIT ? acc : acc / tri_rd<COH>(Lj + (j));
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if (FULL || j < kb) {
        const double* Uj = U + (long)j * ldu;  // row j of U, right-looking
        if (!UNIT) x[j] = x[j] / tri_rd<COH>(Uj + (j));
#pragma unroll
        for (int k = j + 1; k < NB; k++)
          if (FULL || k < kb) x[k] = fma(-x[j], tri_rd<COH>(Uj + (k)), x[k]);
      }
    }
  }
}

}  // namespace fact
