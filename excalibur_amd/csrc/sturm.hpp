// The Sturm count of negative_count (bbtsvd.cpp:7-17), shared by the bisection kernels. Every
// operation is rounded separately (FMA contraction is off for all device code of the library), so
// the counts equal the reference's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

// Number of singular values of the upper bidiagonal B (diagonal a, superdiagonal b) below mu.
__device__ __forceinline__ int sturm_count(int n, const double* __restrict__ a, const double* __restrict__ b,
                                           double mu) {
  const double mu2 = mu * mu;
  double t = -mu2;
  double d = a[0] * a[0] + t;
  int count = d < 0 ? 1 : 0;
  // unrolled so the wave-uniform loads of a and b are issued ahead of the dependent chain
#pragma unroll 8
  for (int i = 1; i < n; i++) {
    const double bb = b[i - 1] * b[i - 1];
    const double aa = a[i] * a[i];
    t = t * (bb / d) - mu2;
    d = aa + t;
    count += d < 0 ? 1 : 0;
  }
  return count;
}
