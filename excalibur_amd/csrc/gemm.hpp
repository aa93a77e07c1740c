// Internal GEMM entry points shared by the factorization kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

// Triangular restriction of one operand (after transposition): the other triangle is treated as
// zero and the zero K range of each output tile is skipped.
enum { EXC_TRI_NONE = 0, EXC_TRI_A_LOWER = 1, EXC_TRI_A_UPPER = 2, EXC_TRI_B_LOWER = 3, EXC_TRI_B_UPPER = 4 };

struct ExcGemmOpts {
  int tri = EXC_TRI_NONE;
  int lower_only = 0;  // square C: compute tiles on and below the diagonal only
  int mirror = 0;      // with lower_only: copy the lower triangle to the upper one afterwards
  void* work = nullptr;  // split-K partials (split-K is used only when given)
  size_t work_bytes = 0;
};

int exc_gemm_opts(int ta, int tb, long M, long N, long K, double alpha, const double* A, long lda, const double* B,
                  long ldb, double beta, double* C, long ldc, const ExcGemmOpts& o, hipStream_t s);
int exc_sgemm_opts(int ta, int tb, long M, long N, long K, double alpha, const float* A, long lda, const float* B,
                   long ldb, double beta, float* C, long ldc, const ExcGemmOpts& o, hipStream_t s);
