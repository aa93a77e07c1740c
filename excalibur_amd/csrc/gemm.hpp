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

// One problem of a grouped product: C = alpha op(A) op(B) + beta C (row-major); lower_only: only
// the tiles on and below the diagonal of a square C; mirror: copy that lower triangle onto the
// upper one.
struct ExcGemmProblem {
  const void* A;
  long lda;
  const void* B;
  long ldb;
  void* C;
  long ldc;
  long M, N, K;
  int tri = EXC_TRI_NONE;
  int lower_only = 0;
  int mirror = 0;
  int beta_zero_hint = 0;
};
// Many independent products (all FP32 when dtype = 4, all FP64 when 8; one transposition pattern)
// in one persistent launch plus one reduction launch. kchunk: split K into chunks of this length
// (0: no split); kchunk < 0 lets the autotuner choose the chunk and the persistent grid size by
// timing candidates on these problems once (beta must be 0), then reuse the choice for problem
// sets of the same shape. grid <= 0: one workgroup per resident slot.
int exc_gemm_grouped_opts(int dtype, int ta, int tb, int nprob, const ExcGemmProblem* probs, double alpha, double beta,
                          long kchunk, int grid, void* work, size_t work_bytes, hipStream_t s);
size_t exc_gemm_grouped_workspace(int dtype, int nprob, const ExcGemmProblem* probs, long kchunk);
// The autotuner's choices so far as a JSON list (extern "C", declared in excalibur_hip.h).
extern "C" int exc_gemm_grouped_tuning(char* buf, size_t cap);
