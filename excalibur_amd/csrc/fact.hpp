// Internal dense-factorization building blocks (row-major FP64 on the GPU).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace fact {

constexpr int NB = 64;  // diagonal block size (LDS for the factorizations, registers for the solves)

// In-place lower Cholesky of the symmetric n x n matrix A (lower triangle read); the strict
// upper triangle is zeroed. info (device int, nullable) receives 0 or the 1-based index of the
// first non-positive pivot.
int potrf_lower(int n, double* A, long lda, int* info, void* work, size_t work_bytes, hipStream_t s);

// X <- op(L)^-1 X from the left; L is n x n lower (upper when upper != 0) triangular, unit
// diagonal when unit != 0; X is n x nrhs. Blocked: GEMM updates on the matrix cores, the
// diagonal blocks solved by substitution (one lane per right-hand side).
int trsm_left(int upper, int unit, int n, long nrhs, const double* L, long ldl, double* X, long ldx, void* work,
              size_t work_bytes, hipStream_t s);

// X <- X U^-1 from the right; U is n x n upper triangular (unit when unit != 0), or
// X <- X L^-T when lower_t != 0 (L lower). X is rows x n.
int trsm_right(int lower_t, int unit, int n, long rows, const double* U, long ldu, double* X, long ldx, void* work,
               size_t work_bytes, hipStream_t s);

// LU without pivoting of A - S with S = diag(+-1) chosen column by column as
// S_jj = -sign(current diagonal) (Householder reconstruction sign rule). A is overwritten by
// L (strict lower, unit diagonal implied) and U (upper); sgn[j] = S_jj. With square_last the
// last column takes S = +sign(diagonal) instead: LAPACK leaves the last column of a square
// matrix unreflected (tau = 0), so its R entry keeps its sign and its pivot is zero.
int lu_sign(int n, double* A, long lda, double* sgn, int square_last, void* work, size_t work_bytes, hipStream_t s);

size_t work_bytes(int n, long wide);

// One-launch tile-parallel forms (tiles.hip), used by potrf_lower and lu_sign for
// 64 < n <= kTileMaxN. Both synchronize the stream to check their bounded waits.
constexpr int kTileMaxN = 4096;
size_t tiles_work_bytes(int n);
int potrf_tiles(int n, double* A, long lda, int* info, void* work, size_t work_bytes, hipStream_t s);
int lu_tiles(int n, double* A, long lda, double* sgn, int square_last, void* work, size_t work_bytes, hipStream_t s);
// Batched tile forms over device arrays of matrix addresses (n <= kTileMaxN), one launch each.
size_t tiles_batched_work_bytes(int n, int count);
int potrf_tiles_batched(int n, double* const* d_As, int count, long lda, int* d_info, void* work, size_t work_bytes,
                        hipStream_t s);
int trinv_tiles_batched(int n, const double* const* d_Ls, long ldl, double* const* d_Xs, long ldx, int count,
                        void* work, size_t work_bytes, hipStream_t s);

}  // namespace fact
