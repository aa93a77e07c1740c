// Internal QR / compact-WY entry points (row-major FP64, transposed tall layout).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

struct QrOut {
  double* rt = nullptr;   // m x m, ld m: R^T (lower), LAPACK signs
  double* qt = nullptr;   // m x r: Q^T (nullable; may alias the input)
  long ldq = 0;
  double* vt = nullptr;   // m x r: Householder vectors, one per row (nullable)
  long ldv = 0;
  double* tau = nullptr;  // m (nullable)
  double* sgn = nullptr;  // m: S of the reconstruction (nullable)
};

namespace qr {
size_t qr_t_workspace(int m, long r, bool want_q);
// QR of X^T for X (m x r, ld ldx, r >= m) in LAPACK's Householder convention.
int qr_t(int m, long r, const double* X, long ldx, const QrOut& o, int square_last, void* work, size_t wb,
         hipStream_t s);
size_t wy_workspace(int k, long L, long rows);
// Mt (rows x L) <- Mt Q, Q = H_0 ... H_{k-1} from reflector rows vt (k x L) and tau: the
// transposed form of dormqr('L', 'T') on the column-stored matrix.
int apply_q_right(long rows, long L, int k, double* Mt, long ldm, const double* vt, long ldv, const double* tau,
                  void* work, size_t wb, hipStream_t s);
// Q^T (k x L) of dorgqr from reflector rows vt and tau.
int orgqr_t(int k, long L, const double* vt, long ldv, const double* tau, double* qt, long ldq, void* work,
            size_t wb, hipStream_t s);
size_t ormqr_house_workspace(int m, long n);
// C (m x n) <- H_0 ... H_{m-1} C, reflector j stored in column j of the row-major m x m house
// (rows j+1.. below an implied unit diagonal): dormqr('L', 'N') of client.cpp:1176.
int ormqr_house_ln(int m, long n, const double* house, const double* tau, double* C, long ldc, void* work,
                   size_t wb, hipStream_t s);
void transpose(long rows, long cols, const double* src, long lds, double* dst, long ldd, hipStream_t s);
void copy2d(long rows, long cols, const double* src, long lds, double* dst, long ldd, hipStream_t s);
}  // namespace qr
