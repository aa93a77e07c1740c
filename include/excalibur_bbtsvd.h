/*
 * excalibur_bbtsvd.h -- host-pointer C ABI of libexcalibur_amd.so with the semantics of the
 * reference's bbtsvd.hpp functions on the bbtsvd path. Arrays are host memory, row-major
 * double, updated in place or written like the reference writes them; every call runs on the
 * GPU (copy in, the device kernels of excalibur_hip.h on the default stream, copy out) and
 * returns when its outputs are on the host. Return value: EXC_OK (0) or a negative EXC_ERR_*
 * code with exc_last_error() describing it (the reference asserts or ignores instead).
 * include/bbtsvd.hpp wraps these in C++ functions with the reference's names and signatures.
 */
#ifndef EXCALIBUR_BBTSVD_H
#define EXCALIBUR_BBTSVD_H

#include <stddef.h>

#include "excalibur_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* negative_count, bbtsvd.hpp:22 / bbtsvd.cpp:7-17: *count = #{sigma(B) < mu}. */
int exc_bb_negative_count(int n, const double* x_diagonal, const double* x_upper_diagonal, double mu, int* count);

/* bisection_singular_values, bbtsvd.hpp:23 / bbtsvd.cpp:19-52: the leaves of the DFS bisection
 * of [a, b] with absolute tolerance tol, descending, at most top_k of them; *sigma_count = 0
 * when top_k <= 0 or nc(a) > nc(b). */
int exc_bb_bisection_singular_values(int n, const double* x_diagonal, const double* x_upper_diagonal, double a,
                                     double b, double tol, int top_k, double* sigma, int* sigma_count);

/* ldl_decomposition, bbtsvd.hpp:24 / bbtsvd.cpp:54-63: l (n-1), d (n). */
int exc_bb_ldl_decomposition(int n, const double* x_diagonal, const double* x_upper_diagonal, double* l, double* d);

/* inverse_iteration, bbtsvd.hpp:25 / bbtsvd.cpp:65-71: v[i*ldv], i < n, from twist k. */
int exc_bb_inverse_iteration(int n, int k, const double* l, const double* u, double* v, int ldv);

/* dstqds, bbtsvd.hpp:26 / bbtsvd.cpp:73-81: l_plus (n-1), d_plus (n), s (n). */
int exc_bb_dstqds(int n, const double* l, const double* d, double tau, double* l_plus, double* d_plus, double* s);

/* dpqds, bbtsvd.hpp:27 / bbtsvd.cpp:83-91: u_plus (n-1), r_plus (n), p (n). */
int exc_bb_dpqds(int n, const double* l, const double* d, double tau, double* u_plus, double* r_plus, double* p);

/* one_pair_singular_vector_twisted_factorization, bbtsvd.hpp:28 / bbtsvd.cpp:93-135: v (n,
 * contiguous) and u (n entries at stride omega_count) for the singular value omega. */
int exc_bb_one_pair_twisted(int n, int omega_count, double omega, const double* l, const double* d,
                            const double* x_diagonal, const double* x_upper_diagonal, double* v, double* u);

/* two_side_singular_vectors_twisted_factorization, bbtsvd.hpp:29 / bbtsvd.cpp:138-147:
 * u (n x omega_count, column i = u_i), v (omega_count x n, row i = v_i). */
int exc_bb_two_side_twisted(int n, int omega_count, const double* omega, const double* x_diagonal,
                            const double* x_upper_diagonal, double* u, double* v);

/* 7-argument bidiagonal_bisection_twisted_svd, bbtsvd.hpp:35 / bbtsvd.cpp:932-968: u (k x top_k),
 * v (top_k x k), sigma (top_k). EXC_ERR_COUNT where the reference's assert at bbtsvd.cpp:938
 * fails. */
int exc_bb_bidiagonal_svd(int k, const double* alpha, const double* beta, double* bidiagonal_u,
                          double* bidiagonal_v, double* sigma, int top_k);

/* givens_mask_cache_optimized_parallel_with_simd, bbtsvd.hpp:32 / bbtsvd.cpp:727-876: x (m x n)
 * in place; reseeds the C library's rand() with seed like the reference. */
int exc_bb_givens_mask(size_t m, size_t n, size_t mask_times, double* x, int mask, int is_left, unsigned int seed);

/* block_shift_mask, bbtsvd.hpp:47 / bbtsvd.cpp:533-669: K <= 10000 rotated lines take the plain
 * sweep (bbtsvd.cpp:533-546), K > 10000 the virtual-matrix mask (bbtsvd.cpp:548-668). */
int exc_bb_block_shift_mask(size_t m, size_t n, size_t mask_times, double* x, int mask, int is_left,
                            unsigned int seed);

/* find_best_convert, bbtsvd.hpp:39 / bbtsvd.cpp:166-188 (host integer search). */
int exc_bb_find_best_convert(size_t k, size_t* out_m, size_t* out_n, size_t* out_remain);

#ifdef __cplusplus
}
#endif

#endif /* EXCALIBUR_BBTSVD_H */
