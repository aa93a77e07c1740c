// Drop-in for the reference's include/bbtsvd.hpp (bbtsvd.hpp:11-47) on the bbtsvd path: the same
// structs and the same C++ function names and signatures, as inline wrappers over the
// host-pointer C ABI of excalibur_bbtsvd.h. Every function runs on the GPU and returns with its
// outputs in the caller's host arrays. Where the reference asserts (bbtsvd.cpp:168, 539, 938)
// these throw std::runtime_error with the library's message; a HIP failure throws too.
//
// Not provided: the pointer-to-pointer overloads of bidiagonal_bisection_twisted_svd
// (bbtsvd.hpp:33-34, never called by the client) and the dense-mask ablation helpers
// (generate_house, house_mask, householder_reflector_mask, naive_givens_mask_parallel,
// givens_mask_cache_optimized_parallel, generate_random_orthogonal_matrix, shuffle_matrix,
// block_mask, shuffle_block_mask, mmap_drot).
#ifndef EXCALIBUR_BBTSVD_HPP
#define EXCALIBUR_BBTSVD_HPP

#include <cstddef>
#include <stdexcept>
#include <string>

#include "excalibur_bbtsvd.h"

struct BISECTION_SV {  // bbtsvd.hpp:11-16
  double low;
  int nc_low;
  double high;
  int nc_high;
};

struct SquareMaskConvert {  // bbtsvd.hpp:18-20
  size_t k, m, n, remain;
};

namespace exc_bbtsvd_detail {
inline void check(int rc, const char* what) {
  if (rc != EXC_OK) throw std::runtime_error(std::string(what) + ": " + exc_last_error());
}
}  // namespace exc_bbtsvd_detail

inline int negative_count(int n, double* x_diagonal, double* x_upper_diagonal, double mu) {
  int c = 0;
  exc_bbtsvd_detail::check(exc_bb_negative_count(n, x_diagonal, x_upper_diagonal, mu, &c), "negative_count");
  return c;
}

inline void bisection_singular_values(int n, double* x_diagonal, double* x_upper_diagonal, double a, double b,
                                      double tol, int top_k, double* sigma, int* sigma_count) {
  exc_bbtsvd_detail::check(
      exc_bb_bisection_singular_values(n, x_diagonal, x_upper_diagonal, a, b, tol, top_k, sigma, sigma_count),
      "bisection_singular_values");
}

inline void ldl_decomposition(int n, double* x_diagonal, double* x_upper_diagonal, double* l, double* d) {
  exc_bbtsvd_detail::check(exc_bb_ldl_decomposition(n, x_diagonal, x_upper_diagonal, l, d), "ldl_decomposition");
}

inline void inverse_iteration(int n, int k, double* l, double* u, double* v, int ldv) {
  exc_bbtsvd_detail::check(exc_bb_inverse_iteration(n, k, l, u, v, ldv), "inverse_iteration");
}

inline void dstqds(int n, double* l, double* d, double tau, double* l_plus, double* d_plus, double* s) {
  exc_bbtsvd_detail::check(exc_bb_dstqds(n, l, d, tau, l_plus, d_plus, s), "dstqds");
}

inline void dpqds(int n, double* l, double* d, double tau, double* u_plus, double* r_plus, double* p) {
  exc_bbtsvd_detail::check(exc_bb_dpqds(n, l, d, tau, u_plus, r_plus, p), "dpqds");
}

inline void one_pair_singular_vector_twisted_factorization(int n, int omege_count, double omega, double* l, double* d,
                                                           double* x_diagonal, double* x_upper_diagonal, double* v,
                                                           double* u) {
  exc_bbtsvd_detail::check(
      exc_bb_one_pair_twisted(n, omege_count, omega, l, d, x_diagonal, x_upper_diagonal, v, u),
      "one_pair_singular_vector_twisted_factorization");
}

inline void two_side_singular_vectors_twisted_factorization(int n, int omege_count, double* omege,
                                                            double* x_diagonal, double* x_upper_diagonal, double* u,
                                                            double* v) {
  exc_bbtsvd_detail::check(exc_bb_two_side_twisted(n, omege_count, omege, x_diagonal, x_upper_diagonal, u, v),
                           "two_side_singular_vectors_twisted_factorization");
}

inline void givens_mask_cache_optimized_parallel_with_simd(size_t m, size_t n, size_t mask_times, double* x,
                                                           bool mask, bool is_left, unsigned int seed) {
  exc_bbtsvd_detail::check(exc_bb_givens_mask(m, n, mask_times, x, mask ? 1 : 0, is_left ? 1 : 0, seed),
                           "givens_mask_cache_optimized_parallel_with_simd");
}

inline void bidiagonal_bisection_twisted_svd(int k, double* alpha, double* beta, double* bidiagonal_u,
                                             double* bidiagonal_v, double* sigma, int top_k) {
  exc_bbtsvd_detail::check(exc_bb_bidiagonal_svd(k, alpha, beta, bidiagonal_u, bidiagonal_v, sigma, top_k),
                           "bidiagonal_bisection_twisted_svd");
}

inline void block_shift_mask(size_t m, size_t n, size_t mask_times, double* x, bool mask, bool is_left,
                             unsigned int seed) {
  exc_bbtsvd_detail::check(exc_bb_block_shift_mask(m, n, mask_times, x, mask ? 1 : 0, is_left ? 1 : 0, seed),
                           "block_shift_mask");
}

inline void find_best_convert(struct SquareMaskConvert* smc) {
  size_t mm = 0, nn = 0, extra = 0;
  exc_bbtsvd_detail::check(exc_bb_find_best_convert(smc->k - smc->remain, &mm, &nn, &extra), "find_best_convert");
  smc->m = mm;
  smc->n = nn;
  smc->remain += extra;
}

#endif  // EXCALIBUR_BBTSVD_HPP
