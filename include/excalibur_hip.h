/*
 * excalibur_hip.h -- C ABI of libexcalibur_amd.so, the gfx950 (MI355X) implementation of
 * Excalibur's per-peer matrix pipeline (reference: include/bbtsvd.hpp, src/bbtsvd.cpp,
 * src/client.cpp).
 *
 * Conventions
 *   - Every matrix is row-major double (or float where the name says f32) unless stated.
 *   - Pointers named d_* are device pointers; the caller owns every buffer. Functions that need
 *     scratch take a caller-provided device workspace and have a *_workspace() size query.
 *   - Work is enqueued on `stream` (NULL = the default stream) and returns without waiting,
 *     except where a function says it synchronizes.
 *   - Return value: EXC_OK (0) or a negative EXC_ERR_* code; exc_last_error() describes it.
 *   - Functions are thread-compatible: one stream per calling thread.
 * The C++ drop-ins for the reference's bbtsvd.hpp (host pointers, same signatures) live in
 * include/bbtsvd.hpp and are built on these entry points.
 */
#ifndef EXCALIBUR_HIP_H
#define EXCALIBUR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* exc_stream_t;

enum {
  EXC_OK = 0,
  EXC_ERR_ARG = -1,     /* invalid argument or shape */
  EXC_ERR_HIP = -2,     /* HIP runtime error */
  EXC_ERR_SPACE = -3,   /* workspace too small */
  EXC_ERR_NUMERIC = -4, /* numerical failure (e.g. non-positive pivot) */
  EXC_ERR_COUNT = -5    /* bisection found fewer singular values than requested */
};

/* ---------------------------------------------------------------- plumbing */
const char* exc_last_error(void);
int exc_version(void);
int exc_device_count(int* n);
int exc_set_device(int dev);
int exc_malloc(void** d_ptr, size_t bytes);
int exc_free(void* d_ptr);
int exc_host_alloc(void** h_ptr, size_t bytes); /* pinned host memory */
int exc_host_free(void* h_ptr);
int exc_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, exc_stream_t stream);
int exc_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, exc_stream_t stream);
int exc_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, exc_stream_t stream);
int exc_memset(void* d_dst, int value, size_t bytes, exc_stream_t stream);
int exc_stream_create(exc_stream_t* stream);
int exc_stream_destroy(exc_stream_t stream);
int exc_stream_sync(exc_stream_t stream);

/* ---------------------------------------------------------------- bidiagonal SVD (FP64) */

/* Sturm counts: d_count[j] = #{singular values of the upper bidiagonal B(diag alpha, super
 * beta) below d_mu[j]}, by the recurrence of negative_count, bbtsvd.cpp:7-17, evaluated with
 * separately rounded operations (bit-exact counts). One lane per shift. */
int exc_sturm_count(int n, const double* d_alpha, const double* d_beta, const double* d_mu, int nmu,
                    int* d_count, exc_stream_t stream);

/* Bisection of bisection_singular_values (bbtsvd.cpp:19-52) from the root [lo, hi]: the same
 * tree nodes, so the same leaves and the same sigma bits, found per singular value by
 * multisection (lanes of a wavefront evaluate a subtree of midpoints per round). Writes the
 * leaves' midpoints in descending order to d_sigma (capacity top_k) and their number to
 * *d_count (device int). d_leaf_lo/d_leaf_hi (nullable, capacity top_k) receive each leaf's
 * interval. Synchronizes the stream; returns EXC_ERR_COUNT never (the count is reported). */
size_t exc_bisection_workspace(int n, int top_k);
int exc_bisection(int n, const double* d_alpha, const double* d_beta, double lo, double hi, double tol,
                  int top_k, double* d_sigma, int* d_count, double* d_leaf_lo, double* d_leaf_hi, void* d_work,
                  size_t work_bytes, exc_stream_t stream);

/* ldl_decomposition (bbtsvd.cpp:54-63): B^T B = L D L^T. d_l has n-1 entries, d_d n. */
int exc_ldl(int n, const double* d_alpha, const double* d_beta, double* d_l, double* d_d, exc_stream_t stream);

/* two_side_singular_vectors_twisted_factorization (bbtsvd.cpp:138-147) for k shifts d_omega:
 * d_u (n x k, column i = u_i), d_v (k x n, row i = v_i). d_twists (nullable, 2k ints) receives
 * the twist indices (v, u) of each pair. */
size_t exc_twisted_workspace(int n, int k);
int exc_twisted_vectors(int n, int k, const double* d_omega, const double* d_alpha, const double* d_beta,
                        double* d_u, double* d_v, int* d_twists, void* d_work, size_t work_bytes,
                        exc_stream_t stream);

/* 7-argument bidiagonal_bisection_twisted_svd (bbtsvd.cpp:932-968): bisection from
 * [0, ||alpha||^2 + ||beta||^2], twisted vectors, the sign rule for v and the "check None" rule.
 * d_u: k x top_k, d_v: top_k x k, d_sigma: top_k. Synchronizes the stream; returns
 * EXC_ERR_COUNT when fewer than top_k singular values were found (the reference asserts). */
size_t exc_bidiag_svd_workspace(int k, int top_k);
int exc_bidiag_svd(int k, const double* d_alpha, const double* d_beta, double* d_u, double* d_v, double* d_sigma,
                   int top_k, int* d_twists, void* d_work, size_t work_bytes, exc_stream_t stream);

/* ---------------------------------------------------------------- Givens masking */

/* Host-side plan of givens_mask_cache_optimized_parallel_with_simd (bbtsvd.cpp:727-779): the
 * index pairs (passes x half x 2, int32) and rotation values (passes x half x 2: c, s), drawn
 * with the C library's srand/rand and the C++ library's std::shuffle over
 * std::default_random_engine(seed) exactly as the reference draws them (mask = 0 gives the
 * reversed plan that removes the mask). half = ceil(K/2). Uses the process-wide rand() state.*/
long exc_givens_half(long K);
int exc_givens_plan(long K, long passes, int mask, unsigned seed, int32_t* h_pairs, double* h_cs);

/* Applies a plan to d_x (m x n row-major) in place: rows are rotated when is_left (K = m),
 * columns otherwise (K = n). x1' = c x1 + s' x2, x2' = -s' x1 + c x2 with s' = (mask ? s : -s),
 * separately rounded (bit-exact with bbtsvd.cpp:796-809). All passes run in one launch with the
 * strip of lines each workgroup owns held in LDS. */
int exc_givens_apply(double* d_x, long m, long n, const int32_t* d_pairs, const double* d_cs, long passes,
                     int mask, int is_left, exc_stream_t stream);
int exc_givens_apply_f32(float* d_x, long m, long n, const int32_t* d_pairs, const double* d_cs, long passes,
                         int mask, int is_left, exc_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* EXCALIBUR_HIP_H */
