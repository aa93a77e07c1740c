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
/* Live timing of the MFMA GEMM and Givens launches with HIP events on their own stream:
 * exc_prof_enable(1) starts recording; exc_prof_report writes a JSON object
 * {"kernel": {"launches", "ms", "flops", "bytes"}} (algorithmic flops and bytes) and clears. */
int exc_prof_enable(int on);
int exc_prof_report(char* buf, size_t cap);

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

/* exc_bidiag_svd without a workspace argument and without any allocation: for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. The scratch is a fixed device array of
 * exc_bidiag_svd_noalloc_capacity() bytes (48 MiB) per device, held until the process ends; its
 * device address is looked up on the first call on each device. Sizes whose workspace
 * (exc_bidiag_svd_workspace(k, top_k)) exceeds it are rejected with EXC_ERR_SPACE (k = top_k =
 * 1024 fits). Every call on one device shares that device's array: calls that overlap on one
 * device (two streams or two threads at once) race and must be serialized by the caller. Results
 * are those of exc_bidiag_svd bit for bit. Synchronizes the stream. */
size_t exc_bidiag_svd_noalloc_capacity(void);
int exc_bidiag_svd_noalloc(int k, const double* d_alpha, const double* d_beta, double* d_u, double* d_v,
                           double* d_sigma, int top_k, int* d_twists, exc_stream_t stream);

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

/* block_shift_mask (bbtsvd.cpp:533-669) on d_x (m x n row-major, in place), K = m (is_left) or n
 * lines rotated. K <= 10000: the plain sweep of exc_givens_apply with the plan of (K, seed). K >
 * 10000: the virtual-matrix mask, each line viewed as a vm x vn matrix plus a remainder
 * (find_best_convert), with the reference's plans of seed + 1000 .. seed + 5000, bit-exact with
 * the reference's loop. Draws the plans with the C library's generator (srand per plan), like the
 * reference. Synchronizes the stream. */
size_t exc_block_shift_mask_workspace(long m, long n, long passes, int is_left);
int exc_block_shift_mask(double* d_x, long m, long n, long passes, int mask, int is_left, unsigned seed, void* d_work,
                         size_t work_bytes, exc_stream_t stream);

/* ---------------------------------------------------------------- dense kernels (MFMA) */

/* C = alpha op(A) op(B) + beta C, row-major, op = transpose when trans_* != 0. FP64 on
 * v_mfma_f64_16x16x4_f64 (the reference's cblas_dgemm calls, client.cpp:1018, 1179, 1190);
 * FP32 on v_mfma_f32_16x16x4_f32 for the mixed-precision form. d_work (nullable) enables
 * split-K for short-and-deep products; exc_gemm_workspace() bounds what it may use. */
size_t exc_gemm_workspace(long M, long N, long K);
int exc_dgemm(int trans_a, int trans_b, long M, long N, long K, double alpha, const double* d_a, long lda,
              const double* d_b, long ldb, double beta, double* d_c, long ldc, void* d_work, size_t work_bytes,
              exc_stream_t stream);
int exc_sgemm(int trans_a, int trans_b, long M, long N, long K, double alpha, const float* d_a, long lda,
              const float* d_b, long ldb, double beta, float* d_c, long ldc, void* d_work, size_t work_bytes,
              exc_stream_t stream);

/* Dense factorizations (FP64, row-major, in place). The workspace (exc_fact_workspace) holds
 * split-K partials of the GEMM updates and the per-tile flags of the one-launch tile forms.
 * exc_potrf_lower and exc_lu_sign synchronize the stream for 64 < n <= 4096 (the tile form
 * checks that its bounded waits completed). */
size_t exc_fact_workspace(int n, long wide);
/* Lower Cholesky of a symmetric n x n matrix (lower triangle read, strict upper zeroed);
 * *d_info = 0 or the 1-based index of the first non-positive pivot. */
int exc_potrf_lower(int n, double* d_a, long lda, int* d_info, void* d_work, size_t work_bytes, exc_stream_t stream);
/* X <- op(T)^-1 X (left; T lower, or upper when upper != 0; unit diagonal when unit != 0). */
int exc_trsm_left(int upper, int unit, int n, long nrhs, const double* d_l, long ldl, double* d_x, long ldx,
                  void* d_work, size_t work_bytes, exc_stream_t stream);
/* X <- X U^-1 (right; U upper), or X <- X L^-T when lower_t != 0. */
int exc_trsm_right(int lower_t, int unit, int n, long rows, const double* d_u, long ldu, double* d_x, long ldx,
                   void* d_work, size_t work_bytes, exc_stream_t stream);
/* LU without pivoting of A - S, S_jj = -sign(current pivot candidate) chosen column by column
 * (the sign rule of Householder reconstruction); d_sgn[j] = S_jj. square_last: the last column
 * takes +sign instead (LAPACK does not reflect the last column of a square matrix). */
int exc_lu_sign(int n, double* d_a, long lda, double* d_sgn, int square_last, void* d_work, size_t work_bytes,
                exc_stream_t stream);

/* ---------------------------------------------------------------- Householder QR (FP64) */

/* QR of X^T (X: m x r row-major, r >= m), i.e. of the column-major r x m matrix the reference
 * passes to LAPACKE_dgeqrf (client.cpp:424, 715), with LAPACK's Householder convention. Outputs:
 * d_rt (m x m, lower) = R^T; d_qt (nullable, m x r, may alias d_x) = Q^T (dorgqr, client.cpp:430,
 * 785); d_vt (nullable, m x r) the reflectors one per row (unit entry at column j, zeros before);
 * d_tau (nullable, m); d_sgn (nullable, m) the reconstruction signs. square_last: treat the last
 * column of a square X^T as LAPACK does (tau = 0). */
size_t exc_qr_t_workspace(int m, long r, int want_q);
int exc_qr_t(int m, long r, const double* d_x, long ldx, double* d_rt, double* d_qt, long ldq, double* d_vt,
             long ldv, double* d_tau, double* d_sgn, int square_last, void* d_work, size_t work_bytes,
             exc_stream_t stream);
/* Mt (rows x L) <- Mt Q with Q = H_0...H_{k-1} from reflector rows d_vt (k x L) and d_tau: the
 * transposed form of LAPACKE_dormqr('L', 'T') at client.cpp:777. Compact WY on the MFMA GEMM,
 * T^-1 = striu(V^T V) + diag(1/tau). */
size_t exc_wy_workspace(int k, long L, long rows);
int exc_apply_q_right(long rows, long L, int k, double* d_mt, long ldm, const double* d_vt, long ldv,
                      const double* d_tau, void* d_work, size_t work_bytes, exc_stream_t stream);
/* Q^T (k x L) = first k rows of (H_0...H_{k-1})^T: LAPACKE_dorgqr at client.cpp:785. */
int exc_orgqr_t(int k, long L, const double* d_vt, long ldv, const double* d_tau, double* d_qt, long ldq,
                void* d_work, size_t work_bytes, exc_stream_t stream);
/* C (m x n) <- H_0...H_{m-1} C with reflector j in column j of the row-major m x m d_house
 * (implied unit diagonal): LAPACKE_dormqr(RowMajor, 'L', 'N') at client.cpp:1176. */
size_t exc_ormqr_house_workspace(int m, long n);
int exc_ormqr_house_ln(int m, long n, const double* d_house, const double* d_tau, double* d_c, long ldc,
                       void* d_work, size_t work_bytes, exc_stream_t stream);

/* ---------------------------------------------------------------- Ralha bidiagonalization */

/* One-sided (Ralha) bidiagonalization of this peer's share F (m x c) of the rows f_j
 * (client.cpp:800-910). F and U are stored column-major here: d_fc[col * m + j] = F[j][col]
 * (a row-major c x m array), likewise d_uc.
 *
 * exc_ralha_step(i) does the column work of step i (0 <= i <= m-3; client.cpp:858-891): from
 * the all-reduced message d_g of step i, dlarfg (LAPACK convention) into column i+1 of d_house
 * (m x m row-major) and d_tau[i+1], F[i+1:] -= tau h (h^T F[i+1:]), f_i -= beta_{i-1} u_{i-1},
 * alpha_i and u_i. It then writes this peer's message for step j = i+1 when j <= m-3
 * (client.cpp:842-855): m+3-j values, f_{j+1+r}.f_j for r < m-1-j, then f_j.u_{j-1},
 * u_{j-1}.u_{j-1}, f_j.f_j, u_{j-1}.f_j when j > 0, zeros after. i = -1 only writes the message
 * of step 0. d_msg may be pinned host memory (exc_host_alloc); it is complete once the stream
 * has reached that point. The workspace is exc_ralha_workspace(m, c) bytes. */
size_t exc_ralha_workspace(int m, int c);
int exc_ralha_step(int i, int m, int c, double* d_fc, double* d_uc, const double* d_g, double* d_house,
                   double* d_tau, double* d_alpha, double* d_beta, double* d_msg, void* d_work, size_t work_bytes,
                   exc_stream_t stream);
/* The same step with the columns streamed from HBM instead of held in registers: exc_ralha_step
 * uses it for m > 4000 and keeps every column in registers below that. Exported so that both
 * forms can be checked against each other at any m. */
int exc_ralha_step_generic(int i, int m, int c, double* d_fc, double* d_uc, const double* d_g, double* d_house,
                           double* d_tau, double* d_alpha, double* d_beta, double* d_msg, void* d_work,
                           size_t work_bytes, exc_stream_t stream);
/* Steps -1 .. m-3 of the same loop (client.cpp:836-910) in ONE launch that runs while the host
 * all-reduces the messages: every workgroup keeps its columns in registers for the whole loop.
 * The messages pass through a link buffer in coherent pinned host memory
 * (exc_ralha_link_alloc). The host drives it like this:
 *
 *   exc_ralha_loop_start(...);                      // synchronizes `stream`, then launches
 *   for (j = 0; j <= m - 3; j++) {
 *     exc_ralha_loop_wait(link, j, timeout, &msg);  // message j: m+3-j doubles, as exc_ralha_step
 *     all-reduce msg in place over the peers;
 *     exc_ralha_loop_post(link, j);
 *   }
 *   exc_ralha_loop_finish(link, stream);            // synchronizes; error if the kernel gave up
 *
 * On any error between start and finish, call exc_ralha_loop_abort(link) and then
 * exc_ralha_loop_finish: every wait of the kernel also ends on the abort flag or after
 * timeout_s, so the grid always drains. The results (house, tau, alpha, beta, U, F) are those of
 * exc_ralha_step for i = -1 .. m-3, with the partial sums of each message added in another fixed
 * order. Applies when exc_ralha_loop_fits(m, c, cols_per_wave) returns 1: m <= 3000, and twice
 * the grid (one workgroup per 4 x cols_per_wave columns, at most 256) fits on the device at
 * once, so two peers' loops can share it. cols_per_wave = 0 picks 1 or 2; other values than
 * 0, 1 and 2 do not fit (the loop declines them). No counterpart in
 * the reference, which runs each step as BLAS calls between blocking all-reduces. */
size_t exc_ralha_loop_workspace(int m, int c);
int exc_ralha_link_alloc(int m, void** h_link);
int exc_ralha_link_free(void* h_link);
int exc_ralha_loop_fits(int m, int c, int cols_per_wave);
int exc_ralha_loop_start(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau, double* d_alpha,
                         double* d_beta, void* h_link, void* d_work, size_t work_bytes, int cols_per_wave,
                         double timeout_s, exc_stream_t stream);
int exc_ralha_loop_wait(void* h_link, int j, double timeout_s, double** h_msg);
int exc_ralha_loop_post(void* h_link, int j);
int exc_ralha_loop_abort(void* h_link);
int exc_ralha_loop_finish(void* h_link, exc_stream_t stream);
/* Number of workgroups of this peer's loop grid (0 when exc_ralha_loop_fits is 0). */
int exc_ralha_loop_grid(int m, int c, int cols_per_wave);
/* Host-only shape of the loop grid (no HIP call): workgroups and LDS bytes per workgroup, both 0
 * when the loop does not apply to (m, c, cols_per_wave). */
int exc_ralha_loop_shape(int m, int c, int cols_per_wave, int* nwg, size_t* lds_bytes);
/* Workgroups of this loop shape that the current device holds at once (occupancy x CUs), 0 when
 * the loop does not apply. */
int exc_ralha_loop_slots(int m, int c, int cols_per_wave);
/* Whether P peers may all run the one-launch loop: each grid nwg[q] fits its device's slots[q],
 * and when the peers share one device (same_device = 1) the sum of all grids fits the fewest
 * slots, since every peer's loop waits on every other's messages and all must be resident
 * together. Host-only, no HIP call. Returns 1 or 0. */
int exc_ralha_loops_colocate(int P, const int* nwg, const int* slots, int same_device);
/* The same loop for peers that share one device: the messages go through per-peer boxes in
 * device memory (exc_ralha_box_bytes(m) each; peer q's box opened in the other peers' processes
 * with exc_ipc_open), and the host does nothing per step. Every peer q's box counter must have
 * been cleared (exc_ralha_box_clear, then a barrier among the peers) before any peer starts;
 * nwg[q] = exc_ralha_loop_grid of peer q. The reduced message is each peer's own plus the others'
 * in peer order, as the host all-reduce adds them, so the results equal the host-link loop's.
 * The host waits with exc_ralha_loop_running (1 while the launch runs) and ends with
 * exc_ralha_loop_finish; exc_ralha_loop_abort ends it early (e.g. when a peer's connection
 * drops). At most 16 peers. */
size_t exc_ralha_box_bytes(int m);
int exc_ralha_box_clear(void* d_box, exc_stream_t stream);
int exc_ralha_loop_start_boxes(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau,
                               double* d_alpha, double* d_beta, void* h_link, void* d_work, size_t work_bytes,
                               int cols_per_wave, double timeout_s, int npeers, int me, void* const* d_boxes,
                               const int* nwg, exc_stream_t stream);
int exc_ralha_loop_running(exc_stream_t stream);
/* Device memory shared between the processes of peers on one device (hipIpcGetMemHandle and
 * hipIpcOpenMemHandle): a handle is EXC_IPC_HANDLE_BYTES opaque bytes. */
#define EXC_IPC_HANDLE_BYTES 64
int exc_ipc_handle(void* d_ptr, void* handle);
int exc_ipc_open(const void* handle, void** d_ptr);
int exc_ipc_close(void* d_ptr);
/* The last two columns and the first (client.cpp:830-833, 897-905): d_out = f_i.u_{i-1};
 * tail_update: f_i -= beta u_{i-1} (beta = reduced value, stored in d_beta[i-1]), d_out =
 * ||f_i||^2 (i = -1: ||f_0||^2 only); tail_scale: alpha_i = sqrt(reduced), u_i = f_i / alpha_i. */
int exc_ralha_dot(int i, int m, int c, const double* d_fc, const double* d_uc, double* d_out, exc_stream_t stream);
int exc_ralha_tail_update(int i, int m, int c, double* d_fc, const double* d_uc, const double* d_breduced,
                          double* d_beta, double* d_out, exc_stream_t stream);
int exc_ralha_tail_scale(int i, int m, int c, const double* d_fc, double* d_uc, const double* d_n2, double* d_alpha,
                         exc_stream_t stream);
/* Lower to upper bidiagonal by drotg (client.cpp:947-954); d_cs = (c_i, s_i) pairs, 2(m-1). */
int exc_lower_to_upper(int m, double* d_alpha, double* d_beta, double* d_cs, exc_stream_t stream);
/* client.cpp:987-995: columns of d_bu (m x k) made to have a positive largest entry, rows of
 * d_bv (k x m) flipped with them; d_flip (k) receives the factors. */
int exc_sign_normalize(int m, int k, double* d_bu, double* d_bv, double* d_flip, exc_stream_t stream);
/* client.cpp:1127-1128: rows (i, i+1) of d_bu (m x k) rotated by (c_i, -s_i), i = m-2..0. */
int exc_rot_back(int m, int k, double* d_bu, const double* d_cs, exc_stream_t stream);
/* d_out2 = {||alpha||^2 + ||beta||^2, 0}: the bisection root bound of client.cpp:966 and the
 * two shifts of its singular-value count. */
int exc_bisection_bound(int n, const double* d_alpha, const double* d_beta, double* d_out2, exc_stream_t stream);

/* One-sided bidiagonalization of an m x c matrix F that one caller holds whole (no peers): the
 * steps -1 .. m-1 of exc_ralha_step with every all-reduce the identity (client.cpp:829-910 for
 * one peer), then lower to upper (client.cpp:947-954) into d_alpha, d_beta, d_cs. The loop kernel
 * runs in one launch when `colocated` copies of its grid (the callers that may run this loop on
 * the device at the same time, >= 1) fit on the device together and h_link (exc_ralha_link_alloc)
 * is given, else one launch per step back to back on the stream; both give the same bits. Layouts
 * as exc_ralha_step. The mixed-precision form uses it for the m x m factor every peer shares. */
size_t exc_bidiag_local_workspace(int m, int c);
int exc_bidiag_local(int m, int c, double* d_fc, double* d_uc, double* d_house, double* d_tau, double* d_alpha,
                     double* d_beta, double* d_cs, void* h_link, int colocated, void* d_work, size_t work_bytes,
                     double timeout_s, exc_stream_t stream);

/* ---------------------------------------------------------------- mixed precision (FP32 bulk) */

/* Precision conversion of a rows x cols row-major block (leading dimensions in elements). */
int exc_convert_f64_f32(long rows, long cols, const double* d_src, long lds, float* d_dst, long ldd,
                        exc_stream_t stream);
int exc_convert_f32_f64(long rows, long cols, const float* d_src, long lds, double* d_dst, long ldd,
                        exc_stream_t stream);
/* d_out[0] = trace(A); A += value * I; d_out[0] = max |A - I| and d_out[1] = ||A - I||_F
 * (n x n, two doubles out). */
/* The lower triangle of an n x n row-major matrix packed by rows (n (n + 1) / 2 values: row i
 * holds columns 0 .. i), and back: unpack writes the triangle and its mirror image. */
int exc_pack_lower_f32(int n, const float* d_a, long lda, float* d_packed, exc_stream_t stream);
int exc_pack_lower_f64(int n, const double* d_a, long lda, double* d_packed, exc_stream_t stream);
int exc_unpack_lower_f64(int n, const double* d_packed, double* d_a, long lda, exc_stream_t stream);
/* One peer's term of the mixed form's Gram sum, added on the device in peer order:
 * d_acc[q] = (first ? 0 : d_acc[q]) + d_part[q] for q < n, d_part FP32 (part_is_f32) or FP64,
 * the sum in FP64. Called for peers 0 .. P-1 in turn it gives the bits of the host loop
 * acc = 0; acc = acc + x_j, on every peer alike. No counterpart in the reference, which sums
 * its exchanged buffers on the host (client.cpp:313-327). */
int exc_acc_packed(long n, const void* d_part, int part_is_f32, int first, double* d_acc, exc_stream_t stream);
int exc_trace(int n, const double* d_a, long lda, double* d_out, exc_stream_t stream);
int exc_add_diag(int n, double* d_a, long lda, double value, exc_stream_t stream);
size_t exc_dev_identity_workspace(int n);
int exc_dev_identity(int n, const double* d_a, long lda, double* d_out, void* d_work, size_t work_bytes,
                     exc_stream_t stream);
/* FP64 Ogita-Aishima refinement of a square SVD L ~ U diag(sigma) W^T (all m x m row-major,
 * L lower triangular when l_lower): up to `iters` steps of U <- U (I + F), W <- W (I + G) from the
 * residual products I - U^T U, I - W^T W and U^T L W on the MFMA GEMM; sigma receives
 * T_ii / (1 - (R_ii + S_ii) / 2) of the last step. When the starting vectors are within 1e-10 of
 * orthonormal (max-entry norm of both deviations), the first step is the last: its error is about
 * the square of the starting one. Deciding that reads two numbers back (a stream synchronization)
 * when iters > 1. Values within relative distance tau form
 * clusters that are orthogonalized symmetrically, and the left vectors of cluster members above
 * floor_rel * max sigma are recomputed as L w / sigma. Ogita and Aishima, J. Comput. Appl. Math.
 * 369 (2020). No counterpart in the reference (its FP64 path has no refinement). */
size_t exc_oa_refine_workspace(int m);
int exc_oa_refine(int m, const double* d_l, long ldl, int l_lower, double* d_u, long ldu, double* d_w, long ldw,
                  double* d_sigma, int iters, double tau, double floor_rel, void* d_work, size_t work_bytes,
                  exc_stream_t stream);

/* ---------------------------------------------------------------- batched peers */

/* The same sweep as exc_givens_apply over nb matrices of one shape (m x n), each with its own
 * plan (device arrays of nb pointers), in one launch. */
int exc_givens_apply_batched(int nb, double* const* d_xs, long m, long n, const int32_t* const* d_pairs,
                             const double* const* d_cs, long passes, int mask, int is_left, exc_stream_t stream);
/* Local QR of nb blocks at once (client.cpp:572-589 for many peers on one device): for every
 * block X_b (m x cols[b] row-major FP64, cols[b] >= m; host array of device pointers h_x) R_b^T
 * (m x m lower, positive diagonal) into h_rt[b] and, when h_qt is given, Q_b^T (m x cols[b]) into
 * h_qt[b]. A shifted CholeskyQR whose first pass runs in FP32 on the MFMA (all blocks' Gram
 * matrices in one grouped persistent launch, its K chunk and grid chosen by the library's
 * autotuner), refined by FP64 passes (grouped launches); the m x m factorizations run one
 * workgroup per block in one launch. h_stats (nullable, 3 nb doubles): per block the FP32 shift,
 * max |Z Z^T - I| before the last FP64 pass and the number of FP64 passes. Synchronizes. No
 * counterpart in the reference, which runs one peer per process. */
size_t exc_batched_local_qr_workspace(int nb, int m, const long* h_cols);
int exc_batched_local_qr(int nb, int m, const long* h_cols, const double* const* h_x, double* const* h_rt,
                         double* const* h_qt, double* h_stats, void* d_work, size_t work_bytes, exc_stream_t stream);
/* The grouped GEMM autotuner's choices so far (JSON list of {dtype, problems, sum_k, kchunk, grid,
 * ms}). */
int exc_gemm_grouped_tuning(char* buf, size_t cap);

/* ---------------------------------------------------------------- learned bisection brackets */

/* nb bidiagonals stored back to back: bidiagonal b has n = d_n[b] values at offset d_off[b] of
 * d_sigma / d_start / d_d / d_e (d_e holds n - 1 values there). exc_gk_bidiagonals runs the
 * Golub-Kahan (Lanczos) bidiagonalization of diag(sigma) from the start vector (n <=
 * exc_gk_max_n()): the bidiagonal of A = Q1 diag(sigma) Q2^T for a random orthogonal pair. */
int exc_gk_max_n(void);
int exc_gk_bidiagonals(int nb, const int* d_n, const long* d_off, const double* d_sigma, const double* d_start,
                       double* d_d, double* d_e, exc_stream_t stream);
/* d_root[b] = ||d||^2 + ||e||^2 as the reference's bisection root (bbtsvd.cpp:936), d_scale[b] =
 * sqrt(sum d^2 + sum e^2). */
int exc_bidiag_scale(int nb, const int* d_n, const long* d_off, const double* d_d, const double* d_e, double* d_root,
                     double* d_scale, exc_stream_t stream);
/* Network features: d_frac[b * K + k] = nc(scale_b * 10^(-k/2)) / n_b (exact Sturm counts). */
int exc_count_features(int nb, int K, const int* d_n, const long* d_off, const double* d_d, const double* d_e,
                       const double* d_scale, float* d_frac, exc_stream_t stream);
/* One lane per target j: the leaf of the reference's bisection tree (bbtsvd.cpp:19-52, root
 * [0, d_root[b]], tolerance tol) holding the d_ti[j]-th largest singular value of bidiagonal
 * d_tb[j]. With brackets (d_blo and d_bhi both given) each end inside the root interval is checked
 * by one exact Sturm count and becomes a lower or an upper bound of the value (an end the value
 * does not lie above bounds it from above); the tree nodes whose midpoints lie beyond a bound are
 * decided without a count. Fallback points (d_flo and d_fhi both given, or both NULL) are counted
 * only for a side the bracket left open. d_iters[j] = Sturm counts evaluated; d_status[j] = 1
 * bounded on both sides, 2 on one side, 0 from the root, -1 the root does not hold the value.
 * Targets whose values share a leaf (closer than its width) get the same leaf; one value per
 * distinct leaf, in order, is the reference's list. */
int exc_bisect_index(int ntg, const int* d_n, const long* d_off, const double* d_d, const double* d_e,
                     const double* d_root, double tol, const int* d_tb, const int* d_ti, const double* d_blo,
                     const double* d_bhi, const double* d_flo, const double* d_fhi, double* d_sigma,
                     double* d_leaf_lo, double* d_leaf_hi, int* d_iters, int* d_status, exc_stream_t stream);
/* The bracket network: in 32 -> 256 -> 256 -> 2, ReLU, parameters in one FP32 array of
 * exc_mlp_param_count() floats (W1 32x256, b1, W2 256x256, b2, W3 256x2, b3; row-major, inputs
 * by rows). Outputs (c, w): the target is c with Laplace scale exp(w). Training in FP32 on the
 * MFMA GEMM: exc_mlp_forward keeps the activations (exc_mlp_acts_floats(B)), exc_mlp_backward
 * writes the gradient of the mean Laplace negative log-likelihood and the loss, exc_mlp_adam
 * updates. Inference in BF16: exc_mlp_to_bf16 converts the parameters (2 bytes each),
 * exc_mlp_infer_bf16 runs a batch. No counterpart in the reference. */
long exc_mlp_param_count(void);
int exc_mlp_dims(int* d_in, int* hidden, int* d_out);
size_t exc_mlp_acts_floats(long B);
size_t exc_mlp_backward_scratch_floats(long B);
int exc_mlp_forward(long B, const float* d_x, const float* d_p, float* d_acts, float* d_out, void* d_work,
                    size_t work_bytes, exc_stream_t stream);
int exc_mlp_backward(long B, const float* d_x, const float* d_y, const float* d_p, float* d_acts, const float* d_out,
                     float* d_grad, float* d_scratch, float* d_loss, void* d_work, size_t work_bytes,
                     exc_stream_t stream);
int exc_mlp_adam(float* d_p, const float* d_grad, float* d_m, float* d_v, float lr, float beta1, float beta2,
                 float eps, int step, exc_stream_t stream);
int exc_mlp_to_bf16(const float* d_p, void* d_pbf16, exc_stream_t stream);
int exc_mlp_infer_bf16(long B, const float* d_x, const void* d_pbf16, float* d_out, exc_stream_t stream);

/* ---------------------------------------------------------------- data movement */
int exc_copy2d(long rows, long cols, const double* d_src, long lds, double* d_dst, long ldd, exc_stream_t stream);
int exc_transpose(long rows, long cols, const double* d_src, long lds, double* d_dst, long ldd, exc_stream_t stream);
int exc_fill2d(long rows, long cols, double value, double* d_dst, long ldd, exc_stream_t stream);
/* Global-QR message of a peer panel (client.cpp:716-724): for each of its c reflectors
 * j = pos..pos+c-1, tau_j then the R-1-j entries of v_j below the diagonal (rows of d_vt). */
long exc_reflector_message_len(long c, long R, long pos);
int exc_pack_reflectors(long c, long R, long pos, const double* d_vt, long ldv, const double* d_tau, double* d_out,
                        exc_stream_t stream);
int exc_unpack_reflectors(long c, long R, long pos, const double* d_in, double* d_vt, long ldv, double* d_tau,
                          exc_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* EXCALIBUR_HIP_H */
