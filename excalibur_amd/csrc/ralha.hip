// One-sided (Ralha) bidiagonalization step kernels (client.cpp:829-910) and the small pipeline
// kernels around the bidiagonal SVD (client.cpp:947-954, 987-995, 1127-1128).
//
// F (m x c, row-major) holds this peer's share of the rows f_j; U (m x c) the u_j; the reductions
// across peers happen on the host between the two kernels of a step:
//   ralha_msg(i):   w_r = f_{i+1+r} . f_i (r < m-1-i) and four dots with u_{i-1}  -> message
//   (host: all-reduce of the message over the peers)
//   ralha_apply(i): dlarfg of the reduced w, F[i+1:] -= tau h (h^T F[i+1:]), f_i -= beta u_{i-1},
//                   alpha_i from the reduced dots, u_i = f_i / alpha_i.
// The message kernel is row-parallel (one wavefront per row, fixed-order reduction); the apply
// kernel is column-parallel (a workgroup owns a slab of columns, so the rank-1 update needs no
// cross-workgroup synchronization). Every workgroup recomputes the reflector from the reduced
// message with the same fixed reduction order, so all of them see the same h and tau.
#include <cfloat>

#include "exc_common.hpp"

namespace {

constexpr int kColsPerWg = 16;  // columns per apply workgroup
constexpr int kRowGroups = 16;  // 16 x 16 = 256 threads

// msg[r] = F[i+1+r] . F[i] for r < m-1-i; if i > 0, msg[m-1-i..m+2-i] = f_i.u_{i-1},
// u_{i-1}.u_{i-1}, f_i.f_i, u_{i-1}.f_i (client.cpp:842-855).
__global__ __launch_bounds__(256) void k_ralha_msg(int i, int m, int c, const double* __restrict__ F,
                                                   const double* __restrict__ U, double* __restrict__ msg) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int nrows = m - 1 - i;
  const int total = nrows + (i > 0 ? 4 : 0);
  if (wave >= total) return;
  const double *x, *y;
  const double* fi = F + (long)i * c;
  if (wave < nrows) {
    x = F + (long)(i + 1 + wave) * c;
    y = fi;
  } else {
    const double* ui = U + (long)(i - 1) * c;
    const int q = wave - nrows;
    x = q == 0 ? fi : (q == 1 ? ui : (q == 2 ? fi : ui));
    y = q == 0 ? ui : (q == 1 ? ui : (q == 2 ? fi : fi));
  }
  double s = 0.0;
  for (int k = lane; k < c; k += 64) s += x[k] * y[k];
  s = wave_sum(s);
  if (lane == 0) msg[wave] = s;
}

__device__ inline double dlapy2_dev(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = fmax(xa, ya), z = fmin(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

// Block-wide sum with a fixed order (same result in every workgroup for the same input).
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += red[q];
  __syncthreads();
  if (threadIdx.x == 0) red[0] = t;
  __syncthreads();
  const double out = red[0];
  __syncthreads();
  return out;
}

// The reduced message g (length m+3-i) -> reflector of LAPACK's dlarfg (n = m-1-i) and the
// update of this workgroup's columns.
__global__ __launch_bounds__(256) void k_ralha_apply(int i, int m, int c, double* __restrict__ F,
                                                     double* __restrict__ U, const double* __restrict__ g,
                                                     double* __restrict__ house, double* __restrict__ tau,
                                                     double* __restrict__ alpha, double* __restrict__ beta) {
  __shared__ double red[8];
  __shared__ double hbuf[4096];
  __shared__ double part[kRowGroups][kColsPerWg];
  __shared__ double vts[kColsPerWg];
  const int n = m - 1 - i;  // reflector length
  const int tid = threadIdx.x;
  // dlarfg: alpha_h = g[0], x = g[1..n-1]
  double ss = 0.0;
  for (int r = 1 + tid; r < n; r += blockDim.x) ss += g[r] * g[r];
  double xnorm = sqrt(block_sum(ss, red));
  double alph = g[0];
  double t = 0.0, scale = 0.0, bet = alph;
  int knt = 0;
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  if (n > 1 && xnorm != 0.0) {
    bet = -copysign(dlapy2_dev(alph, xnorm), alph);
    double xs = 1.0;  // scaling of x during the safmin loop
    if (fabs(bet) < safmin) {
      const double rs = 1.0 / safmin;
      do {
        knt++;
        xs *= rs;
        bet *= rs;
        alph *= rs;
      } while (fabs(bet) < safmin && knt < 20);
      double s2 = 0.0;
      for (int r = 1 + tid; r < n; r += blockDim.x) s2 += (g[r] * xs) * (g[r] * xs);
      xnorm = sqrt(block_sum(s2, red));
      bet = -copysign(dlapy2_dev(alph, xnorm), alph);
    }
    t = (bet - alph) / bet;
    scale = xs * (1.0 / (alph - bet));
  }
  // h[0] = 1, h[r] = x_r * scale; a zero tau leaves F untouched (H = I)
  const bool big = n <= 4096;
  for (int r = tid; r < n && big; r += blockDim.x) hbuf[r] = r == 0 ? 1.0 : g[r] * scale;
  __syncthreads();
  auto h_at = [&](int r) -> double { return big ? hbuf[r] : (r == 0 ? 1.0 : g[r] * scale); };
  if (blockIdx.x == 0) {
    for (int r = tid; r < n; r += blockDim.x) house[(long)(i + 1 + r) * m + (i + 1)] = h_at(r);
    if (tid == 0) {
      tau[i + 1] = t;
      if (i > 0) beta[i - 1] = g[m - i - 1];
    }
  }
  // vt[col] = sum_r F[i+1+r][col] h[r]; F[i+1+r][col] -= tau h[r] vt[col]
  const int cl = tid % kColsPerWg, rg = tid / kColsPerWg;
  const int col = blockIdx.x * kColsPerWg + cl;
  const bool live = col < c;
  double acc = 0.0;
  if (live && t != 0.0)
    for (int r = rg; r < n; r += kRowGroups) acc += F[(long)(i + 1 + r) * c + col] * h_at(r);
  part[rg][cl] = acc;
  __syncthreads();
  if (tid < kColsPerWg) {
    double v = 0.0;
    for (int q = 0; q < kRowGroups; q++) v += part[q][tid];
    vts[tid] = v;
  }
  __syncthreads();
  if (live && t != 0.0) {
    const double vc = vts[cl];
    for (int r = rg; r < n; r += kRowGroups) F[(long)(i + 1 + r) * c + col] += vc * (-t * h_at(r));
  }
  // f_i -= beta_{i-1} u_{i-1}; alpha_i; u_i = f_i / alpha_i (client.cpp:886-890)
  if (i > 0) {
    const double b = g[m - i - 1];
    const double a = sqrt(g[m - i] * (b * b) + g[m - i + 1] - 2 * b * g[m - i + 2]);
    if (blockIdx.x == 0 && tid == 0) alpha[i] = a;
    if (rg == 0 && live) {
      const double f = F[(long)i * c + col] + (-b) * U[(long)(i - 1) * c + col];
      F[(long)i * c + col] = f;
      U[(long)i * c + col] = (1.0 / a) * f;
    }
  }
}

// dot = f_i . u_{i-1} (one workgroup)
__global__ __launch_bounds__(256) void k_ralha_dot(int i, int c, const double* __restrict__ F,
                                                   const double* __restrict__ U, double* __restrict__ out) {
  __shared__ double red[8];
  double s = 0.0;
  for (int k = threadIdx.x; k < c; k += blockDim.x) s += F[(long)i * c + k] * U[(long)(i - 1) * c + k];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[0] = s;
}

// f_i -= beta u_{i-1} (beta from the reduced value), then out = ||f_i||^2 (i = -1: just ||f_0||^2)
__global__ __launch_bounds__(256) void k_ralha_tail_update(int i, int c, double* __restrict__ F,
                                                           const double* __restrict__ U,
                                                           const double* __restrict__ breduced,
                                                           double* __restrict__ beta, double* __restrict__ out) {
  __shared__ double red[8];
  const int row = i < 0 ? 0 : i;
  double s = 0.0;
  const double b = i > 0 ? breduced[0] : 0.0;
  for (int k = threadIdx.x; k < c; k += blockDim.x) {
    double f = F[(long)row * c + k];
    if (i > 0) {
      f = f + (-b) * U[(long)(i - 1) * c + k];
      F[(long)row * c + k] = f;
    }
    s += f * f;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) {
    out[0] = s;
    if (i > 0) beta[i - 1] = b;
  }
}

// alpha_i = sqrt(reduced ||f_i||^2); u_i = f_i / alpha_i
__global__ __launch_bounds__(256) void k_ralha_tail_scale(int i, int c, const double* __restrict__ F,
                                                          double* __restrict__ U, const double* __restrict__ n2,
                                                          double* __restrict__ alpha) {
  const double a = sqrt(n2[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) alpha[i] = a;
  const double inv = 1.0 / a;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < c; k += gridDim.x * blockDim.x)
    U[(long)i * c + k] = inv * F[(long)i * c + k];
}

// Lower to upper bidiagonal with the reference BLAS drotg (client.cpp:947-954), one thread.
__global__ void k_lower_to_upper(int m, double* __restrict__ alpha, double* __restrict__ beta,
                                 double* __restrict__ cs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int i = 0; i < m - 1; i++) {
    double a = alpha[i], b = beta[i];
    double roe = fabs(a) > fabs(b) ? a : b;
    double scl = fabs(a) + fabs(b);
    double c, s, r;
    if (scl == 0.0) {
      c = 1.0; s = 0.0; r = 0.0;
    } else {
      const double as = a / scl, bs = b / scl;
      r = scl * sqrt(as * as + bs * bs);
      r = copysign(1.0, roe) * r;
      c = a / r;
      s = b / r;
    }
    alpha[i] = r;
    cs[2 * i] = c;
    cs[2 * i + 1] = s;
    beta[i] = s * alpha[i + 1];
    alpha[i + 1] = c * alpha[i + 1];
  }
}

// Sign rule of client.cpp:987-995: flip[j] = -1 when the first largest-magnitude entry of
// column j of bu (m x k) is negative.
__global__ __launch_bounds__(256) void k_sign_flags(int m, int k, const double* __restrict__ bu,
                                                    double* __restrict__ flip) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k) return;
  int p = 0;
  double best = -1.0;
  for (int i = 0; i < m; i++) {
    const double a = fabs(bu[(long)i * k + j]);
    if (a > best) { best = a; p = i; }
  }
  flip[j] = bu[(long)p * k + j] < 0 ? -1.0 : 1.0;
}

__global__ __launch_bounds__(256) void k_apply_flags(int m, int k, double* __restrict__ bu, double* __restrict__ bv,
                                                     const double* __restrict__ flip) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)m * k) return;
  const long i = e / k, j = e - i * k;
  bu[e] *= flip[j];            // column j of bu (m x k)
  bv[j * m + i] *= flip[j];    // row j of bv (k x m)
}

// Back-rotation of client.cpp:1127-1128: for i = m-2..0, rows (i, i+1) of bu (m x k) rotated by
// drot with (c_i, -s_i). One lane per column carries row i+1 down the recurrence.
__global__ __launch_bounds__(256) void k_rot_back(int m, int k, double* __restrict__ bu,
                                                  const double* __restrict__ cs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= k || m < 2) return;
  double y = bu[(long)(m - 1) * k + j];
  for (int i = m - 2; i >= 0; i--) {
    const double c = cs[2 * i], s = -cs[2 * i + 1];
    const double x = bu[(long)i * k + j];
    const double xn = c * x + s * y;
    const double yn = c * y - s * x;
    bu[(long)(i + 1) * k + j] = yn;
    y = xn;
  }
  bu[j] = y;
}

}  // namespace

extern "C" {

int exc_ralha_msg(int i, int m, int c, const double* d_f, const double* d_u, double* d_msg, exc_stream_t stream) {
  EXC_REQUIRE(m >= 3 && i >= 0 && i <= m - 3 && c >= 1 && d_f && d_u && d_msg, EXC_ERR_ARG,
              "exc_ralha_msg: bad arguments");
  const int total = (m - 1 - i) + (i > 0 ? 4 : 0);
  hipLaunchKernelGGL(k_ralha_msg, dim3(exc_detail::ceil_div((long)total * 64, 256)), dim3(256), 0,
                     exc_detail::stream(stream), i, m, c, d_f, d_u, d_msg);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_apply(int i, int m, int c, double* d_f, double* d_u, const double* d_g, double* d_house, double* d_tau,
                    double* d_alpha, double* d_beta, exc_stream_t stream) {
  EXC_REQUIRE(m >= 3 && i >= 0 && i <= m - 3 && c >= 1 && d_f && d_u && d_g && d_house && d_tau && d_alpha && d_beta,
              EXC_ERR_ARG, "exc_ralha_apply: bad arguments");
  hipLaunchKernelGGL(k_ralha_apply, dim3(exc_detail::ceil_div(c, kColsPerWg)), dim3(256), 0,
                     exc_detail::stream(stream), i, m, c, d_f, d_u, d_g, d_house, d_tau, d_alpha, d_beta);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_dot(int i, int c, const double* d_f, const double* d_u, double* d_out, exc_stream_t stream) {
  EXC_REQUIRE(i >= 1 && c >= 1, EXC_ERR_ARG, "exc_ralha_dot: bad arguments");
  hipLaunchKernelGGL(k_ralha_dot, dim3(1), dim3(256), 0, exc_detail::stream(stream), i, c, d_f, d_u, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_tail_update(int i, int c, double* d_f, const double* d_u, const double* d_breduced, double* d_beta,
                          double* d_out, exc_stream_t stream) {
  EXC_REQUIRE(i >= -1 && c >= 1, EXC_ERR_ARG, "exc_ralha_tail_update: bad arguments");
  hipLaunchKernelGGL(k_ralha_tail_update, dim3(1), dim3(256), 0, exc_detail::stream(stream), i, c, d_f, d_u,
                     d_breduced, d_beta, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_ralha_tail_scale(int i, int c, const double* d_f, double* d_u, const double* d_n2, double* d_alpha,
                         exc_stream_t stream) {
  EXC_REQUIRE(i >= 0 && c >= 1, EXC_ERR_ARG, "exc_ralha_tail_scale: bad arguments");
  hipLaunchKernelGGL(k_ralha_tail_scale, dim3(exc_detail::ceil_div(c, 256)), dim3(256), 0,
                     exc_detail::stream(stream), i, c, d_f, d_u, d_n2, d_alpha);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_lower_to_upper(int m, double* d_alpha, double* d_beta, double* d_cs, exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && d_alpha && d_beta && d_cs, EXC_ERR_ARG, "exc_lower_to_upper: bad arguments");
  hipLaunchKernelGGL(k_lower_to_upper, dim3(1), dim3(64), 0, exc_detail::stream(stream), m, d_alpha, d_beta, d_cs);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_sign_normalize(int m, int k, double* d_bu, double* d_bv, double* d_flip, exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && k >= 0 && d_bu && d_bv && d_flip, EXC_ERR_ARG, "exc_sign_normalize: bad arguments");
  if (k == 0) return EXC_OK;
  hipStream_t s = exc_detail::stream(stream);
  hipLaunchKernelGGL(k_sign_flags, dim3(exc_detail::ceil_div(k, 256)), dim3(256), 0, s, m, k, d_bu, d_flip);
  hipLaunchKernelGGL(k_apply_flags, dim3(exc_detail::ceil_div((long)m * k, 256)), dim3(256), 0, s, m, k, d_bu, d_bv,
                     d_flip);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_rot_back(int m, int k, double* d_bu, const double* d_cs, exc_stream_t stream) {
  EXC_REQUIRE(m >= 1 && k >= 0 && d_bu && d_cs, EXC_ERR_ARG, "exc_rot_back: bad arguments");
  if (k == 0 || m < 2) return EXC_OK;
  hipLaunchKernelGGL(k_rot_back, dim3(exc_detail::ceil_div(k, 256)), dim3(256), 0, exc_detail::stream(stream), m, k,
                     d_bu, d_cs);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // extern "C"
