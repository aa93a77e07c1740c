// Learned bisection brackets: the data the network learns from (Golub-Kahan bidiagonals of
// matrices with prescribed spectra, their singular values by the reference's bisection), the
// network itself (a 3-layer MLP trained in FP32 and run in BF16), and the bisection that starts
// from a proposed bracket once exact Sturm counts have verified it.
//
// Bisection from a bracket keeps the reference's tree (bbtsvd.cpp:19-52: root [0, ||a||^2 +
// ||b||^2], midpoints (lo + hi) / 2, leaf when the width is <= tol or the midpoint equals an
// endpoint). A verified bracket [blo, bhi) holding the target (nc(blo) <= t < nc(bhi) for the
// t-th smallest value) decides every tree node whose midpoint lies outside it without a count: a
// midpoint at or below blo sends the target right, one at or above bhi sends it left, as the
// reference's counts do when they are monotone. Inside the bracket the counts are evaluated as in
// the reference. The leaf, and so the value, is the reference's. A bracket that fails its two
// counts is dropped and the target is bisected from the root.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>

#include "exc_common.hpp"
#include "gemm.hpp"
#include "sturm.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kLzThreads = 1024;
constexpr int kLzPerThread = 16;  // n <= 16384

__device__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int q = 0; q < nw; q++) s += red[q];
  return s;
}

// Golub-Kahan (Lanczos) bidiagonalization of diag(sig) from the start vector v1 = start/||start||:
// alpha_j u_j = D v_j - beta_{j-1} u_{j-1}, beta_j v_{j+1} = D u_j - alpha_j v_j. The upper
// bidiagonal (alpha, beta) is that of A = Q1 diag(sig) Q2^T run from Q2 v1. One workgroup per
// bidiagonal; each thread holds its elements of u, v and sig in registers.
__global__ __launch_bounds__(kLzThreads) void k_gk_lanczos(const int* __restrict__ ns, const long* __restrict__ off,
                                                           const double* __restrict__ sig,
                                                           const double* __restrict__ start, double* __restrict__ d,
                                                           double* __restrict__ e) {
  __shared__ double red[16];
  const int b = blockIdx.x;
  const int n = ns[b];
  const long o = off[b];
  double s[kLzPerThread], u[kLzPerThread], v[kLzPerThread];
  double ss = 0.0;
#pragma unroll
  for (int q = 0; q < kLzPerThread; q++) {
    const int i = threadIdx.x + q * kLzThreads;
    s[q] = i < n ? sig[o + i] : 0.0;
    v[q] = i < n ? start[o + i] : 0.0;
    u[q] = 0.0;
    ss += v[q] * v[q];
  }
  double nrm = sqrt(block_sum(ss, red));
#pragma unroll
  for (int q = 0; q < kLzPerThread; q++) v[q] /= nrm;
  double beta = 0.0;
  for (int j = 0; j < n; j++) {
    ss = 0.0;
#pragma unroll
    for (int q = 0; q < kLzPerThread; q++) {
      u[q] = s[q] * v[q] - beta * u[q];
      ss += u[q] * u[q];
    }
    const double alpha = fmax(sqrt(block_sum(ss, red)), 1e-300);
#pragma unroll
    for (int q = 0; q < kLzPerThread; q++) u[q] /= alpha;
    if (threadIdx.x == 0) d[o + j] = alpha;
    if (j + 1 == n) break;
    ss = 0.0;
#pragma unroll
    for (int q = 0; q < kLzPerThread; q++) {
      v[q] = s[q] * u[q] - alpha * v[q];
      ss += v[q] * v[q];
    }
    beta = fmax(sqrt(block_sum(ss, red)), 1e-300);
#pragma unroll
    for (int q = 0; q < kLzPerThread; q++) v[q] /= beta;
    if (threadIdx.x == 0) e[o + j] = beta;
  }
}

// Per bidiagonal: the reference's bisection root ||a||^2 + ||b||^2 (sequential sums of squares,
// sqrt, then squared, as bbtsvd.cpp:936 with the reference BLAS nrm2) and the scale
// s = sqrt(sum a^2 + sum b^2). One thread per bidiagonal.
__global__ void k_bidiag_scale(int nb, const int* __restrict__ ns, const long* __restrict__ off,
                               const double* __restrict__ d, const double* __restrict__ e, double* __restrict__ root,
                               double* __restrict__ scale) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int n = ns[b];
  const long o = off[b];
  double sa = 0.0, sb = 0.0;
  for (int i = 0; i < n; i++) sa = sa + d[o + i] * d[o + i];
  for (int i = 0; i + 1 < n; i++) sb = sb + e[o + i] * e[o + i];
  const double na = sqrt(sa), nbv = sqrt(sb);
  root[b] = na * na + nbv * nbv;
  scale[b] = sqrt(sa + sb);
}

// Counting-function features: frac[b * K + k] = nc(scale_b * 10^(-k / 2)) / n_b.
__global__ __launch_bounds__(256) void k_count_features(int nb, int K, const int* __restrict__ ns,
                                                        const long* __restrict__ off, const double* __restrict__ d,
                                                        const double* __restrict__ e,
                                                        const double* __restrict__ scale, float* __restrict__ frac) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long)nb * K) return;
  const int b = (int)(g / K), k = (int)(g - (long)b * K);
  const int n = ns[b];
  const double mu = scale[b] * pow(10.0, -0.5 * k);
  frac[g] = (float)sturm_count(n, d + off[b], e + off[b], mu) / (float)n;
}

// One target per lane: the leaf of the reference's tree that holds the i-th largest singular
// value of bidiagonal tb[j] (t = n - 1 - i from below). A proposed bracket [blo, bhi) is checked
// by one exact count per end that lies strictly inside the root interval: nc(blo) <= t makes blo
// a lower bound and nc(bhi) > t makes bhi an upper bound; an end that fails bounds the value from
// the other side (nc(blo) > t: sigma < blo; nc(bhi) <= t: sigma >= bhi), so a missed bracket
// still narrows one side. Midpoints at or below a lower bound send the target right and those at
// or above an upper bound send it left without a count, as the reference's counts do when they
// are monotone; the others are counted as in the reference. The root's own counts are evaluated
// only for the sides no bound covers. Outputs the leaf midpoint, the leaf's endpoints, the number
// of Sturm counts evaluated and a status (1: bounded on both sides, 2: on one side, 0: from the
// root, -1: the root does not hold the value, as when the reference's root bound is below
// sigma_max). Fallback points flo / fhi (optional) are counted only for a side the bracket left
// open: flo for a missing lower bound, fhi for a missing upper one.
__global__ __launch_bounds__(64) void k_bisect_index(int ntg, const int* __restrict__ ns, const long* __restrict__ off,
                                                     const double* __restrict__ d, const double* __restrict__ e,
                                                     const double* __restrict__ root, double tol,
                                                     const int* __restrict__ tb, const int* __restrict__ ti,
                                                     const double* __restrict__ blo, const double* __restrict__ bhi,
                                                     const double* __restrict__ flo, const double* __restrict__ fhi,
                                                     double* __restrict__ sigma, double* __restrict__ leaf_lo,
                                                     double* __restrict__ leaf_hi, int* __restrict__ iters,
                                                     int* __restrict__ status) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ntg) return;
  const int b = tb[j];
  const int n = ns[b];
  const double* a = d + off[b];
  const double* bb = e + off[b];
  const int t = n - 1 - ti[j];
  int it = 0;
  double lo = 0.0, hi = root[b];
  bool haveL = false, haveH = false;
  double L = 0.0, H = 0.0;
  if (blo && bhi && blo[j] < bhi[j]) {
    const double pl = blo[j], ph = bhi[j];
    if (pl > lo && pl < hi) {
      const bool below = sturm_count(n, a, bb, pl) <= t;
      it++;
      if (below) { L = pl; haveL = true; }
      else { H = pl; haveH = true; }
    }
    if (ph > lo && ph < hi) {
      const bool below = sturm_count(n, a, bb, ph) <= t;
      it++;
      if (!below) {
        if (!haveH) { H = ph; haveH = true; }
      } else if (!haveL || ph > L) {
        L = ph;
        haveL = true;
      }
    }
    if (haveL && haveH && !(L < H)) haveL = haveH = false;  // counts that are not monotone: use neither
  }
  if (flo && !haveL) {
    const double p = flo[j];
    if (p > lo && p < hi && (!haveH || p < H)) {
      it++;
      if (sturm_count(n, a, bb, p) <= t) { L = p; haveL = true; }
    }
  }
  if (fhi && !haveH) {
    const double p = fhi[j];
    if (p > lo && p < hi && (!haveL || p > L)) {
      it++;
      if (t < sturm_count(n, a, bb, p)) { H = p; haveH = true; }
    }
  }
  if (!haveL || !haveH) {
    bool in_tree = true;
    if (!haveL) {
      in_tree = sturm_count(n, a, bb, lo) <= t;
      it++;
    }
    if (!haveH && in_tree) {
      in_tree = t < sturm_count(n, a, bb, hi);
      it++;
    }
    if (!in_tree) {
      sigma[j] = 0.0;
      leaf_lo[j] = leaf_hi[j] = 0.0;
      iters[j] = it;
      status[j] = -1;
      return;
    }
  }
  for (int level = 0; level < 4096; level++) {
    const double mid = (lo + hi) / 2;
    if ((hi - lo) <= tol || mid == lo || mid == hi) break;
    bool right;
    if (haveL && mid <= L) right = true;
    else if (haveH && mid >= H) right = false;
    else {
      right = sturm_count(n, a, bb, mid) <= t;
      it++;
    }
    if (right) lo = mid;
    else hi = mid;
  }
  sigma[j] = (lo + hi) / 2;
  leaf_lo[j] = lo;
  leaf_hi[j] = hi;
  iters[j] = it;
  status[j] = haveL && haveH ? 1 : (haveL || haveH ? 2 : 0);
}

// ------------------------------------------------------------------ the network
// in (32) -> 256 -> 256 -> 2, ReLU; outputs (c, w): y ~ c with Laplace scale exp(w).
constexpr int kIn = 32, kHid = 256, kOut = 2;
constexpr long kW1 = 0, kB1 = kW1 + (long)kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + (long)kHid * kHid,
               kW3 = kB2 + kHid, kB3 = kW3 + (long)kHid * kOut, kParams = kB3 + kOut;

__global__ __launch_bounds__(256) void k_bias_relu(long B, int N, float* __restrict__ Z, const float* __restrict__ bias,
                                                   float* __restrict__ H) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B * N) return;
  const float z = Z[g] + bias[g % N];
  Z[g] = z;
  H[g] = z > 0.f ? z : 0.f;
}

__global__ __launch_bounds__(256) void k_bias(long B, int N, float* __restrict__ Z, const float* __restrict__ bias) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B * N) return;
  Z[g] = Z[g] + bias[g % N];
}

// Laplace negative log-likelihood |y - c| exp(-w) + w, averaged over B; dO its gradient.
__global__ __launch_bounds__(256) void k_head_loss(long B, const float* __restrict__ O, const float* __restrict__ y,
                                                   float* __restrict__ dO, float* __restrict__ loss) {
  __shared__ float red[4];
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (g < B) {
    const float c = O[2 * g], w = O[2 * g + 1];
    const float r = y[g] - c;
    const float ew = expf(-w);
    l = fabsf(r) * ew + w;
    dO[2 * g] = (r > 0.f ? -ew : (r < 0.f ? ew : 0.f)) / (float)B;
    dO[2 * g + 1] = (1.f - fabsf(r) * ew) / (float)B;
  }
  l = wave_sum(l);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) / (float)B);
}

__global__ __launch_bounds__(256) void k_relu_back(long n, float* __restrict__ dH, const float* __restrict__ Z) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n && !(Z[g] > 0.f)) dH[g] = 0.f;
}

// out[j] = sum over the B rows of X[:, j] (fixed order: one thread per column).
__global__ __launch_bounds__(256) void k_colsum(long B, int N, const float* __restrict__ X, float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (long r = 0; r < B; r++) s += X[r * N + j];
  out[j] = s;
}

__global__ __launch_bounds__(256) void k_adam(long n, float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_to_bf16(long n, const float* __restrict__ x, __hip_bfloat16* __restrict__ y) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __float2bfloat16(x[i]);
}

// BF16 inference: weights and activations held in bfloat16, products accumulated in FP32. One
// sample per thread; the first hidden layer passes through LDS transposed (feature-major), so the
// 64 lanes of a wavefront read consecutive words, and the weights are read uniformly by all lanes.
__global__ __launch_bounds__(128) void k_mlp_infer_bf16(long B, const float* __restrict__ X,
                                                        const __hip_bfloat16* __restrict__ P, float* __restrict__ O) {
  __shared__ __hip_bfloat16 h[kHid][128];
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = s < B;
  const int me = threadIdx.x;
  float x[kIn];
#pragma unroll
  for (int k = 0; k < kIn; k++) x[k] = live ? __bfloat162float(__float2bfloat16(X[s * kIn + k])) : 0.f;
  // layer 1
  for (int j0 = 0; j0 < kHid; j0 += 32) {
    float acc[32];
#pragma unroll
    for (int q = 0; q < 32; q++) acc[q] = __bfloat162float(P[kB1 + j0 + q]);
#pragma unroll
    for (int k = 0; k < kIn; k++)
#pragma unroll
      for (int q = 0; q < 32; q++) acc[q] += x[k] * __bfloat162float(P[kW1 + (long)k * kHid + j0 + q]);
#pragma unroll
    for (int q = 0; q < 32; q++) h[j0 + q][me] = __float2bfloat16(acc[q] > 0.f ? acc[q] : 0.f);
  }
  __syncthreads();
  // layer 2 by chunks of 32 outputs, each chunk folded into the output layer at once
  float o0 = __bfloat162float(P[kB3]), o1 = __bfloat162float(P[kB3 + 1]);
  for (int j0 = 0; j0 < kHid; j0 += 32) {
    float acc[32];
#pragma unroll
    for (int q = 0; q < 32; q++) acc[q] = __bfloat162float(P[kB2 + j0 + q]);
    for (int k = 0; k < kHid; k++) {
      const float hk = __bfloat162float(h[k][me]);
#pragma unroll
      for (int q = 0; q < 32; q++) acc[q] += hk * __bfloat162float(P[kW2 + (long)k * kHid + j0 + q]);
    }
#pragma unroll
    for (int q = 0; q < 32; q++) {
      const float hq = __bfloat162float(__float2bfloat16(acc[q] > 0.f ? acc[q] : 0.f));
      o0 += hq * __bfloat162float(P[kW3 + (long)(j0 + q) * kOut]);
      o1 += hq * __bfloat162float(P[kW3 + (long)(j0 + q) * kOut + 1]);
    }
  }
  if (live) {
    O[2 * s] = o0;
    O[2 * s + 1] = o1;
  }
}

unsigned grid_for(long n, int per = 256) { return (unsigned)std::max<long>(1, (n + per - 1) / per); }

}  // namespace

extern "C" {

int exc_gk_bidiagonals(int nb, const int* d_n, const long* d_off, const double* d_sigma, const double* d_start,
                       double* d_d, double* d_e, exc_stream_t stream) {
  EXC_REQUIRE(nb >= 0 && (nb == 0 || (d_n && d_off && d_sigma && d_start && d_d && d_e)), EXC_ERR_ARG,
              "exc_gk_bidiagonals: bad arguments");
  if (nb == 0) return EXC_OK;
  hipLaunchKernelGGL(k_gk_lanczos, dim3(nb), dim3(kLzThreads), 0, exc_detail::stream(stream), d_n, d_off, d_sigma,
                     d_start, d_d, d_e);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_gk_max_n(void) { return kLzThreads * kLzPerThread; }

int exc_bidiag_scale(int nb, const int* d_n, const long* d_off, const double* d_d, const double* d_e, double* d_root,
                     double* d_scale, exc_stream_t stream) {
  EXC_REQUIRE(nb >= 0 && (nb == 0 || (d_n && d_off && d_d && d_e && d_root && d_scale)), EXC_ERR_ARG,
              "exc_bidiag_scale: bad arguments");
  if (nb == 0) return EXC_OK;
  hipLaunchKernelGGL(k_bidiag_scale, dim3(grid_for(nb, 64)), dim3(64), 0, exc_detail::stream(stream), nb, d_n, d_off,
                     d_d, d_e, d_root, d_scale);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_count_features(int nb, int K, const int* d_n, const long* d_off, const double* d_d, const double* d_e,
                       const double* d_scale, float* d_frac, exc_stream_t stream) {
  EXC_REQUIRE(nb >= 0 && K >= 1 && (nb == 0 || (d_n && d_off && d_d && d_e && d_scale && d_frac)), EXC_ERR_ARG,
              "exc_count_features: bad arguments");
  if (nb == 0) return EXC_OK;
  hipLaunchKernelGGL(k_count_features, dim3(grid_for((long)nb * K)), dim3(256), 0, exc_detail::stream(stream), nb, K,
                     d_n, d_off, d_d, d_e, d_scale, d_frac);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_bisect_index(int ntg, const int* d_n, const long* d_off, const double* d_d, const double* d_e,
                     const double* d_root, double tol, const int* d_tb, const int* d_ti, const double* d_blo,
                     const double* d_bhi, const double* d_flo, const double* d_fhi, double* d_sigma,
                     double* d_leaf_lo, double* d_leaf_hi, int* d_iters, int* d_status, exc_stream_t stream) {
  EXC_REQUIRE(ntg >= 0 && (ntg == 0 || (d_n && d_off && d_d && d_e && d_root && d_tb && d_ti && d_sigma &&
                                        d_leaf_lo && d_leaf_hi && d_iters && d_status)) &&
                  (!d_blo) == (!d_bhi) && (!d_flo) == (!d_fhi),
              EXC_ERR_ARG, "exc_bisect_index: bad arguments");
  if (ntg == 0) return EXC_OK;
  hipLaunchKernelGGL(k_bisect_index, dim3(grid_for(ntg, 64)), dim3(64), 0, exc_detail::stream(stream), ntg, d_n,
                     d_off, d_d, d_e, d_root, tol, d_tb, d_ti, d_blo, d_bhi, d_flo, d_fhi, d_sigma, d_leaf_lo,
                     d_leaf_hi, d_iters, d_status);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

long exc_mlp_param_count(void) { return kParams; }
int exc_mlp_dims(int* d_in, int* hidden, int* d_out) {
  if (d_in) *d_in = kIn;
  if (hidden) *hidden = kHid;
  if (d_out) *d_out = kOut;
  return EXC_OK;
}

// Activations of a batch: Z1, H1, Z2, H2 (B x 256 each) and O (B x 2), floats.
size_t exc_mlp_acts_floats(long B) { return (size_t)B * (4 * kHid + kOut); }

int exc_mlp_forward(long B, const float* d_x, const float* d_p, float* d_acts, float* d_out, void* d_work,
                    size_t work_bytes, exc_stream_t stream) {
  EXC_REQUIRE(B >= 1 && d_x && d_p && d_acts && d_out, EXC_ERR_ARG, "exc_mlp_forward: bad arguments");
  hipStream_t s = exc_detail::stream(stream);
  float* Z1 = d_acts;
  float* H1 = Z1 + B * kHid;
  float* Z2 = H1 + B * kHid;
  float* H2 = Z2 + B * kHid;
  ExcGemmOpts o;
  o.work = d_work;
  o.work_bytes = work_bytes;
  int rc = exc_sgemm_opts(0, 0, B, kHid, kIn, 1.0, d_x, kIn, d_p + kW1, kHid, 0.0, Z1, kHid, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bias_relu, dim3(grid_for(B * kHid)), dim3(256), 0, s, B, kHid, Z1, d_p + kB1, H1);
  rc = exc_sgemm_opts(0, 0, B, kHid, kHid, 1.0, H1, kHid, d_p + kW2, kHid, 0.0, Z2, kHid, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bias_relu, dim3(grid_for(B * kHid)), dim3(256), 0, s, B, kHid, Z2, d_p + kB2, H2);
  rc = exc_sgemm_opts(0, 0, B, kOut, kHid, 1.0, H2, kHid, d_p + kW3, kOut, 0.0, d_out, kOut, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bias, dim3(grid_for(B * kOut)), dim3(256), 0, s, B, kOut, d_out, d_p + kB3);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

// Gradient of the mean Laplace negative log-likelihood after exc_mlp_forward on the same batch;
// d_dscratch holds B x (256 + 2) floats. d_loss (one float) receives the loss.
size_t exc_mlp_backward_scratch_floats(long B) { return (size_t)B * (kHid + kOut); }

int exc_mlp_backward(long B, const float* d_x, const float* d_y, const float* d_p, float* d_acts, const float* d_out,
                     float* d_grad, float* d_scratch, float* d_loss, void* d_work, size_t work_bytes,
                     exc_stream_t stream) {
  EXC_REQUIRE(B >= 1 && d_x && d_y && d_p && d_acts && d_out && d_grad && d_scratch && d_loss, EXC_ERR_ARG,
              "exc_mlp_backward: bad arguments");
  hipStream_t s = exc_detail::stream(stream);
  float* Z1 = d_acts;
  float* H1 = Z1 + B * kHid;
  float* Z2 = H1 + B * kHid;
  float* H2 = Z2 + B * kHid;
  float* dO = d_scratch;
  float* dH = dO + B * kOut;
  ExcGemmOpts o;
  o.work = d_work;
  o.work_bytes = work_bytes;
  EXC_CHECK_HIP(hipMemsetAsync(d_loss, 0, sizeof(float), s));
  hipLaunchKernelGGL(k_head_loss, dim3(grid_for(B)), dim3(256), 0, s, B, d_out, d_y, dO, d_loss);
  // layer 3: dW3 = H2^T dO, db3 = colsum dO, dH2 = dO W3^T
  int rc = exc_sgemm_opts(1, 0, kHid, kOut, B, 1.0, H2, kHid, dO, kOut, 0.0, d_grad + kW3, kOut, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, s, B, kOut, dO, d_grad + kB3);
  rc = exc_sgemm_opts(0, 1, B, kHid, kOut, 1.0, dO, kOut, d_p + kW3, kOut, 0.0, dH, kHid, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_relu_back, dim3(grid_for(B * kHid)), dim3(256), 0, s, B * kHid, dH, Z2);
  // layer 2: dW2 = H1^T dZ2, db2, dH1 = dZ2 W2^T (into H1's slot: H1 is not needed afterwards)
  rc = exc_sgemm_opts(1, 0, kHid, kHid, B, 1.0, H1, kHid, dH, kHid, 0.0, d_grad + kW2, kHid, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, s, B, kHid, dH, d_grad + kB2);
  rc = exc_sgemm_opts(0, 1, B, kHid, kHid, 1.0, dH, kHid, d_p + kW2, kHid, 0.0, H1, kHid, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_relu_back, dim3(grid_for(B * kHid)), dim3(256), 0, s, B * kHid, H1, Z1);
  // layer 1: dW1 = X^T dZ1, db1
  rc = exc_sgemm_opts(1, 0, kIn, kHid, B, 1.0, d_x, kIn, H1, kHid, 0.0, d_grad + kW1, kHid, o, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(1), dim3(256), 0, s, B, kHid, H1, d_grad + kB1);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_mlp_adam(float* d_p, const float* d_grad, float* d_m, float* d_v, float lr, float beta1, float beta2,
                 float eps, int step, exc_stream_t stream) {
  EXC_REQUIRE(d_p && d_grad && d_m && d_v && step >= 1, EXC_ERR_ARG, "exc_mlp_adam: bad arguments");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_for(kParams)), dim3(256), 0, exc_detail::stream(stream), kParams, d_p, d_grad,
                     d_m, d_v, lr, beta1, beta2, eps, c1, c2);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_mlp_to_bf16(const float* d_p, void* d_pbf16, exc_stream_t stream) {
  EXC_REQUIRE(d_p && d_pbf16, EXC_ERR_ARG, "exc_mlp_to_bf16: bad arguments");
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(kParams)), dim3(256), 0, exc_detail::stream(stream), kParams, d_p,
                     (__hip_bfloat16*)d_pbf16);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_mlp_infer_bf16(long B, const float* d_x, const void* d_pbf16, float* d_out, exc_stream_t stream) {
  EXC_REQUIRE(B >= 1 && d_x && d_pbf16 && d_out, EXC_ERR_ARG, "exc_mlp_infer_bf16: bad arguments");
  hipLaunchKernelGGL(k_mlp_infer_bf16, dim3(grid_for(B, 128)), dim3(128), 0, exc_detail::stream(stream), B, d_x,
                     (const __hip_bfloat16*)d_pbf16, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // extern "C"
