// Small data-movement kernels of the pipeline: strided copies, transposes, fills, and the packing
// of global-QR reflectors into the reference's message layout (client.cpp:716-724).
#include "exc_common.hpp"
#include "qr.hpp"

namespace {

__global__ __launch_bounds__(256) void k_fill2d(long rows, long cols, double v, double* __restrict__ dst, long ldd) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const long i = e / cols, j = e - i * cols;
  dst[i * ldd + j] = v;
}

// Message of one peer's panel: for each of its c reflectors j = pos + jj, tau_j followed by the
// R-1-j entries of v_j below its unit diagonal (row j of vt from column j+1).
__device__ inline long pack_offset(long jj, long pos, long R) {
  // sum_{q < jj} (1 + R - 1 - (pos + q)) = jj (R - pos) - jj (jj - 1) / 2
  return jj * (R - pos) - jj * (jj - 1) / 2;
}

__global__ __launch_bounds__(256) void k_pack_reflectors(long c, long R, long pos, const double* __restrict__ vt,
                                                         long ldv, const double* __restrict__ tau,
                                                         double* __restrict__ out) {
  const long jj = blockIdx.x;
  if (jj >= c) return;
  const long j = pos + jj;
  const long off = pack_offset(jj, pos, R);
  if (threadIdx.x == 0) out[off] = tau[j];
  for (long q = threadIdx.x; q < R - 1 - j; q += blockDim.x) out[off + 1 + q] = vt[j * ldv + j + 1 + q];
}

__global__ __launch_bounds__(256) void k_unpack_reflectors(long c, long R, long pos, const double* __restrict__ in,
                                                           double* __restrict__ vt, long ldv, double* __restrict__ tau) {
  const long jj = blockIdx.x;
  if (jj >= c) return;
  const long j = pos + jj;
  const long off = pack_offset(jj, pos, R);
  if (threadIdx.x == 0) tau[j] = in[off];
  for (long q = threadIdx.x; q < R; q += blockDim.x) {
    double v = 0.0;
    if (q == j) v = 1.0;
    else if (q > j) v = in[off + 1 + (q - j - 1)];
    vt[j * ldv + q] = v;
  }
}

__global__ __launch_bounds__(256) void k_bound_pair(int n, const double* __restrict__ a, const double* __restrict__ b,
                                                    double* __restrict__ out) {
  __shared__ double buf[2048];
  double sums[2] = {0.0, 0.0};
  for (int which = 0; which < 2; which++) {
    const double* x = which == 0 ? a : b;
    const int len = which == 0 ? n : n - 1;
    double s = 0.0;
    for (int base = 0; base < len; base += 2048) {
      const int cnt = min(2048, len - base);
      for (int i = threadIdx.x; i < cnt; i += blockDim.x) buf[i] = x[base + i];
      __syncthreads();
      if (threadIdx.x == 0)
        for (int i = 0; i < cnt; i++) {
          const double v = buf[i];
          s = s + v * v;
        }
      __syncthreads();
    }
    sums[which] = s;
  }
  if (threadIdx.x == 0) {
    const double na = sqrt(sums[0]), nb = sqrt(sums[1]);
    out[0] = na * na + nb * nb;
    out[1] = 0.0;
  }
}

}  // namespace

extern "C" {

int exc_copy2d(long rows, long cols, const double* d_src, long lds, double* d_dst, long ldd, exc_stream_t stream) {
  EXC_REQUIRE(rows >= 0 && cols >= 0, EXC_ERR_ARG, "exc_copy2d: bad shape");
  qr::copy2d(rows, cols, d_src, lds, d_dst, ldd, exc_detail::stream(stream));
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_transpose(long rows, long cols, const double* d_src, long lds, double* d_dst, long ldd, exc_stream_t stream) {
  EXC_REQUIRE(rows >= 0 && cols >= 0, EXC_ERR_ARG, "exc_transpose: bad shape");
  if (rows * cols == 0) return EXC_OK;
  qr::transpose(rows, cols, d_src, lds, d_dst, ldd, exc_detail::stream(stream));
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_fill2d(long rows, long cols, double value, double* d_dst, long ldd, exc_stream_t stream) {
  EXC_REQUIRE(rows >= 0 && cols >= 0, EXC_ERR_ARG, "exc_fill2d: bad shape");
  if (rows * cols == 0) return EXC_OK;
  hipLaunchKernelGGL(k_fill2d, dim3(exc_detail::ceil_div(rows * cols, 256)), dim3(256), 0, exc_detail::stream(stream),
                     rows, cols, value, d_dst, ldd);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

long exc_reflector_message_len(long c, long R, long pos) { return (2 * R - 1 - pos - (pos + c)) * c / 2 + c; }

int exc_pack_reflectors(long c, long R, long pos, const double* d_vt, long ldv, const double* d_tau, double* d_out,
                        exc_stream_t stream) {
  EXC_REQUIRE(c >= 0 && pos >= 0 && pos + c <= R, EXC_ERR_ARG, "exc_pack_reflectors: bad shape");
  if (c == 0) return EXC_OK;
  hipLaunchKernelGGL(k_pack_reflectors, dim3((unsigned)c), dim3(256), 0, exc_detail::stream(stream), c, R, pos, d_vt,
                     ldv, d_tau, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_unpack_reflectors(long c, long R, long pos, const double* d_in, double* d_vt, long ldv, double* d_tau,
                          exc_stream_t stream) {
  EXC_REQUIRE(c >= 0 && pos >= 0 && pos + c <= R, EXC_ERR_ARG, "exc_unpack_reflectors: bad shape");
  if (c == 0) return EXC_OK;
  hipLaunchKernelGGL(k_unpack_reflectors, dim3((unsigned)c), dim3(256), 0, exc_detail::stream(stream), c, R, pos,
                     d_in, d_vt, ldv, d_tau);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

int exc_bisection_bound(int n, const double* d_alpha, const double* d_beta, double* d_out2, exc_stream_t stream) {
  EXC_REQUIRE(n >= 1 && d_alpha && d_out2 && (n == 1 || d_beta), EXC_ERR_ARG, "exc_bisection_bound: bad arguments");
  hipLaunchKernelGGL(k_bound_pair, dim3(1), dim3(256), 0, exc_detail::stream(stream), n, d_alpha, d_beta, d_out2);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // extern "C"
