// Reductions behind the reference's evaluation report (client.cpp:1215-1309, utils.cpp:15-66):
// mean absolute error, the off-diagonal statistics of an orthogonality test, the RMSE of absolute
// singular vector entries and the upper-triangle sums of the projection distances. Each is a
// two-phase reduction in a fixed order (per-workgroup partials, then one workgroup), so the
// results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "exc_common.hpp"

namespace {

constexpr int kEvalBlocks = 1024;

enum { OP_ABS = 0, OP_ABS_UPPER = 1, OP_SQ_ABSDIFF = 2, OP_OFFDIAG = 3 };

// Partial sums (and maxima) over rows x cols of x (and y): per op, value of element (r, c):
//   OP_ABS: |x|; OP_ABS_UPPER: |x| where c >= row0 + r; OP_SQ_ABSDIFF: (|x| - |y|)^2;
//   OP_OFFDIAG: |x| where r != c (the max of the same set is kept too).
__global__ __launch_bounds__(256) void k_eval_partials(int op, long rows, long cols, const double* __restrict__ x,
                                                       long ldx, const double* __restrict__ y, long ldy, long row0,
                                                       double* __restrict__ part) {
  __shared__ double rs[4], rm[4];
  double s = 0.0, mx = 0.0;
  const long total = rows * cols;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long r = e / cols, c = e - r * cols;
    const double v = x[r * ldx + c];
    double t = 0.0;
    if (op == OP_ABS) t = fabs(v);
    else if (op == OP_ABS_UPPER) t = c >= row0 + r ? fabs(v) : 0.0;
    else if (op == OP_SQ_ABSDIFF) {
      const double dd = fabs(v) - fabs(y[r * ldy + c]);
      t = dd * dd;
    } else t = r != c ? fabs(v) : 0.0;
    s += t;
    mx = fmax(mx, t);
  }
  s = wave_sum(s);
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) {
    rs[threadIdx.x >> 6] = s;
    rm[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rs[0] + rs[1] + rs[2] + rs[3];
    part[2 * blockIdx.x + 1] = fmax(fmax(rm[0], rm[1]), fmax(rm[2], rm[3]));
  }
}

__global__ void k_eval_final(int nparts, const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0, mx = 0.0;
    for (int q = 0; q < nparts; q++) {
      s += part[2 * q];
      mx = fmax(mx, part[2 * q + 1]);
    }
    out[0] = s;
    out[1] = mx;
  }
}

}  // namespace

extern "C" {

size_t exc_eval_workspace(void) { return (size_t)2 * kEvalBlocks * sizeof(double); }

// d_out[0] = the sum, d_out[1] = the max of the op's values (ops as k_eval_partials).
int exc_eval_reduce(int op, long rows, long cols, const double* d_x, long ldx, const double* d_y, long ldy, long row0,
                    double* d_out, void* d_work, size_t work_bytes, exc_stream_t stream) {
  EXC_REQUIRE(op >= 0 && op <= 3 && rows >= 0 && cols >= 0 && d_out && d_work && (rows * cols == 0 || d_x) &&
                  (op != OP_SQ_ABSDIFF || d_y),
              EXC_ERR_ARG, "exc_eval_reduce: bad arguments");
  EXC_REQUIRE(work_bytes >= exc_eval_workspace(), EXC_ERR_SPACE, "exc_eval_reduce: workspace too small");
  hipStream_t s = exc_detail::stream(stream);
  const long total = rows * cols;
  const int blocks = (int)std::max<long>(1, std::min<long>(kEvalBlocks, (total + 255) / 256));
  hipLaunchKernelGGL(k_eval_partials, dim3(blocks), dim3(256), 0, s, op, rows, cols, d_x, ldx, d_y, ldy, row0,
                     (double*)d_work);
  hipLaunchKernelGGL(k_eval_final, dim3(1), dim3(64), 0, s, blocks, (const double*)d_work, d_out);
  EXC_CHECK_HIP(hipGetLastError());
  return EXC_OK;
}

}  // extern "C"
