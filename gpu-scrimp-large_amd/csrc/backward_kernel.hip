// Backward pass of the differentiable join: the gradient of a loss through
// the matrix profile P_i = corr(window i, window I_i) with respect to the
// series, reusing the forward's index I (no join is recomputed; I is piecewise
// constant in the input and carries no gradient).
//
// With x = T[i:i+m], y = T[I_i:I_i+m], x~ / y~ centred, x^ = x~/|x~|,
// y^ = y~/|y~| and r = <x^, y^>:
//     dr/dx = (y^ - r x^) / |x~|,     dr/dy = (x^ - r y^) / |y~|
// (the centring drops out: both vectors sum to zero).  Each profile entry
// adds g_i * dr/dx into T[i .. i+m) and g_i * dr/dy into T[I_i .. I_i+m).
// This path has no counterpart in the reference (SCAMP has no gradients).
#include "device_common.h"
#include "kernels.h"

namespace scamp_amd {
namespace {

constexpr int kBwdBlock = 256;

// One block per (series b, profile entry i).  Means and inverse norms come
// from the statistics kernels (one row of `stride` entries per series); r is
// recomputed in FP64 from the centred windows.
__global__ __launch_bounds__(kBwdBlock) void pearson_backward_kernel(
    const double *__restrict__ T, int64_t n, int64_t m, int64_t ns,
    const double *__restrict__ mean, const double *__restrict__ inorm,
    int64_t stat_stride, const int32_t *__restrict__ index,
    const double *__restrict__ grad_p, double *__restrict__ grad_t) {
  __shared__ double red[kBwdBlock / kWave];
  const int64_t b = blockIdx.y;
  const int64_t i = blockIdx.x;
  const int64_t j = index[b * ns + i];
  const double g = grad_p[b * ns + i];
  if (j < 0 || j >= ns || g == 0.0) return;
  const double *Tb = T + b * n;
  const double mi = mean[b * stat_stride + i], mj = mean[b * stat_stride + j];
  const double ni = inorm[b * stat_stride + i], nj = inorm[b * stat_stride + j];
  if (!(ni == ni) || !(nj == nj)) return;  // NaN / flat window: no gradient
  // r = <x^, y^>
  double part = 0.0;
  for (int64_t k = threadIdx.x; k < m; k += kBwdBlock)
    part = fma((Tb[i + k] - mi) * ni, (Tb[j + k] - mj) * nj, part);
  part = wave_sum(part);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = part;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int w = 0; w < kBwdBlock / kWave; ++w) r += red[w];
  double *Gb = grad_t + b * n;
  for (int64_t k = threadIdx.x; k < m; k += kBwdBlock) {
    const double xh = (Tb[i + k] - mi) * ni;
    const double yh = (Tb[j + k] - mj) * nj;
    atomicAdd(Gb + i + k, g * (yh - r * xh) * ni);
    atomicAdd(Gb + j + k, g * (xh - r * yh) * nj);
  }
}

}  // namespace

hipError_t launch_pearson_backward(const double *T, int64_t n, int64_t m,
                                   int64_t batch, const double *mean,
                                   const double *inorm, int64_t stat_stride,
                                   const int32_t *index, const double *grad_p,
                                   double *grad_t, hipStream_t s) {
  const int64_t ns = n - m + 1;
  if (ns <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(pearson_backward_kernel,
                     dim3(static_cast<unsigned>(ns), static_cast<unsigned>(batch)),
                     dim3(kBwdBlock), 0, s, T, n, m, ns, mean, inorm, stat_stride,
                     index, grad_p, grad_t);
  return hipGetLastError();
}

}  // namespace scamp_amd
