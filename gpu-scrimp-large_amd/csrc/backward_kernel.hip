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
constexpr int kBwdWaves = kBwdBlock / kWave;
constexpr int kBwdChunk = 512;  // entries staged through LDS at a time

// Pass 1, one wavefront per (series b, entry i): r = <x^, y^> from the
// centred windows (means and inverse norms from the statistics kernels, one
// row of `stat_stride` entries per series); the partner half of the gradient,
// g (x^ - r y^) / |y~|, into T[j .. j+m) with FP64 atomics; and the self
// half's coefficients for pass 2, P_i = g/(|x~||y~|) and Q_i = g r/|x~|^2
// (both 0 for an entry without a gradient).
__global__ __launch_bounds__(kBwdBlock) void bwd_partner_kernel(
    const double *__restrict__ T, int64_t n, int64_t m, int64_t ns,
    const double *__restrict__ mean, const double *__restrict__ inorm,
    int64_t stat_stride, const int32_t *__restrict__ index,
    const double *__restrict__ grad_p, double *__restrict__ grad_t,
    double *__restrict__ coef) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * kBwdWaves + threadIdx.x / kWave;
  if (i >= ns) return;
  double *P = coef + 2 * (b * ns + i);
  const int64_t j = index[b * ns + i];
  const double g = grad_p[b * ns + i];
  const double ni = inorm[b * stat_stride + i];
  const double nj = j >= 0 && j < ns ? inorm[b * stat_stride + j] : nan("");
  if (j < 0 || j >= ns || g == 0.0 || !(ni == ni) || !(nj == nj)) {
    if (lane == 0) P[0] = P[1] = 0.0;  // NaN / flat window: no gradient
    return;
  }
  const double *Tb = T + b * n;
  const double mi = mean[b * stat_stride + i], mj = mean[b * stat_stride + j];
  double part = 0.0;
  for (int64_t k = lane; k < m; k += kWave)
    part = fma((Tb[i + k] - mi) * ni, (Tb[j + k] - mj) * nj, part);
  const double r = wave_sum(part);
  double *Gb = grad_t + b * n;
  for (int64_t k = lane; k < m; k += kWave) {
    const double xh = (Tb[i + k] - mi) * ni;
    const double yh = (Tb[j + k] - mj) * nj;
    atomicAdd(Gb + j + k, g * (xh - r * yh) * nj);
  }
  if (lane == 0) {
    P[0] = g * ni * nj;
    P[1] = g * r * ni * ni;
  }
}

// Pass 2, one thread per sample t: the self half, a gather over the windows
// that contain t (entries i in [t-m+1, t]):
//   grad[t] += sum_i P_i (T[j_i + t - i] - mu_j) - Q_i (T[t] - mu_i)
// which is g (y^ - r x^) / |x~| at position t - i of window i.  Each sample is
// written by one thread, after pass 1 on the same stream: no atomics.  The
// block's entries (coefficients, partners, both means) are staged in LDS.
__global__ __launch_bounds__(kBwdBlock) void bwd_self_kernel(
    const double *__restrict__ T, int64_t n, int64_t m, int64_t ns,
    const double *__restrict__ mean, int64_t stat_stride,
    const int32_t *__restrict__ index, const double *__restrict__ coef,
    double *__restrict__ grad_t) {
  __shared__ double sP[kBwdChunk], sQ[kBwdChunk], sMi[kBwdChunk], sMj[kBwdChunk];
  __shared__ int32_t sJ[kBwdChunk];
  const int64_t b = blockIdx.y;
  const int64_t t0 = (int64_t)blockIdx.x * kBwdBlock;
  const int64_t t = t0 + threadIdx.x;
  const double *Tb = T + b * n;
  const double tt = t < n ? Tb[t] : 0.0;
  double acc = 0.0;
  // entries whose windows reach the block's samples
  const int64_t lo = t0 - m + 1 > 0 ? t0 - m + 1 : 0;
  const int64_t hi = t0 + kBwdBlock - 1 < ns - 1 ? t0 + kBwdBlock - 1 : ns - 1;
  for (int64_t c0 = lo; c0 <= hi; c0 += kBwdChunk) {
    const int64_t cn = hi - c0 + 1 < kBwdChunk ? hi - c0 + 1 : kBwdChunk;
    __syncthreads();
    for (int q = threadIdx.x; q < cn; q += kBwdBlock) {
      const int64_t e = c0 + q;
      const double p = coef[2 * (b * ns + e)], qq = coef[2 * (b * ns + e) + 1];
      const int32_t j = index[b * ns + e];
      const bool live = p != 0.0 || qq != 0.0;
      sP[q] = p;
      sQ[q] = qq;
      sJ[q] = live ? j : -1;
      sMi[q] = mean[b * stat_stride + e];
      sMj[q] = live ? mean[b * stat_stride + j] : 0.0;
    }
    __syncthreads();
    if (t < n) {
      // entries i in [t-m+1, t] within this chunk
      int64_t a = t - m + 1 - c0, z = t - c0;
      a = a > 0 ? a : 0;
      z = z < cn - 1 ? z : cn - 1;
      for (int64_t q = a; q <= z; ++q) {
        const int32_t j = sJ[q];
        if (j < 0) continue;
        const int64_t i = c0 + q;
        acc += sP[q] * (Tb[j + t - i] - sMj[q]) - sQ[q] * (tt - sMi[q]);
      }
    }
  }
  if (t < n) grad_t[b * n + t] += acc;
}

}  // namespace

hipError_t launch_pearson_backward(const double *T, int64_t n, int64_t m,
                                   int64_t batch, const double *mean,
                                   const double *inorm, int64_t stat_stride,
                                   const int32_t *index, const double *grad_p,
                                   double *grad_t, double *work, hipStream_t s) {
  const int64_t ns = n - m + 1;
  if (ns <= 0 || batch <= 0) return hipSuccess;
  const unsigned nb = static_cast<unsigned>(batch);
  hipLaunchKernelGGL(bwd_partner_kernel,
                     dim3(static_cast<unsigned>((ns + kBwdWaves - 1) / kBwdWaves), nb),
                     dim3(kBwdBlock), 0, s, T, n, m, ns, mean, inorm, stat_stride,
                     index, grad_p, grad_t, work);
  hipLaunchKernelGGL(bwd_self_kernel,
                     dim3(static_cast<unsigned>((n + kBwdBlock - 1) / kBwdBlock), nb),
                     dim3(kBwdBlock), 0, s, T, n, m, ns, mean, stat_stride, index,
                     work, grad_t);
  return hipGetLastError();
}

}  // namespace scamp_amd
