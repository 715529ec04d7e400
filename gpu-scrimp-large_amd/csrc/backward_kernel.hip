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
//
// Every sample's gradient is gathered by one thread in a fixed order, with no
// atomics, so the result is the same bit for bit from run to run:
//   1. one thread per profile entry: r and three coefficients;
//   2. the entries sorted by partner window (radix sort of (partner, entry):
//      per series in LDS for series up to 8192 windows, else device-wide);
//   3. per sample t, the self half over the windows i that contain t, then
//      the partner half over the entries whose partner window contains t.
#include <rocprim/block/block_radix_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "device_common.h"
#include "kernels.h"

namespace scamp_amd {
namespace {

constexpr int kBwdBlock = 256;
constexpr int kBwdChunk = 512;  // entries staged through LDS at a time

// Per entry e = b * ns + i (partner j = I_i, g = dL/dP_i):
//   P = g / (|x~| |y~|), Q = g r / |x~|^2, R = g r / |y~|^2
// The self half at position k of window i is P (T[j+k] - mu_j) - Q (T[i+k] -
// mu_i); the partner half at position k of window j is P (T[i+k] - mu_i) -
// R (T[j+k] - mu_j).  Entries without a gradient (no match, g = 0, NaN or
// flat window) get P = Q = R = 0 and sort after all others.
struct Coef {
  double P, Q, R;
};

// Pass 1, one thread per (series b, entry i).  Means and inverse norms come
// from the statistics kernels (one row of `stat_stride` entries per series).
// The windows i of a wavefront's lanes overlap, so their loads coalesce; each
// lane's partner window streams through the cache.  (A wavefront per entry
// spreads the m terms over the lanes, but launches one wavefront for every
// profile entry and is bound by wave dispatch.)
__global__ __launch_bounds__(kBwdBlock) void bwd_coef_kernel(
    const double *__restrict__ T, int64_t n, int64_t m, int64_t ns,
    int64_t batch, const double *__restrict__ mean,
    const double *__restrict__ inorm, int64_t stat_stride,
    const int32_t *__restrict__ index, const double *__restrict__ grad_p,
    Coef *__restrict__ coef, uint32_t *__restrict__ key,
    uint32_t *__restrict__ val, int shift) {
  // shift > 0: per-series packed keys (partner << shift | entry, no values)
  const int64_t b = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * kBwdBlock + threadIdx.x;
  if (i >= ns) return;
  const int64_t e = b * ns + i;
  const int64_t j = index[e];
  const double g = grad_p[e];
  const double ni = inorm[b * stat_stride + i];
  const double nj = j >= 0 && j < ns ? inorm[b * stat_stride + j] : nan("");
  if (j < 0 || j >= ns || g == 0.0 || !(ni == ni) || !(nj == nj)) {
    coef[e] = Coef{0.0, 0.0, 0.0};
    // after every real partner
    if (shift > 0) {
      key[e] = (uint32_t)((ns << shift) | i);
    } else {
      key[e] = (uint32_t)(batch * ns);
      val[e] = (uint32_t)e;
    }
    return;
  }
  const double *x = T + b * n + i;
  const double *y = T + b * n + j;
  const double mi = mean[b * stat_stride + i], mj = mean[b * stat_stride + j];
  // four independent chains keep four loads of each window in flight
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t k = 0;
  for (; k + 4 <= m; k += 4) {
    s0 = fma(x[k] - mi, y[k] - mj, s0);
    s1 = fma(x[k + 1] - mi, y[k + 1] - mj, s1);
    s2 = fma(x[k + 2] - mi, y[k + 2] - mj, s2);
    s3 = fma(x[k + 3] - mi, y[k + 3] - mj, s3);
  }
  for (; k < m; ++k) s0 = fma(x[k] - mi, y[k] - mj, s0);
  const double r = ((s0 + s1) + (s2 + s3)) * ni * nj;
  coef[e] = Coef{g * ni * nj, g * r * ni * ni, g * r * nj * nj};
  if (shift > 0) {
    key[e] = (uint32_t)((j << shift) | i);
  } else {
    key[e] = (uint32_t)(b * ns + j);
    val[e] = (uint32_t)e;
  }
}

// Pass 2 for series of at most kSeriesSortMax subsequences: one workgroup
// sorts one series' packed keys in LDS.  The keys are unique, so the order is
// (partner, entry), the same as the device-wide stable sort of (series,
// partner) with entries as values gives.
constexpr int kSeriesSortThreads = 1024;
constexpr int kSeriesSortItems = 8;
constexpr int64_t kSeriesSortMax = kSeriesSortThreads * kSeriesSortItems;

__global__ __launch_bounds__(kSeriesSortThreads) void bwd_sort_series_kernel(
    uint32_t *__restrict__ keys, int64_t ns, int end_bit) {
  using Sort = rocprim::block_radix_sort<uint32_t, kSeriesSortThreads, kSeriesSortItems>;
  __shared__ typename Sort::storage_type storage;
  uint32_t *kb = keys + (int64_t)blockIdx.x * ns;
  uint32_t k[kSeriesSortItems];
  // blocked arrangement: thread t holds positions [t * items, (t + 1) * items)
#pragma unroll
  for (int q = 0; q < kSeriesSortItems; ++q) {
    const int64_t r = (int64_t)threadIdx.x * kSeriesSortItems + q;
    k[q] = r < ns ? kb[r] : 0xFFFFFFFFu;  // padding sorts last
  }
  Sort().sort(k, storage, 0u, static_cast<unsigned>(end_bit));
#pragma unroll
  for (int q = 0; q < kSeriesSortItems; ++q) {
    const int64_t r = (int64_t)threadIdx.x * kSeriesSortItems + q;
    if (r < ns) kb[r] = k[q];
  }
}

// Pass 3a, one thread per sample t: the self half, a gather over the windows
// that contain t (entries i in [t-m+1, t]); writes grad_t.  The block's
// entries (coefficients, partners, both means) are staged in LDS.
__global__ __launch_bounds__(kBwdBlock) void bwd_self_kernel(
    const double *__restrict__ T, int64_t n, int64_t m, int64_t ns,
    const double *__restrict__ mean, int64_t stat_stride,
    const int32_t *__restrict__ index, const Coef *__restrict__ coef,
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
      const Coef cf = coef[b * ns + e];
      const int32_t j = index[b * ns + e];
      const bool live = cf.P != 0.0 || cf.Q != 0.0;
      sP[q] = cf.P;
      sQ[q] = cf.Q;
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
  if (t < n) grad_t[b * n + t] = acc;
}

// First position in keys[0, len) holding a value >= v.
__device__ int64_t lower_bound_u32(const uint32_t *keys, int64_t len, uint32_t v) {
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (keys[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First position in v[0, len) (ascending) holding a value >= x.
__device__ __forceinline__ int first_at_least(const int32_t *v, int len, int64_t x) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) / 2;
    if (v[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Pass 3b, one thread per sample t: the partner half, a gather over the
// entries whose partner window contains t (partners j in [t-m+1, t]), found
// in the entries sorted by (series, partner); adds into grad_t.
__global__ __launch_bounds__(kBwdBlock) void bwd_partner_kernel(
    const double *__restrict__ T, int64_t n, int64_t m, int64_t ns,
    const double *__restrict__ mean, int64_t stat_stride,
    const Coef *__restrict__ coef, const uint32_t *__restrict__ skey,
    const uint32_t *__restrict__ sval, int64_t nent,
    double *__restrict__ grad_t, int shift) {
  // shift > 0: skey holds each series' packed keys (partner << shift | entry)
  // sorted per series; otherwise all (series * ns + partner) keys sorted, with
  // the entries in sval
  __shared__ double sP[kBwdChunk], sR[kBwdChunk], sMi[kBwdChunk], sMj[kBwdChunk];
  __shared__ int32_t sI[kBwdChunk], sJ[kBwdChunk];
  __shared__ int64_t range[2];
  const int64_t b = blockIdx.y;
  const int64_t t0 = (int64_t)blockIdx.x * kBwdBlock;
  const int64_t t = t0 + threadIdx.x;
  const double *Tb = T + b * n;
  const double tt = t < n ? Tb[t] : 0.0;
  // partners reaching the block's samples: [t0-m+1, t0+255] within [0, ns)
  const int64_t jlo = t0 - m + 1 > 0 ? t0 - m + 1 : 0;
  const int64_t jhi = t0 + kBwdBlock - 1 < ns - 1 ? t0 + kBwdBlock - 1 : ns - 1;
  if (shift > 0) skey += b * ns;
  if (threadIdx.x == 0) {
    if (shift > 0) {
      range[0] = lower_bound_u32(skey, ns, (uint32_t)(jlo << shift));
      range[1] = lower_bound_u32(skey, ns, (uint32_t)((jhi + 1) << shift));
    } else {
      range[0] = lower_bound_u32(skey, nent, (uint32_t)(b * ns + jlo));
      range[1] = lower_bound_u32(skey, nent, (uint32_t)(b * ns + jhi + 1));
    }
  }
  __syncthreads();
  const int64_t p0 = range[0], p1 = range[1];
  double acc = 0.0;
  for (int64_t c0 = p0; c0 < p1; c0 += kBwdChunk) {
    const int cn = (int)(p1 - c0 < kBwdChunk ? p1 - c0 : kBwdChunk);
    __syncthreads();
    for (int q = threadIdx.x; q < cn; q += kBwdBlock) {
      int64_t i, j;
      if (shift > 0) {
        const uint32_t k = skey[c0 + q];
        j = k >> shift;
        i = k & ((1u << shift) - 1u);
      } else {
        i = (int64_t)sval[c0 + q] - b * ns;
        j = (int64_t)skey[c0 + q] - b * ns;
      }
      const int64_t e = b * ns + i;
      const Coef cf = coef[e];
      sP[q] = cf.P;
      sR[q] = cf.R;
      sI[q] = (int32_t)i;
      sJ[q] = (int32_t)j;
      sMi[q] = mean[b * stat_stride + i];
      sMj[q] = mean[b * stat_stride + j];
    }
    __syncthreads();
    if (t < n) {
      // the chunk's entries with partner j in [t-m+1, t] (sJ is sorted); the
      // count is known before the loop, so its loads can be issued together
      const int a = first_at_least(sJ, cn, t - m + 1);
      const int z = first_at_least(sJ, cn, t + 1);
#pragma unroll 4
      for (int q = a; q < z; ++q) {
        const int64_t i = sI[q], j = sJ[q];
        acc += sP[q] * (Tb[i + t - j] - sMi[q]) - sR[q] * (tt - sMj[q]);
      }
    }
  }
  if (t < n) grad_t[b * n + t] += acc;
}

// Workspace: coefficients, sort keys and values in and out, the sort's own
// temporary storage.
struct BwdWork {
  Coef *coef;
  uint32_t *key_in, *key_out, *val_in, *val_out;
  void *tmp;
  size_t tmp_bytes;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// The sort is stable (radix), so entries of one partner keep entry order.
// Merge sort is rocPRIM's default below 2^20 items; Onesweep (a few digit
// passes) is several times faster at the sizes of a batch of profiles, so
// the merge-sort limit is set to 0.
using SortConfig = rocprim::radix_sort_config<rocprim::default_config,
                                              rocprim::default_config,
                                              rocprim::default_config, 0>;

hipError_t sort_pairs(void *tmp, size_t &bytes, const uint32_t *kin, uint32_t *kout,
                      const uint32_t *vin, uint32_t *vout, int64_t nent,
                      int end_bit, hipStream_t s) {
  return rocprim::radix_sort_pairs<SortConfig>(tmp, bytes, kin, kout, vin, vout,
                                               static_cast<size_t>(nent), 0u,
                                               static_cast<unsigned>(end_bit), s);
}

hipError_t sort_bytes(int64_t nent, int end_bit, size_t *bytes) {
  return sort_pairs(nullptr, *bytes, nullptr, nullptr, nullptr, nullptr, nent,
                    end_bit, nullptr);
}

int key_bits(int64_t nent) {
  int bits = 1;
  while (bits < 32 && (int64_t(1) << bits) <= nent) ++bits;  // keys <= nent
  return bits;
}

BwdWork carve(void *base, int64_t nent, size_t tmp_bytes) {
  char *p = static_cast<char *>(base);
  BwdWork w;
  w.coef = reinterpret_cast<Coef *>(p);
  p += align256(sizeof(Coef) * nent);
  w.key_in = reinterpret_cast<uint32_t *>(p);
  p += align256(4 * nent);
  w.key_out = reinterpret_cast<uint32_t *>(p);
  p += align256(4 * nent);
  w.val_in = reinterpret_cast<uint32_t *>(p);
  p += align256(4 * nent);
  w.val_out = reinterpret_cast<uint32_t *>(p);
  p += align256(4 * nent);
  w.tmp = p;
  w.tmp_bytes = tmp_bytes;
  return w;
}

}  // namespace

size_t pearson_backward_workspace(int64_t n, int64_t m, int64_t batch) {
  const int64_t nent = batch * (n - m + 1);
  if (nent <= 0) return 0;
  size_t tmp = 0;
  if (sort_bytes(nent, key_bits(nent), &tmp) != hipSuccess) return 0;
  return align256(sizeof(Coef) * nent) + 4 * align256(4 * nent) + align256(tmp);
}

hipError_t launch_pearson_backward(const double *T, int64_t n, int64_t m,
                                   int64_t batch, const double *mean,
                                   const double *inorm, int64_t stat_stride,
                                   const int32_t *index, const double *grad_p,
                                   double *grad_t, void *work, hipStream_t s) {
  const int64_t ns = n - m + 1;
  if (ns <= 0 || batch <= 0) return hipSuccess;
  const int64_t nent = batch * ns;
  const int bits = key_bits(nent);
  size_t tmp = 0;
  hipError_t err = sort_bytes(nent, bits, &tmp);
  if (err != hipSuccess) return err;
  BwdWork w = carve(work, nent, tmp);
  const unsigned nb = static_cast<unsigned>(batch);
  // short series: packed keys sorted per series in LDS (one launch of
  // `batch` workgroups instead of the device-wide sort's passes)
  int shift = 0;
  if (ns <= kSeriesSortMax) {
    shift = 1;
    while ((int64_t(1) << shift) <= ns) ++shift;  // partners 0..ns fit
  }
  hipLaunchKernelGGL(bwd_coef_kernel,
                     dim3(static_cast<unsigned>((ns + kBwdBlock - 1) / kBwdBlock), nb),
                     dim3(kBwdBlock), 0, s, T, n, m, ns, batch, mean, inorm,
                     stat_stride, index, grad_p, w.coef, w.key_in, w.val_in, shift);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  const uint32_t *skey = w.key_out;
  if (shift > 0) {
    hipLaunchKernelGGL(bwd_sort_series_kernel, dim3(nb), dim3(kSeriesSortThreads), 0, s,
                       w.key_in, ns, 2 * shift);
    if ((err = hipGetLastError()) != hipSuccess) return err;
    skey = w.key_in;
  } else {
    err = sort_pairs(w.tmp, w.tmp_bytes, w.key_in, w.key_out, w.val_in, w.val_out,
                     nent, bits, s);
    if (err != hipSuccess) return err;
  }
  const dim3 grid(static_cast<unsigned>((n + kBwdBlock - 1) / kBwdBlock), nb);
  hipLaunchKernelGGL(bwd_self_kernel, grid, dim3(kBwdBlock), 0, s, T, n, m, ns,
                     mean, stat_stride, index, w.coef, grad_t);
  hipLaunchKernelGGL(bwd_partner_kernel, grid, dim3(kBwdBlock), 0, s, T, n, m, ns,
                     mean, stat_stride, w.coef, skey, w.val_out, nent, grad_t, shift);
  return hipGetLastError();
}

}  // namespace scamp_amd
