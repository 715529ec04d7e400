// Statistics, exact QT seeds and profile-format conversions.
//
// Statistics replace compute_statistics_cpu (src/cpu_stats.cpp:123-160): the
// reference runs sequential O(n) recurrences on one host thread before any
// worker starts; here every window is summed directly (O(n*m) work, which the
// GPU finishes in milliseconds) with compensated summation, so the result does
// not accumulate error along the series.  The definitions are SCAMP's:
//   window i has a non-finite sample  -> inorm NaN        (cpu_stats.cpp:11-29)
//   SSD_i <= 1e-13 (flat window)      -> inorm NaN        (cpu_stats.cpp:146)
//   inorm_i = 1/sqrt(SSD_i)                               (cpu_stats.cpp:150)
//   df_i = (T[i+m]-T[i])/2, dg_i = (T[i+m]-mu[i+1])+(T[i]-mu[i]), last = 0
//                                                         (cpu_stats.cpp:154-157)
#include "device_common.h"
#include "kernels.h"

namespace scamp_amd {
namespace {

constexpr int kStatBlock = 256;
constexpr int kStatChunk = 1024;
constexpr double kFlatness = 1e-13;

// The cleaned series replaces each non-finite sample by the nearest finite
// sample before it (after it, for a leading run; 0 beyond kHoldSearch).
// SCAMP writes 0 there (src/cpu_stats.cpp:17-23).  Any value gives the same
// correlations -- every window that contains the sample is masked NaN, and
// the statistics, seeds and recurrence all read this same cleaned series, so
// the covariance of every valid window pair is the exact one -- but holding
// the level avoids update terms as large as the series itself around NaN
// gaps, which the FP32 recurrence could not absorb without losing digits.
constexpr int kHoldSearch = 4096;

// Statistics kernels: one series per blockIdx.y, the input `t_stride` and
// the outputs `clean_len` / `stat_len` entries apart.
__global__ void clean_kernel(const double *__restrict__ T, int64_t n,
                             double *__restrict__ clean, int64_t clean_len,
                             int64_t t_stride) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= clean_len) return;
  T += blockIdx.y * t_stride;
  clean += blockIdx.y * clean_len;
  if (i >= n) {
    clean[i] = 0.0;
    return;
  }
  double v = T[i];
  if (!isfinite(v)) {
    v = 0.0;
    bool found = false;
    for (int64_t j = i - 1; j >= 0 && j >= i - kHoldSearch; --j) {
      if (isfinite(T[j])) {
        v = T[j];
        found = true;
        break;
      }
    }
    for (int64_t j = i + 1; !found && j < n && j <= i + kHoldSearch; ++j) {
      if (isfinite(T[j])) {
        v = T[j];
        found = true;
      }
    }
  }
  clean[i] = v;
}

// One thread per window.  The block's samples are staged through LDS in
// chunks (each sample is read by up to m windows); pass 1 sums the window
// with an error-free two-sum accumulation and flags non-finite samples, pass
// 2 accumulates the squared deviations from the mean.
__global__ __launch_bounds__(kStatBlock) void window_stats_kernel(
    const double *__restrict__ T, const double *__restrict__ clean, int64_t n,
    int64_t m, double *__restrict__ means, double *__restrict__ inorms,
    int64_t ns, int64_t stat_len, int64_t t_stride, int64_t clean_len) {
  T += blockIdx.y * t_stride;
  clean += blockIdx.y * clean_len;
  means += blockIdx.y * stat_len;
  inorms += blockIdx.y * stat_len;
  __shared__ double tile[kStatBlock + kStatChunk];
  __shared__ bool flag[kStatBlock + kStatChunk];
  const int64_t w0 = blockIdx.x * (int64_t)kStatBlock;
  const int64_t w = w0 + threadIdx.x;
  double hi = 0, lo = 0;
  bool bad = false;
  for (int64_t t0 = 0; t0 < m; t0 += kStatChunk) {
    const int64_t tend = (m - t0) < kStatChunk ? (m - t0) : kStatChunk;
    // the samples the block's windows read in this chunk
    const int64_t span = kStatBlock + tend - 1;
    __syncthreads();
    for (int64_t j = threadIdx.x; j < span; j += kStatBlock) {
      int64_t g = w0 + t0 + j;
      // stage the cleaned sample and whether the raw one was non-finite
      const double raw = g < n ? T[g] : 0.0;
      tile[j] = raw;
      if (!isfinite(raw)) tile[j] = clean[g];
      flag[j] = g < n && !isfinite(raw);
    }
    __syncthreads();
    if (w < ns) {
      for (int64_t t = 0; t < tend; ++t) {
        const double v = tile[threadIdx.x + t];
        bad |= flag[threadIdx.x + t];
        double s = hi + v;
        double z = s - hi;
        lo += (hi - (s - z)) + (v - z);
        hi = s;
      }
    }
  }
  const double mu = (hi + lo) / (double)m;
  double ssd = 0;
  for (int64_t t0 = 0; t0 < m; t0 += kStatChunk) {
    const int64_t tend = (m - t0) < kStatChunk ? (m - t0) : kStatChunk;
    const int64_t span = kStatBlock + tend - 1;
    __syncthreads();
    for (int64_t j = threadIdx.x; j < span; j += kStatBlock) {
      int64_t g = w0 + t0 + j;
      tile[j] = g < n ? clean[g] : 0.0;
    }
    __syncthreads();
    if (w < ns) {
      for (int64_t t = 0; t < tend; ++t) {
        double d = tile[threadIdx.x + t] - mu;
        ssd = fma(d, d, ssd);
      }
    }
  }
  if (w < ns) {
    means[w] = mu;
    inorms[w] = (bad || ssd <= kFlatness) ? NAN : 1.0 / sqrt(ssd);
  } else if (w < stat_len) {
    means[w] = 0.0;
    inorms[w] = NAN;
  }
}

__global__ void dfdg_kernel(const double *__restrict__ clean, int64_t m,
                            const double *__restrict__ means, int64_t ns,
                            double *__restrict__ df, double *__restrict__ dg,
                            int64_t stat_len, int64_t clean_len) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= stat_len) return;
  clean += blockIdx.y * clean_len;
  means += blockIdx.y * stat_len;
  df += blockIdx.y * stat_len;
  dg += blockIdx.y * stat_len;
  if (i < ns - 1) {
    double a = clean[i], b = clean[i + m];
    df[i] = (b - a) / 2.0;
    dg[i] = (b - means[i + 1]) + (a - means[i]);
  } else {
    df[i] = 0.0;
    dg[i] = 0.0;
  }
}

__global__ void stage1_factors_kernel(const double *__restrict__ inorm,
                                     float *__restrict__ rq, int64_t len) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) rq[i] = stage1_factor(inorm[i]);
}

__global__ void to_f32_kernel(const double *__restrict__ src,
                              float *__restrict__ dst, int64_t len) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < len) dst[i] = (float)src[i];
}

// Exact-summation seeds: Ogita-Rump-Oishi Dot2 (TwoProduct via FMA +
// TwoSum), the summation the reference's MATLAB notes describe
// (test/dot2s.m); accurate to about one rounding of the true sum of the
// centred products.  Thread j of a task owns diagonal diag0 + j.  The task's
// X samples and centred Y window are staged through LDS, kSeedChunk values of
// t at a time; each thread keeps four Dot2 accumulators over interleaved t
// (independent dependency chains) and merges them with TwoSum at the end.
constexpr int kSeedChunk = 256;

__device__ __forceinline__ void dot2_step(double &s, double &c, double a,
                                          double b) {
  const double p = a * b;
  const double e = fma(a, b, -p);
  const double sn = s + p;
  const double z = sn - s;
  c += (s - (sn - z)) + (p - z) + e;
  s = sn;
}

__global__ __launch_bounds__(256) void seed_kernel(
    const SeedTask *__restrict__ tasks, const double *__restrict__ X,
    const double *__restrict__ muX, const double *__restrict__ Y,
    const double *__restrict__ muY, int64_t m, double *__restrict__ out) {
  __shared__ double sxv[256 + kSeedChunk];
  __shared__ double syv[kSeedChunk];
  const SeedTask task = tasks[blockIdx.x];
  const int j = threadIdx.x;
  const bool live = j < task.count;
  const int64_t col0 = task.row + task.diag0;
  const double mx = live ? muX[col0 + j] : 0.0;
  const double my = muY[task.row];
  double s[4] = {0.0, 0.0, 0.0, 0.0}, c[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t t0 = 0; t0 < m; t0 += kSeedChunk) {
    const int cl = (int)(m - t0 < kSeedChunk ? m - t0 : kSeedChunk);
    __syncthreads();
    // exactly the samples the task's diagonals read
    for (int q = j; q < task.count + cl - 1; q += 256)
      sxv[q] = X[col0 + t0 + q];
    for (int q = j; q < cl; q += 256) syv[q] = Y[task.row + t0 + q] - my;
    __syncthreads();
    if (live) {
      int t = 0;
      for (; t + 4 <= cl; t += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
          dot2_step(s[u], c[u], syv[t + u], sxv[j + t + u] - mx);
      }
      for (; t < cl; ++t) dot2_step(s[0], c[0], syv[t], sxv[j + t] - mx);
    }
  }
  if (!live) return;
  double sum = s[0], err = c[0] + c[1] + c[2] + c[3];
#pragma unroll
  for (int u = 1; u < 4; ++u) {
    const double sn = sum + s[u];
    const double z = sn - sum;
    err += (sum - (sn - z)) + (s[u] - z);
    sum = sn;
  }
  out[task.out_offset + j] = sum + err;
}

__global__ void mp_to_keys_kernel(uint64_t *p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t e = p[i];
  float corr = bits_f32(static_cast<uint32_t>(e));
  uint32_t idx = static_cast<uint32_t>(e >> 32);
  p[i] = make_key(corr, idx);
}

__global__ void keys_to_mp_kernel(uint64_t *p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t k = p[i];
  p[i] = (static_cast<uint64_t>(key_index(k)) << 32) | f32_bits(key_corr(k));
}

// ALL_NEIGHBORS: prior matches -> key lists and list minima.
struct MatchRec {
  float corr;
  uint32_t row;
  uint32_t col;
};

__global__ void matches_to_knn_kernel(const MatchRec *m, int64_t n, int k,
                                      uint64_t init, uint64_t *lists,
                                      uint64_t *keys) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint64_t mn = ~0ull;
  for (int t = 0; t < k; ++t) {
    const MatchRec r = m[j * k + t];
    uint64_t key = init;
    if (r.corr > -2.0f) {
      const uint64_t kk = make_key(r.corr, r.row);
      if (kk > init) key = kk;
    }
    // a prior entry repeated in the list counts once
    for (int u = 0; u < t; ++u)
      if (lists[j * k + u] == key) key = init;
    lists[j * k + t] = key;
    mn = key < mn ? key : mn;
  }
  keys[j] = mn;
}

// Key lists -> matches, best first (selection sort of each short list in
// place; k is the number of neighbours asked for, typically a handful).
__global__ void knn_to_matches_kernel(uint64_t *lists, int64_t n, int k,
                                      uint64_t init, MatchRec *m) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint64_t *l = lists + j * k;
  for (int t = 0; t < k; ++t) {
    int best = t;
    for (int u = t + 1; u < k; ++u)
      if (l[u] > l[best]) best = u;
    const uint64_t key = l[best];
    l[best] = l[t];
    l[t] = key;
    MatchRec r;
    r.col = static_cast<uint32_t>(j);
    if (key > init) {
      r.corr = key_corr(key);
      r.row = key_index(key);
    } else {
      r.corr = -2.0f;
      r.row = 0xFFFFFFFFu;
    }
    m[j * k + t] = r;
  }
}

__global__ void f32_to_keys_kernel(const float *src, uint64_t *dst, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = make_key(src[i], 0xFFFFFFFFu);
}

__global__ void keys_to_f32_kernel(const uint64_t *src, float *dst, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = key_corr(src[i]);
}

// Profile conversions between SCAMP's layouts (n entries per series, series
// back to back) and padded key arrays (`kstride` entries per series); one
// series per blockIdx.y.
__global__ void profile_in_kernel(const void *src, bool mp, uint64_t *keys,
                                  int64_t n, int64_t kstride) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = blockIdx.y;
  uint64_t k;
  if (mp) {
    const uint64_t e = static_cast<const uint64_t *>(src)[q * n + i];
    k = make_key(bits_f32(static_cast<uint32_t>(e)), static_cast<uint32_t>(e >> 32));
  } else {
    k = make_key(static_cast<const float *>(src)[q * n + i], 0xFFFFFFFFu);
  }
  keys[q * kstride + i] = k;
}

__global__ void profile_out_kernel(const uint64_t *keys, int64_t kstride,
                                   void *dst, bool mp, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t q = blockIdx.y;
  const uint64_t k = keys[q * kstride + i];
  if (mp)
    static_cast<uint64_t *>(dst)[q * n + i] =
        (static_cast<uint64_t>(key_index(k)) << 32) | f32_bits(key_corr(k));
  else
    static_cast<float *>(dst)[q * n + i] = key_corr(k);
}

__global__ void f32_to_enc_kernel(float *p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) reinterpret_cast<uint32_t *>(p)[i] = enc_corr(p[i]);
}

__global__ void enc_to_f32_kernel(float *p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = dec_corr(reinterpret_cast<uint32_t *>(p)[i]);
}

__global__ void fill_u64_kernel(uint64_t *p, int64_t n, uint64_t v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

inline unsigned blocks_for(int64_t n, int bs) {
  return static_cast<unsigned>((n + bs - 1) / bs);
}


// Warm start of a self-join's 1NN profile.  Thread i evaluates its window
// against the kWarm windows just past the exclusion zone ahead of it, j = i + d
// for d in [e, e + kWarm), and each pair goes to both of its entries: the
// row side of i (keyRow[i]) and the column side of j (keyCol[j]).  These are
// pairs of the join, so the final profile is unchanged, but the sweep's
// thresholds start near their final values instead of at "no match".
// Covariances:
//   sum_t x_t (T[j + t] - mu_i) - (mu_j - mu_i) sum_t x_t,  x_t = T[i + t] - mu_i
// (both factors centred near the windows' level, so a large offset of the
// series costs no digits); the correlation is formed as
// (cov * inorm_col) * inorm_row, the short-chunk sweep's order (the
// long-chunk two-stage kernels form (cov * inorm_row) * inorm_col; the two
// differ by at most one FP64 rounding).  A block's windows and their partners are
// staged through LDS, kWarmChunk samples of t at a time; each thread slides a
// register window over the partners so that every staged sample feeds kWarm
// FMAs.  The column sides of the block's pairs meet in an LDS table (one key
// per partner window) before one atomicMax per partner; the row sides go out
// with atomicMax too, as both may land in one profile.  One series per
// blockIdx.y (stride entries apart); samples at or past n read as 0 and meet
// x = 0 or an invalid partner.
constexpr int kWarm = (int)kWarmLags;
constexpr int kWarmBlock = 256;
constexpr int kWarmChunk = 256;
constexpr int kWarmStage = kWarmBlock + kWarmChunk + 2 * kWarm;
constexpr int kWarmCols = kWarmBlock + kWarm - 1;  // partners i0 + e + [0, kWarmCols)

__global__ __launch_bounds__(kWarmBlock) void warm_kernel(
    const double *__restrict__ T, const double *__restrict__ mu,
    const double *__restrict__ inorm, int64_t n, int64_t ns, int64_t m,
    int64_t e, int64_t stride, uint64_t *keyCol, uint64_t *keyRow) {
  __shared__ double sx[kWarmStage];  // T[i0 + t0 + q]
  __shared__ double sf[kWarmStage];  // T[i0 + e + t0 + q]
  __shared__ unsigned long long scol[kWarmCols];
  const int j = threadIdx.x;
  const int64_t i0 = blockIdx.x * (int64_t)kWarmBlock;
  const int64_t i = i0 + j;
  const int64_t o = (int64_t)blockIdx.y * stride;
  T += o;
  mu += o;
  inorm += o;
  for (int q = j; q < kWarmCols; q += kWarmBlock) scol[q] = 0;
  const double ini = i < ns ? inorm[i] : nan("");
  const double mi = i < ns ? mu[i] : 0.0;
  double f[kWarm];  // partners i + e + w
#pragma unroll
  for (int w = 0; w < kWarm; ++w) f[w] = 0.0;
  double sxs = 0.0;
  for (int64_t t0 = 0; t0 < m; t0 += kWarmChunk) {
    const int cl = (int)(m - t0 < kWarmChunk ? m - t0 : kWarmChunk);
    __syncthreads();
    for (int q = j; q < kWarmStage; q += kWarmBlock) {
      const int64_t ix = i0 + t0 + q, jf = ix + e;
      sx[q] = ix < n ? T[ix] : 0.0;
      sf[q] = jf < n ? T[jf] : 0.0;
    }
    __syncthreads();
    // fw[q] = T[i + e + tt + q] - mi
    double fw[2 * kWarm - 1];
#pragma unroll
    for (int q = 0; q < kWarm - 1; ++q) fw[kWarm + q] = sf[j + q] - mi;
    for (int tt = 0; tt < cl; tt += kWarm) {
#pragma unroll
      for (int q = 0; q < kWarm - 1; ++q) fw[q] = fw[kWarm + q];
#pragma unroll
      for (int q = 0; q < kWarm; ++q) fw[kWarm - 1 + q] = sf[j + tt + kWarm - 1 + q] - mi;
#pragma unroll
      for (int u = 0; u < kWarm; ++u) {
        const double x = tt + u < cl ? sx[j + tt + u] - mi : 0.0;
        sxs += x;
#pragma unroll
        for (int w = 0; w < kWarm; ++w) f[w] = fma(x, fw[u + w], f[w]);  // T[i+e+w+t] - mi
      }
    }
  }
  uint64_t best_row = 0;
  if (i < ns && ini == ini) {  // a window with matches
#pragma unroll
    for (int w = 0; w < kWarm; ++w) {
      const int64_t jf = i + e + w;
      if (jf >= ns) continue;
      // column jf, row i
      const double c = ((f[w] - (mu[jf] - mi) * sxs) * inorm[jf]) * ini;
      if (!(c == c)) continue;  // NaN or flat partner
      const uint64_t kr = make_key((float)c, (uint32_t)jf);
      best_row = kr > best_row ? kr : best_row;
      if (keyCol)
        atomicMax(&scol[j + w], static_cast<unsigned long long>(
                                    make_key((float)c, (uint32_t)i)));
    }
  }
  if (keyRow && best_row != 0)
    atomicMax(reinterpret_cast<unsigned long long *>(keyRow + o + i),
              static_cast<unsigned long long>(best_row));
  if (!keyCol) return;
  __syncthreads();
  for (int q = j; q < kWarmCols; q += kWarmBlock) {
    const int64_t jf = i0 + e + q;
    const unsigned long long v = scol[q];
    if (v != 0 && jf < ns)
      atomicMax(reinterpret_cast<unsigned long long *>(keyCol + o + jf), v);
  }
}

}  // namespace

hipError_t launch_stats(const double *T, int64_t n, int64_t m, double *clean,
                        int64_t clean_len, double *means, double *inorms,
                        double *df, double *dg, int64_t stat_len,
                        hipStream_t s, int64_t batch, int64_t t_stride) {
  const int64_t ns = n - m + 1;
  const unsigned nb = static_cast<unsigned>(batch);
  hipLaunchKernelGGL(clean_kernel, dim3(blocks_for(clean_len, 256), nb),
                     dim3(256), 0, s, T, n, clean, clean_len, t_stride);
  hipLaunchKernelGGL(window_stats_kernel,
                     dim3(blocks_for(stat_len, kStatBlock), nb), dim3(kStatBlock),
                     0, s, T, clean, n, m, means, inorms, ns, stat_len, t_stride,
                     clean_len);
  hipLaunchKernelGGL(dfdg_kernel, dim3(blocks_for(stat_len, 256), nb), dim3(256),
                     0, s, clean, m, means, ns, df, dg, stat_len, clean_len);
  return hipGetLastError();
}

hipError_t launch_stage1_factors(const double *inorm, float *rq, int64_t len,
                                 hipStream_t s) {
  hipLaunchKernelGGL(stage1_factors_kernel, dim3(blocks_for(len, 256)), dim3(256),
                     0, s, inorm, rq, len);
  return hipGetLastError();
}

hipError_t launch_to_f32(const double *src, float *dst, int64_t len,
                         hipStream_t s) {
  hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_for(len, 256)), dim3(256), 0,
                     s, src, dst, len);
  return hipGetLastError();
}

hipError_t launch_seeds(const SeedTask *tasks, int64_t ntasks,
                        const double *X, const double *muX, const double *Y,
                        const double *muY, int64_t m, double *out,
                        hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(seed_kernel, dim3(static_cast<unsigned>(ntasks)),
                     dim3(256), 0, s, tasks, X, muX, Y, muY, m, out);
  return hipGetLastError();
}

#define SCAMP_SIMPLE_LAUNCH(kernel, n, ...)                                   \
  do {                                                                        \
    if ((n) > 0)                                                              \
      hipLaunchKernelGGL(kernel, dim3(blocks_for((n), 256)), dim3(256), 0, s, \
                         __VA_ARGS__);                                        \
    return hipGetLastError();                                                 \
  } while (0)

hipError_t launch_mp_to_keys(uint64_t *p, int64_t n, hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(mp_to_keys_kernel, n, p, n);
}
hipError_t launch_keys_to_mp(uint64_t *p, int64_t n, hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(keys_to_mp_kernel, n, p, n);
}
hipError_t launch_matches_to_knn(const void *matches, int64_t n, int k,
                                 uint64_t init, uint64_t *lists, uint64_t *keys,
                                 hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(matches_to_knn_kernel, n,
                      static_cast<const MatchRec *>(matches), n, k, init, lists,
                      keys);
}
hipError_t launch_knn_to_matches(uint64_t *lists, int64_t n, int k, uint64_t init,
                                 void *matches, hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(knn_to_matches_kernel, n, lists, n, k, init,
                      static_cast<MatchRec *>(matches));
}
hipError_t launch_f32_to_keys(const float *src, uint64_t *dst, int64_t n,
                              hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(f32_to_keys_kernel, n, src, dst, n);
}
hipError_t launch_keys_to_f32(const uint64_t *src, float *dst, int64_t n,
                              hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(keys_to_f32_kernel, n, src, dst, n);
}
hipError_t launch_profile_in(const void *src, bool mp, uint64_t *keys,
                             int64_t n, int64_t batch, int64_t kstride,
                             hipStream_t s) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(profile_in_kernel,
                     dim3(blocks_for(n, 256), static_cast<unsigned>(batch)),
                     dim3(256), 0, s, src, mp, keys, n, kstride);
  return hipGetLastError();
}
hipError_t launch_profile_out(const uint64_t *keys, int64_t kstride, void *dst,
                              bool mp, int64_t n, int64_t batch, hipStream_t s) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(profile_out_kernel,
                     dim3(blocks_for(n, 256), static_cast<unsigned>(batch)),
                     dim3(256), 0, s, keys, kstride, dst, mp, n);
  return hipGetLastError();
}
hipError_t launch_f32_to_enc(float *p, int64_t n, hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(f32_to_enc_kernel, n, p, n);
}
hipError_t launch_enc_to_f32(float *p, int64_t n, hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(enc_to_f32_kernel, n, p, n);
}
hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s) {
  SCAMP_SIMPLE_LAUNCH(fill_u64_kernel, n, p, n, v);
}


hipError_t launch_warm_profile(const double *clean, const double *mean,
                               const double *inorm, int64_t ns, int64_t m,
                               int64_t e, int64_t batch, int64_t stride,
                               uint64_t *keyCol, uint64_t *keyRow,
                               hipStream_t s) {
  if (ns <= 0 || batch <= 0 || (!keyCol && !keyRow)) return hipSuccess;
  hipLaunchKernelGGL(warm_kernel,
                     dim3(blocks_for(ns, kWarmBlock), static_cast<unsigned>(batch)),
                     dim3(kWarmBlock), 0, s, clean, mean, inorm, ns + m - 1, ns,
                     m, e, stride, keyCol, keyRow);
  return hipGetLastError();
}

}  // namespace scamp_amd
