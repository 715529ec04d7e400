// Device-side helpers shared by the statistics, seed and sweep kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace scamp_amd {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kDiagsPerLane = 8;   // consecutive diagonals held by one lane
constexpr int kBand = kWave * kDiagsPerLane;  // diagonals per wave band
// Column-side arrays (statistics, profiles, clean series) are padded so that
// the sweep's look-ahead loads, the lanes of a partial last band and the
// FP32 form's re-anchoring windows never leave the allocation.
constexpr int64_t kColumnPad = 2 * kBand + 8 * kDiagsPerLane;
// Longest series the sweep takes: (subsequences + padding) * 8 bytes must fit
// the 32-bit byte offsets its loads and atomics use.
constexpr int64_t kMaxSubsequences = (int64_t(1) << 29) - 2 * kColumnPad;

__host__ __device__ inline uint32_t f32_bits(float f) {
  uint32_t b;
  memcpy(&b, &f, sizeof(b));
  return b;
}
__host__ __device__ inline float bits_f32(uint32_t b) {
  float f;
  memcpy(&f, &b, sizeof(f));
  return f;
}

// Profile keys.  A 1NN / 1NN_INDEX entry is held on the device as one uint64
// whose unsigned order is the order matches are ranked in: the FP32
// correlation in an order-preserving encoding in the high word, the
// bit-inverted index in the low word (so equal correlations keep the lower
// index).  One 64-bit atomicMax merges a candidate whatever the order of
// arrival.  The boundary converts to SCAMP's mp_entry layout
// (src/common.h:70-74: correlation in the low word) on the way in and out.
__host__ __device__ inline uint32_t enc_corr(float f) {
  uint32_t b = f32_bits(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float dec_corr(uint32_t e) {
  uint32_t b = (e & 0x80000000u) ? (e & 0x7fffffffu) : ~e;
  return bits_f32(b);
}
__host__ __device__ inline uint64_t make_key(float corr, uint32_t index) {
  return (static_cast<uint64_t>(enc_corr(corr)) << 32) |
         static_cast<uint32_t>(~index);
}
__host__ __device__ inline float key_corr(uint64_t key) {
  return dec_corr(static_cast<uint32_t>(key >> 32));
}
__host__ __device__ inline uint32_t key_index(uint64_t key) {
  return ~static_cast<uint32_t>(key);
}

// Filter thresholds.  A double correlation c can round to an FP32 value
// >= the stored correlation f only if c >= f - 2^-25 (|f| <= 2); testing
// c >= f - 2^-22 therefore never misses a candidate.  The FP32 form of the
// sweep compares rounded values directly.
template <typename T>
struct Thresh;
template <>
struct Thresh<double> {
  __device__ static double from_key_hi(uint32_t hi) {
    return from_corr(dec_corr(hi));
  }
  // the same for a (non-NaN) correlation before its encoding
  __device__ static double from_corr(float f) {
    return static_cast<double>(f) - 0x1p-22;
  }
};
template <>
struct Thresh<float> {
  __device__ static float from_key_hi(uint32_t hi) { return dec_corr(hi); }
  __device__ static float from_corr(float f) { return f; }
};

__device__ inline uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// The FP64 two-stage 1NN filter's first-stage factor of a column with
// inverse norm in_col: one Newton step from the hardware reciprocal
// r = (1 + e) / in_col with 2^-30 taken off inside the step:
// r (2 - in_col r) - 2^-30 r = (1 - e^2 - 2^-30 (1 + e)) / in_col up to two
// roundings of 2^-53, below (1 - 2^-31) / in_col whatever the reciprocal's
// error e (|e| < 1/4), then rounded down to FP32 (FLT_MAX at most).  +inf for
// a column that can never match (in_col NaN).  Computed once per column with
// the statistics (stage1_factors_kernel) and read by the sweep.
__device__ __forceinline__ float stage1_factor(double in_col) {
  const double r = __builtin_amdgcn_rcp(in_col);
  const double q = fma(r, fma(-in_col, r, 1.0 - 0x1p-30), r);
  if (!(in_col == in_col)) return __builtin_inff();
  float f = (float)q;  // nearest; one step down where that rounded up
  if ((double)f > q) f = __int_as_float(__float_as_int(f) - 1);
  return f;
}

}  // namespace scamp_amd
