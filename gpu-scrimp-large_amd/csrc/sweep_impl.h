// The diagonal sweep: the matrix-profile join's hot loop, written for gfx950.
//
// Reference behaviour (what is computed): SCAMP's do_tile kernel
// (src/kernels.cu:22-167, src/kernels_compute.h:412-623) and its CPU twin
// (src/cpu_kernels.cpp:117-199).  For a pair (row i of series Y, column j of
// series X) on diagonal d = j - i the covariance obeys
//     cov(i+1, j+1) = cov(i, j) + df_X[j]*dg_Y[i] + df_Y[i]*dg_X[j]
// (src/cpu_kernels.cpp:166-171), the Pearson correlation is
//     corr = cov(i, j) * inorm_X[j] * inorm_Y[i]        (cpu_kernels.cpp:146)
// and the profile keeps, per column and per row, the best FP32 correlation and
// its index (1NN_INDEX / 1NN), the sum of correlations above a threshold
// (SUM_THRESH), or the max-pooled matrix (MATRIX_SUMMARY).  NaN correlations
// (windows with a non-finite sample, flat windows) never match.
//
// How it runs here (MI355X design, not the reference's):
//  * Work unit = a wave band: 64 lanes x 8 consecutive diagonals = 512
//    diagonals, swept down all their rows by one wavefront.  A persistent grid
//    pulls bands from an atomic counter, longest first.  Band origins are
//    aligned to 8 diagonals (diagonals outside the pass carry NaN).
//  * Each band starts at a pseudo-random row r0 (its FP64 seed computed
//    exactly for that row) and wraps round to row 0, so the bands in flight
//    sit at spread-out rows and each row's visits are spread over the run.
//  * A lane holds its 8 covariances and the terms (update terms, inverse
//    norm, threshold) of the 8 columns its diagonals cross on the current
//    row, in slots indexed by column mod 8.  Going down one row, every lane
//    drops one column and needs one new one -- exactly the column its
//    right-hand neighbour drops, in the same slot.  So one DPP wave shift of
//    that slot per row moves every column one lane down; only lane 63 takes a
//    new column, from an LDS stage filled a block of 8 rows ahead with
//    coalesced loads.  Row terms come through the same stage.  Each column is
//    read from memory once per band.
//  * 1NN / 1NN_INDEX: each correlation is tested against the current best of
//    its row and of its column (thresholds read from the profile itself).
//    Only when some lane passes does the wave take the slow path: exact FP32
//    rounding, lowest-index argmax within the wave, and 64-bit atomicMax of
//    order-preserving keys straight into the profile.  The profile only grows,
//    so a stale threshold costs an extra slow path, never a missed match.
//  * SUM_THRESH: row sums reduced across the wave per row; a column's partial
//    sum travels with the column in its slot and reaches memory once per band,
//    when the column leaves lane 0.
//  * MATRIX_SUMMARY: exact max-pooling; each lane keeps two running cell
//    maxima (its 8 columns straddle at most one cell boundary) and flushes a
//    cell with a 32-bit atomicMax when its columns or the row leave it.
#pragma once

#include "device_common.h"
#include "kernels.h"

namespace scamp_amd {
namespace sweep {

constexpr int K = kDiagsPerLane;
// u-samples per LDS stage of a seed / re-anchor: the FP32 form has its own
// staging, sized (with the staged terms) to 10 KB, which lets 16 one-wave
// workgroups share a CU's 160 KB of LDS (the FP32 short-chunk kernel runs 4
// waves per SIMD); the FP64 form stages through its row and column term
// arrays (idle at a segment start, the only place it seeds), which hold 640
// doubles
constexpr int kAnchorChunk = 184;
constexpr int kAnchorChunk64 = 64;

template <typename DT>
struct Slot {
  DT df, dg, in;   // column update terms and inverse norm
  DT t;            // 1NN: threshold of the column's current best
  // Two-stage 1NN filter (see Sweep::kTwoStage): mt = the first-stage bound
  // min(block row bound, column threshold) / in (row-scaled units); the FP64
  // form keeps its column threshold as the integer image F (enc_thr)
  // instead of t.
  DT mt;
  int32_t F;
  double acc;      // SUM: partial sum of the column, travelling with it
};

// Terms of one row or column, staged through LDS.
template <typename DT>
struct Terms {
  DT df, dg, in;
  DT t;       // 1NN: threshold of the row's / column's current best
  DT b;       // FP32 two-stage filter: the column's first-stage bound
  DT s1;      // two-stage filter, rows: the first stage's scale (in, or 0
              // for a row that can never match)
};
template <>
struct Terms<double> {
  double df, dg, in;
  double t;
  int32_t F;  // two-stage filter: enc_thr of the column's threshold
  double s1;  // two-stage filter, rows: the first stage's scale (in, or 0
              // for a row that can never match)
};

// Integer image of an FP64 threshold t for the two-stage filter: for c >= 0,
// c >= t implies hi32(c) >= enc_thr(t) (the high word of a non-negative
// double orders like the double and truncates it downwards); a negative
// threshold maps to kNegEnc, below every non-negative image, which decodes
// to -1.0625 with a zero low word (the first stage then passes every
// correlation) and which the exact column test treats as "take any
// correlation" (col_passes: negative doubles' high words order the other
// way round as integers); a window that can never match gets kNanEnc, which
// no correlation's high word reaches and which decodes to NaN.
constexpr int32_t kNegEnc = (int32_t)0xBFF10000u;
constexpr int32_t kNanEnc = 0x7FFFFFFF;
__device__ __forceinline__ int32_t enc_thr(double t) {
  return t >= 0.0 ? __double2hiint(t) : kNegEnc;
}
// The FP32 two-stage filter orders thresholds by their bit patterns (one
// integer min, no NaN handling): thresholds enter with every negative value
// replaced by -2, below any correlation, so that the bit patterns of the
// values in play (-2 and non-negative ones) order like the values.
__device__ __forceinline__ float canon_thr(float t) { return t < 0.0f ? -2.0f : t; }
__device__ __forceinline__ float min_thr(float a, float b) {
  const int32_t x = __float_as_int(a), y = __float_as_int(b);
  return __int_as_float(x < y ? x : y);
}
// The first-stage FP64 threshold: the double whose high word is
// min(row bound, column image) and whose low word is 0 (so it never exceeds
// either threshold it stands for).
__device__ __forceinline__ double mt_of(int32_t tlo, int32_t f) {
  return __hiloint2double(tlo < f ? tlo : f, 0);
}
// The FP64 first stage in row-scaled units: a pair passes when the exact
// product s = cov * in_row is at least B = fl(mt_of(tlo, f) * q - 2^-1000),
// q = stage1_factor(in_col) (a column constant, kept per column in memory
// beside the inverse norms, so that a block's refresh of B is one FMA and a
// load issued a block ahead).  q < (1 - 2^-31) / in_col (stage1_factor).
// So for a threshold image m >= 0, B < m / in_col by a relative 2^-31 and an absolute 2^-1000, while
// s * in_col >= threshold >= m gives s >= m / in_col less a relative 2^-52:
// the stage is conservative.  The negative image (-1.0625) gives B below
// -1.0625 (1 - 2^-30) / in_col, under every correlation >= -1.  A column
// that can never match (in_col NaN) has q = +inf and B = +inf, -inf or NaN;
// its pairs may pass the first stage but never the exact one.
__device__ __forceinline__ double stage1_bound(int32_t tlo, int32_t f, float q) {
  return fma(mt_of(tlo, f), (double)q, -0x1p-1000);
}
// Two-stage kernels: the covariance an invalid diagonal (outside the pass,
// at a band's edge) carries instead of NaN, so that its first-stage
// difference is hugely negative rather than NaN (whose sign bit is not
// defined); the exact stage masks such diagonals to NaN.  Far below any
// covariance, and times any inverse norm (< 2^22) still finite.
template <typename DT>
constexpr DT far_cov();
template <>
constexpr double far_cov<double>() { return -0x1p1000; }
template <>
constexpr float far_cov<float>() { return -0x1p100f; }

// The integer whose sign is the sign of x (FP64: the high word).
__device__ __forceinline__ int32_t sign_word(double x) { return __double2hiint(x); }
__device__ __forceinline__ int32_t sign_word(float x) { return __float_as_int(x); }

// The FP32 first stage in row-scaled units (as stage1_bound): B =
// min(tlo, t) / in_col less a relative 2^-16 (far above the FP32 rounding of
// the product and of the hardware reciprocal) and an absolute 2^-100; +inf
// for a column that can never match.  (Carrying a lowered reciprocal with
// the column as the FP64 form does, with the exact stage reading in_col from
// memory, measured slower here: 1193 -> 1284 ms at n = 2^22, m = 1024; the
// FP32 form's 64-row blocks spend half as much per row on the refresh.)
__device__ __forceinline__ float stage1_bound32(float tlo, float t, float in_col) {
  const float b = min_thr(tlo, t) * __builtin_amdgcn_rcpf(in_col);
  const float lowered = fmaf(-fabsf(b), 0x1p-16f, b) - 0x1p-100f;
  return in_col == in_col ? lowered : __builtin_inff();
}

// SUM_THRESH's first stage: the bound thr / in_col in row-scaled units,
// lowered as stage1_bound's (a relative 2^-30 for FP64, 2^-16 for FP32,
// and a tiny absolute amount); +inf for a column that can never match.
__device__ __forceinline__ double sum_bound(double thr, double in_col) {
  double r = __builtin_amdgcn_rcp(in_col);
  r = fma(r, fma(-in_col, r, 1.0), r);
  const double b = thr * r;
  const double lowered = fma(-fabs(b), 0x1p-30, b) - 0x1p-1000;
  return in_col == in_col ? lowered : __builtin_inf();
}
__device__ __forceinline__ float sum_bound(double thr, float in_col) {
  const float b = (float)thr * __builtin_amdgcn_rcpf(in_col);
  const float lowered = fmaf(-fabsf(b), 0x1p-16f, b) - 0x1p-100f;
  return in_col == in_col ? lowered : __builtin_inff();
}

template <typename DT>
__device__ __forceinline__ DT nan_of();
template <>
__device__ __forceinline__ double nan_of<double>() {
  return __builtin_nan("");
}
template <>
__device__ __forceinline__ float nan_of<float>() {
  return __builtin_nanf("");
}

// Wavefront shift by one lane toward lane 0 (DPP wave_shl:1): lane L
// receives lane L+1's value, lane 63 receives `edge`.
__device__ __forceinline__ float shl1(float v, float edge) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      __float_as_int(edge), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ int32_t shl1(int32_t v, int32_t edge) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xf, 0xf, false);
}
__device__ __forceinline__ double shl1(double v, double edge) {
  const long long x = __double_as_longlong(v), e = __double_as_longlong(edge);
  const int lo = __builtin_amdgcn_update_dpp((int)e, (int)x, 0x130, 0xf, 0xf,
                                             false);
  const int hi = __builtin_amdgcn_update_dpp((int)(e >> 32), (int)(x >> 32),
                                             0x130, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Per-lane state of the MATRIX_SUMMARY kinds.
struct CellAcc {
  double lo, hi;       // running maxima of the current cell and the next one
  int64_t lo_cell;     // cell coordinate (along the lane's columns) of `lo`
  int64_t pos;         // (first column of the lane) mod cells-per-step
};

// FRESH (1NN only): a launch whose profile thresholds start far from their
// final values (short series cut into short row chunks), where many rows take
// the merge path.  Its rows merge into an LDS table flushed once per block (no
// wave-wide reduction per row), its columns into a second one (no global
// atomic per record), and a row runs only the side of the merge whose test
// passed.  Kept out of the long-chunk kernel, whose hot loop the extra
// registers would slow down.
template <typename DT, int KIND, bool COLS, bool ROWS, bool FRESH = false>
struct Sweep {
  const SweepArgs &a;
  const int lane;
  const int64_t D0;      // first diagonal of the band (multiple of K)
  const int64_t d0;      // first diagonal of this lane
  // FP64 form: cov[k] is the covariance of diagonal d0+k at the current row.
  // FP32 form: the covariance is hi[k] + cov[k]; hi[k] is its leading part
  // and cov[k] collects the row increments of the last few rows on top of
  // the low part, so FP32 rounding hits the small partial sum every row and
  // the large leading part only once per kRenormRows rows (compensated, by
  // Fast2Sum).  The partial's own rounding grows with its size, which is
  // what sets kRenormRows: large increments (a high-amplitude stretch of one
  // series) leave an absolute error behind that later low-norm windows
  // magnify.
  static constexpr bool kComp = sizeof(DT) == 4;
  // kinds that test each pair against per-row / per-column threshold keys
  static constexpr bool kThr = KIND == kNN || KIND == kKnn;
  static constexpr int kRenormRows = 4;
  // Rows per staged block.  The 1NN kinds take two slot cycles per block,
  // which halves the per-block work (staging, threshold decoding,
  // loop-carried register copies); together with the occupancy targets on
  // sweep_kernel this measured 4.4 % (FP64) and 7.5 % (FP32) faster at
  // n = 2^22, m = 1024 than one cycle per block.  The two-stage kernels
  // (below) take longer blocks still, which spreads the block's staging and
  // row-bound reduction over more rows at the price of a looser row bound
  // and more code.  Measured in one process at n = 2^22, m = 1024 (16 / 32 /
  // 64 rows): FP32 1423 / 1386 / 1329 ms; FP64 (the current kernel, 32 / 64
  // rows) 1289 / 1239 ms, where 64 rows also leave the compiler enough
  // registers to keep the kernel free of spills.  The FP64 column-only
  // kernel (one side of an AB-join) also runs 64-row blocks: 139 registers
  // and no spills against 168 and 12 spilled dwords at 16 rows, AB-join
  // 2^20 x 2^20, m = 256: 197.9 -> 194.8 ms; its FP32 twin and the row-only
  // kernels spill more with longer blocks.  SUM_THRESH runs 64-row blocks
  // (row sums still flushed every 8 rows): self-join n = 2^21, m = 256,
  // 8 / 32 / 64 rows: FP64 467 / 426 / 406 ms, mixed 426 / 392 / 385 ms;
  // A's sums of an AB-join of two 2^20 walks: FP64 218 -> 181 ms, mixed
  // 179 -> 157 ms (8 -> 64 rows).  MATRIX_SUMMARY runs 32-row blocks
  // (self-join n = 2^21, m = 1024, 1024^2 cells, 8 / 32 / 64 rows: FP64
  // 540 / 482 / 488 ms, mixed 416 / 383 / 390 ms; a longer block is less
  // often clear of cell boundaries).  A block is at most one row per lane
  // of the staging arrays.
  static constexpr int kRows =
      KIND == kNN ? (!FRESH && COLS && (ROWS || sizeof(DT) == 8) ? 8 * K : 2 * K)
      : KIND == kSum ? 8 * K
      : KIND == kMatrix ? 4 * K
                        : K;
  static_assert(kRows <= kWave, "one staged row and column per lane");
  // Two-stage 1NN filter (rows and columns, long chunks): every pair is
  // first tested once against min(T_lo, column threshold), T_lo the lowest
  // row threshold of the block; only rows where some pair passes run the
  // exact row and column tests.  The first stage works in row-scaled units:
  // s = cov * in_row against mt = min(T_lo, column threshold) / in_col
  // (a per-column bound, made when the column enters the band or a block
  // starts, that travels with the column), so a pair costs one FMA and a
  // share of a max tree (below); the exact correlation s * in_col is formed
  // only on rows that pass.  With converged thresholds a few percent of rows
  // do.  The FP64 form keeps each column's threshold as an integer image
  // (one register, one DPP move) for its exact column test; its slots carry
  // neither in_col (the exact stage reads it from memory) nor the column's
  // first-stage factor (stage1_factor, read from memory for each block's
  // refresh, a block ahead): per row the hand-off moves df, dg, the bound
  // and the image, 7 DPP moves.
  static constexpr bool kTwoStage = KIND == kNN && !FRESH && COLS && ROWS;
  static constexpr bool kIntCol = kTwoStage && sizeof(DT) == 8;
  // The first stage is tested by sign: d_k = fma(cov_k, in_row, -B_k) for
  // the lane's 8 pairs, and a row passes when the largest high word (FP64)
  // or bit pattern (FP32) of the d_k is >= 0 as an integer, i.e. some
  // d_k >= 0: 8 FMAs and a max3 tree per row in place of 8 multiplies and 8
  // compares, each of which would hand a mask to the scalar unit.  No d_k
  // of a pair that can match may be NaN (its sign bit is not defined):
  // invalid diagonals carry far_cov() instead of NaN (the exact stage masks
  // them with vmask), a row that can never match scales its first stage by 0
  // (Terms::s1) while its exact stage keeps the NaN inverse norm; a column
  // that can never match has the bound +inf (FP32), or +-inf or NaN (FP64:
  // such a column may pass the first stage, never the exact one).
  // SUM_THRESH tests its rows the same way, against a per-column bound
  // thr / in_col (a constant of the column, made when it enters the band and
  // moved with it); the exact sum runs only on rows where some pair passes.
  static constexpr bool kSumStage = KIND == kSum;
  // kernels whose invalid diagonals carry far_cov() (see above)
  static constexpr bool kFarDiag = kTwoStage || kSumStage;
  uint32_t vmask = 0;  // kFarDiag kernels: the lane's valid diagonals
  // the block's row bound: FP32 the threshold itself, FP64 its enc_thr image
  DT tlo_f = (DT)0;
  int32_t tlo_i = 0;
  DT cov[K];
  DT hi[K];
  uint64_t *s_kq = nullptr;  // ALL_NEIGHBORS: per-lane candidate slots (LDS)
  static constexpr bool kTable = KIND == kNN && FRESH;
  uint64_t *s_rt = nullptr;  // FRESH: the block's row table (LDS, kRows keys)
  // FRESH: the block's column table (LDS).  Column i_b + D0 + p, p in
  // [0, kBand + kRows - 1), sits at (p % K) * kColStride + p / K, so the
  // lanes of one merge (p = rr + k + K * lane) hit consecutive entries.
  static constexpr bool kColTable = kTable && COLS;
  static constexpr int kColStride = (kBand + kRows - 1 + K - 1) / K;
  uint64_t *s_ct = nullptr;
  bool ct_dirty = false;     // wave-uniform: some column merged this block
  // SUM_THRESH rows: this lane's partial sum of each row of the current block
  double rsum[K];
  Slot<DT> S[K];         // column slots, indexed by column mod K
  // MATRIX_SUMMARY state
  DT blk0, blk1;         // fast blocks: running max of the lane's pairs
  CellAcc dir, tra;
  int64_t dir_row_cell, dir_row_pos;  // uniform: row cell and row mod rpc
  int64_t tra_col_cell, tra_col_pos;  // uniform: column cell and row mod cpc

  __device__ Sweep(const SweepArgs &args, int ln, int64_t band_d0)
      : a(args), lane(ln), D0(band_d0), d0(band_d0 + (int64_t)ln * K) {}

  __device__ __forceinline__ bool diag_valid(int64_t d) const {
    return d >= a.d_lo && d < a.d_hi;
  }

  // Byte offsets into the padded arrays fit 32 bits, which lets every load
  // use a uniform base address plus one per-lane 32-bit offset.
  template <typename T>
  __device__ __forceinline__ static T ld(const void *base, uint32_t byte_off) {
    return *reinterpret_cast<const T *>(static_cast<const char *>(base) +
                                        byte_off);
  }

  // Loads one column into a slot (segment start only; afterwards columns
  // reach a lane from its right-hand neighbour).
  __device__ __forceinline__ void load_slot(Slot<DT> &s, int64_t col) const {
    const uint32_t off = static_cast<uint32_t>(col) * sizeof(DT);
    s.df = ld<DT>(a.dfX, off);
    s.dg = ld<DT>(a.dgX, off);
    s.in = ld<DT>(a.inX, off);
    if (kThr && COLS) {
      const DT t = Thresh<DT>::from_key_hi(
          ld<uint32_t>(a.keyX, static_cast<uint32_t>(col) * 8u + 4u));
      if constexpr (kIntCol)
        s.F = s.in == s.in ? enc_thr(t) : kNanEnc;
      else if constexpr (kTwoStage)
        s.t = canon_thr(t);
      else
        s.t = t;
    }
    if (KIND == kSum) {
      s.acc = 0.0;
      s.mt = sum_bound(a.thr, s.in);
    }
  }

  // Two-stage filter: the first-stage threshold of a slot from the block's
  // row bound and the column's threshold.
  // q: the slot column's first-stage factor (FP64 form).
  __device__ __forceinline__ void refresh_mt(Slot<DT> &s, float q) const {
    if constexpr (kIntCol)
      s.mt = stage1_bound(tlo_i, s.F, q);
    else
      s.mt = stage1_bound32(tlo_f, s.t, s.in);
  }

  // Terms staged a block ahead: lane l fetches row ib+l (the rows of the
  // block) and column ib+D0+kBand+l (the column entering the band at lane 63
  // on row ib+l); lanes >= kRows fetch clamped / padded entries nobody reads, so
  // the code has no divergent branch.
  // The FP64 two-stage kernel also fetches the first-stage factors of the
  // lane's columns on the block's first row (ib + d0 + k, slot k) for the
  // block's bound refresh.
  struct Pre {
    DT rdf, rdg, rin;
    uint32_t rkey;
    DT cdf, cdg, cin;
    uint32_t ckey;
    float rq[kIntCol ? K : 1];
  };
  __device__ __forceinline__ void prefetch(Pre &p, int64_t ib) const {
    int64_t ri = ib + lane;
    ri = ri < a.nY ? ri : a.nY - 1;
    const uint32_t roff = static_cast<uint32_t>(ri) * sizeof(DT);
    p.rdf = ld<DT>(a.dfY, roff);
    p.rdg = ld<DT>(a.dgY, roff);
    p.rin = ld<DT>(a.inY, roff);
    if (kThr && ROWS)
      p.rkey = ld<uint32_t>(a.keyY, static_cast<uint32_t>(ri) * 8u + 4u);
    const int64_t ci = ib + D0 + kBand + lane;
    const uint32_t coff = static_cast<uint32_t>(ci) * sizeof(DT);
    p.cdf = ld<DT>(a.dfX, coff);
    p.cdg = ld<DT>(a.dgX, coff);
    p.cin = ld<DT>(a.inX, coff);
    if (kThr && COLS)
      p.ckey = ld<uint32_t>(a.keyX, static_cast<uint32_t>(ci) * 8u + 4u);
    if constexpr (kIntCol) {
      // columns ib + d0 + k < ci: inside the padded arrays wherever ci is
      const uint32_t qoff = static_cast<uint32_t>(ib + d0) * 4u;
#pragma unroll
      for (int k = 0; k < K; ++k) p.rq[k] = ld<float>(a.rqX, qoff + 4u * k);
    }
  }
  __device__ __forceinline__ void publish(const Pre &p, Terms<DT> *s_row,
                                          Terms<DT> *s_col) {
    Terms<DT> r, c;
    r.df = p.rdf;
    r.dg = p.rdg;
    r.in = p.rin;
    r.t = (kThr && ROWS) ? Thresh<DT>::from_key_hi(p.rkey) : (DT)0;
    c.df = p.cdf;
    c.dg = p.cdg;
    c.in = p.cin;
    c.t = (kThr && COLS) ? Thresh<DT>::from_key_hi(p.ckey) : (DT)0;
    if constexpr (kIntCol) {
      r.F = 0;
      c.F = c.in == c.in ? enc_thr(c.t) : kNanEnc;
    } else if constexpr (kTwoStage) {
      r.t = canon_thr(r.t);
      c.t = canon_thr(c.t);
    }
    if constexpr (kSumStage) {
      // the entering column's bound travels in its record's t; a row that
      // can never match scales its first stage by 0
      c.t = sum_bound(a.thr, c.in);
      r.s1 = r.in == r.in ? r.in : (DT)0;
    }
    if constexpr (!kTwoStage) {
      s_row[lane] = r;
      s_col[lane] = c;
    }
    if constexpr (kTwoStage) {
      // T_lo: the lowest threshold of the block's rows (lanes < kRows stage
      // them; rows that can never match do not count)
      const bool mine = lane < kRows && r.in == r.in;
      if constexpr (kIntCol) {
        tlo_i = rows_min(mine ? enc_thr(r.t) : kNanEnc);
        // the entering column's first-stage bound travels in its record's t
        c.t = stage1_bound(tlo_i, c.F, stage1_factor(c.in));
        r.s1 = r.in == r.in ? r.in : 0.0;
        s_row[lane] = r;
        s_col[lane] = c;
      } else {
        tlo_f = __int_as_float(rows_min(__float_as_int(mine ? (float)r.t : __builtin_inff())));
        c.b = stage1_bound32(tlo_f, c.t, c.in);
        r.s1 = r.in == r.in ? r.in : 0.0f;
        s_row[lane] = r;
        s_col[lane] = c;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  // The smallest v of lanes [0, kRows) (kRows a multiple of 16), uniform:
  // DPP rotations within each row of 16 lanes (row_ror 8, 4, 2, 1), then one
  // lane read per row; no LDS round trips.
  __device__ __forceinline__ static int32_t rows_min(int32_t v) {
    static_assert(kRows % 16 == 0, "whole DPP rows");
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x122, 0xf, 0xf, false));
    v = min(v, __builtin_amdgcn_update_dpp(v, v, 0x121, 0xf, 0xf, false));
    int32_t t = __builtin_amdgcn_readlane(v, 0);
#pragma unroll
    for (int row = 1; row < kRows / 16; ++row) {
      const int32_t u = __builtin_amdgcn_readlane(v, 16 * row);
      t = u < t ? u : t;
    }
    return t;
  }

  // ---------------------------------------------------------------- seeds
  // Exact-in-FP64 covariance of this lane's 8 diagonals at row i: the seed
  // of every segment (both forms) and the FP32 form's periodic re-anchor,
  //   cov_k = sum_t (X[i+d_k+t] - ref) * yc_t - (muX[i+d_k] - ref) * sum_t yc_t,
  //   yc_t = Y[i+t] - muY[i],  ref = muX[i + D0]
  // (sum_t (X - muX) yc = sum_t (X - ref) yc - (muX - ref) sum_t yc; the
  // reference level keeps the products near the series' local scale, so
  // the FP64 accumulation does not cancel a large offset).  The wave stages
  // the X samples of its 512 diagonals and the centred Y window through LDS
  // in chunks (coalesced loads, one pass), then each lane slides a register
  // window over them so that each X sample feeds 8 FMAs: 8m FMAs per lane.
  __device__ __forceinline__ void anchor(int64_t i, double *s_ax,
                                         double *s_ay) {
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    const double my = a.muY[i];
    const double ref = a.muX[i + D0];
    const int64_t m = a.m;
    const int64_t uend = m + K - 1;  // u = t + k runs over [0, m + K - 1)
    const int xoff = lane * K;       // d0 - D0
    double sy = 0.0;
    double yw[2 * K - 1];            // yc_t for t = u0-K+1 .. u0+K-1
#pragma unroll
    for (int q = 0; q < 2 * K - 1; ++q) yw[q] = 0.0;
    constexpr int kChunk = sizeof(DT) == 8 ? kAnchorChunk64 : kAnchorChunk;
    for (int64_t c0 = 0; c0 < uend; c0 += kChunk) {
      __builtin_amdgcn_wave_barrier();
      const double *xs = a.TX + i + D0 + c0;
      for (int j = lane; j < kBand + kChunk; j += kWave) s_ax[j] = xs[j] - ref;
      for (int j = lane; j < kChunk; j += kWave) {
        const int64_t t = c0 + j;
        s_ay[j] = t < m ? a.TY[i + t] - my : 0.0;
      }
      __builtin_amdgcn_wave_barrier();
      const int64_t cend = uend - c0 < kChunk ? uend - c0 : kChunk;
      for (int u = 0; u < cend; u += K) {
#pragma unroll
        for (int q = 0; q < K - 1; ++q) yw[q] = yw[q + K];
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const double v = s_ay[u + q];
          yw[K - 1 + q] = v;
          sy += v;
        }
        double xv[K];
#pragma unroll
        for (int q = 0; q < K; ++q) xv[q] = s_ax[xoff + u + q];
#pragma unroll
        for (int q = 0; q < K; ++q)
#pragma unroll
          for (int k = 0; k < K; ++k)
            acc[k] = fma(xv[q], yw[q - k + K - 1], acc[k]);
      }
    }
    __builtin_amdgcn_wave_barrier();
    // the FRESH row and column tables share s_ay and s_ax
    if (kTable && lane < kRows) s_rt[lane] = 0;
    if (kColTable)
      for (int q = lane; q < K * kColStride; q += kWave) s_ct[q] = 0;
    if (kFarDiag) vmask = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t d = d0 + k;
      const double v = acc[k] - (a.muX[i + d] - ref) * sy;
      if constexpr (kFarDiag) vmask |= (diag_valid(d) ? 1u : 0u) << k;
      if constexpr (kComp && kFarDiag) {
        const DT h = (DT)v;
        hi[k] = diag_valid(d) ? h : far_cov<DT>();
        cov[k] = diag_valid(d) ? (DT)(v - (double)h) : (DT)0;
      } else if constexpr (kComp) {
        const DT h = (DT)v;
        hi[k] = diag_valid(d) ? h : nan_of<DT>();
        cov[k] = (DT)(v - (double)h);
      } else if constexpr (kFarDiag) {
        cov[k] = diag_valid(d) ? (DT)v : far_cov<DT>();
      } else {
        cov[k] = diag_valid(d) ? (DT)v : nan_of<DT>();
      }
    }
  }

  // FP32 form, every kRenormRows rows: folds the partial sum into the
  // leading part; the rounding error of that addition (exact by Fast2Sum
  // while |hi| >= |partial|) seeds the next partial sum.
  __device__ __forceinline__ void renormalise() {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const DT s = hi[k] + cov[k];
      cov[k] = cov[k] - (s - hi[k]);
      hi[k] = s;
    }
  }

  // ------------------------------------------------------ matrix summary
  __device__ __forceinline__ void cell_flush(double v, int64_t r, int64_t c) {
    if (v > a.thr) {
      const uint32_t e = enc_corr((float)v);
      atomicMax(&a.mat[r * a.mat_w + c], e);
    }
  }
  __device__ __forceinline__ void cells_begin(int64_t i) {
    const int64_t col0 = i + d0;
    if (COLS) {  // direct cells (row i / rpc, column / cpc)
      dir.lo = dir.hi = nan_of<double>();
      dir.lo_cell = col0 / a.cpc;
      dir.pos = col0 - dir.lo_cell * a.cpc;
      dir_row_cell = i / a.rpc;
      dir_row_pos = i - dir_row_cell * a.rpc;
    }
    if (ROWS) {  // transposed cells (column / rpc, row i / cpc)
      tra.lo = tra.hi = nan_of<double>();
      tra.lo_cell = col0 / a.rpc;
      tra.pos = col0 - tra.lo_cell * a.rpc;
      tra_col_cell = i / a.cpc;
      tra_col_pos = i - tra_col_cell * a.cpc;
    }
  }
  __device__ __forceinline__ void cells_flush_all() {
    if (COLS) {
      cell_flush(dir.lo, dir_row_cell, dir.lo_cell);
      cell_flush(dir.hi, dir_row_cell, dir.lo_cell + 1);
      dir.lo = dir.hi = nan_of<double>();
    }
    if (ROWS) {
      cell_flush(tra.lo, tra.lo_cell, tra_col_cell);
      cell_flush(tra.hi, tra.lo_cell + 1, tra_col_cell);
      tra.lo = tra.hi = nan_of<double>();
    }
  }
  // Before row i: a new row cell (direct) / column cell (transposed) starts.
  __device__ __forceinline__ void cells_row_start(bool first) {
    if (COLS && dir_row_pos == 0 && !first) {
      cell_flush(dir.lo, dir_row_cell, dir.lo_cell);
      cell_flush(dir.hi, dir_row_cell, dir.lo_cell + 1);
      dir.lo = dir.hi = nan_of<double>();
      dir_row_cell += 1;
    }
    if (ROWS && tra_col_pos == 0 && !first) {
      cell_flush(tra.lo, tra.lo_cell, tra_col_cell);
      cell_flush(tra.hi, tra.lo_cell + 1, tra_col_cell);
      tra.lo = tra.hi = nan_of<double>();
      tra_col_cell += 1;
    }
  }
  __device__ __forceinline__ void cells_accumulate(const DT (&c)[K]) {
    if (COLS) {
      const int64_t kb = a.cpc - dir.pos;  // first k in the next cell
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double v = (double)c[k];
        const bool in_lo = k < kb;
        dir.lo = fmax(dir.lo, in_lo ? v : nan_of<double>());
        dir.hi = fmax(dir.hi, in_lo ? nan_of<double>() : v);
      }
    }
    if (ROWS) {
      const int64_t kb = a.rpc - tra.pos;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const double v = (double)c[k];
        const bool in_lo = k < kb;
        tra.lo = fmax(tra.lo, in_lo ? v : nan_of<double>());
        tra.hi = fmax(tra.hi, in_lo ? nan_of<double>() : v);
      }
    }
  }
  // After row i: the lane's columns move right by one.
  __device__ __forceinline__ void cells_row_end() {
    if (COLS) {
      if (++dir.pos == a.cpc) {
        cell_flush(dir.lo, dir_row_cell, dir.lo_cell);
        dir.lo = dir.hi;
        dir.hi = nan_of<double>();
        dir.lo_cell += 1;
        dir.pos = 0;
      }
      if (++dir_row_pos == a.rpc) dir_row_pos = 0;
    }
    if (ROWS) {
      if (++tra.pos == a.rpc) {
        cell_flush(tra.lo, tra.lo_cell, tra_col_cell);
        tra.lo = tra.hi;
        tra.hi = nan_of<double>();
        tra.lo_cell += 1;
        tra.pos = 0;
      }
      if (++tra_col_pos == a.cpc) tra_col_pos = 0;
    }
  }

  // Fast blocks: when no cell boundary falls inside any lane's part of the
  // block's kRows rows (columns first .. first + kRows + K - 2 of the lane,
  // rows i_b .. i_b + kRows - 1), every pair of a lane goes to the lane's current cell, so
  // the block keeps one running max per lane and folds it in at the end.
  // Wave-uniform decision; blocks that touch a boundary take the per-pair
  // path above.
  __device__ __forceinline__ bool cells_block_fast() const {
    if (COLS && ROWS) return false;
    bool ok;
    if (COLS)
      ok = dir.pos + kRows + K - 2 < a.cpc && dir_row_pos + kRows <= a.rpc;
    else
      ok = tra.pos + kRows + K - 2 < a.rpc && tra_col_pos + kRows <= a.cpc;
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
  }
  // A fast block needs only the largest correlation of the lane's row:
  // rounding is monotone, so for the row's inverse norm rin > 0,
  // max_k fl(x_k * rin) = fl(max_k x_k * rin) with x_k = cov_k * in_col --
  // one multiply by rin per row instead of one per pair, the same bits (a
  // NaN x_k or rin is ignored by fmax either way).
  __device__ __forceinline__ void cells_fast_accumulate(const DT (&x)[K], DT rin) {
    const DT m0 = fmax(fmax(x[0], x[1]), fmax(x[2], x[3]));
    const DT m1 = fmax(fmax(x[4], x[5]), fmax(x[6], x[7]));
    blk0 = fmax(blk0, fmax(m0, m1) * rin);
  }
  __device__ __forceinline__ void cells_fast_end() {
    const double v = (double)fmax(blk0, blk1);
    if (COLS) {
      dir.lo = fmax(dir.lo, v);
      dir.pos += kRows;
      dir_row_pos += kRows;
      if (dir_row_pos == a.rpc) dir_row_pos = 0;
    } else {
      tra.lo = fmax(tra.lo, v);
      tra.pos += kRows;
      tra_col_pos += kRows;
      if (tra_col_pos == a.cpc) tra_col_pos = 0;
    }
  }

  // -------------------------------------------------- 1NN slow path
  // Row rr of the block (row i): the wave's best pair into the row's entry
  // (FRESH: each lane's best into the block's LDS row table instead).
  __device__ __forceinline__ void nn_row_update(const DT (&c)[K], int64_t i,
                                                int rr, DT trow) {
    float bv = -3.0f;
    int bk = -1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float f = (float)c[k];
      if (f > bv) {
        bv = f;
        bk = k;
      }
    }
    uint64_t key = 0;
    if (bk >= 0 && (DT)bv >= trow)
      key = make_key(bv, static_cast<uint32_t>(i + d0 + bk));
    if (kTable) {
      if (key != 0)
        atomicMax(reinterpret_cast<unsigned long long *>(s_rt + rr),
                  static_cast<unsigned long long>(key));
      return;
    }
    key = wave_max_u64(key);
    if (lane == 0 && key != 0)
      atomicMax(reinterpret_cast<unsigned long long *>(a.keyY + i),
                static_cast<unsigned long long>(key));
  }
  // Row i's pairs that beat their column's best, into the columns' entries
  // (FRESH: into the block's column table).
  __device__ __forceinline__ bool col_passes(DT c, const Slot<DT> &s) const {
    // the FP64 two-stage form tests the integer image (conservative; the
    // NaN of an invalid diagonal passes it and is caught here).  The high
    // word of a negative double orders the other way round as an integer,
    // so a column whose threshold is negative (kNegEnc) takes every
    // correlation, negative ones included.  Non-short-circuit operators: no
    // branch (an exec-mask change) per test.
    if constexpr (kIntCol)
      return (c == c) & ((__double2hiint(c) >= s.F) | (s.F == kNegEnc));
    else
      return c >= s.t;
  }
  __device__ __forceinline__ void nn_col_update(int r, int rr, const DT (&c)[K],
                                                int64_t i) {
    if (kColTable) ct_dirty = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      Slot<DT> &s = S[(r + k) % K];
      if (col_passes(c[k], s)) {
        const float f = (float)c[k];
        const uint64_t key = make_key(f, static_cast<uint32_t>(i));
        if (kColTable)
          atomicMax(reinterpret_cast<unsigned long long *>(
                        s_ct + ((rr + k) % K) * kColStride + (rr + k) / K + lane),
                    static_cast<unsigned long long>(key));
        else
          atomicMax(reinterpret_cast<unsigned long long *>(a.keyX + i + d0 + k),
                    static_cast<unsigned long long>(key));
        // the column's best is now at least f (no need to wait for the
        // atomic's return value); the tighter threshold travels with the
        // column to the next lanes
        const DT tf = Thresh<DT>::from_corr(f);
        if constexpr (kIntCol) {
          const int32_t ef = enc_thr(tf);
          s.F = s.F > ef ? s.F : ef;
        } else if constexpr (kTwoStage) {
          const DT tc = canon_thr(tf);
          s.t = s.t > tc ? s.t : tc;
        } else {
          s.t = s.t > tf ? s.t : tf;
        }
      }
    }
  }

  // FRESH: the block's row table (rows i_b + l, l < kRows) into the profile,
  // and cleared for the next block.
  __device__ __forceinline__ void flush_row_table(int64_t i_b) {
    __builtin_amdgcn_wave_barrier();
    if (lane < kRows) {
      const uint64_t v = s_rt[lane];
      if (v != 0) {
        atomicMax(reinterpret_cast<unsigned long long *>(a.keyY + i_b + lane),
                  static_cast<unsigned long long>(v));
        s_rt[lane] = 0;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }

  // FRESH: the block's column table into the profile, and cleared.
  __device__ __forceinline__ void flush_col_table(int64_t i_b) {
    __builtin_amdgcn_wave_barrier();
    for (int p = lane; p < kBand + kRows - 1; p += kWave) {
      uint64_t *e = s_ct + (p % K) * kColStride + p / K;
      const uint64_t v = *e;
      if (v != 0) {
        atomicMax(reinterpret_cast<unsigned long long *>(a.keyX + i_b + D0 + p),
                  static_cast<unsigned long long>(v));
        *e = 0;
      }
    }
    ct_dirty = false;
    __builtin_amdgcn_wave_barrier();
  }

  // ------------------------------------------------- ALL_NEIGHBORS lists
  // Inserts `key` into the k-entry list at `list` if it beats the list's
  // minimum, by replacing that minimum with a compare-and-swap (retried
  // when another insert changed the list first; every retry follows
  // someone else's successful insert, so the loop ends).  Returns a lower
  // bound of the list's minimum after the call.  Entries are read with
  // device-coherent atomic loads so that a retry sees the newest values.
  __device__ __forceinline__ static uint64_t knn_insert(uint64_t *list, int kk,
                                                        uint64_t key) {
    for (;;) {
      uint64_t mn = ~0ull, second = ~0ull;
      int slot = 0;
      bool present = false;
      for (int t = 0; t < kk; ++t) {
        const uint64_t v = __hip_atomic_load(list + t, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        present |= v == key;  // a pair already in the list (from the prior)
        if (v < mn) {
          second = mn;
          mn = v;
          slot = t;
        } else if (v < second) {
          second = v;
        }
      }
      if (key <= mn || present) return mn;
      const unsigned long long old = atomicCAS(
          reinterpret_cast<unsigned long long *>(list + slot),
          static_cast<unsigned long long>(mn), static_cast<unsigned long long>(key));
      if (old == mn) return key < second ? key : second;
    }
  }

  // ALL_NEIGHBORS slow path of one row: the row's list takes the wave's best
  // candidates (at most k, found by repeated wave maxima so that one lane
  // inserts them); each column's list takes the lane's candidate for it.
  __device__ __forceinline__ void knn_update(int r, const DT (&c)[K], int64_t i,
                                             DT trow) {
    const int kk = a.knn_k;
    if (ROWS) {
      uint32_t taken = 0;
      uint64_t *list = a.knnY + i * kk;
      for (int q = 0; q < kk; ++q) {
        uint64_t best = 0;
        int bk = -1;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!((taken >> k) & 1u) && c[k] >= trow) {
            const uint64_t key =
                make_key((float)c[k], static_cast<uint32_t>(i + d0 + k));
            if (key > best) {
              best = key;
              bk = k;
            }
          }
        }
        const uint64_t wbest = wave_max_u64(best);
        if (wbest == 0) break;
        if (bk >= 0 && best == wbest) taken |= 1u << bk;  // keys are unique
        uint64_t nm = 0;
        if (lane == 0) {
          nm = knn_insert(list, kk, wbest);
          atomicMax(reinterpret_cast<unsigned long long *>(a.keyY + i),
                    static_cast<unsigned long long>(nm));
        }
        // keys in one list are unique, so a candidate that did not go in
        // left a minimum above it, and so would every smaller one
        nm = __shfl(nm, 0);
        if (nm > wbest) break;
      }
    }
    if (COLS) {
      // The lane's candidates go to its own LDS slots first, so that one
      // insert loop (over the set bits) serves all K diagonals: the insert
      // code then exists once per row rather than K times, which keeps the
      // kernel's register count down.  The slots then carry each list's new
      // minimum back to the columns' thresholds.
      uint64_t *q = s_kq + lane * K;
      uint32_t mask = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const bool h = c[k] >= S[(r + k) % K].t;
        q[k] = h ? make_key((float)c[k], static_cast<uint32_t>(i)) : 0;
        mask |= (h ? 1u : 0u) << k;
      }
      for (uint32_t mk = mask; mk != 0; mk &= mk - 1) {
        const int k = __builtin_ctz(mk);
        const int64_t j = i + d0 + k;
        const uint64_t nm = knn_insert(a.knnX + j * kk, kk, q[k]);
        atomicMax(reinterpret_cast<unsigned long long *>(a.keyX + j),
                  static_cast<unsigned long long>(nm));
        q[k] = nm;
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if ((mask >> k) & 1u) {
          Slot<DT> &s = S[(r + k) % K];
          const DT tn = Thresh<DT>::from_key_hi(static_cast<uint32_t>(q[k] >> 32));
          s.t = s.t > tn ? s.t : tn;
        }
      }
    }
  }

  // ------------------------------------------------------------ one row
  // Row i = i_b + r of a block whose first row satisfies (i_b + D0) % K == 0,
  // so diagonal k of every lane crosses the column in slot (r + k) % K.
  // r: slot cycle position (i - i_b) mod K; rr: row of the block (i - i_b).
  // Row i of a two-stage 1NN kernel (see kTwoStage); rs1 is the row's
  // first-stage scale (Terms::s1).
  template <bool MASKED>
  __device__ __forceinline__ void row_two_stage(int r, int rr, int64_t i, DT rin,
                                                DT rs1, DT trow) {
    // first stage: some d_k = cov_k * in_row - B_k >= 0, by the sign of the
    // largest integer image of the d_k (a max3 tree)
    int32_t h[K];
#pragma unroll
    for (int k = 0; k < K; ++k)
      h[k] = sign_word(fma(kComp ? hi[k] + cov[k] : cov[k], rs1, -S[(r + k) % K].mt));
    static_assert(K == 8, "the max tree below takes 8 differences");
    const int32_t m0 = max(max(h[0], h[1]), h[2]);
    const int32_t m1 = max(max(h[3], h[4]), h[5]);
    const int32_t mx = max(max(h[6], h[7]), max(m0, m1));
    // (most rows stop here: the exact stage is laid out off the row's path)
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(mx >= 0) == 0, 1)) return;
    // exact stage: the correlations (cov * in_row) * in_col, NaN where the
    // diagonal is outside the pass or the column past the series.  FP64: the
    // columns' inverse norms come from memory (rows get here rarely; the
    // columns i + d0 + k lie below nX + kBand, inside the padded array).
    DT inc[K];
    const uint32_t coff = static_cast<uint32_t>(i + d0) * sizeof(DT);
#pragma unroll
    for (int k = 0; k < K; ++k)
      inc[k] = kIntCol ? ld<DT>(a.inX, coff + k * (uint32_t)sizeof(DT)) : S[(r + k) % K].in;
    DT c[K];
    uint64_t hit = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      c[k] = (kComp ? hi[k] + cov[k] : cov[k]) * rin * inc[k];
      if (!((vmask >> k) & 1u) || (MASKED && i + d0 + k >= a.nX)) c[k] = nan_of<DT>();
      if constexpr (kIntCol) {
        // col_passes as the masks of single compares (no round trip of a
        // boolean through a register)
        const int32_t F = S[(r + k) % K].F;
        hit |= __builtin_amdgcn_ballot_w64(c[k] == c[k]) &
               (__builtin_amdgcn_ballot_w64(__double2hiint(c[k]) >= F) |
                __builtin_amdgcn_ballot_w64(F == kNegEnc));
      } else {
        hit |= __builtin_amdgcn_ballot_w64(col_passes(c[k], S[(r + k) % K]));
      }
      hit |= __builtin_amdgcn_ballot_w64(c[k] >= trow);
    }
    if (hit) {
      nn_row_update(c, i, rr, trow);
      nn_col_update(r, rr, c, i);
    }
  }

  // Row i of every other kernel: the correlations, then the kind's tests.
  template <bool MASKED, bool FAST>
  __device__ __forceinline__ void row_tests(int r, int rr, int64_t i, DT rin,
                                            DT trow, DT rs1) {
    DT c[K];
    if constexpr (kSumStage) {
      // first stage: some d_k = cov_k * in_row - thr / in_col >= 0
      int32_t h[K];
#pragma unroll
      for (int k = 0; k < K; ++k)
        h[k] = sign_word(fma(kComp ? hi[k] + cov[k] : cov[k], rs1, -S[(r + k) % K].mt));
      static_assert(K == 8, "the max tree below takes 8 differences");
      const int32_t m0 = max(max(h[0], h[1]), h[2]);
      const int32_t m1 = max(max(h[3], h[4]), h[5]);
      const int32_t mx = max(max(h[6], h[7]), max(m0, m1));
      if (__builtin_expect(__builtin_amdgcn_ballot_w64(mx >= 0) == 0, 1)) return;
      // the exact sums, over the correlations (cov * in_col) * in_row
      double rs = 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        DT cv = (kComp ? hi[k] + cov[k] : cov[k]) * S[(r + k) % K].in * rin;
        if (!((vmask >> k) & 1u) || (MASKED && i + d0 + k >= a.nX)) cv = nan_of<DT>();
        const double v = cv > (DT)a.thr ? (double)cv : 0.0;
        if (ROWS) rs += v;
        if (COLS) S[(r + k) % K].acc += v;
      }
      if (ROWS) rsum[r] = rs;
      return;
    }
    if constexpr (KIND == kMatrix && FAST) {
      static_assert(K == 8, "the max tree takes 8 correlations");
#pragma unroll
      for (int k = 0; k < K; ++k)
        c[k] = (kComp ? hi[k] + cov[k] : cov[k]) * S[(r + k) % K].in;
      cells_fast_accumulate(c, rin);
      return;
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const DT cv = kComp ? hi[k] + cov[k] : cov[k];
      c[k] = cv * S[(r + k) % K].in * rin;
      if (MASKED && i + d0 + k >= a.nX) c[k] = nan_of<DT>();
    }
    if (kThr && kTable) {
      // FRESH: the row and column masks are kept apart, so that a row where
      // only one side passes runs only that side's merge.
      uint64_t hit_c = 0, hit_r = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (COLS) hit_c |= __builtin_amdgcn_ballot_w64(c[k] >= S[(r + k) % K].t);
        if (ROWS) hit_r |= __builtin_amdgcn_ballot_w64(c[k] >= trow);
      }
      if (ROWS && hit_r) nn_row_update(c, i, rr, trow);
      if (COLS && hit_c) nn_col_update(r, rr, c, i);
    } else if (kThr) {
      // One compare per test written straight into a wave mask (keeps the
      // compiler from merging the two tests into a NaN-canonicalising min).
      uint64_t hit = 0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (COLS) hit |= __builtin_amdgcn_ballot_w64(c[k] >= S[(r + k) % K].t);
        if (ROWS) hit |= __builtin_amdgcn_ballot_w64(c[k] >= trow);
      }
      if (KIND == kKnn && hit) {
        knn_update(r, c, i, trow);
      } else if (__builtin_expect(hit != 0, 0)) {  // merges off the row's path
        if (ROWS) nn_row_update(c, i, rr, trow);
        if (COLS) nn_col_update(r, rr, c, i);
      }
    } else if (KIND == kSum) {
      // a row with no correlation above the threshold anywhere in the wave
      // costs one compare per pair
      uint64_t hit = 0;
#pragma unroll
      for (int k = 0; k < K; ++k)
        hit |= __builtin_amdgcn_ballot_w64(c[k] > (DT)a.thr);
      if (hit) {
        double rs = 0.0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double v = c[k] > (DT)a.thr ? (double)c[k] : 0.0;
          if (ROWS) rs += v;
          if (COLS) S[(r + k) % K].acc += v;
        }
        if (ROWS) rsum[r] = rs;
      }
    } else if (KIND == kMatrix) {
      cells_accumulate(c);
    } else {  // kMatrixSmall: cells narrower than a lane's 8 columns
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if ((double)c[k] > a.thr) {
          const int64_t j = i + d0 + k;
          const uint32_t e = enc_corr((float)c[k]);
          if (COLS) atomicMax(&a.mat[(i / a.rpc) * a.mat_w + j / a.cpc], e);
          if (ROWS) atomicMax(&a.mat[(j / a.rpc) * a.mat_w + i / a.cpc], e);
        }
      }
    }
  }

  template <bool MASKED, bool FAST>
  __device__ __forceinline__ void row(int r, int rr, int64_t i,
                                      const Terms<DT> *s_row, DT &s1_next) {
    const Terms<DT> rv = s_row[rr];
    const DT rdf = rv.df, rdg = rv.dg, rin = rv.in, trow = rv.t;
    if constexpr (kIntCol) {
      // the FP64 first stage's scale is read a row ahead, so that the row's
      // first FMAs need not wait for LDS (measured 0.7 - 0.9 % faster at
      // n = 2^22, m = 1024; the FP32 form was 1.5 % slower with it)
      const DT s1 = s1_next;
      if (rr + 1 < kRows) s1_next = s_row[rr + 1].s1;
      row_two_stage<MASKED>(r, rr, i, rin, s1, trow);
    } else if constexpr (kTwoStage) {
      row_two_stage<MASKED>(r, rr, i, rin, rv.s1, trow);
    } else if constexpr (kSumStage) {
      row_tests<MASKED, FAST>(r, rr, i, rin, trow, rv.s1);
    } else {
      row_tests<MASKED, FAST>(r, rr, i, rin, trow, rin);
    }
    // advance the recurrence to row i+1
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Slot<DT> &s = S[(r + k) % K];
      cov[k] = fma(s.df, rdg, cov[k]);
      cov[k] = fma(rdf, s.dg, cov[k]);
    }
    if (KIND == kMatrix && !FAST) cells_row_end();
  }

  // Slot r now holds the column that leaves this lane (column i + d0): shift
  // every lane's slot r one lane down.  Lane 0's column leaves the band (its
  // partial sum is flushed); lane 63 takes the entering column.
  __device__ __forceinline__ void shift(int r, int rr, int64_t i,
                                        const Terms<DT> *s_col) {
    Slot<DT> &s = S[r];
    const Terms<DT> e = s_col[rr];
    if (KIND == kSum && COLS) {
      if (lane == 0 && s.acc != 0.0) atomicAdd(a.sumX + i + D0, s.acc);
      s.acc = shl1(s.acc, 0.0);
    }
    s.df = shl1(s.df, e.df);
    s.dg = shl1(s.dg, e.dg);
    if constexpr (!kIntCol) s.in = shl1(s.in, e.in);
    if constexpr (kIntCol) {
      s.F = shl1(s.F, e.F);
      s.mt = shl1(s.mt, e.t);
    } else if constexpr (kSumStage) {
      s.mt = shl1(s.mt, e.t);
    } else if (kThr && COLS) {
      s.t = shl1(s.t, e.t);
      if constexpr (kTwoStage) s.mt = shl1(s.mt, e.b);
    }
  }

  // Flushes what a lane still holds when its segment ends after row i.
  __device__ __forceinline__ void segment_end(int r, int64_t i) {
    if (KIND == kSum && COLS) {
      // slot j holds diagonal (j - r) mod K (indexing by slot keeps the
      // slot array in registers)
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const int64_t k = (j - r + K) % K;
        if (S[j].acc != 0.0) atomicAdd(a.sumX + i + d0 + k, S[j].acc);
      }
    }
    if (KIND == kMatrix) cells_flush_all();
  }

  // SUM_THRESH rows: adds the wave-wide total of rsum[r] to row i_b + r for
  // r < nrows.  Reduce-scatter butterfly: after the xor-32/16/8 exchanges
  // lane L holds a partial of row L >> 3 summed over 8 lanes, and the
  // xor-4/2/1 steps complete it; lanes 0, 8, ..., 56 then add one row each.
  __device__ __forceinline__ void flush_row_sums(int64_t i_b, int64_t nrows) {
    static_assert(K == 8, "the butterfly below reduces 8 rows");
    const bool b5 = lane & 32, b4 = lane & 16, b3 = lane & 8;
    double v4[4], v2[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const double keep = b5 ? rsum[4 + j] : rsum[j];
      const double send = b5 ? rsum[j] : rsum[4 + j];
      v4[j] = keep + __shfl_xor(send, 32);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double keep = b4 ? v4[2 + j] : v4[j];
      const double send = b4 ? v4[j] : v4[2 + j];
      v2[j] = keep + __shfl_xor(send, 16);
    }
    double v = (b3 ? v2[1] : v2[0]) + __shfl_xor(b3 ? v2[0] : v2[1], 8);
    v += __shfl_xor(v, 4);
    v += __shfl_xor(v, 2);
    v += __shfl_xor(v, 1);
    const int row = lane >> 3;
    if ((lane & 7) == 0 && row < nrows && v != 0.0) atomicAdd(a.sumY + i_b + row, v);
  }

  // One block of kRows rows starting at i_b.  `pre` holds this block's staged
  // terms on entry and the next block's on exit.
  template <bool MASKED>
  __device__ __forceinline__ bool block(int64_t i_b, int64_t re, Pre &pre,
                                        Terms<DT> *s_row, Terms<DT> *s_col,
                                        bool first) {
    if (KIND == kMatrix && !MASKED && __builtin_expect(cells_block_fast(), 1))  // most blocks
      return block_rows<false, true>(i_b, re, pre, s_row, s_col, first);
    return block_rows<MASKED, false>(i_b, re, pre, s_row, s_col, first);
  }

  template <bool MASKED, bool FAST>
  __device__ __forceinline__ bool block_rows(int64_t i_b, int64_t re, Pre &pre,
                                             Terms<DT> *s_row, Terms<DT> *s_col,
                                             bool first) {
    publish(pre, s_row, s_col);
    if constexpr (kTwoStage) {
#pragma unroll
      for (int k = 0; k < K; ++k) refresh_mt(S[k], pre.rq[kIntCol ? k : 0]);
    }
    prefetch(pre, i_b + kRows);
    // keep the look-ahead loads where they are issued: the scheduler would
    // otherwise sink them next to their first use a block later
    __builtin_amdgcn_sched_barrier(0);
    if (KIND == kSum && ROWS) {
#pragma unroll
      for (int r = 0; r < K; ++r) rsum[r] = 0.0;
    }
    if (KIND == kMatrix && FAST) {
      cells_row_start(first);
      blk0 = blk1 = nan_of<DT>();
    }
    bool ended = false;
    int64_t fl = i_b;  // SUM_THRESH rows: the first row whose sum is not flushed
    DT s1_next = kIntCol ? s_row[0].s1 : (DT)0;  // the first row's (row() reads ahead)
#pragma clang loop unroll(full)
    for (int rr = 0; rr < kRows; ++rr) {
      const int r = rr % K;
      const int64_t i = i_b + rr;
      if (MASKED && i >= re) {
        ended = true;
        break;
      }
      if (KIND == kMatrix && !FAST) cells_row_start(first && rr == 0);
      row<MASKED, FAST>(r, rr, i, s_row, s1_next);
      if (MASKED && i + 1 == re) {
        segment_end(r, i);
        ended = true;
        break;
      }
      shift(r, rr, i, s_col);
      if (kComp && rr % kRenormRows == kRenormRows - 1) renormalise();
      // SUM_THRESH rows: rsum holds one slot cycle (K rows); flush it at
      // each cycle's end within the block (the last cycle's below)
      if (KIND == kSum && ROWS && r == K - 1 && rr + 1 < kRows) {
        flush_row_sums(fl, K);
#pragma unroll
        for (int q = 0; q < K; ++q) rsum[q] = 0.0;
        fl += K;
      }
    }
    if (KIND == kSum && ROWS) flush_row_sums(fl, re - fl);
    if (kTable && ROWS) flush_row_table(i_b);
    if (kColTable && ct_dirty) flush_col_table(i_b);
    if (KIND == kMatrix && FAST) cells_fast_end();
    if (ended) return true;
    __builtin_amdgcn_wave_barrier();
    return false;
  }

  // Rows [rs, re) of the band (rs % K == 0); all band diagonals are valid
  // for rows < rfull.
  __device__ __forceinline__ void segment(int64_t rs, int64_t re,
                                          int64_t rfull, Terms<DT> *s_row,
                                          Terms<DT> *s_col, double *s_ax,
                                          double *s_ay) {
    if (rs >= re) return;
    Pre pre;
    prefetch(pre, rs);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = 0; k < K; ++k) load_slot(S[k], rs + d0 + k);
    if (KIND == kMatrix) cells_begin(rs);
    int64_t i = rs;
    const int64_t fast_end = re < rfull ? re : rfull;
    bool first = true;
    // Fast region; the last row always falls in the masked tail below,
    // which flushes.
    int64_t to_anchor = a.anchor;
    while (i + kRows < re && i + kRows <= fast_end) {
      if (sizeof(DT) == 4) {
        if (to_anchor <= 0) {
          anchor(i, s_ax, s_ay);
          to_anchor = a.anchor;
        }
        to_anchor -= kRows;
      }
      block<false>(i, re, pre, s_row, s_col, first);
      first = false;
      i += kRows;
    }
    for (;;) {
      if (block<true>(i, re, pre, s_row, s_col, first)) return;
      first = false;
      i += kRows;
    }
  }
};

// Occupancy targets of the 1NN kinds: 3 waves per SIMD (<= 168 VGPRs), and
// 4 (<= 128) for the FP32 short-chunk kernel.  The compiler would otherwise
// use more registers for look-ahead and drop a wave; the few spills the
// targets cost are cheaper than the lost wave (measured at n = 2^22,
// m = 1024: FP64 1465 ms at 3 waves, 1505 at 2; FP32 two-stage 1192 ms at 3,
// 1400 at 4, 1620 at 2; the FP32 short-chunk kernel, n = 2^17, m = 100:
// 2.11 ms at 4, 2.18 at 3).
// ALL_NEIGHBORS in FP32: 3 waves (12 % faster than the 2 the compiler
// picks); in FP64 the same target spills heavily (4x slower), so none.
// SUM_THRESH (both forms) and FP32 MATRIX_SUMMARY: 3 waves instead of the
// compiler's 2 (measured on a 2^21 self-join, m = 1024 / 256: SUM 615 ->
// 462 ms mixed, 612 -> 508 ms FP64; summary 526 -> 440 ms mixed); the
// FP64 summary spills in its row loop at 3 waves (2.5x slower), so none.
template <typename DT, int KIND, bool FRESH>
constexpr int sweep_waves_per_eu() {
  if (KIND == kNN) return sizeof(DT) == 4 && FRESH ? 4 : 3;
  if (KIND == kKnn) return sizeof(DT) == 4 ? 3 : 1;
  if (KIND == kSum) return 3;
  if (KIND == kMatrix) return sizeof(DT) == 4 ? 3 : 1;
  return 1;
}

// The rows one work item sweeps: [rs, re) of a band, then [rs2, re2)
// (empty when re2 <= rs2).
struct ItemRows {
  int64_t band, series, rs, re, rs2, re2;
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// Work item idx of a pass, decoded on the device (no host-built list).
// Items run band-major (band b of every series, then band b + 1), so the low
// diagonals, where a series' best matches mostly lie, go first.  A band
// spans rows [0, rows0) and starts at a hashed row r0 (a multiple of 8 below
// rfull), so that bands in flight sit at unrelated rows' profile entries:
//  * unchunked (nchunks == 1): one item, rows [r0, rows0) then [0, r0);
//  * chunked: the rows are cut at r0 + j R (R from band_chunk_rows: shorter
//    bands, which run later, take shorter chunks), and the band's item j is
//    its j-th chunk counting from the one that starts at r0 (wrapping round
//    to row 0); items beyond the band's chunk count are empty.
// Returns false for an empty item.
__device__ __forceinline__ bool plan_item(const SweepArgs &a, unsigned int idx,
                                          ItemRows &g) {
  const uint64_t per_band = (uint64_t)a.nbatch * (uint64_t)a.nchunks;
  g.band = (int64_t)(idx / per_band);
  const int64_t rem = (int64_t)(idx - (uint64_t)g.band * per_band);
  g.series = rem / a.nchunks;
  const int64_t j = rem - g.series * a.nchunks;
  const int64_t D0 = a.d_base + g.band * kBand;
  const int64_t dfirst = D0 > a.d_lo ? D0 : a.d_lo;
  const int64_t dlast = (D0 + kBand < a.d_hi ? D0 + kBand : a.d_hi) - 1;
  const int64_t rows0 = a.nY < a.nX - dfirst ? a.nY : a.nX - dfirst;
  const int64_t rfull = a.nY < a.nX - dlast ? a.nY : a.nX - dlast;
  if (rows0 <= 0 || rfull <= 0) return false;
  const int64_t slots = (rfull - 1) / kDiagsPerLane + 1;
  const int64_t r0 =
      (int64_t)(mix64((uint64_t)g.band + 0x9e3779b97f4a7c15ULL) % (uint64_t)slots) *
      kDiagsPerLane;
  if (a.nchunks == 1) {
    g.rs = r0;
    g.re = rows0;
    g.rs2 = 0;
    g.re2 = r0;
    return true;
  }
  const int64_t R = band_chunk_rows(a.chunk_rows, a.chunk_min_rows, a.chunk_ref_rows, rows0);
  int64_t first = r0 % R;  // the lowest cut above row 0
  if (first == 0) first = R;
  const int64_t cuts = first < rows0 ? (rows0 - 1 - first) / R + 1 : 0;
  const int64_t n = cuts + 1;
  if (j >= n) return false;
  const int64_t k_r0 = r0 == 0 ? 0 : (r0 - first) / R + 1;
  const int64_t k = (k_r0 + j) % n;
  g.rs = k == 0 ? 0 : first + (k - 1) * R;
  g.re = k == n - 1 ? rows0 : first + k * R;
  g.rs2 = 0;
  g.re2 = 0;
  return true;
}

template <typename DT, int KIND, bool COLS, bool ROWS, bool FRESH>
__global__ __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu(sweep_waves_per_eu<DT, KIND, FRESH>())))
void sweep_kernel(SweepArgs a) {
  using SW = Sweep<DT, KIND, COLS, ROWS, FRESH>;
  __shared__ Terms<DT> s_terms[2 * kWave];  // staged row / column terms
  Terms<DT> *const s_row = s_terms;
  Terms<DT> *const s_col = s_terms + kWave;
  // FP64 staging of segment seeds: the term arrays (no block is in flight at
  // a segment start); FP32 (which also re-anchors mid-segment): its own
  constexpr bool kF32 = sizeof(DT) == 4;
  static_assert(kF32 || (kBand + 2 * kAnchorChunk64) * sizeof(double) <= sizeof(s_terms),
                "the FP64 seed staging fits in the term arrays");
  __shared__ double s_anchor_own[kF32 ? kBand + 2 * kAnchorChunk : 1];
  double *const s_ax = kF32 ? s_anchor_own : reinterpret_cast<double *>(s_terms);
  double *const s_ay = kF32 ? s_anchor_own + kBand + kAnchorChunk
                            : reinterpret_cast<double *>(s_terms) + kBand + kAnchorChunk64;
  __shared__ uint64_t s_kq[KIND == kKnn ? kBand : 1];
  // FRESH: best key of each row of the current block, and of each column
  // the block crosses (zero = none).  The FP32 form keeps both tables in its
  // anchor staging, which is only in use between blocks (anchor() clears
  // them again when it is done).
  constexpr bool kTable = KIND == kNN && FRESH;
  __shared__ uint64_t s_rt_own[kTable && !kF32 ? SW::kRows : 1];
  __shared__ uint64_t s_ct_own[SW::kColTable && !kF32 ? kDiagsPerLane * SW::kColStride : 1];
  static_assert(kDiagsPerLane * SW::kColStride <= kBand + kAnchorChunk,
                "the FP32 column table fits in the anchor staging");
  static_assert(SW::kRows <= kAnchorChunk, "the FP32 row table fits in the anchor staging");
  uint64_t *const s_ct = kF32 ? reinterpret_cast<uint64_t *>(s_ax) : s_ct_own;
  uint64_t *const s_rt = kF32 ? reinterpret_cast<uint64_t *>(s_ay) : s_rt_own;
  const int lane = threadIdx.x;
  if (kTable && lane < SW::kRows) s_rt[lane] = 0;
  if (SW::kColTable)
    for (int q = lane; q < kDiagsPerLane * SW::kColStride; q += kWave) s_ct[q] = 0;
  const unsigned int nitems = (unsigned int)a.nitems;
  for (;;) {
    unsigned int b = 0;
    if (lane == 0) b = atomicAdd(a.counter, 1u);
    b = (unsigned int)__builtin_amdgcn_readfirstlane((int)b);
    if (b >= nitems) break;
    ItemRows g;
    if (!plan_item(a, b, g)) continue;
    // batched joins: the item's series, seen through offset pointers
    SweepArgs la = a;
    if (a.nbatch > 1) {
      const int64_t o = g.series * a.batch_stride;
      la.dfX = static_cast<const DT *>(a.dfX) + o;
      la.dgX = static_cast<const DT *>(a.dgX) + o;
      la.inX = static_cast<const DT *>(a.inX) + o;
      la.rqX = a.rqX + o;
      la.dfY = static_cast<const DT *>(a.dfY) + o;
      la.dgY = static_cast<const DT *>(a.dgY) + o;
      la.inY = static_cast<const DT *>(a.inY) + o;
      la.TX = a.TX + o;
      la.TY = a.TY + o;
      la.muX = a.muX + o;
      la.muY = a.muY + o;
      la.keyX = a.keyX + o;
      la.keyY = a.keyY + o;
    }
    const int64_t D0 = la.d_base + g.band * kBand;
    Sweep<DT, KIND, COLS, ROWS, FRESH> sw(la, lane, D0);
    sw.s_kq = s_kq;
    if (kTable) sw.s_rt = s_rt;
    if (SW::kColTable) sw.s_ct = s_ct;
    const int64_t dlast = D0 + kBand < la.d_hi ? D0 + kBand - 1 : la.d_hi - 1;
    // all of the band's diagonals are valid on rows below rfull
    const int64_t rfull = la.nY < la.nX - dlast ? la.nY : la.nX - dlast;
    for (int seg = 0; seg < 2; ++seg) {
      const int64_t rs = seg == 0 ? g.rs : g.rs2;
      const int64_t re = seg == 0 ? g.re : g.re2;
      if (rs >= re) continue;
      sw.anchor(rs, s_ax, s_ay);
      sw.segment(rs, re, rfull, s_row, s_col, s_ax, s_ay);
    }
  }
}

template <typename DT, int KIND, bool FRESH = false>
hipError_t launch_kind(const SweepArgs &a, bool cols, bool rows, int grid,
                       hipStream_t s) {
  if (cols && rows) {
    hipLaunchKernelGGL((sweep_kernel<DT, KIND, true, true, FRESH>), dim3(grid),
                       dim3(kWave), 0, s, a);
  } else if (cols) {
    hipLaunchKernelGGL((sweep_kernel<DT, KIND, true, false, FRESH>), dim3(grid),
                       dim3(kWave), 0, s, a);
  } else if (rows) {
    hipLaunchKernelGGL((sweep_kernel<DT, KIND, false, true, FRESH>), dim3(grid),
                       dim3(kWave), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace sweep
}  // namespace scamp_amd
