// Host-side launchers of the HIP kernels (implemented in prep_kernels.hip and
// sweep_kernel.hip); used by the C ABI in scamp_capi.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace scamp_amd {

// ---- statistics (prep_kernels.hip) ----------------------------------------
// T[n] raw series -> clean[n] (non-finite -> 0); means/inorms/df/dg[n-m+1].
// Output arrays may be longer (padding): the pad is filled with mean 0,
// inorm NaN, df 0, dg 0, clean 0.  batch > 1: series q at T + q * t_stride,
// its outputs at q * clean_len (clean) and q * stat_len (the others).
hipError_t launch_stats(const double *T, int64_t n, int64_t m, double *clean,
                        int64_t clean_len, double *means, double *inorms,
                        double *df, double *dg, int64_t stat_len,
                        hipStream_t s, int64_t batch = 1, int64_t t_stride = 0);

// rq[i] = stage1_factor(inorm[i]) for i < len (the FP64 two-stage 1NN
// filter's per-column factor)
hipError_t launch_stage1_factors(const double *inorm, float *rq, int64_t len,
                                 hipStream_t s);
// FP64 -> FP32 copies of the per-window update terms and inverse norms.
hipError_t launch_to_f32(const double *src, float *dst, int64_t len,
                         hipStream_t s);

// ---- exact QT seeds (prep_kernels.hip) ------------------------------------
// One task = `count` (<= 256) consecutive diagonals starting at diag0, all at
// row `row`:  out[j] = sum_t (X[row+diag0+j+t]-muX[row+diag0+j]) *
//                            (Y[row+t]-muY[row])
// with compensated (Dot2) accumulation.
struct SeedTask {
  int64_t row;
  int64_t diag0;
  int64_t out_offset;
  int32_t count;
  int32_t pad;
};
hipError_t launch_seeds(const SeedTask *tasks, int64_t ntasks,
                        const double *X, const double *muX, const double *Y,
                        const double *muY, int64_t m, double *out,
                        hipStream_t s);

// ---- profile conversions (prep_kernels.hip) -------------------------------
// SCAMP mp_entry (corr low word, index high word) <-> sortable keys.
hipError_t launch_mp_to_keys(uint64_t *p, int64_t n, hipStream_t s);
hipError_t launch_keys_to_mp(uint64_t *p, int64_t n, hipStream_t s);
// float profile (1NN) <-> keys (index unused)
hipError_t launch_f32_to_keys(const float *src, uint64_t *dst, int64_t n,
                              hipStream_t s);
hipError_t launch_keys_to_f32(const uint64_t *src, float *dst, int64_t n,
                              hipStream_t s);
// float matrix (MATRIX_SUMMARY) <-> order-preserving uint32
hipError_t launch_f32_to_enc(float *p, int64_t n, hipStream_t s);
hipError_t launch_enc_to_f32(float *p, int64_t n, hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, int64_t n, uint64_t v, hipStream_t s);
// batch profiles of n entries, back to back, in SCAMP's layout (mp: uint64
// mp_entry, else float) <-> keys with kstride entries between series
hipError_t launch_profile_in(const void *src, bool mp, uint64_t *keys,
                             int64_t n, int64_t batch, int64_t kstride,
                             hipStream_t s);
hipError_t launch_profile_out(const uint64_t *keys, int64_t kstride, void *dst,
                              bool mp, int64_t n, int64_t batch, hipStream_t s);
// Self-join 1NN warm start: each subsequence's entries take the best of its
// correlations with the kWarmLags windows just past the exclusion zone e on
// either side (forward partners into keyRow, backward ones into keyCol;
// either may be null; equal pointers merge both).  Every pair on diagonals
// [e, e + kWarmLags) is merged into both of its entries, so the sweep that
// follows starts at diagonal e + kWarmLags.  batch series, stride entries
// apart.
constexpr int64_t kWarmLags = 8;
hipError_t launch_warm_profile(const double *clean, const double *mean,
                               const double *inorm, int64_t ns, int64_t m,
                               int64_t e, int64_t batch, int64_t stride,
                               uint64_t *keyCol, uint64_t *keyRow,
                               hipStream_t s);

// ---- FP64 refinement of mixed 1NN_INDEX profiles (refine_kernel.hip) -----
// Pairs (u, v) with u an index of the profile's series U and v of V are in
// the join unless zone > 0 and |a - b + delta| < zone, where a and b are
// the A and B indices of the pair (u_is_a: u is the A index).  A self-join
// passes zone = m/4, delta = 0, u_is_a = true, and `side` says which of
// its pairs the profile holds: 0 both, -1 only partners v <= u - zone (the
// profile of columns: pairs on diagonals j - i >= zone reach column j from
// rows i below it), +1 only partners v >= u + zone (the profile of rows).
struct RefineExclusion {
  int64_t zone;
  int64_t delta;
  bool u_is_a;
  int32_t side;
};
// `drift` (one counter) counts the refined entries whose FP32 sweep value
// differs from the FP64 correlation of the same pair by more than
// `drift_tol`: the FP32 recurrence went beyond its error bound there.
hipError_t launch_refine(uint64_t *key, const uint64_t *prior, int64_t nsU,
                         const double *TU, const double *muU, const double *inU,
                         const double *dfU, const double *dgU, int64_t nsV,
                         const double *TV, const double *muV, const double *inV,
                         const double *dfV, const double *dgV, int64_t m,
                         RefineExclusion ex, int64_t batch, int64_t stride,
                         uint64_t *drift, double drift_tol, hipStream_t s);

// ---- backward of the differentiable join (backward_kernel.hip) ------------
// grad_t[b, :] = sum over profile entries i of grad_p[b, i] * d corr(i, I_i)
// / d T[b, :] for `batch` series of length n stored back to back; mean/inorm
// rows have `stat_stride` entries per series; `work` holds
// pearson_backward_workspace(n, m, batch) bytes (batch * (n - m + 1) must be
// below 2^31).  Deterministic: no atomics.
size_t pearson_backward_workspace(int64_t n, int64_t m, int64_t batch);
hipError_t launch_pearson_backward(const double *T, int64_t n, int64_t m,
                                   int64_t batch, const double *mean,
                                   const double *inorm, int64_t stat_stride,
                                   const int32_t *index, const double *grad_p,
                                   double *grad_t, void *work, hipStream_t s);

// ---- ALL_NEIGHBORS list conversions --------------------------------------
// prior matches (scamp_hip_match[n * k], corr > -2 = a match) -> key lists
// (entries not above `init` become `init`) and each list's minimum in keys[j]
hipError_t launch_matches_to_knn(const void *matches, int64_t n, int k,
                                 uint64_t init, uint64_t *lists, uint64_t *keys,
                                 hipStream_t s);
// key lists -> matches sorted best first (entries not above `init` become
// empty: corr -2, row 0xFFFFFFFF); sorts the lists in place
hipError_t launch_knn_to_matches(uint64_t *lists, int64_t n, int k, uint64_t init,
                                 void *matches, hipStream_t s);

// ---- the diagonal sweep (sweep_kernel.hip) --------------------------------
enum SweepKind { kNN = 0, kSum = 1, kMatrix = 2, kMatrixSmall = 3, kKnn = 4 };

// Chunk length of a band of rows0 rows in a chunked pass: R for the longest
// band (ref_rows rows), and shorter in proportion for shorter bands (never
// below rmin, a multiple of 16), so that every band has about as many items.
// Items run band by band, longest bands first, so the items that start last
// are short and the launch does not end on a few long items while the rest
// of the GPU idles.
__host__ __device__ inline int64_t band_chunk_rows(int64_t R, int64_t rmin,
                                                   int64_t ref_rows, int64_t rows0) {
  if (rows0 >= ref_rows) return R;
  int64_t r = (R * rows0 + ref_rows - 1) / ref_rows;
  r = (r + 15) / 16 * 16;
  return r < rmin ? rmin : (r > R ? R : r);
}

struct SweepArgs {
  // column side X / row side Y update terms and inverse norms (FP64 or FP32,
  // padded by kColumnPad entries)
  const void *dfX, *dgX, *inX;
  const void *dfY, *dgY, *inY;
  // column side X: each column's FP64 first-stage factor (stage1_factor of
  // its FP64 inverse norm; padded like inX), read by the FP64 two-stage sweep
  const float *rqX;
  // clean series and means (the in-kernel FP64 seeds and re-anchors)
  const double *TX, *TY, *muX, *muY;
  // work items, taken in order through `counter` and decoded in the kernel
  // (plan_item): nitems = bands x nbatch x nchunks; when nchunks > 1 a band
  // of rows0 rows is cut into chunks of band_chunk_rows(chunk_rows,
  // chunk_min_rows, chunk_ref_rows, rows0) rows
  int64_t nitems;
  int64_t nchunks, chunk_rows, chunk_min_rows, chunk_ref_rows;
  int64_t nX, nY;        // subsequence counts of X and Y
  int64_t d_lo, d_hi;    // diagonal range of this pass
  int64_t d_base;        // d_lo rounded down to a multiple of 8: band b
                         // covers diagonals d_base + 512 b + [0, 512)
  int64_t m;
  int64_t anchor;        // FP32 form: rows between re-anchors (multiple of 16)
  uint64_t *keyX, *keyY; // 1NN / 1NN_INDEX profiles (columns, rows)
  double *sumX, *sumY;   // SUM_THRESH profiles
  uint32_t *mat;         // MATRIX_SUMMARY cells (order-preserving uint32)
  int64_t mat_w, rpc, cpc;
  // ALL_NEIGHBORS: per subsequence a list of knn_k order-preserving keys (the
  // k best matches so far; keyX / keyY then hold each list's minimum)
  uint64_t *knnX, *knnY;
  int32_t knn_k;
  double thr;
  unsigned int *counter; // work counter, zeroed before launch
  // Batched self-joins: nbatch series of one length, every per-subsequence
  // array with stride batch_stride entries between series (an item's
  // `series` selects them).  nbatch = 1: a single join.
  int64_t nbatch, batch_stride;
};

// cols/rows: which profiles the pass updates (for kMatrix: direct cells
// (row, col) and transposed cells (col, row) respectively).  fresh (kNN):
// the launch's thresholds start far from final (short series cut into row
// chunks); selects the kernel variant that makes the merge path cheaper.
hipError_t launch_sweep(const SweepArgs &a, bool fp32, SweepKind kind,
                        bool cols, bool rows, int grid, hipStream_t s,
                        bool fresh);

}  // namespace scamp_amd
