/*
 * scamp_hip.h -- C ABI of the MI355X (gfx950) matrix-profile join engine.
 *
 * This is the boundary the host side (the C++ SCAMP::do_SCAMP entry point,
 * the pyscamp module and the torch extension) calls.  Plain pointers and
 * sizes only; every function returns an int status (scamp_hip_status) and
 * never throws; the message of the last failure on the calling thread is
 * available from scamp_hip_last_error().
 *
 * Profile buffers use SCAMP's layouts bit for bit (src/common.h:70-74,
 * src/common.cpp:283-319):
 *   1NN_INDEX      uint64 mp_entry: low 32 bits = FP32 Pearson correlation,
 *                  high 32 bits = index; "no match" = {-2.0f, 0xFFFFFFFF}
 *   1NN            float correlation, "no match" = -2
 *   SUM_THRESH     double sum of correlations above the threshold, init 0
 *   MATRIX_SUMMARY float[height * width] row-major, init -2; cell (r, c)
 *                  holds the largest correlation of the pairs (row window
 *                  in band r, column window in band c); a self-join fills
 *                  the pairs i < j only (cells below the diagonal keep
 *                  their prior), as SCAMP's upper-triangular tiles do
 *   ALL_NEIGHBORS  scamp_hip_match[n_sub * k]: for subsequence j its k best
 *                  matches, entries j*k .. j*k+k-1, best first; an empty
 *                  entry has corr -2 and row 0xFFFFFFFF (k =
 *                  max_matches_per_column)
 * Profiles are read as the prior best-so-far and updated in place, as
 * SCAMP's tiles do (src/tile.cpp:322-389).
 */
#ifndef SCAMP_HIP_H
#define SCAMP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  0..6 keep the values of SCAMPError_t (src/common.h:359-367);
 * the rest are this library's. */
typedef enum {
  SCAMP_HIP_OK = 0,
  SCAMP_HIP_UNIMPLEMENTED = 1,   /* SCAMP_FUNCTIONALITY_UNIMPLEMENTED */
  SCAMP_HIP_ILLEGAL_TYPE = 2,    /* SCAMP_TILE_ILLEGAL_TYPE */
  SCAMP_HIP_DEVICE_ERROR = 3,    /* SCAMP_CUDA_ERROR: a HIP runtime error */
  SCAMP_HIP_DIM_INCOMPATIBLE = 6,/* SCAMP_DIM_INCOMPATIBLE */
  SCAMP_HIP_INVALID_ARGS = 16,   /* argument validation failed */
  SCAMP_HIP_OUT_OF_MEMORY = 17,
  SCAMP_HIP_NO_DEVICE = 18
} scamp_hip_status;

/* Profile types: the SCAMPProfileType values (src/common.h:34-44). */
enum {
  SCAMP_HIP_1NN_INDEX = 1,
  SCAMP_HIP_SUM_THRESH = 2,
  SCAMP_HIP_1NN = 6,
  SCAMP_HIP_ALL_NEIGHBORS = 7,
  SCAMP_HIP_MATRIX_SUMMARY = 8
};

/* One ALL_NEIGHBORS match: Pearson correlation of subsequence `col` with
 * subsequence `row` of the other series (or of the same one). */
typedef struct {
  float corr;
  uint32_t row;
  uint32_t col;
} scamp_hip_match;

/* Precisions: the SCAMPPrecisionType values (src/common.h:52-58).
 * DOUBLE (and ULTRA, which maps to it: the statistics are direct window sums
 * either way) = all-FP64 recurrence;
 * MIXED (and SINGLE, which names the same form) = compensated FP32
 * recurrence re-anchored with FP64 every anchor_interval rows; its 1NN_INDEX
 * entries that the join improved are then re-evaluated in FP64 over the
 * matched lag and the two lags on either side (no counterpart in the
 * reference), so the stored correlation is the FP64 one and the index the
 * FP64 form's wherever that lies within two lags.  The same pass audits the
 * FP32 sweep: when a matched pair's FP32 correlation is off by more than
 * 5e-6, the join is recomputed in FP64 (scamp_hip_last_precision). */
enum {
  SCAMP_HIP_SINGLE = 1,
  SCAMP_HIP_MIXED = 2,
  SCAMP_HIP_DOUBLE = 3,
  SCAMP_HIP_ULTRA = 4
};

typedef struct scamp_hip_ctx scamp_hip_ctx;

/* Join options: the fields of SCAMPArgs (src/common.h:164-188) that select
 * what is computed, plus the mixed form's re-anchoring interval. */
typedef struct {
  int64_t window;              /* m, >= 3 */
  int32_t profile_type;        /* SCAMP_HIP_1NN_INDEX ... */
  int32_t precision;           /* SCAMP_HIP_DOUBLE / _MIXED / ... */
  int32_t has_b;               /* 0 = self-join of a */
  int32_t computing_rows;      /* rows of the distance matrix (series b) */
  int32_t computing_columns;   /* columns (series a) */
  int32_t keep_rows_separate;  /* row profile goes to profile_b */
  int32_t is_aligned;          /* AB-join of aligned series: apply m/4 zone */
  int32_t silent_mode;
  int64_t distributed_start_row;  /* -1 = unset */
  int64_t distributed_start_col;
  double distance_threshold;   /* SUM_THRESH / MATRIX_SUMMARY threshold */
  int64_t matrix_height;
  int64_t matrix_width;
  int64_t max_tile_size;       /* validated as SCAMP does; the sweep is untiled */
  int64_t anchor_interval;     /* MIXED only: rows between FP64 re-anchors,
                                  0 = library default (2 * window, within
                                  [512, 16384]) */
  int64_t max_matches_per_column; /* ALL_NEIGHBORS: k, 1..256 */
} scamp_hip_params;

/* Replaces: the per-worker device state of SCAMP's Tile/ExecInfo
 * (src/tile.cpp:141-283, src/common.cpp:109-138).  Binds a HIP device and
 * creates one stream on it.  Calls on one context run one after another
 * (the context holds a lock during each call); separate contexts are
 * independent. */
int scamp_hip_ctx_create(int device, scamp_hip_ctx **out);
int scamp_hip_ctx_destroy(scamp_hip_ctx *ctx);

/* No counterpart in the reference (whose tiles own their streams,
 * src/common.cpp:109-124): puts the context's work on `stream`, a hipStream_t
 * of the context's device that the caller owns (NULL: back to the context's
 * own stream), so that its joins order with the caller's own work on that
 * stream (a torch stream, say) without host synchronisation.  While a
 * caller's stream is set, scamp_hip_join_device, scamp_hip_selfjoin_batch and
 * scamp_hip_pearson_backward return once their work is queued: their outputs
 * are ready in stream order, and their inputs must stay valid until then.  A
 * MIXED/SINGLE 1NN_INDEX join still waits for its FP64 audit, and the
 * no-allocation join still finishes before it returns (it shares the GPU's
 * scratch array); the host-buffer calls always block.
 * scamp_hip_last_timing then waits for the last join to end.  Switching
 * streams first waits for the work on the old one. */
int scamp_hip_ctx_set_stream(scamp_hip_ctx *ctx, void *stream);

/* Number of visible HIP devices (0 when none), as SCAMP's do_SCAMP counts
 * CUDA devices (src/SCAMP.cpp:261-275). */
int scamp_hip_device_count(int *count);

/* Copies the last error message of the calling thread into buf. */
int scamp_hip_last_error(char *buf, size_t len);

/* The join.  Replaces SCAMP_Operation::do_join (src/SCAMP.cpp:174-259) with
 * its tile queue, per-tile QT seeds (src/qt_helper.cpp:75-121), the
 * gpu_kernel_* tile kernels (src/kernels.h:9-12) and the host merges
 * (src/common.cpp:364-415).  a/b/profiles are HOST buffers; profile_b is
 * used only with keep_rows_separate.  Blocks until the profiles are written. */
int scamp_hip_join(scamp_hip_ctx *ctx, const double *a, size_t na,
                   const double *b, size_t nb, const scamp_hip_params *p,
                   void *profile_a, void *profile_b);

/* The same join on DEVICE buffers already resident in HBM (series as float64
 * device pointers, profiles in the layouts above).  Work is queued on the
 * context's stream; the call returns after the stream has finished (after
 * queueing it, on a stream set by scamp_hip_ctx_set_stream). */
int scamp_hip_join_device(scamp_hip_ctx *ctx, const double *d_a, size_t na,
                          const double *d_b, size_t nb,
                          const scamp_hip_params *p, void *d_profile_a,
                          void *d_profile_b);

/* Many self-joins of equal-length series in one launch (no counterpart in
 * the reference, whose callers loop over do_SCAMP per series,
 * python/SCAMP_python.cpp:67-110).  d_T holds `batch` float64 series of n
 * samples back to back (DEVICE); d_profiles holds batch * (n - m + 1)
 * entries of a 1NN_INDEX (uint64 mp_entry) or 1NN (float) profile, read as
 * the prior best-so-far and updated in place; computing_rows/columns apply
 * as in a self-join, keep_rows_separate is refused (one profile per series).
 * Equals `batch` calls of scamp_hip_join_device with has_b = 0 (bit for bit
 * while the batch is small
 * enough that its bands are cut into row chunks like a single series'; up to
 * FP64 rounding of the recurrence beyond); fills the GPU when each series
 * alone has too few diagonals to.  Blocks until done (queues the work, on a
 * stream set by scamp_hip_ctx_set_stream). */
int scamp_hip_selfjoin_batch(scamp_hip_ctx *ctx, const double *d_T, size_t n,
                             int64_t batch, const scamp_hip_params *p,
                             void *d_profiles);

/* Statistics of one series on the GPU (replaces compute_statistics_cpu,
 * src/cpu_stats.cpp:123-160, with direct compensated per-window sums).
 * Host buffers: T[n]; outputs means/inorms/df/dg have n - m + 1 entries. */
int scamp_hip_stats(scamp_hip_ctx *ctx, const double *T, size_t n,
                    int64_t window, double *means, double *inorms, double *df,
                    double *dg);

/* Exact-summation FP64 QT seeds (replaces qt_compute_helper::compute_QT,
 * src/qt_helper.cpp:75-121, and its cuFFT kernels, src/qt_kernels.cu:3-61):
 *   qt[j] = sum_{t<m} (X[row+diag0+j+t] - mean(X window)) *
 *                     (Y[row+t] - mean(Y window)),  j < count
 * on host buffers X[nx], Y[ny], read through the same cleaning as the join
 * (a non-finite sample takes the value of the nearest finite one). */
int scamp_hip_qt_seed(scamp_hip_ctx *ctx, const double *X, size_t nx,
                      const double *Y, size_t ny, int64_t window, int64_t row,
                      int64_t diag0, int64_t count, double *qt);

/* Backward of the differentiable join (no counterpart in the reference):
 * for `batch` series of length n stored back to back in d_T (DEVICE, float64)
 * and the forward's 1NN indices d_index[batch][n-m+1] (int32, -1 = none),
 *   d_grad_T[b][t] = sum_i d_grad_corr[b][i] * d corr(T[i:i+m], T[I_i:I_i+m])
 *                                             / d T[t]
 * with the index held fixed.  d_grad_T is overwritten; the result does not
 * depend on scheduling (no atomics: the same bits on every call).  At most
 * 2^31 - 1 profile entries (batch * (n - m + 1)) per call.  Blocks until
 * done (queues the work, on a stream set by scamp_hip_ctx_set_stream). */
int scamp_hip_pearson_backward(scamp_hip_ctx *ctx, const double *d_T, size_t n,
                               int64_t window, int64_t batch,
                               const int32_t *d_index,
                               const double *d_grad_corr, double *d_grad_T);

/* Timing of the last join on this context, from HIP events on its stream:
 * total, statistics+seeds, and the diagonal sweep, in milliseconds (waits
 * for that join to end when it was queued on a caller's stream). */
int scamp_hip_last_timing(scamp_hip_ctx *ctx, double *total_ms,
                          double *prep_ms, double *sweep_ms);

/* The precision the last join on this context ran in (a scamp_hip_precision
 * value) and the number of 1NN_INDEX entries whose FP32 sweep correlation
 * failed the mixed form's FP64 audit (0 unless MIXED/SINGLE).  A nonzero
 * count means the join was recomputed in FP64 and precision reads DOUBLE.
 * No counterpart in the reference, whose mixed mode has no audit
 * (docs/source/cli.rst:127 lists inputs on which it fails). */
int scamp_hip_last_precision(scamp_hip_ctx *ctx, int *precision,
                             uint64_t *drift_entries);

/* Working memory of scamp_hip_join_device_noalloc: one array of this many
 * bytes per GPU, held from the library's first context on that GPU until the
 * process ends. */
#define SCAMP_HIP_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)

/* scamp_hip_join_device without allocating: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, semantics and results as scamp_hip_join_device; every working
 * buffer (statistics, working profiles, counters) is carved from the
 * device's scratch array instead of hipMalloc, and the call makes no
 * hipMalloc, hipFree or event creation.  A join whose buffers exceed
 * SCAMP_HIP_NOALLOC_SCRATCH_BYTES is refused with SCAMP_HIP_INVALID_ARGS
 * (about 2.5 million subsequences for an FP64 1NN_INDEX self-join, 1.8
 * million for MIXED).  All calls on one GPU share its array, which keeps the
 * last call's contents (each call initialises what it reads): calls on one
 * GPU take turns under a per-device lock inside the library, and any other
 * code that overlapped them on the same array would race.  Blocks until
 * done. */
int scamp_hip_join_device_noalloc(scamp_hip_ctx *ctx, const double *d_a,
                                  size_t na, const double *d_b, size_t nb,
                                  const scamp_hip_params *p, void *d_profile_a,
                                  void *d_profile_b);

#ifdef __cplusplus
}
#endif
#endif
