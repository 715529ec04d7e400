/*
 * Oracle: a CPU restatement of SCAMP's CPU double path (the matrix-profile
 * join: statistics -> per-tile QT seed -> diagonal tile recurrence -> merge).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this code;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
 * and only to check results against.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the GPU-SCRIMP-LARGE source tree).
 */
#ifndef SCAMP_ORACLE_H
#define SCAMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Profile type codes, same values as SCAMPProfileType (src/common.h:34-44). */
enum { ORA_1NN_INDEX = 1, ORA_SUM_THRESH = 2, ORA_1NN = 6 };

/* Statistics of one series (src/cpu_stats.cpp:11-29, 50-75, 103-160).
 * T has n samples; outputs have n_sub = n - m + 1 entries.  clean (length n)
 * receives T with non-finite values replaced by 0.  Returns 0 on success. */
int ora_stats(const double *T, int64_t n, int64_t m, double *clean,
              double *means, double *norms, double *df, double *dg);

/* Direct centred sliding dot products of one tile (src/qt_helper.cpp:52-72):
 * QT[i] = sum_j (T[i+j] - mu_i)(Q[j] - qmean), i < nq, mu_i a rolling mean. */
void ora_qt_seed(const double *T, const double *Q, int64_t nq, int64_t m,
                 double *QT);

/* Options of one join (the fields of SCAMPArgs, src/common.h:164-188, that
 * the CPU double path reads). */
typedef struct {
  int64_t window;
  int64_t max_tile_size;
  int32_t profile_type;      /* ORA_* */
  int32_t has_b;             /* 0 = self-join */
  int32_t computing_rows;
  int32_t computing_cols;
  int32_t keep_rows_separate;
  int32_t is_aligned;
  int64_t global_start_row;  /* distributed_start_row, -1 = unset */
  int64_t global_start_col;
  double threshold;
  int32_t num_workers;       /* only sizes the tiles (src/common.cpp:56-57) */
  int32_t fp64_dist;         /* 1 = keep correlations in double (no FP32
                                rounding of stored values); 0 = SCAMP's FP32 */
} ora_args;

/* The join.  Profiles are read as the prior and written in place, in the
 * oracle's own wide layout:
 *   1NN_INDEX : corr[n_sub] (double) + idx[n_sub] (int64, -1 = none)
 *   1NN       : corr[n_sub]
 *   SUM_THRESH: corr[n_sub] (the sums)
 * profile A has n_sub(a) entries; profile B (only with keep_rows_separate)
 * n_sub(b).  Returns 0, or a nonzero SCAMPError_t-like code. */
int ora_join(const double *a, int64_t na, const double *b, int64_t nb,
             const ora_args *args, double *corr_a, int64_t *idx_a,
             double *corr_b, int64_t *idx_b);

#ifdef __cplusplus
}
#endif
#endif
