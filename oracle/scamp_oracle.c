/*
 * Oracle: CPU restatement of SCAMP's CPU double path.
 *
 * TEST INFRASTRUCTURE ONLY -- results are checked against it; it is never the thing measured or
 * shipped.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it.
 *
 * It follows, operation for operation, the reference's
 *   statistics      src/cpu_stats.cpp:11-160
 *   tile QT seed    src/qt_helper.cpp:52-72
 *   tile kernel     src/cpu_kernels.cpp:20-199 (256-diagonal blocks)
 *   dispatch        src/cpu_kernels.cpp:201-311, src/kernel_common.cpp:5-38
 *   exclusion zones src/tile.cpp:12-138
 *   tile execution  src/tile.cpp:546-707
 *   tile order      src/SCAMP.cpp:25-90
 *   worker loop     src/SCAMP.cpp:97-169 (one worker, so merges are ordered)
 *   tile sizing     src/common.cpp:31-107
 *   merges          src/common.cpp:167-197, 364-415
 *   init values     src/common.cpp:283-319
 * Multiply-adds that the reference's build contracts into FMAs
 * (-ffp-contract=fast, CMakeLists.txt:129-143) are written as fma() here and
 * the file is built with -ffp-contract=off, so its rounding is fixed by the
 * source rather than by compiler choices.
 */
#include "scamp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FLATNESS_EPSILON 1e-13 /* src/cpu_stats.cpp:6 */
#define DIAG_BLOCK 256         /* src/cpu_kernels.cpp:11 */

/* ---------------------------------------------------------------- stats */

int ora_stats(const double *T, int64_t n, int64_t m, double *clean,
              double *means, double *norms, double *df, double *dg) {
  if (m < 1 || n < m) return 1;
  const int64_t ns = n - m + 1;
  unsigned char *has_nan = (unsigned char *)calloc((size_t)ns, 1);
  if (!has_nan) return 1;

  /* Non-finite samples become 0; a window is flagged when a non-finite
   * sample lies inside it (src/cpu_stats.cpp:11-29). */
  int64_t since_nan = m;
  for (int64_t i = 0; i < n; ++i) {
    if (isfinite(T[i])) {
      clean[i] = T[i];
    } else {
      since_nan = 0;
      clean[i] = 0;
    }
    if (i >= m - 1) has_nan[i - m + 1] = since_nan < m;
    since_nan++;
  }

  /* Sliding mean with an error-free-transformation running compensation
   * (src/cpu_stats.cpp:50-75). */
  double hi = clean[0], lo = 0;
  for (int64_t i = 1; i < m; ++i) {
    double s = hi + clean[i];
    double t = s - hi;
    lo += (hi - (s - t)) + (clean[i] - t);
    hi = s;
  }
  means[0] = (hi + lo) / (double)m;
  for (int64_t i = m; i < n; ++i) {
    double s = hi - clean[i - m];
    double t = s - hi;
    lo += (hi - (s - t)) - (clean[i - m] + t);
    hi = s;
    s = hi + clean[i];
    t = s - hi;
    lo += (hi - (s - t)) + (clean[i] - t);
    hi = s;
    means[i - m + 1] = (hi + lo) / (double)m;
  }

  /* Sum of squared deviations: first window directly, then the one-step
   * update (src/cpu_stats.cpp:103-120). */
  double acc = 0;
  for (int64_t k = 0; k < m; ++k) {
    double v = clean[k] - means[0];
    acc = fma(v, v, acc);
  }
  norms[0] = acc;
  for (int64_t i = 1; i < ns; ++i) {
    norms[i] = fma((clean[i - 1] - means[i - 1]) + (clean[i + m - 1] - means[i]),
                   clean[i + m - 1] - clean[i - 1], norms[i - 1]);
  }

  /* Inverse norms; NaN for windows with a non-finite sample or a flat window
   * (src/cpu_stats.cpp:140-152). */
  for (int64_t i = 0; i < ns; ++i) {
    if (has_nan[i] || norms[i] <= FLATNESS_EPSILON) {
      norms[i] = NAN;
    } else {
      norms[i] = 1.0 / sqrt(norms[i]);
    }
  }

  /* Update terms of the covariance recurrence (src/cpu_stats.cpp:154-157);
   * the last entry stays 0. */
  for (int64_t i = 0; i < ns - 1; ++i) {
    df[i] = (clean[i + m] - clean[i]) / 2.0;
    dg[i] = (clean[i + m] - means[i + 1]) + (clean[i] - means[i]);
  }
  df[ns - 1] = 0;
  dg[ns - 1] = 0;
  free(has_nan);
  return 0;
}

/* -------------------------------------------------------------- QT seed */

void ora_qt_seed(const double *T, const double *Q, int64_t nq, int64_t m,
                 double *QT) {
  /* src/qt_helper.cpp:52-72: rolling (uncompensated) mean of T, direct mean
   * of Q, then a direct centred dot product per offset. */
  double roll = 0, qmean = 0;
  for (int64_t k = 0; k < m; ++k) {
    roll += T[k];
    qmean += Q[k];
  }
  qmean /= (double)m;
  for (int64_t i = 0; i < nq; ++i) {
    double mu = roll / (double)m;
    double acc = 0;
    for (int64_t k = 0; k < m; ++k) acc = fma(T[i + k] - mu, Q[k] - qmean, acc);
    QT[i] = acc;
    if (i + 1 < nq) roll = roll - T[i] + T[i + m];
  }
}

/* --------------------------------------------------------- tile kernel */

typedef struct {
  double *cov;
  const double *dfa, *dfb, *dga, *dgb, *na, *nb;
  int64_t n_x, n_y;
  int64_t excl_lo, excl_hi;
  double thresh;
} tile_args;

/* A tile-local profile in the wide layout. */
typedef struct {
  double *corr;
  int64_t *idx;
} wide_prof;

static inline double round_dist(double v, int fp64) {
  return fp64 ? v : (double)(float)v;
}

/* update_mp (src/cpu_kernels.cpp:37-67) */
static inline void update_entry(int ptype, wide_prof p, double corr,
                                int64_t match, int64_t pos, double thresh) {
  switch (ptype) {
    case ORA_SUM_THRESH:
      if (corr > thresh) p.corr[pos] += corr;
      break;
    case ORA_1NN_INDEX:
      if (corr > p.corr[pos]) {
        p.corr[pos] = corr;
        p.idx[pos] = match;
      }
      break;
    default: /* ORA_1NN */
      if (!(p.corr[pos] >= corr)) p.corr[pos] = corr;
      break;
  }
}

/* do_tile (src/cpu_kernels.cpp:117-199): 256-diagonal blocks, a fast path
 * over the rows where the whole block is inside the tile, then a tail. */
static void ora_do_tile(const tile_args *a, int ptype, int fp64, int do_rows,
                        int do_cols, wide_prof pa, wide_prof pb) {
  /* SUM_THRESH keeps double distances (src/cpu_kernels.cpp:261-265). */
  const int exact = fp64 || ptype == ORA_SUM_THRESH;
  const double init = ptype == ORA_SUM_THRESH ? 0.0 : -2.0;
  const int64_t num_diags = a->n_x - a->excl_hi + 1;
  double corr[DIAG_BLOCK];
  int64_t cidx[DIAG_BLOCK];
  for (int64_t td = a->excl_lo; td < num_diags; td += DIAG_BLOCK) {
    int64_t row_iters = a->n_x - td < a->n_y ? a->n_x - td : a->n_y;
    int64_t full_rows;
    if (td + DIAG_BLOCK >= num_diags) {
      full_rows = 0;
    } else {
      full_rows = a->n_x - td - DIAG_BLOCK + 1;
      if (full_rows > a->n_y) full_rows = a->n_y;
      if (full_rows < 0) full_rows = 0;
    }
    for (int64_t row = 0; row < full_rows; ++row) {
      for (int ld = 0; ld < DIAG_BLOCK; ++ld) {
        int64_t col = td + ld + row;
        double c = round_dist(a->cov[td + ld] * a->na[col] * a->nb[row], exact);
        corr[ld] = isfinite(c) ? c : init;
      }
      if (do_cols) {
        for (int ld = 0; ld < DIAG_BLOCK; ++ld)
          update_entry(ptype, pa, corr[ld], row, td + ld + row, a->thresh);
      }
      if (do_rows) {
        /* reduce_row (src/cpu_kernels.cpp:69-115) */
        if (ptype == ORA_SUM_THRESH) {
          double s = 0;
          for (int i = 0; i < DIAG_BLOCK; ++i)
            if (corr[i] > a->thresh) s += corr[i];
          /* the block sum passes through update_mp's threshold test */
          update_entry(ptype, pb, s, 0, row, a->thresh);
        } else {
          for (int i = 0; i < DIAG_BLOCK / 2; ++i) {
            int keep = corr[i] >= corr[i + DIAG_BLOCK / 2];
            cidx[i] = keep ? i : i + DIAG_BLOCK / 2;
            corr[i] = keep ? corr[i] : corr[i + DIAG_BLOCK / 2];
          }
          for (int off = DIAG_BLOCK / 4; off > 0; off /= 2) {
            for (int i = 0; i < off; ++i) {
              int keep = corr[i] >= corr[i + off];
              cidx[i] = keep ? cidx[i] : cidx[i + off];
              corr[i] = keep ? corr[i] : corr[i + off];
            }
          }
          update_entry(ptype, pb, corr[0], cidx[0] + td + row, row, a->thresh);
        }
      }
      for (int ld = 0; ld < DIAG_BLOCK; ++ld) {
        int64_t col = td + ld + row;
        a->cov[td + ld] = fma(a->dfa[col], a->dgb[row], a->cov[td + ld]);
        a->cov[td + ld] = fma(a->dfb[row], a->dga[col], a->cov[td + ld]);
      }
    }
    for (int64_t row = full_rows; row < row_iters; ++row) {
      int64_t dmax = a->n_x - a->excl_hi + 1;
      if (a->n_x - row < dmax) dmax = a->n_x - row;
      if (td + DIAG_BLOCK < dmax) dmax = td + DIAG_BLOCK;
      for (int64_t d = td; d < dmax; ++d) {
        int64_t col = d + row;
        double c = round_dist(a->cov[d] * a->na[col] * a->nb[row], exact);
        c = isfinite(c) ? c : init;
        if (do_cols) update_entry(ptype, pa, c, row, col, a->thresh);
        if (do_rows) update_entry(ptype, pb, c, col, row, a->thresh);
      }
      for (int64_t d = td; d < dmax; ++d) {
        int64_t col = d + row;
        a->cov[d] = fma(a->dfa[col], a->dgb[row], a->cov[d]);
        a->cov[d] = fma(a->dga[col], a->dfb[row], a->cov[d]);
      }
    }
  }
}

/* ------------------------------------------------------- tile executor */

typedef struct {
  /* operation (OpInfo, src/common.h:216-273) */
  int64_t lenA, lenB, m, tile_ts, tile_w, tile_h;
  int self_join, ptype, fp64, rows, cols, keep_rows, aligned;
  int64_t gstart_row, gstart_col;
  double thresh;
  /* global stats */
  const double *cleanA, *cleanB;
  const double *meanA, *normA, *dfA, *dgA, *meanB, *normB, *dfB, *dgB;
  /* tile position */
  int64_t trow, tcol, width, height;
  /* tile buffers */
  double *QT;
  wide_prof ta, tb;
} op_state;

/* get_exclusion (src/tile.cpp:12-19) */
static int64_t exclusion_at(int64_t m, int64_t start_row, int64_t start_col) {
  int64_t e = m / 4;
  if (start_col >= start_row && start_col < start_row + e) return e;
  return 0;
}

/* src/tile.cpp:21-38 */
static void excl_self(const op_state *s, int upper, int64_t *lo, int64_t *hi) {
  int64_t extra = s->ptype == ORA_SUM_THRESH ? 1 : 0;
  if (upper) {
    *lo = exclusion_at(s->m, s->trow, s->tcol);
    *hi = 0;
    return;
  }
  int64_t h = s->height - s->m + 1;
  *lo = extra;
  *hi = exclusion_at(s->m, s->tcol, s->trow + h);
}

/* src/tile.cpp:44-138 */
static void excl_ab(const op_state *s, int upper, int64_t *lo, int64_t *hi) {
  int64_t elo = 0, ehi = 0, alt = 0;
  if (upper && s->ptype == ORA_SUM_THRESH) alt += 1;
  if (!s->aligned) {
    *lo = elo > alt ? elo : alt;
    *hi = ehi;
    return;
  }
  int64_t h = s->height - s->m + 1, w = s->width - s->m + 1;
  int64_t sc = s->tcol, sr = s->trow;
  if (s->gstart_col >= 0 && s->gstart_row >= 0) {
    sc += s->gstart_col;
    sr += s->gstart_row;
  }
  if (upper) {
    elo = exclusion_at(s->m, sr, sc);
    ehi = sr > sc ? exclusion_at(s->m, sr, sc + w) : 0;
  } else {
    elo = exclusion_at(s->m, sc, sr);
    ehi = sr <= sc ? exclusion_at(s->m, sc, sr + h) : 0;
  }
  *lo = elo > alt ? elo : alt;
  *hi = ehi;
}

/* One kernel call on the current tile (src/kernel_common.cpp:5-38,
 * src/cpu_kernels.cpp:254-311). */
static void run_half(op_state *s, int transpose, int ab) {
  tile_args a;
  const int64_t ra = s->tcol, rb = s->trow;
  a.cov = s->QT;
  a.dfa = transpose ? s->dfB + rb : s->dfA + ra;
  a.dfb = transpose ? s->dfA + ra : s->dfB + rb;
  a.dga = transpose ? s->dgB + rb : s->dgA + ra;
  a.dgb = transpose ? s->dgA + ra : s->dgB + rb;
  a.na = transpose ? s->normB + rb : s->normA + ra;
  a.nb = transpose ? s->normA + ra : s->normB + rb;
  a.n_x = (transpose ? s->height : s->width) - s->m + 1;
  a.n_y = (transpose ? s->width : s->height) - s->m + 1;
  if (ab) excl_ab(s, !transpose, &a.excl_lo, &a.excl_hi);
  else excl_self(s, !transpose, &a.excl_lo, &a.excl_hi);
  a.thresh = s->thresh;
  if (a.excl_lo + a.excl_hi >= a.n_x) return;
  if (!transpose) {
    ora_do_tile(&a, s->ptype, s->fp64, s->rows, s->cols, s->ta, s->tb);
  } else {
    ora_do_tile(&a, s->ptype, s->fp64, s->cols, s->rows, s->tb, s->ta);
  }
}

/* Tile::execute (src/tile.cpp:546-707) */
static void execute_tile(op_state *s, int diag_tile) {
  const double *TA = s->cleanA + s->tcol, *TB = s->cleanB + s->trow;
  int64_t nx = s->width - s->m + 1, ny = s->height - s->m + 1;
  ora_qt_seed(TA, TB, nx, s->m, s->QT);
  run_half(s, 0, !s->self_join);
  if (s->self_join && diag_tile) return;
  ora_qt_seed(TB, TA, ny, s->m, s->QT);
  run_half(s, 1, !s->self_join);
}

/* MergeTileToProfile (src/common.cpp:364-415) */
static void merge(const op_state *s, wide_prof g, wide_prof t, int64_t pos,
                  int64_t len, int64_t index_start) {
  for (int64_t i = 0; i < len; ++i) {
    switch (s->ptype) {
      case ORA_SUM_THRESH:
        g.corr[pos + i] += t.corr[i];
        break;
      case ORA_1NN_INDEX:
        if (g.corr[pos + i] < t.corr[i]) {
          g.corr[pos + i] = t.corr[i];
          g.idx[pos + i] = t.idx[i] + index_start;
        }
        break;
      default:
        if (g.corr[pos + i] < t.corr[i]) g.corr[pos + i] = t.corr[i];
        break;
    }
  }
}

/* Tile::InitProfile (src/tile.cpp:322-389) for the profile types the CPU
 * double path supports. */
static void init_tile_profile(op_state *s, wide_prof ga, wide_prof gb) {
  int64_t w = s->width - s->m + 1, h = s->height - s->m + 1;
  if (s->ptype == ORA_SUM_THRESH) {
    memset(s->ta.corr, 0, sizeof(double) * (size_t)w);
    memset(s->tb.corr, 0, sizeof(double) * (size_t)h);
    return;
  }
  memcpy(s->ta.corr, ga.corr + s->tcol, sizeof(double) * (size_t)w);
  if (s->ptype == ORA_1NN_INDEX) {
    for (int64_t i = 0; i < w; ++i) s->ta.idx[i] = ga.idx[s->tcol + i];
    if (s->rows && s->keep_rows) {
      memcpy(s->tb.corr, gb.corr + s->trow, sizeof(double) * (size_t)h);
      for (int64_t i = 0; i < h; ++i) s->tb.idx[i] = gb.idx[s->trow + i];
    } else if (s->self_join) {
      memcpy(s->tb.corr, ga.corr + s->trow, sizeof(double) * (size_t)h);
      for (int64_t i = 0; i < h; ++i) s->tb.idx[i] = ga.idx[s->trow + i];
    }
  } else {
    if (s->self_join) {
      memcpy(s->tb.corr, ga.corr + s->trow, sizeof(double) * (size_t)h);
    } else if (s->rows && s->keep_rows) {
      memcpy(s->tb.corr, gb.corr + s->trow, sizeof(double) * (size_t)h);
    }
  }
}

int ora_join(const double *a, int64_t na, const double *b, int64_t nb,
             const ora_args *args, double *corr_a, int64_t *idx_a,
             double *corr_b, int64_t *idx_b) {
  const int64_t m = args->window;
  const int self_join = !args->has_b;
  if (self_join) {
    b = a;
    nb = na;
  }
  /* SCAMPArgs::validate (src/common.cpp:417-436) */
  if (m < 3 || args->max_tile_size < 1024 || args->max_tile_size / 2 < m ||
      na < m || nb < m)
    return 7;
  if (args->profile_type != ORA_1NN_INDEX && args->profile_type != ORA_1NN &&
      args->profile_type != ORA_SUM_THRESH)
    return 1; /* SCAMP_FUNCTIONALITY_UNIMPLEMENTED */

  op_state s;
  memset(&s, 0, sizeof(s));
  s.lenA = na;
  s.lenB = nb;
  s.m = m;
  s.self_join = self_join;
  s.ptype = args->profile_type;
  s.fp64 = args->fp64_dist;
  s.rows = args->computing_rows;
  s.cols = args->computing_cols;
  s.keep_rows = args->keep_rows_separate;
  s.aligned = args->is_aligned;
  s.gstart_row = args->global_start_row;
  s.gstart_col = args->global_start_col;
  s.thresh = args->threshold;

  /* OpInfo tile sizing (src/common.cpp:56-73) */
  int workers = args->num_workers > 0 ? args->num_workers : 1;
  int64_t ts = (na > nb ? na : nb) / workers;
  if (ts > args->max_tile_size) ts = args->max_tile_size;
  if (ts < 2 * m) ts = 2 * m;
  s.tile_ts = ts;
  s.tile_w = ts - m + 1;
  s.tile_h = s.tile_w;

  const int64_t nsa = na - m + 1, nsb = nb - m + 1;
  double *cleanA = malloc(sizeof(double) * (size_t)na);
  double *statA = malloc(sizeof(double) * 4 * (size_t)nsa);
  double *cleanB = NULL, *statB = NULL;
  if (!cleanA || !statA) return 8;
  if (ora_stats(a, na, m, cleanA, statA, statA + nsa, statA + 2 * nsa,
                statA + 3 * nsa))
    return 7;
  if (self_join) {
    cleanB = cleanA;
    statB = statA;
  } else {
    cleanB = malloc(sizeof(double) * (size_t)nb);
    statB = malloc(sizeof(double) * 4 * (size_t)nsb);
    if (!cleanB || !statB) return 8;
    if (ora_stats(b, nb, m, cleanB, statB, statB + nsb, statB + 2 * nsb,
                  statB + 3 * nsb))
      return 7;
  }
  s.cleanA = cleanA;
  s.cleanB = cleanB;
  s.meanA = statA;
  s.normA = statA + nsa;
  s.dfA = statA + 2 * nsa;
  s.dgA = statA + 3 * nsa;
  s.meanB = statB;
  s.normB = statB + nsb;
  s.dfB = statB + 2 * nsb;
  s.dgB = statB + 3 * nsb;

  s.QT = malloc(sizeof(double) * (size_t)ts);
  s.ta.corr = malloc(sizeof(double) * (size_t)s.tile_w);
  s.ta.idx = malloc(sizeof(int64_t) * (size_t)s.tile_w);
  s.tb.corr = malloc(sizeof(double) * (size_t)s.tile_h);
  s.tb.idx = malloc(sizeof(int64_t) * (size_t)s.tile_h);
  for (int64_t i = 0; i < s.tile_h; ++i) {
    s.tb.corr[i] = s.ptype == ORA_SUM_THRESH ? 0.0 : -2.0;
    s.tb.idx[i] = -1;
  }
  wide_prof ga = {corr_a, idx_a}, gb = {corr_b, idx_b};

  /* get_tiles (src/SCAMP.cpp:25-90) */
  int64_t nrows = (int64_t)ceil((double)(nsb) / (double)s.tile_h);
  int64_t ncols = (int64_t)ceil((double)(nsa) / (double)s.tile_w);
  int64_t cap = nrows * ncols + 4, nt = 0;
  int64_t *tr = malloc(sizeof(int64_t) * (size_t)cap);
  int64_t *tc = malloc(sizeof(int64_t) * (size_t)cap);
#define PUSH(r, c) \
  do {             \
    tr[nt] = (r);  \
    tc[nt] = (c);  \
    ++nt;          \
  } while (0)
  if (self_join) {
    for (int64_t off = 0; off < nrows - 1; ++off)
      for (int64_t d = 0; d < ncols - 1 - off; ++d) PUSH(d, d + off);
    for (int64_t i = 0; i < nrows; ++i) PUSH(i, ncols - 1);
  } else {
    for (int64_t d = 0; d < ncols - 1; ++d)
      for (int64_t off = 0; off + d < ncols - 1 && off < nrows - 1; ++off)
        PUSH(off, d + off);
    for (int64_t d = 1; d < nrows - 1; ++d)
      for (int64_t off = 0; off + d < nrows - 1 && off < ncols - 1; ++off)
        PUSH(off + d, off);
    PUSH(nrows - 1, ncols - 1);
    int64_t x = 0, y = 0;
    while (x < ncols - 1 && y < nrows - 1) {
      PUSH(y, ncols - 1);
      PUSH(nrows - 1, x);
      ++x;
      ++y;
    }
    while (x < ncols - 1) {
      PUSH(nrows - 1, x);
      ++x;
    }
    while (y < nrows - 1) {
      PUSH(y, ncols - 1);
      ++y;
    }
  }
#undef PUSH

  /* Worker loop (src/SCAMP.cpp:106-165) and Tile::MergeProfile
   * (src/tile.cpp:473-544). */
  int err = 0;
  for (int64_t t = 0; t < nt && !err; ++t) {
    s.trow = tr[t] * s.tile_h;
    s.tcol = tc[t] * s.tile_w;
    s.width = s.tile_ts < na - s.tcol ? s.tile_ts : na - s.tcol;
    s.height = s.tile_ts < nb - s.trow ? s.tile_ts : nb - s.trow;
    if (m > s.width || m > s.height) {
      err = 6; /* SCAMP_DIM_INCOMPATIBLE (src/tile.cpp:608-613) */
      break;
    }
    init_tile_profile(&s, ga, gb);
    execute_tile(&s, self_join && tr[t] == tc[t]);
    int64_t w = s.width - m + 1, h = s.height - m + 1;
    merge(&s, ga, s.ta, s.tcol, w, s.trow);
    if (s.rows && s.keep_rows) {
      merge(&s, gb, s.tb, s.trow, h, s.tcol);
    } else if (self_join && s.rows) {
      merge(&s, ga, s.tb, s.trow, h, s.tcol);
    }
  }

  free(tr);
  free(tc);
  free(s.QT);
  free(s.ta.corr);
  free(s.ta.idx);
  free(s.tb.corr);
  free(s.tb.idx);
  if (!self_join) {
    free(cleanB);
    free(statB);
  }
  free(cleanA);
  free(statA);
  return err;
}
