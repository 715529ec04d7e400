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
// (windows with a non-finite sample or flat windows) never match.
//
// How it runs here (MI355X design, not the reference's):
//  * Work unit = a wave band: 64 lanes x 8 consecutive diagonals = 512
//    diagonals, swept down all their rows by one wavefront.  A persistent grid
//    pulls bands from an atomic counter, longest first.
//  * Each band starts at a pseudo-random row r0 (its FP64 seed computed
//    exactly for that row) and wraps round to row 0, so the bands in flight
//    sit at spread-out rows and each row's visits are spread over the run.
//  * Per row a lane evaluates its 8 pairs from registers: the 8 covariances,
//    a sliding window of 8 columns' update terms / inverse norms / thresholds
//    (two register sets in ping-pong, refilled one column per row with loads
//    issued a block of 8 rows before use) and the row's terms (uniform).
//  * 1NN / 1NN_INDEX: instead of a max+argmax reduction for every pair, each
//    correlation is tested against the current best of its row and of its
//    column (kept in the profile itself, read as lowered thresholds).  Only
//    when some lane passes does the wave take the slow path: exact FP32
//    rounding, lowest-index argmax within the wave, and 64-bit atomicMax of
//    order-preserving keys straight into the profile.  The profile only grows,
//    so stale thresholds only cost extra slow paths, never a missed match.
//  * SUM_THRESH: row sums reduced across the wave per row; column partial
//    sums carried lane to lane (a column leaves lane L into lane L-1) so each
//    column reaches memory once per band.
//  * MATRIX_SUMMARY: exact max-pooling; each lane keeps two running cell
//    maxima (its 8 columns straddle at most one cell boundary) and flushes a
//    cell with a 32-bit atomicMax when its columns or the row leave it.
#include "device_common.h"
#include "kernels.h"

namespace scamp_amd {
namespace {

constexpr int K = kDiagsPerLane;

template <typename DT>
struct Slot {
  DT df, dg, in;   // column update terms and inverse norm
  DT t;            // 1NN: lowered threshold of the column's current best
  uint32_t raw;    // 1NN: key high word as loaded, decoded one row later
  double acc;      // SUM: this lane's partial column sum
};

template <typename DT>
struct Cols {
  Slot<DT> s[K];
};

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

// Per-lane state of the MATRIX_SUMMARY kinds.
struct CellAcc {
  double lo, hi;       // running maxima of the current cell and the next one
  int64_t lo_cell;     // cell coordinate (along the lane's columns) of `lo`
  int64_t pos;         // (first column of the lane) mod cells-per-step
};

template <typename DT, int KIND, bool COLS, bool ROWS>
struct Sweep {
  const SweepArgs &a;
  const int lane;
  const int64_t D0;      // first diagonal of the band
  const int64_t d0;      // first diagonal of this lane
  DT cov[K];
  // MATRIX_SUMMARY state
  CellAcc dir, tra;
  int64_t dir_row_cell, dir_row_pos;  // uniform: row cell and row mod rpc
  int64_t tra_col_cell, tra_col_pos;  // uniform: column cell and row mod cpc

  __device__ Sweep(const SweepArgs &args, int ln, int64_t band_d0)
      : a(args), lane(ln), D0(band_d0), d0(band_d0 + (int64_t)ln * K) {}

  __device__ __forceinline__ const DT *dfX() const { return (const DT *)a.dfX; }
  __device__ __forceinline__ const DT *dgX() const { return (const DT *)a.dgX; }
  __device__ __forceinline__ const DT *inX() const { return (const DT *)a.inX; }
  __device__ __forceinline__ const DT *dfY() const { return (const DT *)a.dfY; }
  __device__ __forceinline__ const DT *dgY() const { return (const DT *)a.dgY; }
  __device__ __forceinline__ const DT *inY() const { return (const DT *)a.inY; }

  // Issue the loads of one column into a slot (the column index is always
  // inside the padded arrays).
  __device__ __forceinline__ void load_slot(Slot<DT> &s, int64_t col) const {
    s.df = dfX()[col];
    s.dg = dgX()[col];
    s.in = inX()[col];
    if (KIND == kNN && COLS)
      s.raw = reinterpret_cast<const uint32_t *>(a.keyX)[2 * col + 1];
    if (KIND == kSum) s.acc = 0.0;
  }
  __device__ __forceinline__ void decode_slot(Slot<DT> &s) const {
    if (KIND == kNN && COLS) s.t = Thresh<DT>::from_key_hi(s.raw);
  }

  // ---------------------------------------------------------------- seeds
  // FP32 form: exact FP64 covariance of this lane's 8 diagonals at row i,
  //   cov_k = sum_t X[i+d_k+t]*yc_t - muX[i+d_k]*sum_t yc_t,
  //   yc_t = Y[i+t] - muY[i]
  // swept as a sliding window: each loaded X sample feeds 8 products.
  __device__ __forceinline__ void anchor(int64_t i) {
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    const double *x = a.TX + i + d0;
    const double *y = a.TY + i;
    const double my = a.muY[i];
    const int64_t m = a.m;
    double sy = 0.0;
    double yw[2 * K - 1];  // yc_t for t = u0-K+1 .. u0+K-1
#pragma unroll
    for (int q = 0; q < 2 * K - 1; ++q) yw[q] = 0.0;
    for (int64_t u0 = 0; u0 < m + K - 1; u0 += K) {
#pragma unroll
      for (int q = 0; q < K - 1; ++q) yw[q] = yw[q + K];
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const int64_t t = u0 + q;
        const double v = t < m ? y[t] - my : 0.0;
        yw[K - 1 + q] = v;
        sy += v;
      }
      double xv[K];
#pragma unroll
      for (int q = 0; q < K; ++q) xv[q] = x[u0 + q];
#pragma unroll
      for (int q = 0; q < K; ++q)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = fma(xv[q], yw[q - k + K - 1], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t d = d0 + k;
      const double v = acc[k] - a.muX[i + d] * sy;
      cov[k] = d < a.d_hi ? (DT)v : nan_of<DT>();
    }
  }

  __device__ __forceinline__ void seed_from(const double *src) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int64_t d = d0 + k;
      cov[k] = d < a.d_hi ? (DT)src[k] : nan_of<DT>();
    }
  }

  // ------------------------------------------------------ matrix summary
  __device__ __forceinline__ void cell_flush(double v, int64_t r, int64_t c) {
    if (v > a.thr)
      atomicMax(&a.mat[r * a.mat_w + c], enc_corr((float)v));
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

  // -------------------------------------------------- 1NN slow path
  __device__ __forceinline__ void nn_row_update(const DT (&c)[K], int64_t i,
                                                DT trow) {
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
    key = wave_max_u64(key);
    if (lane == 0 && key != 0)
      atomicMax(reinterpret_cast<unsigned long long *>(a.keyY + i),
                static_cast<unsigned long long>(key));
  }

  // ------------------------------------------------------------ one row
  template <bool MASKED>
  __device__ __forceinline__ void row(int r, int64_t i, Cols<DT> &cur,
                                      Cols<DT> &nxt, DT trow) {
    // the slot first used on the next row: decode its threshold now
    decode_slot(nxt.s[r]);
    const DT rdf = dfY()[i], rdg = dgY()[i], rin = inY()[i];
    DT c[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Slot<DT> &S = (r + k < K) ? cur.s[r + k] : nxt.s[r + k - K];
      c[k] = cov[k] * S.in * rin;
      if (MASKED && i + d0 + k >= a.nX) c[k] = nan_of<DT>();
    }
    if (KIND == kNN) {
      bool hit = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const Slot<DT> &S = (r + k < K) ? cur.s[r + k] : nxt.s[r + k - K];
        if (COLS) hit |= c[k] >= S.t;
        if (ROWS) hit |= c[k] >= trow;
      }
      if (__builtin_amdgcn_ballot_w64(hit)) {
        if (ROWS) nn_row_update(c, i, trow);
        if (COLS) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            Slot<DT> &S = (r + k < K) ? cur.s[r + k] : nxt.s[r + k - K];
            if (c[k] >= S.t) {
              const float f = (float)c[k];
              const uint64_t key = make_key(f, static_cast<uint32_t>(i));
              const unsigned long long old = atomicMax(
                  reinterpret_cast<unsigned long long *>(a.keyX + i + d0 + k),
                  static_cast<unsigned long long>(key));
              const uint64_t best = old > key ? old : key;
              S.t = Thresh<DT>::from_key_hi(static_cast<uint32_t>(best >> 32));
            }
          }
        }
      }
    } else if (KIND == kSum) {
      double rs = 0.0;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        Slot<DT> &S = (r + k < K) ? cur.s[r + k] : nxt.s[r + k - K];
        const double v = c[k] > (DT)a.thr ? (double)c[k] : 0.0;
        if (ROWS) rs += v;
        if (COLS) S.acc += v;
      }
      if (ROWS && __builtin_amdgcn_ballot_w64(rs != 0.0)) {
        rs = wave_sum(rs);
        if (lane == 0) atomicAdd(a.sumY + i, rs);
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
    // advance the recurrence to row i+1
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Slot<DT> &S = (r + k < K) ? cur.s[r + k] : nxt.s[r + k - K];
      cov[k] = fma(S.df, rdg, cov[k]);
      cov[k] = fma(rdf, S.dg, cov[k]);
    }
    // cur.s[r] held column i + d0 (diagonal 0 of the lane): it leaves this
    // lane now.
    if (KIND == kSum && COLS) {
      const double ex = cur.s[r].acc;
      const double rcv = __shfl_down(ex, 1);
      nxt.s[r].acc = lane == kWave - 1 ? 0.0 : rcv;
      if (lane == 0 && ex != 0.0) atomicAdd(a.sumX + i + D0, ex);
    }
    if (KIND == kMatrix) cells_row_end();
  }

  // Flushes what a lane still holds when its segment ends after row i_last
  // of the block that started at i_b (cur/nxt as in that block).
  __device__ __forceinline__ void segment_end(int64_t i_b, int r_last,
                                              Cols<DT> &cur, Cols<DT> &nxt) {
    if (KIND == kSum && COLS) {
#pragma unroll
      for (int s = 0; s < K; ++s) {
        if (s > r_last && cur.s[s].acc != 0.0)
          atomicAdd(a.sumX + i_b + d0 + s, cur.s[s].acc);
        if (s <= r_last && nxt.s[s].acc != 0.0)
          atomicAdd(a.sumX + i_b + K + d0 + s, nxt.s[s].acc);
      }
    }
    if (KIND == kMatrix) cells_flush_all();
  }

  // One block of K rows starting at i_b; cur/nxt swap roles between blocks.
  template <bool MASKED>
  __device__ __forceinline__ bool block(int64_t i_b, int64_t re, Cols<DT> &cur,
                                        Cols<DT> &nxt, DT *s_rowt,
                                        bool first) {
    if (KIND == kNN && ROWS) {
      if (lane < K) {
        int64_t ri = i_b + lane;
        ri = ri < a.nY ? ri : a.nY - 1;
        s_rowt[lane] = Thresh<DT>::from_key_hi(
            reinterpret_cast<const uint32_t *>(a.keyY)[2 * ri + 1]);
      }
      __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
      const int64_t i = i_b + r;
      if (MASKED && i >= re) return true;
      if (KIND == kMatrix) cells_row_start(first && r == 0);
      const DT trow = (KIND == kNN && ROWS) ? s_rowt[r] : (DT)0;
      row<MASKED>(r, i, cur, nxt, trow);
      if (MASKED && i + 1 == re) {
        segment_end(i_b, r, cur, nxt);
        return true;
      }
      // refill: cur.s[r] becomes slot r of the next block's look-ahead set
      load_slot(cur.s[r], i_b + 2 * K + d0 + r);
    }
    if (KIND == kNN && ROWS) __builtin_amdgcn_wave_barrier();
    return false;
  }

  // Rows [rs, re) of the band; all band diagonals are valid for rows < rfull.
  __device__ __forceinline__ void segment(int64_t rs, int64_t re,
                                          int64_t rfull, DT *s_rowt) {
    if (rs >= re) return;
    Cols<DT> A, B;
#pragma unroll
    for (int s = 0; s < K; ++s) {
      load_slot(A.s[s], rs + d0 + s);
      load_slot(B.s[s], rs + d0 + K + s);
    }
#pragma unroll
    for (int s = 0; s < K; ++s) decode_slot(A.s[s]);
    if (KIND == kMatrix) cells_begin(rs);
    int64_t i = rs;
    const int64_t fast_end = re < rfull ? re : rfull;
    bool first = true;
    // The last row always falls in the masked tail below, which flushes.
    while (i + 2 * K < re && i + 2 * K <= fast_end) {
      if (sizeof(DT) == 4 && i != rs && ((i - rs) % a.anchor) == 0) anchor(i);
      block<false>(i, re, A, B, s_rowt, first);
      first = false;
      block<false>(i + K, re, B, A, s_rowt, false);
      i += 2 * K;
    }
    for (;;) {
      if (block<true>(i, re, A, B, s_rowt, first)) return;
      first = false;
      i += K;
      if (block<true>(i, re, B, A, s_rowt, false)) return;
      i += K;
    }
  }
};

template <typename DT, int KIND, bool COLS, bool ROWS>
__global__ __launch_bounds__(kWave) void sweep_kernel(SweepArgs a) {
  __shared__ DT s_rowt[K];
  const int lane = threadIdx.x;
  for (;;) {
    unsigned int b = 0;
    if (lane == 0) b = atomicAdd(a.counter, 1u);
    b = (unsigned int)__builtin_amdgcn_readfirstlane((int)b);
    if (b >= a.nbands) break;
    const int64_t D0 = a.d_lo + (int64_t)b * kBand;
    Sweep<DT, KIND, COLS, ROWS> sw(a, lane, D0);
    const int64_t dlast = D0 + kBand < a.d_hi ? D0 + kBand - 1 : a.d_hi - 1;
    const int64_t rows0 = a.nY < a.nX - D0 ? a.nY : a.nX - D0;
    const int64_t rfull0 = a.nY < a.nX - dlast ? a.nY : a.nX - dlast;
    const int64_t r0 = a.band_r0[b];
    // segment 0: rows [r0, rows0) from the band's own seed at row r0;
    // segment 1: rows [0, r0) from the row-0 seeds (all diagonals valid).
    for (int seg = 0; seg < 2; ++seg) {
      const int64_t rs = seg == 0 ? r0 : 0;
      const int64_t re = seg == 0 ? rows0 : r0;
      const int64_t rfull = seg == 0 ? rfull0 : r0;
      if (rs >= re) continue;
      if (sizeof(DT) == 8) {
        sw.seed_from(seg == 0 ? a.seedB + (int64_t)b * kBand + (int64_t)lane * K
                              : a.seed0 + (sw.d0 - a.d_lo));
      } else {
        sw.anchor(rs);
      }
      sw.segment(rs, re, rfull, s_rowt);
    }
  }
}

template <typename DT, int KIND>
hipError_t launch_kind(const SweepArgs &a, bool cols, bool rows, int grid,
                       hipStream_t s) {
  if (cols && rows) {
    hipLaunchKernelGGL((sweep_kernel<DT, KIND, true, true>), dim3(grid),
                       dim3(kWave), 0, s, a);
  } else if (cols) {
    hipLaunchKernelGGL((sweep_kernel<DT, KIND, true, false>), dim3(grid),
                       dim3(kWave), 0, s, a);
  } else if (rows) {
    hipLaunchKernelGGL((sweep_kernel<DT, KIND, false, true>), dim3(grid),
                       dim3(kWave), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_sweep(const SweepArgs &a, bool fp32, SweepKind kind,
                        bool cols, bool rows, int grid, hipStream_t s) {
  if (!cols && !rows) return hipSuccess;
  if (fp32) {
    switch (kind) {
      case kNN: return launch_kind<float, kNN>(a, cols, rows, grid, s);
      case kSum: return launch_kind<float, kSum>(a, cols, rows, grid, s);
      case kMatrix: return launch_kind<float, kMatrix>(a, cols, rows, grid, s);
      default: return launch_kind<float, kMatrixSmall>(a, cols, rows, grid, s);
    }
  }
  switch (kind) {
    case kNN: return launch_kind<double, kNN>(a, cols, rows, grid, s);
    case kSum: return launch_kind<double, kSum>(a, cols, rows, grid, s);
    case kMatrix: return launch_kind<double, kMatrix>(a, cols, rows, grid, s);
    default: return launch_kind<double, kMatrixSmall>(a, cols, rows, grid, s);
  }
}

}  // namespace scamp_amd
