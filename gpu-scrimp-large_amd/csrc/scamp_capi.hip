// C ABI of the engine (include/scamp_hip.h): argument validation, device
// buffers, the statistics / seed / sweep launches of one join, and the
// conversion of profiles between SCAMP's boundary layouts and the device keys.
//
// One join replaces SCAMP_Operation::do_join (src/SCAMP.cpp:174-259): where
// the reference precomputes statistics on one host thread, cuts the distance
// matrix into tiles, seeds each tile with cuFFT and merges each tile's
// profile on the host under a mutex, this runs the statistics on the GPU,
// seeds every wave band exactly, sweeps the whole triangle (self-join) or both
// triangles (AB-join) with one persistent kernel per diagonal range, and
// merges in place with atomics; the profile stays in HBM until the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/scamp_hip.h"
#include "device_common.h"
#include "kernels.h"

using namespace scamp_amd;

// the prep kernels' MatchRec mirrors this layout
static_assert(sizeof(scamp_hip_match) == 12, "scamp_hip_match is 12 bytes");

// Working memory of scamp_hip_join_device_noalloc: one copy per GPU, made
// when the library's code object loads on that GPU and held until the
// process ends.  In host code this name is not a device address:
// scamp_hip_ctx_create looks up each device's copy with hipGetSymbolAddress,
// and the join uses that pointer.
constexpr size_t kScratchBytes = SCAMP_HIP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Calls on one context run one after another (each holds `mu`); separate
// contexts, on the same device or not, run independently.
struct scamp_hip_ctx {
  int device = 0;
  // the stream the context's work goes on: its own (`own`), or the caller's
  // after scamp_hip_ctx_set_stream, which also makes the device-buffer calls
  // return once their work is queued (`async`)
  hipStream_t stream = nullptr;
  hipStream_t own = nullptr;
  bool async = false;
  int num_cus = 0;
  float total_ms = 0, prep_ms = 0, sweep_ms = 0;
  // async: the last join's events are recorded but not yet read
  bool timing_pending = false;
  // the precision the last join ran in, and the entries whose FP32 sweep
  // value failed the mixed form's audit (then the join ran again in FP64)
  int last_precision = 0;
  uint64_t last_drift = 0;
  // a join's timing events (start, statistics done, end, sweep start, sweep
  // end), made once by scamp_hip_ctx_create
  hipEvent_t ev[5] = {};
  // this device's copy of the no-allocation join's scratch array (null when
  // its address could not be obtained; scratch_status holds why)
  void *scratch = nullptr;
  hipError_t scratch_status = hipSuccess;
  std::mutex mu;
  // Device blocks kept from earlier calls (bytes, pointer), reused by later
  // ones: a short join would otherwise spend a noticeable part of its time
  // in hipMalloc / hipFree.  Only blocks up to kCacheBlock bytes, at most
  // kCacheTotal bytes in all; freed by scamp_hip_ctx_destroy.
  std::vector<std::pair<size_t, void *>> cache;
  size_t cached = 0;
  static constexpr size_t kCacheBlock = size_t(256) << 20;
  static constexpr size_t kCacheTotal = size_t(2) << 30;
};

namespace {

thread_local std::string g_last_error;

// Each device's scratch address, looked up once (with that device current)
// by the first context made on it.  Joins through the scratch array on one
// device take turns under that device's lock: they share its memory.
constexpr int kMaxScratchDevices = 64;
struct ScratchSlot {
  std::once_flag once;
  void *ptr = nullptr;
  hipError_t status = hipErrorNotInitialized;
  std::mutex mu;
};
ScratchSlot g_scratch[kMaxScratchDevices];

// The mixed form's audit: an entry whose FP32 sweep correlation is further
// than this from the FP64 correlation of the same pair (checked by the
// refinement pass) means the FP32 recurrence left its error bound, and the
// join is recomputed in FP64.  While every pair's FP32 value is within
// kDriftTol of its FP64 one, the match the sweep keeps is within 2 kDriftTol
// = 1e-5 of the true best, and the refinement stores its exact value.  The
// compensated sweep stays below 1e-6 on well-conditioned series at the
// default anchor interval.
constexpr double kDriftTol = 5e-6;

// Rows between the mixed form's FP64 re-anchors when the caller leaves it
// to the engine: 2 m, within [512, 16384].  An anchor costs m / R FP64 FMAs
// per pair (at most 0.5 here); the FP32 drift grows with the rows since the
// last one, and series with loud stretches (the planted high-amplitude
// motifs of an AB-join, for one) need short intervals to stay within the
// audit's bound.  Callers that tune R for their data pass it explicitly.
int64_t anchor_rows(const scamp_hip_params *p) {
  int64_t r = p->anchor_interval > 0
                  ? p->anchor_interval
                  : std::min<int64_t>(16384, std::max<int64_t>(512, 2 * p->window));
  return std::max<int64_t>(16, (r + 15) / 16 * 16);
}
// join_device / selfjoin_batch: the audit failed; run again in FP64
constexpr int kRerunFp64 = -1;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

struct HipFailure {
  hipError_t err;
  const char *what;
};

#define HIP_CHECK(call)                                    \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) throw HipFailure{e_, #call};     \
  } while (0)

// Every working buffer is a whole number of 256-byte granules: blocks are
// reused across nearby sizes, and carved blocks stay 256-byte aligned.
template <typename T>
size_t granules(int64_t count) {
  return (sizeof(T) * (size_t)std::max<int64_t>(count, 1) + 255) / 256 * 256;
}

// Device allocations owned by one call (made with the context locked);
// returned to the context's cache, or freed, on scope exit.  Work that still
// reads a returned block is queued on the context's stream, ahead of any
// later call's use of it.
struct DevArena {
  scamp_hip_ctx *ctx;
  std::vector<std::pair<size_t, void *>> ptrs;
  explicit DevArena(scamp_hip_ctx *c) : ctx(c) {}
  DevArena(const DevArena &) = delete;
  DevArena &operator=(const DevArena &) = delete;
  template <typename T>
  T *alloc(int64_t count) {
    const size_t bytes = granules<T>(count);
    void *p = nullptr;
    // the smallest cached block that fits and wastes at most half of it
    size_t best = ctx->cache.size();
    for (size_t i = 0; i < ctx->cache.size(); ++i) {
      const size_t b = ctx->cache[i].first;
      if (b >= bytes && b <= 2 * bytes &&
          (best == ctx->cache.size() || b < ctx->cache[best].first))
        best = i;
    }
    size_t got = bytes;
    if (best < ctx->cache.size()) {
      got = ctx->cache[best].first;
      p = ctx->cache[best].second;
      ctx->cached -= got;
      ctx->cache.erase(ctx->cache.begin() + best);
    } else {
      HIP_CHECK(hipMalloc(&p, bytes));
    }
    ptrs.emplace_back(got, p);
    return static_cast<T *>(p);
  }
  // Blocks go back to the cache for the context's next call, which runs
  // after this one on the same stream; an asynchronous context waits for its
  // stream before it frees one that does not fit the cache.
  ~DevArena() {
    bool waited = !ctx->async;
    for (auto &b : ptrs) {
      if (b.first <= scamp_hip_ctx::kCacheBlock &&
          ctx->cached + b.first <= scamp_hip_ctx::kCacheTotal) {
        ctx->cache.push_back(b);
        ctx->cached += b.first;
      } else {
        if (!waited) {
          (void)hipStreamSynchronize(ctx->stream);
          waited = true;
        }
        (void)hipFree(b.second);
      }
    }
  }
};

// The no-allocation join's buffers: consecutive blocks of one device's
// scratch array, handed out in the order the join asks for them.
struct ScratchArena {
  char *base;
  size_t cap, used = 0;
  ScratchArena(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
  template <typename T>
  T *alloc(int64_t count) {
    const size_t bytes = granules<T>(count);
    // the entry point checks the join's size first (SizeArena); this is
    // the last guard against writing past the array
    if (bytes > cap - used)
      throw HipFailure{hipErrorOutOfMemory, "scratch array exhausted"};
    T *p = reinterpret_cast<T *>(base + used);
    used += bytes;
    return p;
  }
};

// Counts the bytes a join's buffers take, through the same allocation
// sequence (alloc_join) that hands them out.
struct SizeArena {
  size_t bytes = 0;
  template <typename T>
  T *alloc(int64_t count) {
    bytes += granules<T>(count);
    return nullptr;
  }
};

// Statistics of one series, in the padded device layout the sweep reads.
struct SeriesStats {
  int64_t n = 0, ns = 0;
  double *clean = nullptr, *mean = nullptr, *inorm = nullptr, *df = nullptr,
         *dg = nullptr;
  float *f_inorm = nullptr, *f_df = nullptr, *f_dg = nullptr;
  // the FP64 sweep's per-column first-stage factors (both precisions: a
  // mixed join whose audit fails runs again in FP64)
  float *rq = nullptr;
};

template <class Arena>
SeriesStats alloc_stats(Arena &ar, int64_t n, int64_t m, bool fp32) {
  SeriesStats st;
  st.n = n;
  st.ns = n - m + 1;
  const int64_t stat_len = st.ns + kColumnPad;
  st.clean = ar.template alloc<double>(n + kColumnPad);
  st.mean = ar.template alloc<double>(stat_len);
  st.inorm = ar.template alloc<double>(stat_len);
  st.df = ar.template alloc<double>(stat_len);
  st.dg = ar.template alloc<double>(stat_len);
  st.rq = ar.template alloc<float>(stat_len);
  if (fp32) {
    st.f_inorm = ar.template alloc<float>(stat_len);
    st.f_df = ar.template alloc<float>(stat_len);
    st.f_dg = ar.template alloc<float>(stat_len);
  }
  return st;
}

void compute_stats(const SeriesStats &st, const double *d_T, int64_t m,
                   hipStream_t s) {
  const int64_t stat_len = st.ns + kColumnPad;
  HIP_CHECK(launch_stats(d_T, st.n, m, st.clean, st.n + kColumnPad, st.mean,
                         st.inorm, st.df, st.dg, stat_len, s));
  HIP_CHECK(launch_stage1_factors(st.inorm, st.rq, stat_len, s));
  if (st.f_inorm) {
    HIP_CHECK(launch_to_f32(st.inorm, st.f_inorm, stat_len, s));
    HIP_CHECK(launch_to_f32(st.df, st.f_df, stat_len, s));
    HIP_CHECK(launch_to_f32(st.dg, st.f_dg, stat_len, s));
  }
}

SeriesStats make_stats(DevArena &ar, const double *d_T, int64_t n, int64_t m,
                       bool fp32, hipStream_t s) {
  SeriesStats st = alloc_stats(ar, n, m, fp32);
  compute_stats(st, d_T, m, s);
  return st;
}

// [lo, hi) minus the open interval (xlo, xhi): at most two ranges.
struct Ranges {
  int count = 0;
  int64_t lo[2] = {}, hi[2] = {};
  void add(int64_t l, int64_t h) {
    lo[count] = l;
    hi[count] = h;
    ++count;
  }
};
Ranges cut(int64_t lo, int64_t hi, int64_t xlo, int64_t xhi) {
  Ranges out;
  if (lo >= hi) return out;
  // excluded integers: xlo < d < xhi
  int64_t ex0 = xlo + 1, ex1 = xhi;  // [ex0, ex1)
  if (ex0 >= ex1 || ex1 <= lo || ex0 >= hi) {
    out.add(lo, hi);
    return out;
  }
  if (lo < ex0) out.add(lo, ex0);
  if (ex1 < hi) out.add(ex1, hi);
  return out;
}

uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

// End of a join: a blocking context waits for the end event and records
// the join's times; an asynchronous one (scamp_hip_ctx_set_stream) records
// the event and leaves reading the times to scamp_hip_last_timing.  Returns
// whether the times are known.  ev[0] / ev[2] bracket the join, ev[3] /
// ev[4] its sweep when asynchronous (a blocking join adds up its passes'
// sweep times as they finish).
bool finish_join(scamp_hip_ctx *ctx, float sweep_ms) {
  HIP_CHECK(hipEventRecord(ctx->ev[2], ctx->stream));
  if (ctx->async) {
    ctx->timing_pending = true;
    return false;
  }
  HIP_CHECK(hipEventSynchronize(ctx->ev[2]));
  float total = 0;
  HIP_CHECK(hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[2]));
  ctx->total_ms = total;
  ctx->sweep_ms = sweep_ms;
  ctx->prep_ms = total - sweep_ms;
  ctx->timing_pending = false;
  return true;
}

// Reads the times an asynchronous join left pending (waits for its end).
void resolve_timing(scamp_hip_ctx *ctx) {
  if (!ctx->timing_pending) return;
  HIP_CHECK(hipEventSynchronize(ctx->ev[2]));
  float total = 0, sweep = 0;
  HIP_CHECK(hipEventElapsedTime(&total, ctx->ev[0], ctx->ev[2]));
  HIP_CHECK(hipEventElapsedTime(&sweep, ctx->ev[3], ctx->ev[4]));
  ctx->total_ms = total;
  ctx->sweep_ms = sweep;
  ctx->prep_ms = total - sweep;
  ctx->timing_pending = false;
}

struct Pass {
  const SeriesStats *X, *Y;
  int64_t d_lo, d_hi;
  bool cols, rows;
  uint64_t *keyX, *keyY;
  double *sumX, *sumY;
  uint64_t *knnX, *knnY;  // ALL_NEIGHBORS lists
};

// Row chunks up to this length run the 1NN sweep's short-chunk variant.
constexpr int64_t kFreshMaxRows = 16384;

// Passes of one join: a self-join has one; an AB-join up to two per
// triangle (an aligned join's exclusion band can split a diagonal range).
constexpr int kMaxPasses = 4;

// Runs one pass (one diagonal range of one orientation); `counter` is its
// work counter (one word, cleared here).  sweep_ms: adds the sweep's
// HIP-event time (null: no events, no wait for the sweep).
void run_pass(scamp_hip_ctx *ctx, uint64_t *counter, const Pass &p, int64_t m,
              bool fp32, SweepKind kind, uint32_t *mat, int64_t mat_w,
              int64_t rpc, int64_t cpc, double thr, int64_t anchor, int knn_k,
              float *sweep_ms, int64_t nbatch = 1, int64_t stride = 0,
              int64_t ns_each = 0) {
  // nbatch > 1: nbatch self-joins of series of ns_each subsequences stored
  // with `stride` entries between series (p.X == p.Y holds all of them)
  hipStream_t s = ctx->stream;
  const int64_t nX = nbatch > 1 ? ns_each : p.X->ns;
  const int64_t nY = nbatch > 1 ? ns_each : p.Y->ns;
  const int64_t ndiag = p.d_hi - p.d_lo;
  if (ndiag <= 0 || (!p.cols && !p.rows)) return;
  // Bands start on multiples of kDiagsPerLane (the sweep indexes its column
  // slots by column mod 8); diagonals below d_lo in the first band are NaN.
  const int64_t d_base = p.d_lo - p.d_lo % kDiagsPerLane;
  const int64_t nbands = (p.d_hi - d_base + kBand - 1) / kBand;

  // Work items (decoded by the sweep itself, plan_item in sweep_impl.h).
  // Band b spans rows [0, rows0) (rows with at least one valid diagonal);
  // all of its diagonals are valid below rfull.  Its rows are cut at
  // r0 + j * R: r0 is a hashed multiple of 8 below rfull, so that items
  // running at the same time work on different rows' profile entries; R is
  // unbounded when the pass has enough bands to fill the GPU, otherwise
  // short enough that the items do, and for the longest band never below
  // max(512, 4m) rows so the segment seeds stay cheap next to the rows they
  // start; shorter bands take proportionally shorter chunks
  // (band_chunk_rows).
  auto band_rows = [&](int64_t b, int64_t *rfull) {
    const int64_t D0 = d_base + b * kBand;
    const int64_t dfirst = std::max(D0, p.d_lo);
    const int64_t dlast = std::min(D0 + kBand, p.d_hi) - 1;
    *rfull = std::min(nY, nX - dlast);
    return std::min(nY, nX - dfirst);
  };
  int64_t total_rows = 0;
  for (int64_t b = 0; b < nbands; ++b) {
    int64_t rf;
    total_rows += std::max<int64_t>(0, band_rows(b, &rf));
  }
  // (a batch fills the GPU with its series' bands alone once there are
  // enough of them; the chunk length itself depends on one series only)
  const int64_t wave_slots = (int64_t)ctx->num_cus * 12;
  int64_t R = std::numeric_limits<int64_t>::max();
  // SCAMP_HIP_NO_CHUNKS (tests): never cut bands into chunks, so that short
  // passes run the kernels of long ones
  const char *no_chunks = std::getenv("SCAMP_HIP_NO_CHUNKS");
  const bool may_chunk = no_chunks == nullptr || *no_chunks == '\0' || *no_chunks == '0';
  if (may_chunk && nbands * nbatch < 2 * wave_slots) {
    const int64_t floor_rows = std::max<int64_t>(512, 4 * m);
    R = std::max(floor_rows, (total_rows + 4 * wave_slots - 1) / (4 * wave_slots));
    R = (R + 15) / 16 * 16;
  }
  const bool chunked = R != std::numeric_limits<int64_t>::max();
  int64_t nchunks = 1;
  // shorter bands take proportionally shorter chunks, down to rmin rows
  // (at least the window, so a chunk's seed stays cheap next to its rows)
  const int64_t rmin = std::min<int64_t>(R, (std::max<int64_t>(128, m) + 15) / 16 * 16);
  int64_t ref_rows = 0;
  if (chunked) {
    for (int64_t b = 0; b < nbands; ++b) {
      int64_t rfull;
      ref_rows = std::max(ref_rows, band_rows(b, &rfull));
    }
    // the largest chunk count of any band (the kernel skips the items a
    // shorter band does not have); the same cuts as plan_item
    for (int64_t b = 0; b < nbands; ++b) {
      int64_t rfull;
      const int64_t rows0 = band_rows(b, &rfull);
      if (rows0 <= 0 || rfull <= 0) continue;
      const int64_t slots = (rfull - 1) / kDiagsPerLane + 1;
      const int64_t r0 = (int64_t)(mix64((uint64_t)b + 0x9e3779b97f4a7c15ULL) %
                                   (uint64_t)slots) * kDiagsPerLane;
      const int64_t Rb = band_chunk_rows(R, rmin, ref_rows, rows0);
      int64_t first = r0 % Rb;
      if (first == 0) first = Rb;
      const int64_t n = (first < rows0 ? (rows0 - 1 - first) / Rb + 1 : 0) + 1;
      nchunks = std::max(nchunks, n);
    }
  }
  const int64_t nitems = nbands * nbatch * nchunks;
  if (nitems >= (int64_t)0xFFFFFFFFu - (int64_t)ctx->num_cus * 16)
    throw HipFailure{hipErrorInvalidValue, "too many sweep work items"};

  // the work counter, cleared by a fill kernel (no memset node)
  HIP_CHECK(launch_fill_u64(counter, 1, 0, s));

  SweepArgs a{};
  if (fp32) {
    a.dfX = p.X->f_df; a.dgX = p.X->f_dg; a.inX = p.X->f_inorm;
    a.dfY = p.Y->f_df; a.dgY = p.Y->f_dg; a.inY = p.Y->f_inorm;
  } else {
    a.dfX = p.X->df; a.dgX = p.X->dg; a.inX = p.X->inorm;
    a.dfY = p.Y->df; a.dgY = p.Y->dg; a.inY = p.Y->inorm;
  }
  a.rqX = p.X->rq;
  a.TX = p.X->clean; a.TY = p.Y->clean; a.muX = p.X->mean; a.muY = p.Y->mean;
  a.nitems = nitems; a.nchunks = nchunks; a.chunk_rows = chunked ? R : 0;
  a.chunk_min_rows = rmin; a.chunk_ref_rows = ref_rows;
  a.nX = nX; a.nY = nY; a.d_lo = p.d_lo; a.d_hi = p.d_hi; a.d_base = d_base;
  a.m = m; a.anchor = anchor;
  a.keyX = p.keyX; a.keyY = p.keyY; a.sumX = p.sumX; a.sumY = p.sumY;
  a.knnX = p.knnX; a.knnY = p.knnY; a.knn_k = knn_k;
  a.mat = mat; a.mat_w = mat_w; a.rpc = rpc; a.cpc = cpc; a.thr = thr;
  a.counter = reinterpret_cast<unsigned int *>(counter);
  a.nbatch = nbatch;
  a.batch_stride = stride;

  const int grid = (int)std::min<int64_t>(nitems, (int64_t)ctx->num_cus * 16);
  // Short row chunks start with the profile far below its final values, so
  // many of their rows merge: the 1NN sweep then runs its variant with
  // block-level merge tables (the long-chunk kernel keeps fewer registers).
  const bool fresh = chunked && R <= kFreshMaxRows;
  if (sweep_ms == nullptr) {
    HIP_CHECK(launch_sweep(a, fp32, kind, p.cols, p.rows, grid, s, fresh));
    return;
  }
  HIP_CHECK(hipEventRecord(ctx->ev[3], s));
  HIP_CHECK(launch_sweep(a, fp32, kind, p.cols, p.rows, grid, s, fresh));
  HIP_CHECK(hipEventRecord(ctx->ev[4], s));
  HIP_CHECK(hipEventSynchronize(ctx->ev[4]));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev[3], ctx->ev[4]));
  *sweep_ms += ms;
}

const char *validate(const scamp_hip_params *p, int64_t na, int64_t nb) {
  // SCAMPArgs::validate (src/common.cpp:417-436)
  if (p->window < 3) return "Error: Subsequence length must be at least 3";
  if (p->max_tile_size < 1024)
    return "Error: max tile size must be at least 1024";
  if (p->max_tile_size / 2 < p->window)
    return "Error: Tile length and width must be at least 2x larger than the "
           "window size";
  if (na < p->window || (p->has_b && nb < p->window))
    return "Error: Input time series must be at least 'subesequence window "
           "size' in length";
  // The sweep addresses a column's terms and profile key with 32-bit byte
  // offsets (8-byte entries, padded arrays).
  if (p->profile_type == SCAMP_HIP_ALL_NEIGHBORS &&
      (p->max_matches_per_column < 1 || p->max_matches_per_column > 256))
    return "Error: max_matches_per_column must be between 1 and 256";
  if (na - p->window + 1 > kMaxSubsequences ||
      (p->has_b && nb - p->window + 1 > kMaxSubsequences))
    return "Error: time series too long for this engine (at most 536868736 "
           "subsequences per series)";
  return nullptr;
}

// What one join's buffers depend on.
struct JoinShape {
  int64_t na = 0, nb = 0, m = 0;
  bool self = true, fp32 = false, sep_rows = false, refine = false;
  int pt = 0, knn_k = 0;
  int64_t H = 0, W = 0;
};

// Checks the parameters a join adds to validate() and fills `js` (nb = na
// for a self-join).  Returns SCAMP_HIP_OK, or fail()'s code and message.
int join_shape(const scamp_hip_params *p, int64_t na, int64_t nb, JoinShape *js) {
  const int pt = p->profile_type;
  if (pt != SCAMP_HIP_1NN_INDEX && pt != SCAMP_HIP_1NN &&
      pt != SCAMP_HIP_SUM_THRESH && pt != SCAMP_HIP_MATRIX_SUMMARY &&
      pt != SCAMP_HIP_ALL_NEIGHBORS)
    return fail(SCAMP_HIP_UNIMPLEMENTED,
                "profile type not supported by this engine");
  const int prec = p->precision;
  if (prec != SCAMP_HIP_SINGLE && prec != SCAMP_HIP_MIXED &&
      prec != SCAMP_HIP_DOUBLE && prec != SCAMP_HIP_ULTRA)
    return fail(SCAMP_HIP_INVALID_ARGS, "invalid precision type");
  if (pt == SCAMP_HIP_MATRIX_SUMMARY &&
      (p->matrix_height <= 0 || p->matrix_width <= 0))
    return fail(SCAMP_HIP_INVALID_ARGS, "matrix height and width must be > 0");
  js->m = p->window;
  js->self = !p->has_b;
  js->na = na;
  js->nb = js->self ? na : nb;
  js->fp32 = prec == SCAMP_HIP_MIXED || prec == SCAMP_HIP_SINGLE;
  js->pt = pt;
  js->sep_rows = p->computing_rows && p->keep_rows_separate;
  // The mixed form's 1NN_INDEX profiles are refined in FP64 after the sweep
  // (refine_kernel.hip); entries still equal to the prior are left alone.
  js->refine = js->fp32 && pt == SCAMP_HIP_1NN_INDEX;
  js->knn_k = (int)p->max_matches_per_column;
  js->H = p->matrix_height;
  js->W = p->matrix_width;
  return SCAMP_HIP_OK;
}

// The working buffers of one join, taken from `ar` in one fixed order
// before any kernel writes the caller's memory.
struct JoinBuffers {
  SeriesStats A, B;
  // working profiles (padded); ALL_NEIGHBORS lists; the matrix summary
  uint64_t *keyA = nullptr, *keyB = nullptr, *knnA = nullptr, *knnB = nullptr;
  uint32_t *mat = nullptr;
  // the mixed form's priors and audit counter; one work counter per pass
  uint64_t *priorA = nullptr, *priorB = nullptr, *drift = nullptr;
  uint64_t *counters = nullptr;
};

template <class Arena>
JoinBuffers alloc_join(Arena &ar, const JoinShape &js) {
  JoinBuffers b;
  const int64_t nsA = js.na - js.m + 1, nsB = js.nb - js.m + 1;
  b.A = alloc_stats(ar, js.na, js.m, js.fp32);
  b.B = js.self ? b.A : alloc_stats(ar, js.nb, js.m, js.fp32);
  if (js.pt == SCAMP_HIP_ALL_NEIGHBORS) {
    b.keyA = ar.template alloc<uint64_t>(nsA + kColumnPad);
    b.knnA = ar.template alloc<uint64_t>(nsA * js.knn_k);
    if (js.sep_rows) {
      b.keyB = ar.template alloc<uint64_t>(nsB + kColumnPad);
      b.knnB = ar.template alloc<uint64_t>(nsB * js.knn_k);
    }
  } else if (js.pt == SCAMP_HIP_1NN_INDEX || js.pt == SCAMP_HIP_1NN) {
    b.keyA = ar.template alloc<uint64_t>(nsA + kColumnPad);
    if (js.sep_rows) b.keyB = ar.template alloc<uint64_t>(nsB + kColumnPad);
  } else if (js.pt == SCAMP_HIP_MATRIX_SUMMARY) {
    b.mat = ar.template alloc<uint32_t>(js.H * js.W);
  }
  if (js.refine) {
    b.drift = ar.template alloc<uint64_t>(1);
    b.priorA = ar.template alloc<uint64_t>(nsA);
    if (js.sep_rows) b.priorB = ar.template alloc<uint64_t>(nsB);
  }
  b.counters = ar.template alloc<uint64_t>(kMaxPasses);
  return b;
}

// Bytes of device memory a join's working buffers take.
size_t join_bytes(const JoinShape &js) {
  SizeArena sz;
  (void)alloc_join(sz, js);
  return sz.bytes;
}

template <class Arena>
int join_device(scamp_hip_ctx *ctx, Arena &ar, const double *d_a, int64_t na,
                const double *d_b, int64_t nb, const scamp_hip_params *p,
                void *d_prof_a, void *d_prof_b) {
  JoinShape js;
  if (int rc = join_shape(p, na, nb, &js)) return rc;
  const int64_t m = js.m;
  const bool self = js.self;
  if (self) d_b = d_a;
  const int pt = js.pt;
  const int prec = p->precision;
  const bool fp32 = js.fp32;
  const int64_t anchor = anchor_rows(p);

  hipStream_t s = ctx->stream;
  hipEvent_t ev_start = ctx->ev[0], ev_prep = ctx->ev[1], ev_end = ctx->ev[2];
  HIP_CHECK(hipEventRecord(ev_start, s));

  JoinBuffers bf = alloc_join(ar, js);
  const SeriesStats &A = bf.A, &B = bf.B;
  compute_stats(A, d_a, m, s);
  if (!self) compute_stats(B, d_b, m, s);
  const int64_t nsA = A.ns, nsB = B.ns;
  const bool rows_out = p->computing_rows && (self || p->keep_rows_separate);
  const bool cols_out = p->computing_columns != 0;
  const bool sep_rows = js.sep_rows;

  // Working profiles (padded) initialised from the caller's prior.
  uint64_t *keyA = bf.keyA, *keyB = bf.keyB, *knnA = bf.knnA, *knnB = bf.knnB;
  double *sumA = nullptr, *sumB = nullptr;
  uint32_t *mat = bf.mat;
  const int64_t H = js.H, W = js.W;
  const int knn_k = js.knn_k;
  // ALL_NEIGHBORS: a match enters a list when its key beats `knn_init`,
  // i.e. when its FP32 correlation is above the threshold
  const uint64_t knn_init = make_key((float)p->distance_threshold, 0u);
  if (pt == SCAMP_HIP_ALL_NEIGHBORS) {
    HIP_CHECK(launch_fill_u64(keyA + nsA, kColumnPad, knn_init, s));
    HIP_CHECK(launch_matches_to_knn(d_prof_a, nsA, knn_k, knn_init, knnA, keyA, s));
    if (sep_rows) {
      HIP_CHECK(launch_fill_u64(keyB + nsB, kColumnPad, knn_init, s));
      HIP_CHECK(launch_matches_to_knn(d_prof_b, nsB, knn_k, knn_init, knnB, keyB, s));
    }
  } else if (pt == SCAMP_HIP_1NN_INDEX || pt == SCAMP_HIP_1NN) {
    const uint64_t init = make_key(-2.0f, 0xFFFFFFFFu);
    HIP_CHECK(launch_fill_u64(keyA + nsA, kColumnPad, init, s));
    if (pt == SCAMP_HIP_1NN_INDEX) {
      HIP_CHECK(hipMemcpyAsync(keyA, d_prof_a, sizeof(uint64_t) * nsA,
                               hipMemcpyDeviceToDevice, s));
      HIP_CHECK(launch_mp_to_keys(keyA, nsA, s));
    } else {
      HIP_CHECK(launch_f32_to_keys((const float *)d_prof_a, keyA, nsA, s));
    }
    if (sep_rows) {
      HIP_CHECK(launch_fill_u64(keyB + nsB, kColumnPad, init, s));
      if (pt == SCAMP_HIP_1NN_INDEX) {
        HIP_CHECK(hipMemcpyAsync(keyB, d_prof_b, sizeof(uint64_t) * nsB,
                                 hipMemcpyDeviceToDevice, s));
        HIP_CHECK(launch_mp_to_keys(keyB, nsB, s));
      } else {
        HIP_CHECK(launch_f32_to_keys((const float *)d_prof_b, keyB, nsB, s));
      }
    }
  } else if (pt == SCAMP_HIP_SUM_THRESH) {
    // SCAMP adds tile sums into the caller's profile (src/common.cpp:167-173)
    sumA = static_cast<double *>(d_prof_a);
    if (sep_rows) sumB = static_cast<double *>(d_prof_b);
  } else {
    HIP_CHECK(hipMemcpyAsync(mat, d_prof_a, sizeof(float) * H * W,
                             hipMemcpyDeviceToDevice, s));
    HIP_CHECK(launch_f32_to_enc(reinterpret_cast<float *>(mat), H * W, s));
  }
  const bool refine = js.refine;
  uint64_t *priorA = bf.priorA, *priorB = bf.priorB, *drift = bf.drift;
  if (refine) {
    HIP_CHECK(launch_fill_u64(drift, 1, 0, s));
    HIP_CHECK(hipMemcpyAsync(priorA, keyA, sizeof(uint64_t) * nsA,
                             hipMemcpyDeviceToDevice, s));
    if (sep_rows)
      HIP_CHECK(hipMemcpyAsync(priorB, keyB, sizeof(uint64_t) * nsB,
                               hipMemcpyDeviceToDevice, s));
  }
  HIP_CHECK(hipEventRecord(ev_prep, s));

  // Passes.  Self-join: diagonals [m/4, nsA) of the upper triangle, each pair
  // updating its row and its column in the same profile (src/tile.cpp:12-38
  // keeps |i - j| < m/4 out).  AB-join: the upper triangle (columns = A,
  // rows = B, diagonals >= 0) and the lower one transposed (columns = B,
  // rows = A, diagonals >= 1), as do_ab_join_full does per tile
  // (src/tile.cpp:645-707); aligned AB-joins drop |global col - global row| <
  // m/4 (src/tile.cpp:44-138).
  const int64_t e = m / 4;
  SweepKind kind = kNN;
  int64_t rpc = 1, cpc = 1;
  if (pt == SCAMP_HIP_SUM_THRESH) kind = kSum;
  if (pt == SCAMP_HIP_ALL_NEIGHBORS) kind = kKnn;
  if (pt == SCAMP_HIP_MATRIX_SUMMARY) {
    cpc = (nsA + W - 1) / W;
    rpc = (nsB + H - 1) / H;
    kind = (cpc >= kDiagsPerLane && rpc >= kDiagsPerLane) ? kMatrix
                                                          : kMatrixSmall;
  }
  // Self-join 1NN profiles start from each subsequence's best match among
  // the windows just past the exclusion zone (pairs of the join, so the
  // result is the same; the sweep's thresholds then start near their final
  // values, which short series would otherwise spend much of their rows on).
  // Those diagonals are then done: were the sweep to evaluate them again,
  // every row of its first band would meet its warm-start match exactly at
  // its threshold and take the merge path.
  const bool warm = self && kind == kNN && (cols_out || rows_out);
  if (warm)
    HIP_CHECK(launch_warm_profile(A.clean, A.mean, A.inorm, nsA, m, e, 1, 0,
                                  cols_out ? keyA : nullptr,
                                  rows_out ? (sep_rows ? keyB : keyA) : nullptr, s));
  Pass passes[kMaxPasses];
  int npass = 0;
  if (self) {
    const Ranges r = cut(warm ? e + kWarmLags : e, nsA, 0, 0);
    for (int i = 0; i < r.count; ++i) {
      Pass ps{&A, &A, r.lo[i], r.hi[i], cols_out, rows_out, keyA,
              sep_rows ? keyB : keyA, sumA, sep_rows ? sumB : sumA,
              knnA, sep_rows ? knnB : knnA};
      // MATRIX_SUMMARY of a self-join: the cells of the pairs (row i, column
      // j > i), on and above the diagonal, as the reference's
      // upper-triangular tiles give them with rows switched off
      // (src/SCAMP.cpp:34-43, src/common.cpp:98-106); cells below stay empty
      if (pt == SCAMP_HIP_MATRIX_SUMMARY) {
        ps.cols = true;
        ps.rows = false;
      }
      passes[npass++] = ps;
    }
  } else {
    int64_t delta = 0;
    if (p->distributed_start_col >= 0 && p->distributed_start_row >= 0)
      delta = p->distributed_start_col - p->distributed_start_row;
    const bool al = p->is_aligned != 0 && e > 0;
    // upper: d = j - i; excluded when |d + delta| < e
    const Ranges up = al ? cut(0, nsA, -e - delta, e - delta) : cut(0, nsA, 0, 0);
    for (int i = 0; i < up.count; ++i) {
      Pass ps{&A, &B, up.lo[i], up.hi[i], cols_out, sep_rows, keyA, keyB,
              sumA, sumB, knnA, knnB};
      if (pt == SCAMP_HIP_MATRIX_SUMMARY) {
        ps.cols = true;
        ps.rows = false;
      }
      passes[npass++] = ps;
    }
    // lower (transposed): d = i - j >= 1; excluded when |d - delta| < e
    const Ranges lo = al ? cut(1, nsB, delta - e, delta + e) : cut(1, nsB, 0, 0);
    for (int i = 0; i < lo.count; ++i) {
      Pass ps{&B, &A, lo.lo[i], lo.hi[i], sep_rows, cols_out, keyB, keyA,
              sumB, sumA, knnB, knnA};
      if (pt == SCAMP_HIP_MATRIX_SUMMARY) {
        ps.cols = false;
        ps.rows = true;
      }
      passes[npass++] = ps;
    }
  }
  float sweep_ms = 0;
  if (ctx->async) HIP_CHECK(hipEventRecord(ctx->ev[3], s));
  for (int i = 0; i < npass; ++i)
    run_pass(ctx, bf.counters + i, passes[i], m, fp32, kind, mat, W, rpc, cpc,
             p->distance_threshold, anchor, knn_k, ctx->async ? nullptr : &sweep_ms);
  if (ctx->async) HIP_CHECK(hipEventRecord(ctx->ev[4], s));

  if (refine) {
    // profile A holds A's windows matched in B (self: in A); profile B the
    // reverse.  Excluded pairs as in the passes above.
    int64_t zone = e, delta = 0;
    if (!self) {
      if (p->distributed_start_col >= 0 && p->distributed_start_row >= 0)
        delta = p->distributed_start_col - p->distributed_start_row;
      zone = (p->is_aligned != 0 && e > 0) ? e : 0;
    }
    // a self-join's profile A holds column pairs (partners below), row pairs
    // (partners above), or both, as the passes above wrote it
    int32_t side_a = 0, side_b = 0;
    if (self) {
      side_a = cols_out && rows_out && !sep_rows ? 0 : (cols_out ? -1 : 1);
      side_b = 1;
    }
    HIP_CHECK(launch_refine(keyA, priorA, nsA, A.clean, A.mean, A.inorm, A.df,
                            A.dg, nsB, B.clean, B.mean, B.inorm, B.df, B.dg, m,
                            RefineExclusion{zone, delta, true, side_a}, 1, 0, drift,
                            kDriftTol, s));
    if (sep_rows)
      HIP_CHECK(launch_refine(keyB, priorB, nsB, B.clean, B.mean, B.inorm, B.df,
                              B.dg, nsA, A.clean, A.mean, A.inorm, A.df, A.dg, m,
                              RefineExclusion{zone, delta, false, side_b}, 1, 0, drift,
                              kDriftTol, s));
    uint64_t h_drift = 0;
    HIP_CHECK(hipMemcpyAsync(&h_drift, drift, sizeof(h_drift), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    ctx->last_drift = h_drift;
    if (h_drift > 0) {
      // the caller's profiles are untouched so far (the result lives in the
      // working keys): the join runs again in FP64
      float spent = 0;
      HIP_CHECK(hipEventRecord(ev_end, s));
      HIP_CHECK(hipEventSynchronize(ev_end));
      HIP_CHECK(hipEventElapsedTime(&spent, ev_start, ev_end));
      ctx->total_ms = spent;
      ctx->sweep_ms = sweep_ms;
      if (!p->silent_mode)
        std::printf("scamp_hip: the mixed form's FP32 sweep drifted beyond %.0e on "
                    "%llu entries; recomputing in FP64\n",
                    kDriftTol, (unsigned long long)h_drift);
      return kRerunFp64;
    }
  }

  // Back to SCAMP's layouts.
  if (pt == SCAMP_HIP_1NN_INDEX) {
    HIP_CHECK(launch_keys_to_mp(keyA, nsA, s));
    HIP_CHECK(hipMemcpyAsync(d_prof_a, keyA, sizeof(uint64_t) * nsA,
                             hipMemcpyDeviceToDevice, s));
    if (sep_rows) {
      HIP_CHECK(launch_keys_to_mp(keyB, nsB, s));
      HIP_CHECK(hipMemcpyAsync(d_prof_b, keyB, sizeof(uint64_t) * nsB,
                               hipMemcpyDeviceToDevice, s));
    }
  } else if (pt == SCAMP_HIP_1NN) {
    HIP_CHECK(launch_keys_to_f32(keyA, (float *)d_prof_a, nsA, s));
    if (sep_rows) HIP_CHECK(launch_keys_to_f32(keyB, (float *)d_prof_b, nsB, s));
  } else if (pt == SCAMP_HIP_ALL_NEIGHBORS) {
    HIP_CHECK(launch_knn_to_matches(knnA, nsA, knn_k, knn_init, d_prof_a, s));
    if (sep_rows)
      HIP_CHECK(launch_knn_to_matches(knnB, nsB, knn_k, knn_init, d_prof_b, s));
  } else if (pt == SCAMP_HIP_MATRIX_SUMMARY) {
    HIP_CHECK(launch_enc_to_f32(reinterpret_cast<float *>(mat), H * W, s));
    HIP_CHECK(hipMemcpyAsync(d_prof_a, mat, sizeof(float) * H * W,
                             hipMemcpyDeviceToDevice, s));
  }
  ctx->last_precision = prec;
  if (!fp32) ctx->last_drift = 0;
  if (finish_join(ctx, sweep_ms) && !p->silent_mode)
    std::printf("scamp_hip: join of %lld x %lld subsequences in %.3f ms "
                "(sweep %.3f ms)\n",
                (long long)nsA, (long long)nsB, ctx->total_ms, sweep_ms);
  return SCAMP_HIP_OK;
}

// `batch` independent self-joins of series of length n stored back to back
// in d_T, in one persistent sweep whose work items are (series, band).  Each
// series gets its own padded segment of `seg` entries in every per-series
// array (cleaned samples, statistics, profile keys), so the sweep's reads and
// writes past a series' last subsequence land in that series' padding.
int selfjoin_batch(scamp_hip_ctx *ctx, const double *d_T, int64_t n,
                   int64_t batch, const scamp_hip_params *p, void *d_prof) {
  const int64_t m = p->window;
  const int pt = p->profile_type;
  const int prec = p->precision;
  const bool fp32 = prec == SCAMP_HIP_MIXED || prec == SCAMP_HIP_SINGLE;
  const int64_t anchor = anchor_rows(p);
  const int64_t ns = n - m + 1;
  const int64_t seg = (n + kColumnPad + 63) / 64 * 64;
  hipStream_t s = ctx->stream;
  DevArena ar(ctx);
  hipEvent_t ev_start = ctx->ev[0], ev_end = ctx->ev[2];
  HIP_CHECK(hipEventRecord(ev_start, s));

  SeriesStats A;
  A.n = n;
  A.ns = ns;
  A.clean = ar.alloc<double>(batch * seg);
  A.mean = ar.alloc<double>(batch * seg);
  A.inorm = ar.alloc<double>(batch * seg);
  A.df = ar.alloc<double>(batch * seg);
  A.dg = ar.alloc<double>(batch * seg);
  HIP_CHECK(launch_stats(d_T, n, m, A.clean, seg, A.mean, A.inorm, A.df, A.dg,
                         seg, s, batch, n));
  A.rq = ar.alloc<float>(batch * seg);
  HIP_CHECK(launch_stage1_factors(A.inorm, A.rq, batch * seg, s));
  if (fp32) {
    A.f_inorm = ar.alloc<float>(batch * seg);
    A.f_df = ar.alloc<float>(batch * seg);
    A.f_dg = ar.alloc<float>(batch * seg);
    HIP_CHECK(launch_to_f32(A.inorm, A.f_inorm, batch * seg, s));
    HIP_CHECK(launch_to_f32(A.df, A.f_df, batch * seg, s));
    HIP_CHECK(launch_to_f32(A.dg, A.f_dg, batch * seg, s));
  }

  // keys[q * seg + j]: subsequence j of series q, from the prior profiles;
  // the padding holds "no match"
  uint64_t *key = ar.alloc<uint64_t>(batch * seg);
  HIP_CHECK(launch_fill_u64(key, batch * seg, make_key(-2.0f, 0xFFFFFFFFu), s));
  const bool mp = pt == SCAMP_HIP_1NN_INDEX;
  HIP_CHECK(launch_profile_in(d_prof, mp, key, ns, batch, seg, s));
  // mixed 1NN_INDEX: FP64 refinement after the sweep, as in a single join
  const bool refine = fp32 && mp;
  uint64_t *prior = nullptr, *drift = nullptr;
  if (refine) {
    drift = ar.alloc<uint64_t>(1);
    HIP_CHECK(launch_fill_u64(drift, 1, 0, s));
    prior = ar.alloc<uint64_t>(batch * seg);
    HIP_CHECK(hipMemcpyAsync(prior, key, sizeof(uint64_t) * batch * seg,
                             hipMemcpyDeviceToDevice, s));
  }

  // diagonals [m/4, ns) of each series (src/tile.cpp:12-38 keeps |i - j| <
  // m/4 out), rows and columns into one profile as the flags ask, after the
  // same warm start as a single self-join
  const bool cols = p->computing_columns != 0, rows = p->computing_rows != 0;
  HIP_CHECK(launch_warm_profile(A.clean, A.mean, A.inorm, ns, m, m / 4, batch,
                                seg, cols ? key : nullptr, rows ? key : nullptr,
                                s));
  uint64_t *counters = ar.alloc<uint64_t>(kMaxPasses);
  float sweep_ms = 0;
  // the warm start has done diagonals [m/4, m/4 + kWarmLags)
  const Ranges r = cut(m / 4 + ((cols || rows) ? kWarmLags : 0), ns, 0, 0);
  if (ctx->async) HIP_CHECK(hipEventRecord(ctx->ev[3], s));
  for (int i = 0; i < r.count; ++i) {
    Pass ps{&A, &A, r.lo[i], r.hi[i], cols, rows, key, key,
            nullptr, nullptr, nullptr, nullptr};
    run_pass(ctx, counters + i, ps, m, fp32, kNN, nullptr, 0, 1, 1, 0.0, anchor,
             0, ctx->async ? nullptr : &sweep_ms, batch, seg, ns);
  }
  if (ctx->async) HIP_CHECK(hipEventRecord(ctx->ev[4], s));
  if (refine) {
    HIP_CHECK(launch_refine(key, prior, ns, A.clean, A.mean, A.inorm, A.df, A.dg,
                            ns, A.clean, A.mean, A.inorm, A.df, A.dg, m,
                            RefineExclusion{m / 4, 0, true,
                                            cols && rows ? 0 : (cols ? -1 : 1)},
                            batch, seg, drift, kDriftTol, s));
    uint64_t h_drift = 0;
    HIP_CHECK(hipMemcpyAsync(&h_drift, drift, sizeof(h_drift), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    ctx->last_drift = h_drift;
    if (h_drift > 0) {
      float spent = 0;
      HIP_CHECK(hipEventRecord(ev_end, s));
      HIP_CHECK(hipEventSynchronize(ev_end));
      HIP_CHECK(hipEventElapsedTime(&spent, ev_start, ev_end));
      ctx->total_ms = spent;
      ctx->sweep_ms = sweep_ms;
      if (!p->silent_mode)
        std::printf("scamp_hip: the mixed form's FP32 sweep drifted beyond %.0e on "
                    "%llu entries; recomputing in FP64\n",
                    kDriftTol, (unsigned long long)h_drift);
      return kRerunFp64;
    }
  }

  HIP_CHECK(launch_profile_out(key, seg, d_prof, mp, ns, batch, s));
  ctx->last_precision = prec;
  if (!fp32) ctx->last_drift = 0;
  if (finish_join(ctx, sweep_ms) && !p->silent_mode)
    std::printf("scamp_hip: %lld self-joins of %lld subsequences in %.3f ms "
                "(sweep %.3f ms)\n",
                (long long)batch, (long long)ns, ctx->total_ms, sweep_ms);
  return SCAMP_HIP_OK;
}

// Runs `run(params)`; when the mixed form's audit asks for it, runs again
// with the precision set to DOUBLE, and reports the time of both runs.
template <typename Run>
int with_fp64_rerun(scamp_hip_ctx *ctx, const scamp_hip_params *p, Run run) {
  int rc = run(p);
  if (rc != kRerunFp64) return rc;
  const float t0 = ctx->total_ms, s0 = ctx->sweep_ms;
  const uint64_t drift = ctx->last_drift;
  scamp_hip_params p64 = *p;
  p64.precision = SCAMP_HIP_DOUBLE;
  rc = run(&p64);
  resolve_timing(ctx);  // an asynchronous rerun's times (reruns are rare)
  ctx->total_ms += t0;
  ctx->sweep_ms += s0;
  ctx->prep_ms = ctx->total_ms - ctx->sweep_ms;
  ctx->last_drift = drift;
  return rc;
}

size_t profile_elem_bytes(int pt) {
  switch (pt) {
    case SCAMP_HIP_1NN_INDEX: return sizeof(uint64_t);
    case SCAMP_HIP_SUM_THRESH: return sizeof(double);
    case SCAMP_HIP_ALL_NEIGHBORS: return sizeof(scamp_hip_match);
    default: return sizeof(float);
  }
}

}  // namespace

extern "C" {

int scamp_hip_device_count(int *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  if (count) *count = n;
  return SCAMP_HIP_OK;
}

int scamp_hip_ctx_create(int device, scamp_hip_ctx **out) {
  if (!out) return fail(SCAMP_HIP_INVALID_ARGS, "null output pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    return fail(SCAMP_HIP_NO_DEVICE, "no HIP device is visible");
  if (device < 0 || device >= n)
    return fail(SCAMP_HIP_INVALID_ARGS, "device id out of range");
  auto *ctx = new scamp_hip_ctx();
  ctx->device = device;
  try {
    HIP_CHECK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, device));
    ctx->num_cus = prop.multiProcessorCount;
    HIP_CHECK(hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking));
    ctx->stream = ctx->own;
    for (hipEvent_t &e : ctx->ev) HIP_CHECK(hipEventCreate(&e));
  } catch (const HipFailure &f) {
    scamp_hip_ctx_destroy(ctx);
    return fail(SCAMP_HIP_DEVICE_ERROR,
                std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  // this device's scratch array (its own copy, at its own address)
  if (device < kMaxScratchDevices) {
    ScratchSlot &sl = g_scratch[device];
    std::call_once(sl.once, [&sl] {
      void *ptr = nullptr;
      sl.status = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (sl.status == hipSuccess) sl.ptr = ptr;
    });
    ctx->scratch = sl.ptr;
    ctx->scratch_status = sl.status;
  } else {
    ctx->scratch_status = hipErrorInvalidDevice;
  }
  *out = ctx;
  return SCAMP_HIP_OK;
}

int scamp_hip_ctx_destroy(scamp_hip_ctx *ctx) {
  if (!ctx) return SCAMP_HIP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &b : ctx->cache) (void)hipFree(b.second);
  for (hipEvent_t e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->own) (void)hipStreamDestroy(ctx->own);
  delete ctx;
  return SCAMP_HIP_OK;
}

int scamp_hip_ctx_set_stream(scamp_hip_ctx *ctx, void *stream) {
  if (!ctx) return fail(SCAMP_HIP_INVALID_ARGS, "null context");
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t want = stream ? static_cast<hipStream_t>(stream) : ctx->own;
    if (want != ctx->stream) {
      // the cached blocks may still be in use by work on the old stream
      HIP_CHECK(hipStreamSynchronize(ctx->stream));
      resolve_timing(ctx);
      ctx->stream = want;
    }
    ctx->async = stream != nullptr;
  } catch (const HipFailure &f) {
    return fail(SCAMP_HIP_DEVICE_ERROR,
                std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  return SCAMP_HIP_OK;
}

int scamp_hip_last_error(char *buf, size_t len) {
  if (buf && len) {
    std::strncpy(buf, g_last_error.c_str(), len - 1);
    buf[len - 1] = 0;
  }
  return SCAMP_HIP_OK;
}

int scamp_hip_join_device(scamp_hip_ctx *ctx, const double *d_a, size_t na,
                          const double *d_b, size_t nb,
                          const scamp_hip_params *p, void *d_profile_a,
                          void *d_profile_b) {
  if (!ctx || !p || !d_a || !d_profile_a)
    return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  if (p->has_b && !d_b) return fail(SCAMP_HIP_INVALID_ARGS, "missing series b");
  if (const char *msg = validate(p, (int64_t)na, (int64_t)nb))
    return fail(SCAMP_HIP_INVALID_ARGS, msg);
  if (p->computing_rows && p->keep_rows_separate && !d_profile_b &&
      p->profile_type != SCAMP_HIP_MATRIX_SUMMARY)
    return fail(SCAMP_HIP_INVALID_ARGS, "missing profile b");
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    return with_fp64_rerun(ctx, p, [&](const scamp_hip_params *q) {
      DevArena ar(ctx);
      return join_device(ctx, ar, d_a, (int64_t)na, d_b, (int64_t)nb, q,
                         d_profile_a, d_profile_b);
    });
  } catch (const HipFailure &f) {
    (void)hipStreamSynchronize(ctx->stream);
    int code = f.err == hipErrorOutOfMemory ? SCAMP_HIP_OUT_OF_MEMORY
                                            : SCAMP_HIP_DEVICE_ERROR;
    return fail(code, std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
}

int scamp_hip_join_device_noalloc(scamp_hip_ctx *ctx, const double *d_a,
                                  size_t na, const double *d_b, size_t nb,
                                  const scamp_hip_params *p, void *d_profile_a,
                                  void *d_profile_b) {
  if (!ctx || !p || !d_a || !d_profile_a)
    return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  if (p->has_b && !d_b) return fail(SCAMP_HIP_INVALID_ARGS, "missing series b");
  if (const char *msg = validate(p, (int64_t)na, (int64_t)nb))
    return fail(SCAMP_HIP_INVALID_ARGS, msg);
  if (p->computing_rows && p->keep_rows_separate && !d_profile_b &&
      p->profile_type != SCAMP_HIP_MATRIX_SUMMARY)
    return fail(SCAMP_HIP_INVALID_ARGS, "missing profile b");
  JoinShape js;
  if (int rc = join_shape(p, (int64_t)na, (int64_t)nb, &js)) return rc;
  char msg[192];
  if (!ctx->scratch) {
    std::snprintf(msg, sizeof msg,
                  "no-allocation join: no scratch array on device %d (%s)",
                  ctx->device, hipGetErrorString(ctx->scratch_status));
    return fail(SCAMP_HIP_DEVICE_ERROR, msg);
  }
  // the requested precision's buffers; an FP64 rerun after a failed audit
  // needs fewer (no FP32 terms, no priors)
  const size_t need = join_bytes(js);
  if (need > kScratchBytes) {
    std::snprintf(msg, sizeof msg,
                  "no-allocation join: needs %zu bytes of working memory, the "
                  "scratch array holds %zu",
                  need, kScratchBytes);
    return fail(SCAMP_HIP_INVALID_ARGS, msg);
  }
  try {
    std::lock_guard<std::mutex> dev_lock(g_scratch[ctx->device].mu);
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    const int rc = with_fp64_rerun(ctx, p, [&](const scamp_hip_params *q) {
      ScratchArena ar(ctx->scratch, kScratchBytes);
      return join_device(ctx, ar, d_a, (int64_t)na, d_b, (int64_t)nb, q,
                         d_profile_a, d_profile_b);
    });
    // every call on this GPU shares the scratch array: even on a context
    // bound to a caller's stream, the join is finished before the lock goes
    if (ctx->async) HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return rc;
  } catch (const HipFailure &f) {
    (void)hipStreamSynchronize(ctx->stream);
    int code = f.err == hipErrorOutOfMemory ? SCAMP_HIP_OUT_OF_MEMORY
                                            : SCAMP_HIP_DEVICE_ERROR;
    return fail(code, std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
}

int scamp_hip_selfjoin_batch(scamp_hip_ctx *ctx, const double *d_T,
                              size_t n, int64_t batch,
                              const scamp_hip_params *p, void *d_profiles) {
  if (!ctx || !p || !d_T || !d_profiles)
    return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  if (batch < 1) return fail(SCAMP_HIP_INVALID_ARGS, "batch must be >= 1");
  if (p->has_b)
    return fail(SCAMP_HIP_INVALID_ARGS, "a batched join is a self-join");
  if (p->keep_rows_separate)
    return fail(SCAMP_HIP_INVALID_ARGS,
                "a batched self-join keeps rows and columns in one profile");
  if (p->profile_type != SCAMP_HIP_1NN_INDEX &&
      p->profile_type != SCAMP_HIP_1NN)
    return fail(SCAMP_HIP_UNIMPLEMENTED,
                "batched joins support 1NN_INDEX and 1NN profiles");
  if (p->precision != SCAMP_HIP_SINGLE && p->precision != SCAMP_HIP_MIXED &&
      p->precision != SCAMP_HIP_DOUBLE && p->precision != SCAMP_HIP_ULTRA)
    return fail(SCAMP_HIP_INVALID_ARGS, "invalid precision type");
  if (const char *msg = validate(p, (int64_t)n, 0))
    return fail(SCAMP_HIP_INVALID_ARGS, msg);
  // (series, band) work items are counted in 32 bits by the sweep
  const int64_t bands = ((int64_t)n - p->window + 1) / kBand + 2;
  if (batch > (int64_t)(0x7FFFFFFF / bands))
    return fail(SCAMP_HIP_INVALID_ARGS, "batch too large for one launch");
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    return with_fp64_rerun(ctx, p, [&](const scamp_hip_params *q) {
      return selfjoin_batch(ctx, d_T, (int64_t)n, batch, q, d_profiles);
    });
  } catch (const HipFailure &f) {
    (void)hipStreamSynchronize(ctx->stream);
    int code = f.err == hipErrorOutOfMemory ? SCAMP_HIP_OUT_OF_MEMORY
                                            : SCAMP_HIP_DEVICE_ERROR;
    return fail(code, std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
}

int scamp_hip_join(scamp_hip_ctx *ctx, const double *a, size_t na,
                   const double *b, size_t nb, const scamp_hip_params *p,
                   void *profile_a, void *profile_b) {
  if (!ctx || !p || !a || !profile_a)
    return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  if (p->has_b && !b) return fail(SCAMP_HIP_INVALID_ARGS, "missing series b");
  if (const char *msg = validate(p, (int64_t)na, (int64_t)nb))
    return fail(SCAMP_HIP_INVALID_ARGS, msg);
  const int64_t m = p->window;
  const size_t eb = profile_elem_bytes(p->profile_type);
  const bool mat = p->profile_type == SCAMP_HIP_MATRIX_SUMMARY;
  // entries per subsequence (ALL_NEIGHBORS: k matches)
  const size_t per = p->profile_type == SCAMP_HIP_ALL_NEIGHBORS
                         ? (size_t)p->max_matches_per_column : 1;
  const size_t len_a =
      mat ? (size_t)(p->matrix_height * p->matrix_width) : (na - m + 1) * per;
  const bool sep = !mat && p->computing_rows && p->keep_rows_separate;
  const size_t len_b = sep ? ((p->has_b ? nb : na) - m + 1) * per : 0;
  if (sep && !profile_b) return fail(SCAMP_HIP_INVALID_ARGS, "missing profile b");
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    DevArena ar(ctx);
    hipStream_t s = ctx->stream;
    double *d_a = ar.alloc<double>((int64_t)na);
    HIP_CHECK(hipMemcpyAsync(d_a, a, sizeof(double) * na, hipMemcpyHostToDevice, s));
    double *d_b = nullptr;
    if (p->has_b) {
      d_b = ar.alloc<double>((int64_t)nb);
      HIP_CHECK(hipMemcpyAsync(d_b, b, sizeof(double) * nb, hipMemcpyHostToDevice, s));
    }
    void *d_pa = ar.alloc<char>((int64_t)(eb * len_a));
    HIP_CHECK(hipMemcpyAsync(d_pa, profile_a, eb * len_a, hipMemcpyHostToDevice, s));
    void *d_pb = nullptr;
    if (sep) {
      d_pb = ar.alloc<char>((int64_t)(eb * len_b));
      HIP_CHECK(hipMemcpyAsync(d_pb, profile_b, eb * len_b, hipMemcpyHostToDevice, s));
    }
    int rc = with_fp64_rerun(ctx, p, [&](const scamp_hip_params *q) {
      DevArena work(ctx);
      return join_device(ctx, work, d_a, (int64_t)na, d_b, (int64_t)nb, q, d_pa,
                         d_pb);
    });
    if (rc != SCAMP_HIP_OK) return rc;
    HIP_CHECK(hipMemcpyAsync(profile_a, d_pa, eb * len_a, hipMemcpyDeviceToHost, s));
    if (sep)
      HIP_CHECK(hipMemcpyAsync(profile_b, d_pb, eb * len_b, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (const HipFailure &f) {
    (void)hipStreamSynchronize(ctx->stream);
    int code = f.err == hipErrorOutOfMemory ? SCAMP_HIP_OUT_OF_MEMORY
                                            : SCAMP_HIP_DEVICE_ERROR;
    return fail(code, std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  return SCAMP_HIP_OK;
}

int scamp_hip_stats(scamp_hip_ctx *ctx, const double *T, size_t n,
                    int64_t window, double *means, double *inorms, double *df,
                    double *dg) {
  if (!ctx || !T || !means || !inorms || !df || !dg)
    return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  if (window < 1 || (int64_t)n < window)
    return fail(SCAMP_HIP_INVALID_ARGS, "series shorter than the window");
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    DevArena ar(ctx);
    hipStream_t s = ctx->stream;
    double *d_T = ar.alloc<double>((int64_t)n);
    HIP_CHECK(hipMemcpyAsync(d_T, T, sizeof(double) * n, hipMemcpyHostToDevice, s));
    SeriesStats st = make_stats(ar, d_T, (int64_t)n, window, false, s);
    const size_t bytes = sizeof(double) * st.ns;
    HIP_CHECK(hipMemcpyAsync(means, st.mean, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(inorms, st.inorm, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(df, st.df, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(dg, st.dg, bytes, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (const HipFailure &f) {
    return fail(SCAMP_HIP_DEVICE_ERROR,
                std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  return SCAMP_HIP_OK;
}

int scamp_hip_qt_seed(scamp_hip_ctx *ctx, const double *X, size_t nx,
                      const double *Y, size_t ny, int64_t window, int64_t row,
                      int64_t diag0, int64_t count, double *qt) {
  if (!ctx || !X || !Y || !qt) return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  const int64_t m = window;
  if (m < 1 || (int64_t)nx < m || (int64_t)ny < m || row < 0 || diag0 < 0 ||
      count < 0 || row + m > (int64_t)ny ||
      row + diag0 + count - 1 + m > (int64_t)nx)
    return fail(SCAMP_HIP_INVALID_ARGS, "seed window outside the series");
  if (count == 0) return SCAMP_HIP_OK;
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    DevArena ar(ctx);
    hipStream_t s = ctx->stream;
    double *dX = ar.alloc<double>((int64_t)nx);
    double *dY = ar.alloc<double>((int64_t)ny);
    HIP_CHECK(hipMemcpyAsync(dX, X, sizeof(double) * nx, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(dY, Y, sizeof(double) * ny, hipMemcpyHostToDevice, s));
    SeriesStats sx = make_stats(ar, dX, (int64_t)nx, m, false, s);
    SeriesStats sy = make_stats(ar, dY, (int64_t)ny, m, false, s);
    std::vector<SeedTask> tasks;
    for (int64_t j = 0; j < count; j += 256)
      tasks.push_back(SeedTask{row, diag0 + j, j,
                               (int32_t)std::min<int64_t>(256, count - j), 0});
    SeedTask *d_tasks = ar.alloc<SeedTask>((int64_t)tasks.size());
    HIP_CHECK(hipMemcpyAsync(d_tasks, tasks.data(), sizeof(SeedTask) * tasks.size(),
                             hipMemcpyHostToDevice, s));
    double *d_out = ar.alloc<double>(count);
    HIP_CHECK(launch_seeds(d_tasks, (int64_t)tasks.size(), sx.clean, sx.mean,
                           sy.clean, sy.mean, m, d_out, s));
    HIP_CHECK(hipMemcpyAsync(qt, d_out, sizeof(double) * count,
                             hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  } catch (const HipFailure &f) {
    return fail(SCAMP_HIP_DEVICE_ERROR,
                std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  return SCAMP_HIP_OK;
}

int scamp_hip_pearson_backward(scamp_hip_ctx *ctx, const double *d_T, size_t n,
                               int64_t window, int64_t batch,
                               const int32_t *d_index,
                               const double *d_grad_corr, double *d_grad_T) {
  if (!ctx || !d_T || !d_index || !d_grad_corr || !d_grad_T)
    return fail(SCAMP_HIP_INVALID_ARGS, "null argument");
  const int64_t m = window;
  if (m < 3 || (int64_t)n < m || batch < 1)
    return fail(SCAMP_HIP_INVALID_ARGS, "invalid window, length or batch");
  if (batch > (int64_t)0x7FFFFFFF / ((int64_t)n - m + 1))
    return fail(SCAMP_HIP_INVALID_ARGS,
                "backward: at most 2^31 - 1 profile entries per call");
  try {
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_CHECK(hipSetDevice(ctx->device));
    DevArena ar(ctx);
    hipStream_t s = ctx->stream;
    const int64_t ns = (int64_t)n - m + 1;
    const int64_t stat_len = ns + kColumnPad;
    const int64_t clean_len = (int64_t)n + kColumnPad;
    double *mean = ar.alloc<double>(batch * stat_len);
    double *inorm = ar.alloc<double>(batch * stat_len);
    double *clean = ar.alloc<double>(batch * clean_len);
    double *df = ar.alloc<double>(batch * stat_len);
    double *dg = ar.alloc<double>(batch * stat_len);
    HIP_CHECK(launch_stats(d_T, (int64_t)n, m, clean, clean_len, mean, inorm,
                           df, dg, stat_len, s, batch, (int64_t)n));
    const size_t work_bytes = pearson_backward_workspace((int64_t)n, m, batch);
    if (work_bytes == 0) throw HipFailure{hipErrorInvalidValue, "backward workspace"};
    char *work = ar.alloc<char>((int64_t)work_bytes);
    HIP_CHECK(launch_pearson_backward(d_T, (int64_t)n, m, batch, mean, inorm,
                                      stat_len, d_index, d_grad_corr, d_grad_T,
                                      work, s));
    if (!ctx->async) HIP_CHECK(hipStreamSynchronize(s));
  } catch (const HipFailure &f) {
    (void)hipStreamSynchronize(ctx->stream);
    return fail(SCAMP_HIP_DEVICE_ERROR,
                std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  return SCAMP_HIP_OK;
}

int scamp_hip_last_precision(scamp_hip_ctx *ctx, int *precision,
                             uint64_t *drift_entries) {
  if (!ctx) return fail(SCAMP_HIP_INVALID_ARGS, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (precision) *precision = ctx->last_precision;
  if (drift_entries) *drift_entries = ctx->last_drift;
  return SCAMP_HIP_OK;
}

int scamp_hip_last_timing(scamp_hip_ctx *ctx, double *total_ms,
                          double *prep_ms, double *sweep_ms) {
  if (!ctx) return fail(SCAMP_HIP_INVALID_ARGS, "null context");
  std::lock_guard<std::mutex> lock(ctx->mu);
  try {
    HIP_CHECK(hipSetDevice(ctx->device));
    resolve_timing(ctx);
  } catch (const HipFailure &f) {
    return fail(SCAMP_HIP_DEVICE_ERROR,
                std::string(f.what) + ": " + hipGetErrorString(f.err));
  }
  if (total_ms) *total_ms = ctx->total_ms;
  if (prep_ms) *prep_ms = ctx->prep_ms;
  if (sweep_ms) *sweep_ms = ctx->sweep_ms;
  return SCAMP_HIP_OK;
}

}  // extern "C"
