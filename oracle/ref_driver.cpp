// Driver that lets tests call SCAMP's own CPU double path, compiled from the
// reference's sources (src/common.cpp, cpu_stats.cpp, cpu_kernels.cpp,
// kernel_common.cpp, tile.cpp, qt_helper.cpp, SCAMP.cpp, scamp_utils.cpp) by
// oracle/Makefile into oracle/_ref/.  The reference's CLI (src/main.cpp) needs
// gflags, which is not vendored, so this small C entry point takes its place:
// it fills SCAMPArgs the way the CLI and pyscamp do
// (python/SCAMP_python.cpp:67-89, 179-193; src/scamp_utils.cpp:226-245) and
// calls SCAMP::do_SCAMP(args, devices={}, num_threads)
// (src/SCAMP.cpp:279-334).
//
// TEST INFRASTRUCTURE ONLY (the tests' reference results and the CPU baseline).
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "SCAMP.h"
#include "common.h"
#include "scamp_utils.h"

extern "C" int ref_join(const double *a, int64_t na, const double *b,
                        int64_t nb, int64_t m, int ptype, int64_t max_tile,
                        int rows, int cols, int keep_rows, int aligned,
                        int64_t grow, int64_t gcol, double thr, int nthreads,
                        int precision, double *corr_a, int64_t *idx_a,
                        double *corr_b, int64_t *idx_b) {
  try {
    SCAMP::SCAMPArgs args;
    args.timeseries_a.assign(a, a + na);
    args.has_b = b != nullptr;
    if (args.has_b) {
      args.timeseries_b.assign(b, b + nb);
    } else {
      args.timeseries_b = args.timeseries_a;
    }
    args.window = m;
    args.max_tile_size = max_tile;
    args.distributed_start_row = grow;
    args.distributed_start_col = gcol;
    args.distance_threshold = thr;
    // SCAMPPrecisionType values (src/common.h:53-58); the CPU path runs
    // DOUBLE and ULTRA (src/cpu_kernels.cpp:201-252)
    args.precision_type = static_cast<SCAMP::SCAMPPrecisionType>(precision);
    args.profile_type = static_cast<SCAMP::SCAMPProfileType>(ptype);
    args.computing_rows = rows != 0;
    args.computing_columns = cols != 0;
    args.keep_rows_separate = keep_rows != 0;
    args.is_aligned = aligned != 0;
    args.silent_mode = true;
    args.max_matches_per_column = 5;
    args.matrix_height = 50;
    args.matrix_width = 50;
    args.profile_a.type = args.profile_type;
    args.profile_b.type = args.profile_type;
    if (!InitProfileMemory(&args)) return 7;
    SCAMP::do_SCAMP(&args, std::vector<int>(), nthreads);

    auto unpack = [&](SCAMP::Profile &p, double *corr, int64_t *idx) {
      if (corr == nullptr) return;
      auto &d = p.data[0];
      switch (args.profile_type) {
        case SCAMP::PROFILE_TYPE_1NN_INDEX:
          for (size_t i = 0; i < d.uint64_value.size(); ++i) {
            SCAMP::mp_entry e;
            e.ulong = d.uint64_value[i];
            corr[i] = e.floats[0];
            idx[i] = e.ints[1] == 0xFFFFFFFFu ? -1 : (int64_t)e.ints[1];
          }
          break;
        case SCAMP::PROFILE_TYPE_1NN:
          for (size_t i = 0; i < d.float_value.size(); ++i)
            corr[i] = d.float_value[i];
          break;
        case SCAMP::PROFILE_TYPE_SUM_THRESH:
          for (size_t i = 0; i < d.double_value.size(); ++i)
            corr[i] = d.double_value[i];
          break;
        default:
          break;
      }
    };
    unpack(args.profile_a, corr_a, idx_a);
    if (args.keep_rows_separate) unpack(args.profile_b, corr_b, idx_b);
  } catch (const std::exception &) {
    return 1;
  }
  return 0;
}
