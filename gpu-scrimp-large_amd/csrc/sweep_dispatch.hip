// Dispatch of the diagonal sweep to its per-kind instantiations.
#include "kernels.h"

namespace scamp_amd {

hipError_t launch_sweep_nn(const SweepArgs &, bool, bool, bool, int, hipStream_t);
hipError_t launch_sweep_nn_fresh(const SweepArgs &, bool, bool, bool, int,
                                 hipStream_t);
hipError_t launch_sweep_sum(const SweepArgs &, bool, bool, bool, int, hipStream_t);
hipError_t launch_sweep_matrix(const SweepArgs &, bool, bool, bool, int, hipStream_t);
hipError_t launch_sweep_matrix_small(const SweepArgs &, bool, bool, bool, int,
                                     hipStream_t);
hipError_t launch_sweep_knn(const SweepArgs &, bool, bool, bool, int, hipStream_t);

hipError_t launch_sweep(const SweepArgs &a, bool fp32, SweepKind kind,
                        bool cols, bool rows, int grid, hipStream_t s,
                        bool fresh) {
  if (!cols && !rows) return hipSuccess;
  switch (kind) {
    case kNN:
      return fresh ? launch_sweep_nn_fresh(a, fp32, cols, rows, grid, s)
                   : launch_sweep_nn(a, fp32, cols, rows, grid, s);
    case kSum: return launch_sweep_sum(a, fp32, cols, rows, grid, s);
    case kMatrix: return launch_sweep_matrix(a, fp32, cols, rows, grid, s);
    case kKnn: return launch_sweep_knn(a, fp32, cols, rows, grid, s);
    default: return launch_sweep_matrix_small(a, fp32, cols, rows, grid, s);
  }
}

}  // namespace scamp_amd
