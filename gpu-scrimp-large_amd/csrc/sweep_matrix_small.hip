// Instantiations of the diagonal sweep (sweep_impl.h) for profile kind
// kMatrixSmall, FP64 and FP32 forms.  One translation unit per kind keeps the build
// parallel.
#include "sweep_impl.h"

namespace scamp_amd {

hipError_t launch_sweep_matrix_small(const SweepArgs &a, bool fp32, bool cols,
                              bool rows, int grid, hipStream_t s) {
  if (fp32) return sweep::launch_kind<float, kMatrixSmall>(a, cols, rows, grid, s);
  return sweep::launch_kind<double, kMatrixSmall>(a, cols, rows, grid, s);
}

}  // namespace scamp_amd
