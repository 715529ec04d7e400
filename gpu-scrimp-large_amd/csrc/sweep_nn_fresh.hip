// Instantiations of the diagonal sweep (sweep_impl.h) for profile kind kNN in
// its FRESH variant (launches of short series cut into row chunks), FP64 and
// FP32 forms.
#include "sweep_impl.h"

namespace scamp_amd {

hipError_t launch_sweep_nn_fresh(const SweepArgs &a, bool fp32, bool cols,
                                 bool rows, int grid, hipStream_t s) {
  if (fp32) return sweep::launch_kind<float, kNN, true>(a, cols, rows, grid, s);
  return sweep::launch_kind<double, kNN, true>(a, cols, rows, grid, s);
}

}  // namespace scamp_amd
