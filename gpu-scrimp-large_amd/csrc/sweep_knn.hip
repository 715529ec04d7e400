// Instantiations of the diagonal sweep (sweep_impl.h) for profile kind
// kKnn (ALL_NEIGHBORS: the k best matches of every subsequence), FP64 and
// FP32 forms.
#include "sweep_impl.h"

namespace scamp_amd {

hipError_t launch_sweep_knn(const SweepArgs &a, bool fp32, bool cols, bool rows,
                            int grid, hipStream_t s) {
  if (fp32) return sweep::launch_kind<float, kKnn>(a, cols, rows, grid, s);
  return sweep::launch_kind<double, kKnn>(a, cols, rows, grid, s);
}

}  // namespace scamp_amd
