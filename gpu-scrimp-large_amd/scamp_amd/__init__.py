"""scamp_amd: MI355X (gfx950) matrix-profile join engine.

* ``scamp_amd._capi``    -- ctypes binding of the C ABI (include/scamp_hip.h)
* ``scamp_amd.tuning``   -- re-anchoring interval autotune of the mixed form
* ``scamp_amd.autograd`` -- differentiable batched self-join (torch)
* ``pyscamp`` (a sibling module in gpu-scrimp-large_amd/) -- the pyscamp
  drop-in (selfjoin / abjoin / *_sum / *_matrix / gpu_supported) over the C++
  ``SCAMP::do_SCAMP`` entry point.
"""
from . import _capi  # noqa: F401  (raises if the HIP library was not built)

__all__ = ["_capi"]
