import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gpu-scrimp-large_amd"), os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)

# torch's ROCm wheel ships its own libamdhip64.so.7, and the engine links the
# system one of the same soname: a process maps whichever is loaded first for
# both.  torch initialises only on the runtime it was built with, so it is
# loaded before any test opens the engine's libraries.
try:
    import torch  # noqa: F401
except ImportError:
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")
