"""Python loader for the MI355X Barnes-Hut C-ABI library (libbhgpu.so).

The library itself is C/C++/HIP; this package only wraps it with ctypes for
tests, the benchmark and Python callers. See include/bh_gpu.h for the ABI.
"""
from .layout import FORCE_PARAMS, PARTICLE, WALK_NODE, force_params, make_particles  # noqa: F401
