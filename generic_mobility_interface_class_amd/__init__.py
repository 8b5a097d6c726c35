"""MI355X-native RPY grand-mobility hot path (gfx950 HIP behind a C ABI).

The compute lives in libgmi.so (csrc/ + include/gmi.h).  This package holds
its ctypes binding (capi), the build recipe (build) and the synthetic
workload geometry (geometry).
"""
from . import geometry  # noqa: F401

__all__ = ["geometry"]
