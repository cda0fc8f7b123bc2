"""ear_amd -- MI355X-native acoustic ray-tracing core with E.A.R's semantics.

The product is libear_hip.so (HIP kernels for gfx950 behind the C ABI in
include/ear.h) and the host C++ layer (Scene / Mesh / Material / Recorder and
the EAR CLI).  This package holds the Python front end over that ABI
(`api`), seeded procedural scenes (`scenes`) and acoustic metrics
(`acoustics`).
"""
from . import _ffi  # noqa: F401

__all__ = ["api", "scenes", "acoustics", "build"]
