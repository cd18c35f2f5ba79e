"""MI355X-native gammaray hot path: SAH BVH traversal, Disney BSDF path tracing
and film accumulation as hand-written gfx950 HIP kernels behind a C ABI
(include/gammaray.h).  This package is the Python host over that ABI."""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
