"""helio_amd -- MI355X (gfx950) implementation of Helio's deferred-lighting hot path.

Native code: libhelio_hip.so (HIP kernels + C ABI, include/helio_hip/helio_hip.h)
and libhelio_synth.so (synthetic workloads).  This package is the host-side
mirror of the reference passes over that ABI; it has no CPU compute path.
"""
from .abi import HelioError  # noqa: F401

__all__ = ["HelioError"]
__version__ = "0.1.0"
