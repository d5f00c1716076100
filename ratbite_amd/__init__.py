"""ratbite_amd — MI355X-native spectral path-tracing hot path.

The compute path is libratbite_hip.so (hand-written HIP kernels for gfx950
behind the C ABI in include/ratbite_hip.h). This package holds the ctypes
binding, the host-side mirror of the reference's scene / material / render
settings interface, and the seeded synthetic workloads.
"""
from . import _lib
from ._lib import LaunchParams, RbhError, lib
from .context import Context

__all__ = ["Context", "LaunchParams", "RbhError", "lib", "_lib"]
