"""delta_amd — MI355X-native contact detection and explicit DEM time stepping.

The compute path is libdelta_hip.so (hand-written HIP kernels for gfx950
behind a C ABI, include/delta_hip.h).  This package is the Python host layer
over that ABI: device buffers, the reference-shaped contact API
(delta_amd.contact), the force/dynamics API and the synthetic workloads.
Importing the package does not touch the GPU; the library is loaded on first
use and a missing or unbuilt library raises instead of falling back to CPU.
"""
from ._lib import lib, check, DhError  # noqa: F401
from .device import DeviceArray  # noqa: F401

__all__ = ["lib", "check", "DhError", "DeviceArray"]
