"""MI355X (gfx950) explicit finite-difference step for SymPhas-style models.

The product is the HIP library symphas_amd/lib/libsps_fd.so behind the C ABI
in include/sps_fd.h; this package is its Python host side (ctypes binding,
model descriptors, seeded inputs).  The C++ drop-in solver adapter lives in
symphas_amd/host.
"""
from . import abi, models, inputs  # noqa: F401
from .solver import FDBatch, FDPlan, device_count  # noqa: F401

__all__ = ["abi", "models", "inputs", "FDPlan", "FDBatch", "device_count"]
