"""MI355X-native FEM-H1 K-means gamma/theta path (SPG-QP over per-time-step simplices).

The product is libpasc_hip.so (HIP kernels for gfx950 behind the C ABI in include/pasc_hip.h);
this package only loads it.  See capi.py for the binding and solver.py for Python handles.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
