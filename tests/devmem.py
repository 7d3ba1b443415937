"""Device buffers for the kernel-level tests, through the HIP runtime libpasc_hip.so links.

torch bundles its own HIP runtime; mixing it with the system one in one process leaves torch
without a device, so the kernel-level tests allocate and copy with hipMalloc/hipMemcpy here.
"""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from pasc_inference_amd import capi
        capi.lib()  # loads /opt/rocm/lib/libamdhip64.so.7 as a dependency
        _hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceArray:
    """A device copy of a numpy array (float64 or float32)."""

    def __init__(self, host: np.ndarray):
        self.dtype = host.dtype
        self.shape = host.shape
        self.nbytes = host.nbytes
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), max(self.nbytes, 8))
        assert rc == 0, f"hipMalloc failed ({rc})"
        self.ptr = p
        self.upload(host)

    def upload(self, host: np.ndarray):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.nbytes == self.nbytes
        assert hip().hipMemcpy(self.ptr, h.ctypes.data_as(C.c_void_p), self.nbytes, 1) == 0

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, 2) == 0
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
