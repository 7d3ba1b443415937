"""ctypes binding of libpasc_hip.so (include/pasc_hip.h).

The library is the product: this module only loads it and mirrors its structs.  If the shared
library is missing the import of :func:`lib` raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpasc_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "pasc_hip.h")

PASC_OK, PASC_ERR_ARG, PASC_ERR_HIP, PASC_ERR_ALLOC, PASC_ERR_STATE, PASC_ERR_NUMERIC = range(6)
LAYOUT_TMAJOR, LAYOUT_KMAJOR = 0, 1
FP64, FP32_BULK = 0, 1
FEM_NONE, FEM_CONSTANT, FEM_HAT = 0, 1, 2
DTYPE_F64, DTYPE_F32 = 0, 1
STOP_DIFFF, STOP_NORMGP, STOP_NORMGP_NORMB, STOP_ANORMGP, STOP_ANORMGP_NORMB, STOP_KKT = (
    1, 2, 4, 8, 16, 32)
REASONS = ["maxit", "difff", "normgp", "normgp_normb", "Anormgp", "Anormgp_normb", "kkt",
           "stationary", "numeric"]
REASON_KKT = REASONS.index("kkt")
MAX_K = 8


class PascError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pasc error {code}: {msg}")
        self.code = code


class ProblemDesc(C.Structure):
    _fields_ = [
        ("T", C.c_int64), ("R", C.c_int64), ("K", C.c_int32), ("xdim", C.c_int32),
        ("scalef", C.c_int32), ("precision", C.c_int32), ("x", C.POINTER(C.c_double)),
        ("nbr_ptr", C.POINTER(C.c_int64)), ("nbr_idx", C.POINTER(C.c_int32)),
        ("heldout", C.POINTER(C.c_uint8)), ("grid_width", C.c_int32),
        ("fem_reduce", C.c_double), ("fem_type", C.c_int32),
    ]


class SpgParams(C.Structure):
    _fields_ = [
        ("maxit", C.c_int32), ("eps", C.c_double), ("m", C.c_int32), ("gamma", C.c_double),
        ("sigma1", C.c_double), ("sigma2", C.c_double), ("alphainit", C.c_double),
        ("stop_mask", C.c_uint32), ("kkt_tol", C.c_double), ("check_every", C.c_int32),
        ("use_graph", C.c_int32), ("fp32_switch", C.c_double), ("stall_window", C.c_int32),
        ("polish_maxit", C.c_int32), ("kkt_every", C.c_int32),
    ]


class SpgResult(C.Structure):
    _fields_ = [
        ("it", C.c_int32), ("hessmult", C.c_int32), ("status", C.c_int32),
        ("stop_reason", C.c_int32), ("it_fp32", C.c_int32), ("it_fp64", C.c_int32),
        ("fx", C.c_double), ("dd", C.c_double), ("dAd", C.c_double), ("kkt", C.c_double),
        ("normb", C.c_double), ("time_s", C.c_double), ("pass_time_fp32_s", C.c_double),
        ("pass_time_fp64_s", C.c_double), ("passes_fp32", C.c_int32), ("passes_fp64", C.c_int32),
    ]


class TsParams(C.Structure):
    _fields_ = [("maxit", C.c_int32), ("eps", C.c_double), ("theta_fixed", C.c_int32),
                ("skip_first_theta", C.c_int32), ("annealing", C.c_int32),
                ("restart_seed", C.c_uint64)]


class TsResult(C.Structure):
    _fields_ = [("it", C.c_int32), ("spg_it_sum", C.c_int32), ("L", C.c_double),
                ("deltaL", C.c_double), ("increased", C.c_int32), ("aic", C.c_double),
                ("it_sum", C.c_int32), ("best_run", C.c_int32)]


class SeriesDesc(C.Structure):
    _fields_ = [("T", C.c_int64), ("K", C.c_int32), ("x", C.POINTER(C.c_double)),
                ("theta", C.POINTER(C.c_double))]


class PassShape(C.Structure):
    _fields_ = [("block", C.c_int32), ("chunk", C.c_int32), ("prefetch", C.c_int32),
                ("max_slots", C.c_int32), ("min_tiles", C.c_int32), ("us_per_pass", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LaunchShape(C.Structure):
    _fields_ = [("main", PassShape), ("polish", PassShape)]


class InstanceResult(C.Structure):
    _fields_ = [("it", C.c_int32), ("it_fp32", C.c_int32), ("it_fp64", C.c_int32),
                ("stop_reason", C.c_int32), ("status", C.c_int32), ("fx", C.c_double),
                ("kkt", C.c_double)]


vp = C.c_void_p
_SIGS = {
    "pasc_last_error": ([], C.c_char_p),
    "pasc_version": ([], C.c_int),
    "pasc_ctx_create": ([C.c_int, C.POINTER(vp)], C.c_int),
    "pasc_ctx_destroy": ([vp], C.c_int),
    "pasc_ctx_set_stream": ([vp, vp], C.c_int),
    "pasc_ctx_synchronize": ([vp], C.c_int),
    "pasc_problem_create": ([vp, C.POINTER(ProblemDesc), C.POINTER(vp)], C.c_int),
    "pasc_problem_destroy": ([vp], C.c_int),
    "pasc_set_theta": ([vp, C.POINTER(C.c_double)], C.c_int),
    "pasc_get_theta": ([vp, C.POINTER(C.c_double)], C.c_int),
    "pasc_set_gamma": ([vp, C.POINTER(C.c_double), C.c_int], C.c_int),
    "pasc_set_gamma_uniform": ([vp], C.c_int),
    "pasc_get_gamma": ([vp, C.POINTER(C.c_double), C.c_int], C.c_int),
    "pasc_set_epssqr": ([vp, C.c_double], C.c_int),
    "pasc_spg_default_params": ([C.POINTER(SpgParams)], None),
    "pasc_gamma_solve": ([vp, C.POINTER(SpgParams), C.POINTER(SpgResult)], C.c_int),
    "pasc_theta_update": ([vp], C.c_int),
    "pasc_set_linear_term": ([vp, C.POINTER(C.c_double), C.c_int], C.c_int),
    "pasc_set_data": ([vp, C.POINTER(C.c_double)], C.c_int),
    "pasc_fem_reduced_T": ([C.c_int64, C.c_double], C.c_int64),
    "pasc_fem_reduce": ([vp, vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_int64, C.c_int64,
                         C.c_int32], C.c_int),
    "pasc_fem_prolongate": ([vp, vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_int64, C.c_int64,
                             C.c_int32], C.c_int),
    "pasc_theta_update_noalloc": ([vp], C.c_int),
    "pasc_ts_default_params": ([C.POINTER(TsParams)], None),
    "pasc_ts_solve": ([vp, C.POINTER(TsParams), C.POINTER(SpgParams), C.POINTER(TsResult)],
                      C.c_int),
    "pasc_objective": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "pasc_abserr_reconstructed": ([vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "pasc_get_labels": ([vp, C.POINTER(C.c_int32)], C.c_int),
    "pasc_project_simplex": ([vp, vp, C.c_int64, C.c_int32, C.c_int, C.c_int, C.c_int64],
                             C.c_int),
    "pasc_h1_matvec": ([vp, vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_int64, C.c_double,
                        C.c_int32, C.c_int], C.c_int),
    "pasc_h1_matvec_csr": ([vp, vp, vp, C.c_int64, C.c_int64, C.c_int32, C.c_int64, C.c_double,
                            C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.c_int], C.c_int),
    "pasc_batch_create": ([vp, C.c_int32, C.POINTER(SeriesDesc), C.c_int32,
                           C.POINTER(C.c_double), C.c_int32, C.POINTER(vp)], C.c_int),
    "pasc_batch_destroy": ([vp], C.c_int),
    "pasc_batch_create_problem": ([vp, C.POINTER(ProblemDesc), C.POINTER(C.c_double), C.c_int32,
                                   C.POINTER(C.c_double), C.POINTER(vp)], C.c_int),
    "pasc_batch_solve": ([vp, C.POINTER(SpgParams), C.POINTER(InstanceResult),
                          C.POINTER(C.c_double)], C.c_int),
    "pasc_batch_get_gamma": ([vp, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_int],
                             C.c_int),
    "pasc_adjoint": ([vp, C.POINTER(C.c_uint8), C.c_double, C.c_int32, C.POINTER(C.c_double),
                      C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int32)],
                     C.c_int),
    "pasc_set_profiling": ([vp, C.c_int], C.c_int),
    "pasc_device_copy_bandwidth": ([vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "pasc_autotune": ([vp, C.c_char_p, C.POINTER(LaunchShape)], C.c_int),
    "pasc_set_launch_shape": ([vp, C.POINTER(LaunchShape)], C.c_int),
    "pasc_batch_autotune": ([vp, C.c_char_p, C.POINTER(LaunchShape)], C.c_int),
    "pasc_batch_set_launch_shape": ([vp, C.POINTER(LaunchShape)], C.c_int),
    "pasc_generate_markov": ([C.c_uint64, C.c_int64, C.c_int32, C.POINTER(C.c_double),
                              C.c_double, C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_int32)], C.c_int),
    "pasc_generate_mask": ([C.c_uint64, C.c_int64, C.c_double, C.POINTER(C.c_uint8)], C.c_int),
    "pasc_generate_uniform": ([C.c_uint64, C.c_int64, C.POINTER(C.c_double)], C.c_int),
    "pasc_generate_rect_image": ([C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                  C.POINTER(C.c_int32)], C.c_int),
    "pasc_generate_batch_params": ([C.c_uint64, C.c_int32, C.c_double, C.c_double, C.c_int32,
                                    C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double), C.POINTER(C.c_uint64)], C.c_int),
}

_lib = None


def lib():
    """Load libpasc_hip.so; raises when it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (the HIP extension is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int):
    if rc != PASC_OK:
        msg = lib().pasc_last_error()
        raise PascError(rc, msg.decode() if msg else "")


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def u8ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def default_params(**kw) -> SpgParams:
    p = SpgParams()
    lib().pasc_spg_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def header_symbols() -> list[str]:
    """Every function declared in include/pasc_hip.h."""
    import re
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(pasc_[a-z0-9_]+)\s*\(", txt)))
