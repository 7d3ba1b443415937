"""ctypes binding of the FP64 CPU oracle (liboracle.so, built from pasc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg of
bench.py, never by the product package.  Every function takes and returns numpy float64 arrays in
the reference's t-major layout (idx = t*R*K + r*K + k, decomposition.h:488-492).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_P = np.ctypeslib.ndpointer
_f64 = _P(dtype=np.float64, flags="C_CONTIGUOUS")
_i64 = _P(dtype=np.int64, flags="C_CONTIGUOUS")
_i32 = _P(dtype=np.int32, flags="C_CONTIGUOUS")
_u8 = _P(dtype=np.uint8, flags="C_CONTIGUOUS")


class SpgParams(C.Structure):
    _fields_ = [
        ("maxit", C.c_int), ("eps", C.c_double), ("m", C.c_int), ("gamma", C.c_double),
        ("sigma1", C.c_double), ("sigma2", C.c_double), ("alphainit", C.c_double),
        ("stop_difff", C.c_int), ("stop_normgp", C.c_int), ("stop_normgp_normb", C.c_int),
        ("stop_Anormgp", C.c_int), ("stop_Anormgp_normb", C.c_int), ("stop_kkt", C.c_int),
        ("kkt_tol", C.c_double), ("kkt_scale", C.c_double), ("nthreads", C.c_int),
        ("fx_identity", C.c_int), ("gd_centered", C.c_int), ("safeguard", C.c_int),
        ("g_recompute", C.c_int), ("gd_certified", C.c_int), ("dad_edges", C.c_int),
    ]


class SpgResult(C.Structure):
    _fields_ = [
        ("it", C.c_int), ("hessmult", C.c_int), ("fx", C.c_double), ("dd", C.c_double),
        ("dAd", C.c_double), ("kkt", C.c_double), ("normb", C.c_double),
        ("stop_reason", C.c_int),
    ]


class TsParams(C.Structure):
    _fields_ = [
        ("maxit", C.c_int), ("eps", C.c_double), ("theta_fixed", C.c_int),
        ("skip_first_theta", C.c_int), ("scalef", C.c_int), ("fem_reduce", C.c_double),
        ("fem_type", C.c_int), ("annealing", C.c_int), ("restart_seed", C.c_uint64),
    ]


class TsResult(C.Structure):
    _fields_ = [
        ("it", C.c_int), ("spg_it_sum", C.c_int), ("L", C.c_double), ("deltaL", C.c_double),
        ("increased", C.c_int), ("aic", C.c_double), ("it_sum", C.c_int), ("best_run", C.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_sort_bubble.argtypes = [_f64, C.c_int]
        L.or_project_simplex.argtypes = [_f64, C.c_int64, C.c_int, C.c_int]
        L.or_h1_wsum.argtypes = [C.c_int64, C.c_int64, C.c_int]
        L.or_h1_wsum.restype = C.c_int
        L.or_h1_matvec.argtypes = [_f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_double,
                                   C.c_void_p, C.c_void_p, C.c_int]
        L.or_assemble_b.argtypes = [_f64, _f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                    C.c_int, C.c_void_p, C.c_int]
        L.or_theta_update.argtypes = [_f64, _f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                      C.c_void_p, C.c_int]
        L.or_kkt_residual.argtypes = [_f64, _f64, C.c_int64, C.c_int, C.c_double, C.c_int]
        L.or_kkt_residual.restype = C.c_double
        L.or_spg_default_params.argtypes = [C.POINTER(SpgParams)]
        L.or_spg_solve.argtypes = [_f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_double,
                                   C.c_void_p, C.c_void_p, C.POINTER(SpgParams),
                                   C.POINTER(SpgResult)]
        L.or_spg_solve.restype = C.c_int
        L.or_ts_solve.argtypes = [_f64, _f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_int,
                                  C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.POINTER(TsParams), C.POINTER(SpgParams), C.POINTER(TsResult)]
        L.or_ts_solve.restype = C.c_int
        L.or_abserr_reconstructed.argtypes = [_f64, _f64, _f64, C.c_int64, C.c_int64, C.c_int]
        L.or_abserr_reconstructed.restype = C.c_double
        L.or_uniform.argtypes = [C.c_uint64, C.c_int64, _f64]
        L.or_uniform.restype = None
        L.or_objective.argtypes = [_f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_double,
                                   C.c_void_p, C.c_void_p, C.c_int]
        L.or_objective.restype = C.c_double
        L.or_host_copy_bandwidth.argtypes = [C.c_int64, C.c_int, C.c_int]
        L.or_host_copy_bandwidth.restype = C.c_double
        L.or_max_threads.restype = C.c_int
        L.or_fem_reduced_T.argtypes = [C.c_int64, C.c_double]
        L.or_fem_reduced_T.restype = C.c_int64
        L.or_fem_reduce.argtypes = [_f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_int]
        L.or_fem_prolongate.argtypes = [_f64, _f64, C.c_int64, C.c_int64, C.c_int, C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _graph(graph):
    if graph is None:
        return None, None, (None, None)
    ptr = np.ascontiguousarray(graph[0], dtype=np.int64)
    idx = np.ascontiguousarray(graph[1], dtype=np.int32)
    return _ptr(ptr), _ptr(idx), (ptr, idx)


def max_threads() -> int:
    return int(lib().or_max_threads())


def sort_bubble(y):
    y = np.array(y, dtype=np.float64)
    lib().or_sort_bubble(y, y.size)
    return y


def project_simplex(x, K: int, nthreads: int = 1):
    """Project every consecutive K-block of x (t-major) onto the unit simplex; returns a copy."""
    y = np.array(x, dtype=np.float64).ravel().copy()
    assert y.size % K == 0
    lib().or_project_simplex(y, y.size // K, K, nthreads)
    return y


def h1_wsum(t: int, T: int, deg: int) -> int:
    return int(lib().or_h1_wsum(t, T, deg))


def h1_matvec(x, T: int, R: int, K: int, alpha: float, graph=None, nthreads: int = 1):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    y = np.empty_like(x)
    gp, gi, keep = _graph(graph)
    lib().or_h1_matvec(x, y, T, R, K, alpha, gp, gi, nthreads)
    del keep
    return y


def assemble_b(x, theta, T: int, R: int, K: int, xdim: int = 1, scalef: bool = True,
               heldout=None, nthreads: int = 1):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    theta = np.ascontiguousarray(theta, dtype=np.float64).ravel()
    b = np.empty(T * R * K, dtype=np.float64)
    ho = None if heldout is None else np.ascontiguousarray(heldout, dtype=np.uint8)
    lib().or_assemble_b(x, theta, b, T, R, K, xdim, int(scalef), _ptr(ho), nthreads)
    return b


def theta_update(x, gamma, T: int, R: int, K: int, xdim: int = 1, heldout=None,
                 nthreads: int = 1):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    gamma = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
    theta = np.empty(K * xdim, dtype=np.float64)
    ho = None if heldout is None else np.ascontiguousarray(heldout, dtype=np.uint8)
    lib().or_theta_update(x, gamma, theta, T, R, K, xdim, _ptr(ho), nthreads)
    return theta


def kkt_residual(gamma, g, K: int, s: float, nthreads: int = 1) -> float:
    gamma = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
    g = np.ascontiguousarray(g, dtype=np.float64).ravel()
    return float(lib().or_kkt_residual(gamma, g, gamma.size // K, K, s, nthreads))


def spg_params(**kw) -> SpgParams:
    p = SpgParams()
    lib().or_spg_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def spg_solve(x0, b, T: int, R: int, K: int, alpha: float, graph=None, **params):
    """SPG-QP (spgqpsolver.cpp:55-312).  Returns (x, SpgResult)."""
    x = np.array(x0, dtype=np.float64).ravel().copy()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    p = spg_params(**params)
    r = SpgResult()
    gp, gi, keep = _graph(graph)
    rc = lib().or_spg_solve(x, b, T, R, K, alpha, gp, gi, C.byref(p), C.byref(r))
    del keep
    if rc != 0:
        raise MemoryError("or_spg_solve failed")
    return x, r


def objective(x, b, T: int, R: int, K: int, alpha: float, graph=None, nthreads: int = 1):
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    gp, gi, keep = _graph(graph)
    v = float(lib().or_objective(x, b, T, R, K, alpha, gp, gi, nthreads))
    del keep
    return v


def gradient(x, b, T: int, R: int, K: int, alpha: float, graph=None):
    """g = alpha*A0*x - b."""
    return h1_matvec(x, T, R, K, alpha, graph) - np.asarray(b, dtype=np.float64).ravel()


def ts_solve(x, gamma0, theta0, T: int, R: int, K: int, xdim: int, epssqr: float, graph=None,
             heldout=None, ts=None, **spg):
    """Outer alternation (tssolver.h:491-715).  Returns (gamma, theta, TsResult)."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    gamma = np.array(gamma0, dtype=np.float64).ravel().copy()
    theta = np.array(theta0, dtype=np.float64).ravel().copy()
    tp = TsParams(maxit=300, eps=1e-6, theta_fixed=0, skip_first_theta=0, scalef=1,
                  fem_reduce=1.0, fem_type=0, annealing=1, restart_seed=13)
    for k, v in (ts or {}).items():
        setattr(tp, k, v)
    sp = spg_params(**spg)
    r = TsResult()
    ho = None if heldout is None else np.ascontiguousarray(heldout, dtype=np.uint8)
    gp, gi, keep = _graph(graph)
    rc = lib().or_ts_solve(x, gamma, theta, T, R, K, xdim, epssqr, gp, gi, _ptr(ho),
                           C.byref(tp), C.byref(sp), C.byref(r))
    del keep
    if rc != 0:
        raise MemoryError("or_ts_solve failed")
    return gamma, theta, r


def abserr_reconstructed(gamma, theta, solution, T: int, R: int, K: int) -> float:
    return float(lib().or_abserr_reconstructed(
        np.ascontiguousarray(gamma, dtype=np.float64).ravel(),
        np.ascontiguousarray(theta, dtype=np.float64).ravel(),
        np.ascontiguousarray(solution, dtype=np.float64).ravel(), T, R, K))


def uniform(seed: int, n: int):
    """n uniforms in [0, 1) of the library's SplitMix64 generator (the restarts' random gamma)."""
    out = np.empty(n, dtype=np.float64)
    lib().or_uniform(seed, n, out)
    return out


def host_copy_bandwidth(nbytes: int = 1 << 28, reps: int = 3, nthreads: int = 1) -> float:
    return float(lib().or_host_copy_bandwidth(nbytes, reps, nthreads))


FEM_CONSTANT, FEM_HAT = 1, 2


def fem_reduced_T(T1: int, fem_reduce: float) -> int:
    return int(lib().or_fem_reduced_T(T1, fem_reduce))


def fem_reduce(x, T1: int, T2: int, K: int, fem_type: int):
    """Fine -> coarse (fem.cpp:28-95 / femhat.cpp:26-114), t-major [T*K] arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros(T2 * K)
    lib().or_fem_reduce(x, out, T1, T2, K, fem_type)
    return out


def fem_prolongate(x, T1: int, T2: int, K: int, fem_type: int):
    """Coarse -> fine (fem.cpp:97-166 / femhat.cpp:116-217), t-major [T*K] arrays."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.zeros(T1 * K)
    lib().or_fem_prolongate(x, out, T1, T2, K, fem_type)
    return out
