"""Thin Python handles over the C ABI (used by tests, the smoke check and bench.py).

Names follow the reference: a :class:`Problem` is the GraphH1FEMModel + TSData of one series
(or image), :meth:`Problem.gamma_solve` is the SPGQPSolver gamma step, :meth:`Problem.theta_update`
the theta step and :meth:`Problem.ts_solve` the TSSolver alternation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, dptr, iptr, lib, u8ptr


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().pasc_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pasc_ctx_destroy(self.h)
            self.h = None

    def synchronize(self):
        check(lib().pasc_ctx_synchronize(self.h))

    def copy_bandwidth(self, nbytes: int = 1 << 30, reps: int = 5) -> float:
        """Best GB/s (read + write) of a plain device copy kernel over nbytes."""
        g = C.c_double()
        check(lib().pasc_device_copy_bandwidth(self.h, nbytes, reps, C.byref(g)))
        return g.value

    def set_stream(self, stream_ptr: int | None):
        check(lib().pasc_ctx_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _problem_desc(x, K, T, R, xdim, scalef, precision, heldout, grid_width, fem_reduce,
                  fem_type, graph):
    """A pasc_problem_desc and the arrays it points into (kept alive by the caller).
    x: data [T][R][xdim] (None: no data term, T and R given).  The graph of the R nodes is the
    4-neighbour grid of width grid_width (T defaults to 1, an image), a CSR adjacency
    graph = (nbr_ptr[R+1], nbr_idx), or none (R unconnected time chains)."""
    keep = {}
    if graph is not None:
        keep["graph"] = (np.ascontiguousarray(graph[0], dtype=np.int64),
                         np.ascontiguousarray(graph[1], dtype=np.int32))
        R = keep["graph"][0].size - 1
    if x is None:
        if T is None:
            raise ValueError("x=None needs T (and R / grid_width)")
        keep["x"] = None
        n = T * R
    else:
        keep["x"] = np.ascontiguousarray(x, dtype=np.float64).ravel()
        n = keep["x"].size // xdim
        if grid_width > 0 and graph is None:
            T = 1 if T is None else T
            R = n // T
        elif T is None:
            T = n // R
    assert T * R == n, "data size must be T*R*xdim"
    keep["heldout"] = None if heldout is None else np.ascontiguousarray(heldout, dtype=np.uint8)
    d = capi.ProblemDesc()
    d.T, d.R, d.K, d.xdim = T, R, K, xdim
    d.scalef, d.precision = int(scalef), precision
    d.x = dptr(keep["x"]) if keep["x"] is not None else None
    d.heldout = u8ptr(keep["heldout"]) if keep["heldout"] is not None else None
    d.grid_width = grid_width
    d.fem_reduce, d.fem_type = float(fem_reduce), int(fem_type)
    if graph is not None:
        d.nbr_ptr = keep["graph"][0].ctypes.data_as(C.POINTER(C.c_int64))
        d.nbr_idx = keep["graph"][1].ctypes.data_as(C.POINTER(C.c_int32))
    return d, keep, T, R, n


class Problem:
    def __init__(self, ctx: Context, x, K: int, T: int | None = None, R: int = 1, xdim: int = 1,
                 scalef: bool = True, precision: int = capi.FP64, heldout=None,
                 grid_width: int = 0, epssqr: float = 1.0, fem_reduce: float = 1.0,
                 fem_type: int = capi.FEM_NONE, graph=None):
        """One gamma QP (GraphH1FEMModel + TSData); arguments as in _problem_desc."""
        self.ctx = ctx
        d, self._keep, T, R, n = _problem_desc(x, K, T, R, xdim, scalef, precision, heldout,
                                               grid_width, fem_reduce, fem_type, graph)
        self.x = self._keep["x"]
        self.heldout = self._keep["heldout"]
        self.T, self.R, self.K, self.xdim = T, R, K, xdim
        self.nodes = n
        self.grid_width = grid_width
        h = C.c_void_p()
        check(lib().pasc_problem_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.set_epssqr(epssqr)

    def close(self):
        if self.h:
            lib().pasc_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_epssqr(self, eps: float):
        check(lib().pasc_set_epssqr(self.h, float(eps)))

    def set_theta(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float64).ravel()
        assert th.size == self.K * self.xdim
        check(lib().pasc_set_theta(self.h, dptr(th)))

    def get_theta(self):
        th = np.empty(self.K * self.xdim)
        check(lib().pasc_get_theta(self.h, dptr(th)))
        return th

    def set_gamma(self, gamma, layout: int = capi.LAYOUT_TMAJOR):
        g = np.ascontiguousarray(gamma, dtype=np.float64).ravel()
        assert g.size == self.nodes * self.K
        check(lib().pasc_set_gamma(self.h, dptr(g), layout))

    def set_gamma_uniform(self):
        check(lib().pasc_set_gamma_uniform(self.h))

    def get_gamma(self, layout: int = capi.LAYOUT_TMAJOR):
        g = np.empty(self.nodes * self.K)
        check(lib().pasc_get_gamma(self.h, dptr(g), layout))
        return g

    def gamma_solve(self, **params) -> capi.SpgResult:
        p = capi.default_params(**params)
        r = capi.SpgResult()
        check(lib().pasc_gamma_solve(self.h, C.byref(p), C.byref(r)))
        return r

    def set_linear_term(self, b, layout: int = capi.LAYOUT_TMAJOR):
        """Explicit b of the QP (None returns to the data term built from x and theta)."""
        if b is None:
            check(lib().pasc_set_linear_term(self.h, None, layout))
            return
        self._b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        assert self._b.size == self.nodes * self.K
        check(lib().pasc_set_linear_term(self.h, dptr(self._b), layout))

    def set_data(self, x):
        """Replace the data (same layout and size as at creation), as TSData::cutdata /
        scaledata edit the datavector; theta, gamma and epsilon are kept."""
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        assert x.size == self.x.size
        check(lib().pasc_set_data(self.h, dptr(x)))
        self.x = x
        self._keep["x"] = x

    def theta_update(self):
        check(lib().pasc_theta_update(self.h))
        return self.get_theta()

    def ts_solve(self, ts=None, **spg) -> capi.TsResult:
        tp = capi.TsParams()
        lib().pasc_ts_default_params(C.byref(tp))
        for k, v in (ts or {}).items():
            setattr(tp, k, v)
        p = capi.default_params(**spg)
        r = capi.TsResult()
        check(lib().pasc_ts_solve(self.h, C.byref(tp), C.byref(p), C.byref(r)))
        return r

    def objective(self):
        f, k = C.c_double(), C.c_double()
        check(lib().pasc_objective(self.h, C.byref(f), C.byref(k)))
        return f.value, k.value

    def abserr(self, solution):
        s = np.ascontiguousarray(solution, dtype=np.float64)
        a = C.c_double()
        check(lib().pasc_abserr_reconstructed(self.h, dptr(s), C.byref(a)))
        return a.value

    def labels(self):
        lab = np.empty(self.nodes, dtype=np.int32)
        check(lib().pasc_get_labels(self.h, iptr(lab)))
        return lab

    def adjoint(self, mask, cg_tol: float = 1e-10, cg_maxit: int = 10000):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        loss, dle = C.c_double(), C.c_double()
        dth = np.empty(self.K * self.xdim)
        it = C.c_int32()
        check(lib().pasc_adjoint(self.h, u8ptr(m), cg_tol, cg_maxit, C.byref(loss), C.byref(dle),
                                 dptr(dth), C.byref(it)))
        return loss.value, dle.value, dth, it.value

    def set_profiling(self, enable: bool = True):
        check(lib().pasc_set_profiling(self.h, int(enable)))

    def autotune(self, cache_path: str | None = None) -> dict:
        """Tunes the launch shapes; returns {"main": {...}, "polish": {...}} (see LaunchShape)."""
        sh = capi.LaunchShape()
        check(lib().pasc_autotune(self.h, cache_path.encode() if cache_path else None,
                                  C.byref(sh)))
        return {"main": sh.main.as_dict(), "polish": sh.polish.as_dict()}

    def set_launch_shape(self, block: int, chunk: int, prefetch: int = 0, max_slots: int = 0,
                         min_tiles: int = 0, polish=None):
        """main shape (FP64 problems: every pass; FP32-bulk: the FP32 phase); polish: a dict
        for the FP64 polish of FP32-bulk problems (default: same shape, chunk / 2).  Images
        (grid_width, T = 1): block 64, chunk = image rows per wave tile, prefetch = rows loaded
        ahead (1..3); the polish defaults to the same rows."""
        sh = capi.LaunchShape()
        sh.main = capi.PassShape(block, chunk, prefetch, max_slots, min_tiles, 0.0)
        if polish:
            sh.polish = capi.PassShape(polish["block"], polish["chunk"], polish.get("prefetch", 0),
                                       polish.get("max_slots", 0), polish.get("min_tiles", 0), 0.0)
        check(lib().pasc_set_launch_shape(self.h, C.byref(sh)))


class Batch:
    """n_series x n_eps independent gamma QPs (uneven T and K) solved together."""

    def __init__(self, ctx: Context, xs, Ks, thetas, eps, precision: int = capi.FP32_BULK):
        self.ctx = ctx
        self.xs = [np.ascontiguousarray(x, dtype=np.float64) for x in xs]
        self.thetas = [np.ascontiguousarray(t, dtype=np.float64) for t in thetas]
        self.Ks = list(Ks)
        self.eps = np.ascontiguousarray(eps, dtype=np.float64)
        n = len(self.xs)
        descs = (capi.SeriesDesc * n)()
        for i in range(n):
            descs[i].T = self.xs[i].size
            descs[i].K = self.Ks[i]
            descs[i].x = dptr(self.xs[i])
            descs[i].theta = dptr(self.thetas[i])
        h = C.c_void_p()
        check(lib().pasc_batch_create(ctx.h, n, descs, self.eps.size, dptr(self.eps), precision,
                                      C.byref(h)))
        self.h = h
        self.n_series, self.n_eps = n, self.eps.size

    def close(self):
        if self.h:
            lib().pasc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, **params):
        p = capi.default_params(**params)
        res = (capi.InstanceResult * (self.n_series * self.n_eps))()
        t = C.c_double()
        check(lib().pasc_batch_solve(self.h, C.byref(p), res, C.byref(t)))
        return list(res), t.value

    def get_gamma(self, s: int, e: int, layout: int = capi.LAYOUT_TMAJOR):
        g = np.empty(self.xs[s].size * self.Ks[s])
        check(lib().pasc_batch_get_gamma(self.h, s, e, dptr(g), layout))
        return g

    def autotune(self, cache_path: str | None = None) -> dict:
        sh = capi.LaunchShape()
        check(lib().pasc_batch_autotune(self.h, cache_path.encode() if cache_path else None,
                                        C.byref(sh)))
        return {"main": sh.main.as_dict(), "polish": sh.polish.as_dict()}


class ProblemBatch:
    """One problem (any graph, xdim, held-out mask) solved for several epsilon values at once."""

    def __init__(self, ctx: Context, x, K: int, theta, eps, precision: int = capi.FP32_BULK,
                 **problem_args):
        self.ctx = ctx
        d, self._keep, T, R, n = _problem_desc(x, K, problem_args.pop("T", None),
                                               problem_args.pop("R", 1),
                                               problem_args.pop("xdim", 1),
                                               problem_args.pop("scalef", True), precision,
                                               problem_args.pop("heldout", None),
                                               problem_args.pop("grid_width", 0), 1.0,
                                               capi.FEM_NONE, problem_args.pop("graph", None))
        assert not problem_args, f"unknown arguments {list(problem_args)}"
        self.theta = np.ascontiguousarray(theta, dtype=np.float64).ravel()
        self.eps = np.ascontiguousarray(eps, dtype=np.float64).ravel()
        self.K, self.nodes = K, n
        h = C.c_void_p()
        check(lib().pasc_batch_create_problem(ctx.h, C.byref(d), dptr(self.theta), self.eps.size,
                                              dptr(self.eps), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().pasc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, **params):
        p = capi.default_params(**params)
        res = (capi.InstanceResult * self.eps.size)()
        t = C.c_double()
        check(lib().pasc_batch_solve(self.h, C.byref(p), res, C.byref(t)))
        return list(res), t.value

    def get_gamma(self, e: int, layout: int = capi.LAYOUT_TMAJOR):
        g = np.empty(self.nodes * self.K)
        check(lib().pasc_batch_get_gamma(self.h, 0, e, dptr(g), layout))
        return g


# ---- synthetic data --------------------------------------------------------------------------

def generate_markov(seed: int, T: int, levels, p_switch: float, sigma: float):
    lv = np.ascontiguousarray(levels, dtype=np.float64)
    x = np.empty(T)
    lab = np.empty(T, dtype=np.int32)
    check(lib().pasc_generate_markov(seed, T, lv.size, dptr(lv), p_switch, sigma, dptr(x),
                                     iptr(lab)))
    return x, lab


def generate_mask(seed: int, n: int, fraction: float):
    m = np.empty(n, dtype=np.uint8)
    check(lib().pasc_generate_mask(seed, n, fraction, u8ptr(m)))
    return m


def generate_uniform(seed: int, n: int):
    """n uniforms in [0, 1) of the library's SplitMix64 generator (the annealing restarts)."""
    u = np.empty(n)
    check(lib().pasc_generate_uniform(seed, n, dptr(u)))
    return u


def generate_rect_image(seed: int, width: int, height: int, K: int, n_rect: int, sigma: float):
    x = np.empty(width * height * 3)
    col = np.empty(K * 3)
    lab = np.empty(width * height, dtype=np.int32)
    check(lib().pasc_generate_rect_image(seed, width, height, K, n_rect, sigma, dptr(x),
                                         dptr(col), iptr(lab)))
    return x, col, lab


def generate_batch_params(seed: int, n_series: int, lg_tmin=4.0, lg_tmax=6.0, kmin=2, kmax=4,
                          level_max=5.0, pmin=1e-4, pmax=1e-2, smin=0.2, smax=1.5):
    T = np.empty(n_series, dtype=np.int64)
    K = np.empty(n_series, dtype=np.int32)
    lv = np.empty(n_series * capi.MAX_K)
    p = np.empty(n_series)
    s = np.empty(n_series)
    sd = np.empty(n_series, dtype=np.uint64)
    check(lib().pasc_generate_batch_params(
        seed, n_series, lg_tmin, lg_tmax, kmin, kmax, level_max, pmin, pmax, smin, smax,
        T.ctypes.data_as(C.POINTER(C.c_int64)), iptr(K), dptr(lv), dptr(p), dptr(s),
        sd.ctypes.data_as(C.POINTER(C.c_uint64))))
    return T, K, lv.reshape(n_series, capi.MAX_K), p, s, sd


# Workload definitions (sizes and seeds of the benchmark and parity configurations).
SIGNAL_C1 = dict(seed=0, T=10_000, levels=[1.0, 2.0], p_switch=0.002, sigma=0.5, eps=10.0,
                 theta0=[0.8, 2.2])
SIGNAL_C2 = dict(seed=1, T=10_000_000, levels=[0.0, 1.0, 3.0], p_switch=1e-4, sigma=1.0,
                 eps=100.0)
SIGNAL_C4 = dict(seed=3, T=1_000_000, levels=[0.0, 1.5, 3.0], p_switch=1e-3, sigma=0.8,
                 eps=30.0, heldout=0.1)
SCAN_EPS = 10.0 ** (-1.0 + 4.0 * np.arange(16) / 15.0)
IMAGE_C5 = dict(seed=4, width=1024, height=1024, K=4, n_rect=64, sigma=0.15,
                eps=10.0 ** (3.0 * np.arange(8) / 7.0))
