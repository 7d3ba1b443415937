"""The outer alternation on the device (pasc_ts_solve, TSSolver::solve tssolver.h:491-715)
against the oracle: the outer iteration counter (it_last: the index of the breaking
iteration), QPSolver::it_sum, L, AIC = 2 log L + K (graphh1fem.h:443-447) and the annealing
restarts (tssolver.h:522-692)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPG_O = dict(maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12, gd_centered=1,
             gd_certified=1, dad_edges=1, g_recompute=1, safeguard=1, nthreads=8)


def _signal(seed, T, levels, sigma):
    rng = np.random.default_rng(seed)
    lab = np.repeat(rng.integers(0, len(levels), T // 25 + 1), 25)[:T]
    return np.asarray(levels)[lab] + sigma * rng.normal(size=T)


def test_ts_counters_and_aic_match_oracle(ctx, oracle):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x = _signal(7, 800, [1.0, 2.0], 0.45)
    T, K, eps = x.size, 2, 10.0
    ts = dict(maxit=50, eps=1e-7, skip_first_theta=1)
    g_o, th_o, r_o = oracle.ts_solve(x, np.full(T * K, 0.5), [0.8, 2.2], T, 1, K, 1, eps,
                                     ts=ts, kkt_scale=T, **SPG_O)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P.set_theta([0.8, 2.2])
    P.set_gamma_uniform()
    r = P.ts_solve(ts=ts, maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-12)
    g, th = P.get_gamma(), P.get_theta()
    P.close()
    print(f"oracle it={r_o.it} spg={r_o.spg_it_sum} L={r_o.L!r}; "
          f"device it={r.it} spg={r.spg_it_sum} L={r.L!r}", flush=True)
    assert r.it == r_o.it
    assert r.it_sum == r.it and r.best_run == 0
    assert r.L == pytest.approx(r_o.L, rel=1e-10)
    assert r.aic == pytest.approx(2.0 * np.log(r_o.L) + K, rel=1e-10)
    # it + 1 gamma solves; their SPG counts agree to within a few iterations per solve
    assert abs(r.spg_it_sum - r_o.spg_it_sum) <= 0.1 * r_o.spg_it_sum
    np.testing.assert_allclose(th, th_o, atol=1e-8)
    assert np.max(np.abs(g - g_o)) <= 1e-6


def test_ts_annealing_keeps_the_best_run(ctx, oracle):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x = _signal(11, 500, [0.0, 1.0, 3.0], 0.5)
    T, K, eps = x.size, 3, 5.0
    th0 = [2.9, 3.0, 3.1]  # a poor start: two centres on one level
    ts = dict(maxit=40, eps=1e-9, skip_first_theta=1, annealing=3)
    g_o, th_o, r_o = oracle.ts_solve(x, np.full(T * K, 1.0 / K), th0, T, 1, K, 1, eps, ts=ts,
                                     kkt_scale=T, **SPG_O)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P.set_theta(th0)
    P.set_gamma_uniform()
    r = P.ts_solve(ts=ts, maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-12)
    g, th = P.get_gamma(), P.get_theta()
    P.close()
    print(f"oracle best={r_o.best_run} it={r_o.it} it_sum={r_o.it_sum} L={r_o.L!r}; device "
          f"best={r.best_run} it={r.it} it_sum={r.it_sum} L={r.L!r}", flush=True)
    assert r.best_run == r_o.best_run == 1
    assert r.it == r_o.it and r.it_sum == r_o.it_sum
    assert r.L == pytest.approx(r_o.L, rel=1e-10)
    assert r.aic == pytest.approx(r_o.aic, rel=1e-10)
    np.testing.assert_allclose(th, th_o, atol=1e-8)
    assert np.max(np.abs(g - g_o)) <= 1e-6
