"""Pins the oracle's FEM coarsening (fem.cpp:28-182, femhat.cpp:26-217, one rank) on the CPU.

Fem: closed-form cells (round(t2*T1/T2) boundaries), cell sums and copies.  FemHat: the
prolongation is linear interpolation (exact on linear data, nodal values kept); the reduction's
weights are the hat function sampled at the fine steps, plus the two as-written details of
femhat.cpp:68-101 -- the step floor(left) enters with the weight (floor(left)-left)/diff (< 0
when left is not an integer) and steps at or beyond floor(diff*(T2-1)) are left out of the right
half of a hat.  The alternation with FEM is checked for its bookkeeping (coarse grid, scaling)
against a restatement built from the oracle's own reduce / SPG / prolongate pieces."""
import math

import numpy as np
import pytest


def _cells(T1, T2):
    d = T1 / T2
    return [(int(math.floor(t * d + 0.5)), int(math.floor((t + 1) * d + 0.5))) for t in range(T2)]


@pytest.mark.parametrize("T1,r", [(23, 0.3), (100, 0.1), (1000, 0.07), (37, 0.5)])
def test_fem_constant_cells(oracle, T1, r):
    T2 = oracle.fem_reduced_T(T1, r)
    assert T2 == math.ceil(T1 * r)
    K = 3
    x = np.random.default_rng(T1).random(T1 * K)
    red = oracle.fem_reduce(x, T1, T2, K, oracle.FEM_CONSTANT).reshape(T2, K)
    pro = oracle.fem_prolongate(red.ravel(), T1, T2, K, oracle.FEM_CONSTANT).reshape(T1, K)
    xs = x.reshape(T1, K)
    cells = _cells(T1, T2)
    assert cells[0][0] == 0 and cells[-1][1] == T1
    for t2, (a, b) in enumerate(cells):
        np.testing.assert_allclose(red[t2], xs[a:b].sum(0), rtol=1e-14)
        assert np.all(pro[a:b] == red[t2])
    assert abs(red.sum() - x.sum()) <= 1e-10 * x.sum()


@pytest.mark.parametrize("T1,T2", [(41, 11), (101, 21), (57, 9), (10, 4), (1000, 73)])
def test_femhat_prolongation_is_linear_interpolation(oracle, T1, T2):
    K = 2
    diff = (T1 - 1) / (T2 - 1)
    coarse = np.stack([3.0 - 0.25 * np.arange(T2), 1.0 + 2.0 * np.arange(T2)], 1)
    fine = oracle.fem_prolongate(coarse.ravel(), T1, T2, K, oracle.FEM_HAT).reshape(T1, K)
    t = np.arange(T1) / diff
    np.testing.assert_allclose(fine[:, 0], 3.0 - 0.25 * t, atol=1e-12)
    np.testing.assert_allclose(fine[:, 1], 1.0 + 2.0 * t, atol=1e-11)
    # nodal values are kept where a fine step sits on a coarse node
    for t2 in range(T2):
        t1 = t2 * diff
        if abs(t1 - round(t1)) < 1e-12:
            np.testing.assert_allclose(fine[int(round(t1))], coarse[t2], atol=1e-12)


def _hat_weights_written(T1, T2):
    """The weights of femhat.cpp:68-101 as written, restated: W[t2, t1]."""
    diff = (T1 - 1) / (T2 - 1)
    right_idx = math.floor(diff * (T2 - 1))
    W = np.zeros((T2, T1))
    for t2 in range(T2):
        c, l, r = t2 * diff, (t2 - 1) * diff, (t2 + 1) * diff
        t1 = math.floor(l)
        while t1 <= c:
            if t1 >= 0:
                W[t2, t1] += (t1 - l) / (c - l)
            t1 += 1
        while t1 < r:
            if t1 < right_idx:
                W[t2, t1] += (t1 - r) / (c - r)
            t1 += 1
    return W


@pytest.mark.parametrize("T1,T2", [(41, 11), (101, 21), (57, 9), (30, 7)])
def test_femhat_reduction_weights(oracle, T1, T2):
    K = 1
    diff = (T1 - 1) / (T2 - 1)
    W = np.zeros((T2, T1))
    for t1 in range(T1):
        e = np.zeros(T1)
        e[t1] = 1.0
        W[:, t1] = oracle.fem_reduce(e, T1, T2, K, oracle.FEM_HAT)
    np.testing.assert_allclose(W, _hat_weights_written(T1, T2), atol=1e-13)
    # away from the quirks the weights are the hat function at the fine steps
    for t2 in range(1, T2 - 1):
        c = t2 * diff
        for t1 in range(T1):
            hat = max(0.0, 1.0 - abs(t1 - c) / diff)
            l = (t2 - 1) * diff
            if t1 == math.floor(l) and l != math.floor(l):
                assert W[t2, t1] < 0 and abs(W[t2, t1] - (t1 - l) / diff) < 1e-12
            elif t1 < math.floor(diff * (T2 - 1)) or t1 <= c:
                assert abs(W[t2, t1] - hat) < 1e-12, (t2, t1)
    if abs(diff - round(diff)) < 1e-12:
        # integer spacing: interior hats sum to diff, no negative weight
        assert np.all(W >= -1e-15)
        np.testing.assert_allclose(W[1:-1].sum(1), diff, rtol=1e-12)


def test_fem_alternation_bookkeeping(oracle):
    """or_ts_solve with FEM == an outer loop assembled from the oracle's reduce, spg_solve and
    prolongate (scalef: A = eps/T * A0 on the coarse grid, b = -reduce(residual)/T2)."""
    from pasc_inference_amd.solver import generate_markov
    T, K, eps, r = 2000, 2, 10.0, 0.1
    x, truth = generate_markov(0, T, [1.0, 2.0], 0.01, 0.4)
    theta0 = np.array([0.8, 2.2])
    spg = dict(maxit=20000, stop_difff=0, stop_kkt=1, kkt_tol=1e-10, kkt_scale=1.0, nthreads=4,
               gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    for ftype in (oracle.FEM_CONSTANT, oracle.FEM_HAT):
        g0 = np.full(T * K, 1.0 / K)
        ts = dict(maxit=4, eps=0.0, skip_first_theta=1, fem_reduce=r, fem_type=ftype)
        g, th, res = oracle.ts_solve(x, g0, theta0, T, 1, K, 1, eps, ts=ts, **spg)
        # restatement
        T2 = oracle.fem_reduced_T(T, r)
        gam, tht = g0.copy(), theta0.copy()
        Ls = []
        for it in range(4):
            if it > 0:
                tht = oracle.theta_update(x, gam, T, 1, K)
            resid = ((x[:, None] - tht[None, :]) ** 2).ravel()
            b2 = -oracle.fem_reduce(resid, T, T2, K, ftype) / T2
            g2 = oracle.fem_reduce(gam, T, T2, K, ftype)
            g2, rr = oracle.spg_solve(g2, b2, T2, 1, K, eps / T, **spg)
            gam = oracle.fem_prolongate(g2, T, T2, K, ftype)
            Ls.append(rr.fx)
        print(ftype, res.it, res.L, Ls[-1])
        assert res.it == 4
        np.testing.assert_allclose(g, gam, atol=1e-12)
        np.testing.assert_allclose(th, tht, atol=1e-12)
        assert abs(res.L - Ls[-1]) <= 1e-12 * abs(Ls[-1])
        # the coarse solution still separates the two levels
        lab = g.reshape(T, K).argmax(1)
        assert np.mean(lab == truth) > 0.95
