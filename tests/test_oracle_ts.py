"""Pins of the oracle's outer alternation, its absolute error and its AIC (CPU only).

or_ts_solve restates TSSolver::solve (tssolver.h:491-715) and or_abserr_reconstructed restates
Signal1DData::compute_abserr_reconstructed (signal1Ddata.cpp:133-191).  Each check below is
chosen so that a plausible slip fails it:
  * order of the steps -- with a uniform gamma0 and no skipped theta step, the first theta step
    collapses every theta_k to mean(x), and the symmetric gamma step keeps gamma uniform; a
    gamma step first would move theta away from mean(x);
  * the alternation itself -- at epsilon = 0 the gamma step is the nearest-centre assignment, so
    the alternation is Lloyd's k-means, written here from its definition in numpy;
  * the stop -- the loop breaks at the first |L - L_old| < eps, before its it++ (it_last is the
    index of the breaking iteration), found here from the L sequence of runs of fixed length;
  * AIC = 2 log L + K (graphh1fem.h:443-447), and the annealing restarts (tssolver.h:522-692);
  * abserr -- an L1 norm written out by hand, the t-major index of gamma, and the identity with
    the reference's own recovered-signal file (recovered = sum_k theta_k gamma_k).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

SPG_TIGHT = dict(maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12, gd_centered=1,
                 gd_certified=1, dad_edges=1, g_recompute=1, safeguard=1, nthreads=4)


def _two_level_signal(seed, T, levels, sigma):
    rng = np.random.default_rng(seed)
    lab = np.repeat(rng.integers(0, len(levels), T // 25 + 1), 25)[:T]
    return np.asarray(levels)[lab] + sigma * rng.normal(size=T), lab


# ---------------------------------------------------------------------------------------------
# abserr (signal1Ddata.cpp:133-191)
# ---------------------------------------------------------------------------------------------

def test_abserr_is_the_l1_norm_of_the_reconstruction_error(oracle):
    """abserr = sum_t |sum_k theta_k gamma_tk - solution_t|, gamma t-major (K innermost)."""
    rng = np.random.default_rng(3)
    T, R, K = 7, 3, 3
    n = T * R
    gamma = rng.random(n * K)
    theta = np.array([0.5, -1.25, 2.0])
    sol = rng.normal(size=n)
    by_hand = 0.0
    for i in range(n):
        rec = 0.0
        for k in range(K):
            rec += theta[k] * gamma[i * K + k]
        by_hand += abs(rec - sol[i])
    assert oracle.abserr_reconstructed(gamma, theta, sol, T, R, K) == pytest.approx(by_hand, rel=1e-15)
    # a k-major reading of gamma gives a different value (the layout is pinned, not just the norm)
    km = np.abs(gamma.reshape(K, n).T @ theta - sol).sum()
    assert abs(km - by_hand) > 1e-3
    # one-hot gamma reconstructs theta of the label exactly
    lab = rng.integers(0, K, n)
    onehot = np.eye(K)[lab].ravel()
    assert oracle.abserr_reconstructed(onehot, theta, theta[lab], T, R, K) == 0.0


def test_abserr_on_the_reference_output_fixture(oracle):
    """The reference's gamma and recovered-signal files of signal1D_id1_idSigma1 (theta = {1, 2}):
    abserr of that gamma equals ||recovered - solution||_1 on the true signal."""
    from pasc_inference_amd import petsc_io
    d = os.path.join(GOLD, "reference_data")
    gref = petsc_io.read_vec(os.path.join(d, "signal1D_gauss_id1_idSigma_1_gamma.bin"))
    rec = petsc_io.read_vec(os.path.join(d, "signal1D_gauss_id1_idSigma_1_recovered.bin"))
    truth = petsc_io.read_vec(os.path.join(d, "samplesignal_solution.bin"))
    T, K = rec.size, 2
    a = oracle.abserr_reconstructed(gref, np.array([1.0, 2.0]), truth, T, 1, K)
    assert a == pytest.approx(float(np.abs(rec - truth).sum()), rel=1e-13)
    # with theta = {1, 2} and gamma on the simplex, a step whose argmax label is wrong
    # reconstructs at least 0.5 away from the true level
    wrong = np.sum(gref.reshape(T, K).argmax(1) != (truth > 1.5))
    assert a >= 0.5 * wrong


# ---------------------------------------------------------------------------------------------
# TSSolver::solve (tssolver.h:491-715)
# ---------------------------------------------------------------------------------------------

def test_ts_theta_step_comes_first(oracle):
    """Uniform gamma0, no skipped theta step: after one outer iteration every theta_k is mean(x)
    and gamma is still uniform (the symmetric fixed point of tssolver.h:548-596)."""
    x, _ = _two_level_signal(1, 400, [1.0, 2.0], 0.3)
    T, K = x.size, 2
    g, th, r = oracle.ts_solve(x, np.full(T * K, 0.5), [0.8, 2.2], T, 1, K, 1, 10.0,
                               ts=dict(maxit=1, eps=0.0), kkt_scale=T, **SPG_TIGHT)
    np.testing.assert_allclose(th, [x.mean(), x.mean()], rtol=1e-14)
    np.testing.assert_allclose(g, 0.5, atol=1e-14)
    assert r.it == 1  # the loop ran out (maxit), no break
    # with the first theta step skipped, the gamma step uses theta0 and moves gamma
    g2, th2, _ = oracle.ts_solve(x, np.full(T * K, 0.5), [0.8, 2.2], T, 1, K, 1, 10.0,
                                 ts=dict(maxit=1, eps=0.0, skip_first_theta=1), kkt_scale=T,
                                 **SPG_TIGHT)
    np.testing.assert_array_equal(th2, [0.8, 2.2])
    assert np.abs(g2 - 0.5).max() > 0.4


def _lloyd(x, theta0, maxit, eps):
    """Lloyd's k-means in the reference loop's shape: gamma step (nearest centre) with the current
    theta, L = mean of the squared distances, stop at the first |L - L_old| < eps, then the
    theta step (cluster means) at the top of the next iteration."""
    th = np.array(theta0, dtype=float)
    L = np.finfo(float).max
    ths, Ls = [], []
    for it in range(maxit):
        if it > 0:
            for k in range(th.size):
                if np.any(lab == k):
                    th[k] = x[lab == k].mean()
                else:
                    th[k] = 0.0
        lab = np.argmin((x[:, None] - th[None]) ** 2, 1)
        L_old, L = L, float(((x - th[lab]) ** 2).mean())
        ths.append(th.copy())
        Ls.append(L)
        if abs(L - L_old) < eps:
            return it, th, lab, Ls, ths
    return maxit, th, lab, Ls, ths


def test_ts_epsilon_zero_is_lloyd(oracle):
    """epsilon = 0: the alternation reproduces Lloyd's k-means -- the theta sequence, L and the
    outer iteration count -- from theta0 with the first theta step skipped.  With A = 0 the BB
    step restarts at alphainit on every iteration (dAd = 0), so a large alphainit makes every
    gamma step the exact nearest-centre assignment within a few iterations, near-ties included."""
    x, _ = _two_level_signal(5, 600, [0.0, 1.0, 3.0], 0.6)
    T, K = x.size, 3
    theta0 = [0.2, 1.4, 1.9]  # far enough from the levels that Lloyd needs several steps
    it_l, th_l, lab_l, Ls, _ = _lloyd(x, theta0, 100, 1e-12)
    assert it_l >= 3
    spg = dict(SPG_TIGHT, alphainit=1e12)
    g, th, r = oracle.ts_solve(x, np.full(T * K, 1.0 / K), theta0, T, 1, K, 1, 0.0,
                               ts=dict(maxit=100, eps=1e-12, skip_first_theta=1), kkt_scale=T,
                               **spg)
    assert r.it == it_l
    np.testing.assert_allclose(th, th_l, rtol=1e-13)
    np.testing.assert_array_equal(g.reshape(T, K).argmax(1), lab_l)
    assert r.L == pytest.approx(Ls[-1], rel=1e-12)
    # the same iterations, one at a time: the theta of every outer iteration
    _, _, _, _, ths = _lloyd(x, theta0, it_l + 1, 1e-12)
    for m in range(1, it_l + 1):
        _, th_m, _ = oracle.ts_solve(x, np.full(T * K, 1.0 / K), theta0, T, 1, K, 1, 0.0,
                                     ts=dict(maxit=m, eps=0.0, skip_first_theta=1), kkt_scale=T,
                                     **spg)
        np.testing.assert_allclose(th_m, ths[m - 1], rtol=1e-13, err_msg=f"outer iteration {m}")


def test_ts_stops_at_the_first_small_deltaL(oracle):
    """The stop fires at the first outer iteration with |L - L_old| < eps and reports that
    iteration's index (the reference breaks before it++); spg_it_sum counts every gamma solve."""
    x, _ = _two_level_signal(7, 800, [1.0, 2.0], 0.45)
    T, K = x.size, 2
    run = lambda maxit, eps: oracle.ts_solve(  # noqa: E731
        x, np.full(T * K, 0.5), [0.8, 2.2], T, 1, K, 1, 10.0,
        ts=dict(maxit=maxit, eps=eps, skip_first_theta=1), kkt_scale=T, **SPG_TIGHT)
    # the L sequence from runs of fixed length (eps = 0 never stops)
    Ls, spg = [], []
    for m in range(1, 7):
        _, _, r = run(m, 0.0)
        assert r.it == m
        Ls.append(r.L)
        spg.append(r.spg_it_sum)
    dL = np.abs(np.diff(Ls))
    assert np.all(dL[:-1] > 0)
    for j in (2, 3, 4):
        # eps just above |L_j - L_{j-1}| and below every earlier difference
        eps = dL[j - 1] * (1 + 1e-9)
        if np.any(dL[: j - 1] < eps):
            continue
        _, _, r = run(50, eps)
        assert r.it == j
        assert r.L == Ls[j]
        assert r.deltaL == pytest.approx(dL[j - 1], rel=1e-15)
        assert r.spg_it_sum == spg[j]  # j + 1 gamma solves


def test_ts_aic_and_annealing_restarts(oracle):
    """AIC = 2 log L + K (graphh1fem.h:443-447).  annealing = 3: run 0 from gamma0, runs 1 and 2
    from the same random vector (the reference's set_random re-seeds with 13 every time), each
    restart beginning with a theta step from the theta the previous run ended with; the run with
    the smallest AIC is kept, it is the last run's it_last and it_sum adds the runs up."""
    x, _ = _two_level_signal(11, 500, [0.0, 1.0, 3.0], 0.5)
    T, K = x.size, 3
    ts = dict(maxit=40, eps=1e-10, skip_first_theta=1)
    g0 = np.full(T * K, 1.0 / K)
    th0 = [2.9, 3.0, 3.1]  # a poor start: two centres on one level
    g, th, r = oracle.ts_solve(x, g0, th0, T, 1, K, 1, 5.0, ts=dict(ts, annealing=3),
                               kkt_scale=T, **SPG_TIGHT)
    assert r.aic == pytest.approx(2 * np.log(r.L) + K, rel=1e-15)
    # the same runs one by one
    u = oracle.uniform(13, T * K)
    np.testing.assert_array_equal(u, oracle.uniform(13, T * K))
    ga, tha, ra = oracle.ts_solve(x, g0, th0, T, 1, K, 1, 5.0, ts=ts, kkt_scale=T, **SPG_TIGHT)
    gb, thb, rb = oracle.ts_solve(x, u, tha, T, 1, K, 1, 5.0, ts=dict(ts, skip_first_theta=0),
                                  kkt_scale=T, **SPG_TIGHT)
    gc, thc, rc = oracle.ts_solve(x, u, thb, T, 1, K, 1, 5.0, ts=dict(ts, skip_first_theta=0),
                                  kkt_scale=T, **SPG_TIGHT)
    runs = [(ga, tha, ra), (gb, thb, rb), (gc, thc, rc)]
    aics = [2 * np.log(q[2].L) + K for q in runs]
    best = int(np.argmin(aics))
    assert r.best_run == best
    np.testing.assert_array_equal(g, runs[best][0])
    np.testing.assert_array_equal(th, runs[best][1])
    assert r.L == runs[best][2].L
    assert r.it == rc.it
    assert r.it_sum == ra.it + rb.it + rc.it
    assert r.spg_it_sum == ra.spg_it_sum + rb.spg_it_sum + rc.spg_it_sum
    # the restart found the three levels that the poor start missed
    assert aics[1] < aics[0]
    np.testing.assert_allclose(np.sort(th), [0.0, 1.0, 3.0], atol=0.1)
