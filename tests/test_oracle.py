"""Pins the FP64 CPU oracle (CPU only).

Each routine of oracle/pasc_oracle.c is checked against something it does not share code with:
closed forms and invariants, a brute-force support enumeration, an independently constructed
H1 matrix (edge-weighted graph Laplacian), exact minimisers of tiny QPs (tests/golden/tiny_qp.json,
brute force cross-checked with scipy), the reference's Matlab restatement of SPG-QP
(Matlab/signal1D_fem_matlab/compute_femh1/sqpqp.m) and the reference's own data files
(tests/golden/reference_data).
"""
import itertools
import json
import os

import numpy as np
import pytest

from helpers import grid_csr

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


# ---------------------------------------------------------------------------------------------
# projection (simplex_local.h:153-249)
# ---------------------------------------------------------------------------------------------

def _brute_projection(y):
    """Euclidean projection onto the simplex by enumerating the support."""
    K = y.size
    best, bd = None, np.inf
    for m in range(1, K + 1):
        for S in itertools.combinations(range(K), m):
            tau = (y[list(S)].sum() - 1.0) / m
            z = np.maximum(y - tau, 0.0)
            if abs(z.sum() - 1.0) > 1e-12:
                continue
            d = np.sum((z - y) ** 2)
            if d < bd:
                best, bd = z, d
    return best


def _duchi_projection(y):
    """Sort-descending threshold of Duchi et al. (2008)."""
    u = np.sort(y)[::-1]
    css = np.cumsum(u)
    j = np.arange(1, y.size + 1)
    rho = np.nonzero(u - (css - 1.0) / j > 0)[0][-1]
    tau = (css[rho] - 1.0) / (rho + 1)
    return np.maximum(y - tau, 0.0)


def test_projection_closed_forms(oracle):
    for K in [1, 2, 3, 5, 8]:
        # K = 1 gives 1
        if K == 1:
            np.testing.assert_array_equal(oracle.project_simplex([-3.7], 1), [1.0])
            np.testing.assert_array_equal(oracle.project_simplex([12.0], 1), [1.0])
            continue
        # equal entries give 1/K
        np.testing.assert_allclose(oracle.project_simplex(np.full(K, 0.37), K), 1.0 / K, rtol=1e-15)
        # an entry exceeding all others by >= 1 gives a one-hot vector
        y = np.linspace(-1.0, 0.0, K)
        y[K // 2] = 2.5
        z = oracle.project_simplex(y, K)
        e = np.zeros(K)
        e[K // 2] = 1.0
        np.testing.assert_array_equal(z, e)
        # exactly feasible input is returned unchanged (bit-for-bit)
        f = np.array([0.5] + [0.5 / (K - 1)] * (K - 1))
        if f.sum() == 1.0:
            np.testing.assert_array_equal(oracle.project_simplex(f, K), f)


def test_projection_matches_brute_force_and_duchi(oracle):
    rng = np.random.default_rng(1)
    for K in range(1, 7):
        Y = rng.normal(size=(400, K)) * rng.choice([0.01, 1.0, 30.0], size=(400, 1))
        Z = oracle.project_simplex(Y.ravel(), K).reshape(-1, K)
        for y, z in zip(Y, Z):
            np.testing.assert_allclose(z, _brute_projection(y), atol=1e-12)
            np.testing.assert_allclose(z, _duchi_projection(y), atol=1e-12)


def test_projection_invariants(oracle):
    rng = np.random.default_rng(2)
    K = 6
    Y = rng.normal(size=(300, K)) * 3
    Z = oracle.project_simplex(Y.ravel(), K).reshape(-1, K)
    # shift invariance P(y + c 1) = P(y)
    Zs = oracle.project_simplex((Y + 4.25).ravel(), K).reshape(-1, K)
    np.testing.assert_allclose(Zs, Z, atol=1e-13)
    # permutation equivariance
    perm = rng.permutation(K)
    Zp = oracle.project_simplex(Y[:, perm].ravel(), K).reshape(-1, K)
    np.testing.assert_allclose(Zp, Z[:, perm], atol=1e-14)
    # KKT of the projection: feasible, y - z = tau on the support, <= tau off it
    assert Z.min() >= 0.0
    np.testing.assert_allclose(Z.sum(1), 1.0, atol=1e-14)
    for y, z in zip(Y, Z):
        S = z > 0
        tau = (y - z)[S]
        assert np.ptp(tau) < 1e-12
        assert np.all(y[~S] <= tau[0] + 1e-12)


def test_sort_bubble(oracle):
    rng = np.random.default_rng(3)
    for n in [1, 2, 5, 17]:
        y = rng.normal(size=n)
        np.testing.assert_array_equal(oracle.sort_bubble(y), np.sort(y))


# ---------------------------------------------------------------------------------------------
# H1 matrix (blockgraphsparse.cpp:48-109, 151-194)
# ---------------------------------------------------------------------------------------------

def _laplacian_reference_graph(T, R, nbrs):
    import sys
    sys.path.insert(0, GOLD)
    from make_golden import laplacian_a0
    return laplacian_a0(T, R, nbrs)


@pytest.mark.parametrize("T,R,graph", [(1, 1, None), (2, 1, None), (7, 1, None), (1, 6, "path"),
                                       (4, 3, "path"), (3, 4, "grid"), (1, 12, "grid")])
def test_matvec_columns_equal_weighted_graph_laplacian(oracle, T, R, graph):
    """A*e_j columns (the reference's own check, test_blockgraphsparsematrix.cu:169-220) against
    A0 built as the Laplacian of the edge-weighted (t, r) graph."""
    K = 2
    if graph is None:
        nbrs = [[] for _ in range(R)]
        g = None
    elif graph == "path":
        nbrs = [[s for s in (r - 1, r + 1) if 0 <= s < R] for r in range(R)]
        ptr = np.cumsum([0] + [len(n) for n in nbrs])
        g = (ptr, np.concatenate([np.array(n, dtype=np.int32) for n in nbrs]))
    else:
        w = 2 if R == 4 else 4
        h = R // w
        g = grid_csr(w, h)
        nbrs = [list(g[1][g[0][r]:g[0][r + 1]]) for r in range(R)]
    L = _laplacian_reference_graph(T, R, nbrs)
    Afull = np.kron(L, np.eye(K))
    n = T * R * K
    cols = np.stack([oracle.h1_matvec(np.eye(n)[j], T, R, K, 1.0, graph=g) for j in range(n)], 1)
    np.testing.assert_array_equal(cols, Afull)
    # A*1 = 0, symmetry, and the scaling by alpha
    np.testing.assert_array_equal(oracle.h1_matvec(np.ones(n), T, R, K, 1.0, graph=g), 0.0)
    np.testing.assert_array_equal(cols, cols.T)
    x = np.random.default_rng(4).normal(size=n)
    np.testing.assert_allclose(oracle.h1_matvec(x, T, R, K, 0.25, graph=g), 0.25 * Afull @ x,
                               rtol=1e-14, atol=1e-14)


def test_h1_diagonal_entries(oracle):
    """Wsum of blockgraphsparse.cpp:57-73 (diagonal of A0) from the Laplacian degrees."""
    for T in [1, 2, 5]:
        for deg in [0, 1, 4]:
            nbrs = [list(range(1, deg + 1))] + [[0] for _ in range(deg)]
            L = _laplacian_reference_graph(T, deg + 1, nbrs)
            for t in range(T):
                assert oracle.h1_wsum(t, T, deg) == L[t * (deg + 1), t * (deg + 1)]


# ---------------------------------------------------------------------------------------------
# b, theta, objective
# ---------------------------------------------------------------------------------------------

def test_b_and_objective_on_one_hot_gamma(oracle):
    """For one-hot gamma and epsilon = 0 the objective is the k-means energy / T."""
    rng = np.random.default_rng(5)
    T, K = 50, 3
    x = rng.normal(size=T) * 2
    theta = np.array([-1.0, 0.5, 2.0])
    lab = rng.integers(0, K, T)
    g = np.eye(K)[lab].ravel()
    b = oracle.assemble_b(x, theta, T, 1, K)
    f = oracle.objective(g, b, T, 1, K, 0.0)
    np.testing.assert_allclose(f, np.sum((x - theta[lab]) ** 2) / T, rtol=1e-14)
    # held-out steps carry no data term; scalef = False leaves b unscaled
    ho = (rng.random(T) < 0.3).astype(np.uint8)
    bh = oracle.assemble_b(x, theta, T, 1, K, heldout=ho).reshape(T, K)
    assert np.all(bh[ho == 1] == 0.0)
    bu = oracle.assemble_b(x, theta, T, 1, K, scalef=False)
    np.testing.assert_allclose(bu, b * T, rtol=1e-14)
    # xdim = 3: squared Euclidean distance to the cluster colour
    X3 = rng.random((T, 3))
    th3 = rng.random((K, 3))
    b3 = oracle.assemble_b(X3.ravel(), th3.ravel(), T, 1, K, xdim=3).reshape(T, K)
    np.testing.assert_allclose(-b3 * T, ((X3[:, None, :] - th3[None]) ** 2).sum(-1), rtol=1e-13)


def test_theta_update_is_weighted_mean(oracle):
    rng = np.random.default_rng(6)
    T, K = 200, 4
    x = rng.normal(size=T)
    lab = rng.integers(0, K - 1, T)  # cluster K-1 stays empty -> theta 0 (graphh1fem.cpp:375-379)
    g = np.eye(K)[lab].ravel()
    th = oracle.theta_update(x, g, T, 1, K)
    for k in range(K - 1):
        np.testing.assert_allclose(th[k], x[lab == k].mean(), rtol=1e-13)
    assert th[K - 1] == 0.0
    # fractional gamma: weighted means
    G = oracle.project_simplex(rng.random(T * K), K).reshape(T, K)
    th = oracle.theta_update(x, G.ravel(), T, 1, K)
    np.testing.assert_allclose(th, (G * x[:, None]).sum(0) / G.sum(0), rtol=1e-13)


def test_kkt_residual_zero_iff_stationary(oracle):
    """The natural residual vanishes at the exact minimiser of a tiny QP and not elsewhere."""
    case = json.load(open(os.path.join(GOLD, "tiny_qp.json")))["cases"][5]
    T, R, K, a = case["T"], case["R"], case["K"], case["alpha"]
    b = np.array(case["b"])
    gam = np.array(case["gamma"])
    g = oracle.gradient(gam, b, T, R, K, a)
    assert oracle.kkt_residual(gam, g, K, 1.0) < 1e-10
    other = np.full_like(gam, 1.0 / K)
    assert oracle.kkt_residual(other, oracle.gradient(other, b, T, R, K, a), K, 1.0) > 1e-3


# ---------------------------------------------------------------------------------------------
# SPG-QP (spgqpsolver.cpp:55-367) against exact minimisers
# ---------------------------------------------------------------------------------------------

def test_spg_reaches_exact_minimisers_of_tiny_qps(oracle):
    cases = json.load(open(os.path.join(GOLD, "tiny_qp.json")))["cases"]
    assert len(cases) >= 20
    for c in cases:
        T, R, K, a = c["T"], c["R"], c["K"], c["alpha"]
        graph = None
        if c["graph_idx"]:
            graph = (np.array(c["graph_ptr"]), np.array(c["graph_idx"], dtype=np.int32))
        b = np.array(c["b"])
        x, r = oracle.spg_solve(np.full(T * R * K, 1.0 / K), b, T, R, K, a, graph=graph,
                                maxit=20000, stop_difff=0, stop_kkt=1, kkt_tol=1e-13,
                                kkt_scale=1.0, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1, safeguard=1)
        assert abs(r.fx - c["f"]) <= 1e-10 * max(1.0, abs(c["f"])), c
        assert abs(oracle.objective(x, b, T, R, K, a, graph=graph) - c["f"]) <= 1e-10
        # the minimiser is unique when A is definite on the feasible face; compare where the
        # brute force and SPG agree on the support
        np.testing.assert_allclose(x, c["gamma"], atol=1e-6)


@pytest.mark.parametrize("R,graph", [(1, None), (6, "grid")])
def test_robust_dot_forms_equal_op_by_op(oracle, R, graph):
    """The minimiser options (certified <g,d>, d'Ad over edges, g recomputed) change only the
    rounding: one iteration from the same start gives the same dAd and f as the op-by-op
    formulas, and both runs reach the same minimiser."""
    T, K, eps = 400, 3, 5.0
    rng = np.random.default_rng(3)
    csr = grid_csr(3, 2) if graph else None
    b = -rng.random(T * R * K) / T
    g0 = rng.random(T * R * K)
    base = dict(stop_difff=0, nthreads=2)
    rob = dict(base, gd_certified=1, dad_edges=1, g_recompute=1)
    for maxit in (1, 2):
        _, r0 = oracle.spg_solve(g0, b, T, R, K, eps / T, graph=csr, maxit=maxit, **base)
        _, r1 = oracle.spg_solve(g0, b, T, R, K, eps / T, graph=csr, maxit=maxit, **rob)
        assert r1.dAd == pytest.approx(r0.dAd, rel=1e-11)
        assert r1.dd == pytest.approx(r0.dd, rel=1e-12)
        assert r1.fx == pytest.approx(r0.fx, rel=1e-13)
    x0, r0 = oracle.spg_solve(g0, b, T, R, K, eps / T, graph=csr, maxit=200000, stop_kkt=1,
                              kkt_tol=1e-10, kkt_scale=T, **dict(base, gd_centered=1))
    x1, r1 = oracle.spg_solve(g0, b, T, R, K, eps / T, graph=csr, maxit=200000, stop_kkt=1,
                              kkt_tol=1e-12, kkt_scale=T, **rob)
    assert r0.kkt <= 1e-10 and r1.kkt <= 1e-12
    assert np.max(np.abs(x0 - x1)) <= 1e-7
    assert r1.fx == pytest.approx(r0.fx, rel=1e-12)


def test_reference_divides_by_zero_when_A_vanishes(oracle):
    """T = 1 without edges makes A = 0: the reference's step (spgqpsolver.cpp:182-216) divides
    by dAd = 0 and the BB step becomes infinite; the faithful oracle reproduces the non-finite
    iterate, the safeguarded variant (what the device path does) solves the LP."""
    b = np.array([-0.2, -1.5, -0.7])
    x, r = oracle.spg_solve(np.full(3, 1 / 3), b, 1, 1, 3, 0.5, maxit=5, stop_difff=0)
    # the infinite BB step drives the iterate off the simplex
    assert not np.all(np.isfinite(x)) or abs(x.sum() - 1.0) > 0.5
    x, r = oracle.spg_solve(np.full(3, 1 / 3), b, 1, 1, 3, 0.5, maxit=50, stop_difff=0,
                            safeguard=1)
    np.testing.assert_allclose(x, [1.0, 0.0, 0.0], atol=1e-15)  # smallest -b_k gets all weight


def test_spg_reference_iteration_invariants(oracle):
    """Feasibility, the f identity f_new = f + beta*gd + beta^2*dAd/2 (spgqpsolver.h:835-842,
    the generic twin of get_fx), and that the default |df| < eps stop terminates."""
    from pasc_inference_amd.solver import generate_markov
    T, K, eps = 3000, 3, 50.0
    x, _ = generate_markov(3, T, [0.0, 1.0, 3.0], 2e-3, 0.8)
    b = oracle.assemble_b(x, [0.0, 1.0, 3.0], T, 1, K)
    prev = None
    for maxit in [1, 2, 3, 10, 40]:
        g, r = oracle.spg_solve(np.full(T * K, 1 / 3), b, T, 1, K, eps / T, maxit=maxit,
                                stop_difff=0)
        assert g.min() >= -1e-15
        np.testing.assert_allclose(g.reshape(T, K).sum(1), 1.0, atol=1e-12)
        assert r.it == maxit and r.hessmult == maxit + 1
        prev = r
    gi, ri = oracle.spg_solve(np.full(T * K, 1 / 3), b, T, 1, K, eps / T, maxit=40, stop_difff=0,
                              fx_identity=1)
    assert abs(ri.fx - prev.fx) < 1e-12
    g, r = oracle.spg_solve(np.full(T * K, 1 / 3), b, T, 1, K, eps / T)
    assert r.stop_reason == 1 and r.it < 1000


def _sqpqp_matlab(A, b, x0, K, maxit):
    """Matlab/signal1D_fem_matlab/compute_femh1/sqpqp.m (lines 17-69), k-major vectors as in the
    Matlab code, with its FIFO fs window and without the stopping test (fixed iteration count)."""
    def proj(v):
        T = v.size // K
        out = v.copy()
        for i in range(T):
            idx = np.arange(i, i + K * T, T)
            y = np.sort(v[idx])
            n = K
            t_hat = 0.0
            j = n - 1
            while j >= 1:
                ti = (y[j:].sum() - 1) / (n - j)
                if ti >= y[j - 1]:
                    t_hat = ti
                    break
                j -= 1
            if j == 0:
                t_hat = (y.sum() - 1) / n
            out[idx] = np.maximum(v[idx] - t_hat, 0)
        return out
    x = proj(x0)
    m, gamma, sigma2 = 20, 0.9, 1.0
    g = A @ x - b
    f = 0.5 * (g - b) @ x
    fs = f * np.ones(m)
    a = 2.0
    for _ in range(maxit):
        d = proj(x - a * g) - x
        dd = d @ d
        Ad = A @ d
        dAd = Ad @ d
        xi = (fs.max() - f) / dAd
        bb = -(g @ d) / dAd
        bh = gamma * bb + np.sqrt(gamma ** 2 * bb ** 2 + 2 * xi)
        beta = min(sigma2, bh)
        x = x + beta * d
        g = g + beta * Ad
        f = 0.5 * (g - b) @ x
        fs[:-1] = fs[1:]
        fs[-1] = f
        a = dd / dAd
    return x, f


def test_spg_matches_reference_matlab_restatement(oracle):
    rng = np.random.default_rng(8)
    T, K, eps = 60, 3, 20.0
    x = rng.normal(size=T) + np.repeat([0.0, 2.0, 1.0], T // 3)
    th = np.array([0.0, 1.0, 2.0])
    b = oracle.assemble_b(x, th, T, 1, K)
    A = (eps / T) * np.stack([oracle.h1_matvec(np.eye(T * K)[j], T, 1, K, 1.0)
                              for j in range(T * K)], 1)
    # Matlab works k-major; permute
    P = np.arange(T * K).reshape(T, K).T.ravel()
    Ak = A[np.ix_(P, P)]
    bk = b[P]
    for maxit in [1, 5, 25]:
        xm, fm = _sqpqp_matlab(Ak, bk, np.full(T * K, 1.0 / K), K, maxit)
        xo, ro = oracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, eps / T, maxit=maxit,
                                  stop_difff=0)
        np.testing.assert_allclose(xo[P], xm, atol=1e-12)
        assert abs(ro.fx - fm) < 1e-13


def test_lloyd_limit_epsilon_zero(oracle):
    """epsilon -> 0: gamma becomes the one-hot nearest-centre assignment (k-means step)."""
    rng = np.random.default_rng(9)
    T, K = 500, 3
    th = np.array([0.0, 1.0, 3.0])
    x = th[rng.integers(0, K, T)] + 0.3 * rng.normal(size=T)
    b = oracle.assemble_b(x, th, T, 1, K)
    g, r = oracle.spg_solve(np.full(T * K, 1 / 3), b, T, 1, K, 0.0, maxit=200000, stop_difff=0,
                            stop_kkt=1, kkt_tol=1e-12, kkt_scale=T, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1, safeguard=1)
    assert r.stop_reason == 6
    lab = np.argmin((x[:, None] - th[None]) ** 2, 1)
    np.testing.assert_allclose(g.reshape(T, K), np.eye(K)[lab], atol=1e-9)


def test_large_epsilon_limit_constant_gamma(oracle):
    """epsilon -> infinity: every gamma_k becomes constant in t (zero H1 energy)."""
    rng = np.random.default_rng(10)
    T, K = 300, 2
    th = np.array([0.0, 1.0])
    x = rng.normal(size=T) * 0.5 + 0.3
    b = oracle.assemble_b(x, th, T, 1, K)
    g, r = oracle.spg_solve(np.full(T * K, 0.5), b, T, 1, K, 1e6 / T, maxit=200000,
                            stop_difff=0, stop_kkt=1, kkt_tol=1e-10, kkt_scale=T, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    G = g.reshape(T, K)
    assert np.ptp(G, axis=0).max() < 1e-6
    # the constant is the LP optimum: all weight on the cluster with the smaller data cost
    best = np.argmin(((x[:, None] - th[None]) ** 2).sum(0))
    np.testing.assert_allclose(G[:, best], 1.0, atol=1e-6)


def test_ts_alternation_monotone_on_reference_fixture(oracle):
    """TSSolver alternation (theta then gamma, tssolver.h:542-650) on the reference's
    samplesignal-family data: L never increases and gamma recovers the true levels."""
    from pasc_inference_amd import petsc_io
    d = os.path.join(GOLD, "reference_data")
    x = petsc_io.read_vec(os.path.join(d, "signal1D_id1_idSigma1.bin"))
    truth = petsc_io.read_vec(os.path.join(d, "samplesignal_solution.bin"))
    T, K = x.size, 2
    g, th, r = oracle.ts_solve(x, np.full(T * K, 0.5), [0.8, 2.2], T, 1, K, 1, 10.0,
                               ts=dict(maxit=50, eps=1e-9, skip_first_theta=1),
                               maxit=20000, stop_difff=0, stop_kkt=1, kkt_tol=1e-9,
                               kkt_scale=T, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1, nthreads=4)
    assert r.increased == 0
    np.testing.assert_allclose(th, [1.0, 2.0], atol=0.05)  # clusters keep the order of theta0
    lab = g.reshape(T, K).argmax(1)
    assert np.mean(lab == (truth > 1.5)) > 0.99


def test_labels_match_reference_output_fixture(oracle):
    """The reference's own gamma output for signal1D_id1_idSigma1 (an older build, t-major, K = 2,
    theta = {1, 2}, epsilon chosen by its best-abserr rule, test_signal1D.cpp:346-351) against the
    oracle at the best-abserr epsilon of a grid.  Label-level agreement only: the epsilon of the
    reference run is not recorded."""
    from pasc_inference_amd import petsc_io
    d = os.path.join(GOLD, "reference_data")
    x = petsc_io.read_vec(os.path.join(d, "signal1D_id1_idSigma1.bin"))
    gref = petsc_io.read_vec(os.path.join(d, "signal1D_gauss_id1_idSigma_1_gamma.bin"))
    rec = petsc_io.read_vec(os.path.join(d, "signal1D_gauss_id1_idSigma_1_recovered.bin"))
    truth = petsc_io.read_vec(os.path.join(d, "samplesignal_solution.bin"))
    T, K = x.size, 2
    th = np.array([1.0, 2.0])
    np.testing.assert_allclose(gref.reshape(T, K) @ th, rec, atol=1e-12)
    b = oracle.assemble_b(x, th, T, 1, K)
    best = None
    for eps in [0.1, 1.0, 10.0, 100.0]:
        g, r = oracle.spg_solve(np.full(T * K, 0.5), b, T, 1, K, eps / T, maxit=50000,
                                stop_difff=0, stop_kkt=1, kkt_tol=1e-8, kkt_scale=T,
                                gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1, nthreads=4)
        ae = oracle.abserr_reconstructed(g, th, truth, T, 1, K)
        if best is None or ae < best[0]:
            best = (ae, g)
    lab = best[1].reshape(T, K).argmax(1)
    agree = np.mean(lab == gref.reshape(T, K).argmax(1))
    assert agree >= 0.995, agree


def test_petsc_vec_roundtrip(tmp_path):
    from pasc_inference_amd import petsc_io
    v = np.random.default_rng(0).normal(size=37)
    p = str(tmp_path / "v.bin")
    petsc_io.write_vec(p, v)
    np.testing.assert_array_equal(petsc_io.read_vec(p), v)
    raw = open(p, "rb").read()
    assert raw[:4] == (1211214).to_bytes(4, "big")
    g = np.arange(12.0)
    np.testing.assert_array_equal(petsc_io.tmajor_to_kmajor(petsc_io.kmajor_to_tmajor(g, 3), 3), g)


def test_unrolled_spg_reference_matches_central_differences(oracle):
    """Pins the torch restatement of SPG used as the unrolled-gradient reference of the adjoint
    test: its objective matches the oracle and its tangent matches central differences."""
    import torch
    from torch.func import jvp
    from test_gpu_adjoint import _case, _torch_spg_loss
    T = 600
    x, mask, theta, eps = _case(T, seed=11)
    K = theta.size

    def J_oracle(le, th):
        b = oracle.assemble_b(x, th, T, 1, K, heldout=mask)
        g, _ = oracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, np.exp(le) / T,
                                maxit=100000, stop_difff=0, stop_kkt=1, kkt_tol=1e-13,
                                kkt_scale=T, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
        return float(np.sum(((x - g.reshape(T, K) @ th)[mask == 1]) ** 2))

    le = np.log(eps)
    f = lambda a, b: _torch_spg_loss(x, mask, b, a, T, K, 600)  # noqa: E731
    J, d = jvp(f, (torch.tensor(le, dtype=torch.float64), torch.tensor(theta)),
               (torch.tensor(1.0, dtype=torch.float64), torch.zeros(K, dtype=torch.float64)))
    assert abs(J.item() - J_oracle(le, theta)) <= 1e-9 * J.item()
    h = 1e-4
    fd = (J_oracle(le + h, theta) - J_oracle(le - h, theta)) / (2 * h)
    assert abs(d.item() - fd) <= 1e-5 * max(abs(fd), 1e-3 * J.item())
