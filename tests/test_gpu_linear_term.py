"""The explicit-b boundary: pasc_set_linear_term hands the QP's linear term over as the
reference's QPData holds it (qpdata->get_b(), spgqpsolver.cpp:60), so the device SPG solves
min 1/2 g'Ag - b'g over the simplices for any b, not only the k-means data term."""
import numpy as np
import pytest

from helpers import grid_csr, label_agreement

pytestmark = pytest.mark.gpu


def _rand_b(seed, T, K, scale):
    """A linear term that no (x, theta) pair produces: smooth random clusters scores."""
    rng = np.random.default_rng(seed)
    walk = np.cumsum(rng.normal(0.0, 0.05, size=(T, K)), axis=0)
    return (-(walk - walk.mean(0)) ** 2 - rng.uniform(0, 0.2, size=(T, K))) * scale


@pytest.mark.parametrize("prec_name,K,eps", [("FP64", 3, 10.0), ("FP64", 5, 1.0),
                                             ("FP32_BULK", 3, 10.0), ("FP32_BULK", 4, 30.0)])
def test_explicit_b_matches_oracle_minimiser(ctx, oracle, prec_name, K, eps):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    prec = getattr(capi, prec_name)
    T = 12000
    b = _rand_b(K * 7 + int(eps), T, K, 1.0 / T).ravel()  # t-major, 1/T-scaled like the reference
    g0 = np.full(T * K, 1.0 / K)
    ref, rr = oracle.spg_solve(g0, b, T, 1, K, eps / T, maxit=400000, stop_difff=0, stop_kkt=1,
                               kkt_tol=1e-12, kkt_scale=T, nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    assert rr.kkt <= 1e-12
    x = np.zeros(T)  # the data are ignored once b is given
    P = Problem(ctx, x, K, precision=prec, epssqr=eps)
    P.set_linear_term(b)
    P.set_gamma_uniform()
    r = P.gamma_solve(maxit=400000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    f, kkt = P.objective()
    print(f"{prec_name} K={K} eps={eps}: it={r.it} ({r.it_fp32}+{r.it_fp64}) kkt={kkt:.2e} "
          f"f={f!r} oracle f={rr.fx!r} err={np.max(np.abs(g - ref)):.2e}", flush=True)
    assert r.stop_reason == capi.REASON_KKT and kkt <= 1e-11 * 1.001
    assert np.max(np.abs(g - ref)) <= 1e-6
    assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)
    assert label_agreement(g, ref, K) >= 0.9999


def test_explicit_b_equals_data_term_and_layouts(ctx, oracle):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_markov
    T, K, eps = 15000, 3, 20.0
    levels = [0.0, 1.0, 3.0]
    x, _ = generate_markov(4, T, levels, 1e-3, 0.7)
    theta = np.array(levels)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P.set_theta(theta)
    P.set_gamma_uniform()
    r0 = P.gamma_solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g_data = P.get_gamma()
    f_data, _ = P.objective()
    b = oracle.assemble_b(x, theta, T, 1, K)
    for layout, bl in [(capi.LAYOUT_TMAJOR, b),
                       (capi.LAYOUT_KMAJOR, b.reshape(T, K).T.copy().ravel())]:
        P.set_linear_term(bl, layout)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        g = P.get_gamma()
        f, _ = P.objective()
        print(f"layout={layout}: it={r.it} vs data term it={r0.it}; "
              f"gamma diff={np.max(np.abs(g - g_data)):.2e} f={f!r} vs {f_data!r}", flush=True)
        assert np.max(np.abs(g - g_data)) <= 1e-8
        assert abs(f - f_data) <= 1e-12 * abs(f_data)
    # the alternation and the adjoint need the data term
    with pytest.raises(capi.PascError):
        P.ts_solve()
    with pytest.raises(capi.PascError):
        P.adjoint(np.zeros(T, dtype=np.uint8))
    # NULL returns to the data term
    P.set_linear_term(None)
    P.set_gamma_uniform()
    P.gamma_solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    assert np.max(np.abs(P.get_gamma() - g_data)) <= 1e-12


def test_explicit_b_on_grid(ctx, oracle):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    w, h, K, eps = 40, 32, 3, 2.0
    R = w * h
    b = _rand_b(21, R, K, 1.0).ravel()
    graph = grid_csr(w, h)
    ref, rr = oracle.spg_solve(np.full(R * K, 1.0 / K), b, 1, R, K, eps, graph=graph,
                               maxit=400000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                               kkt_scale=1.0, nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    for prec in [capi.FP64, capi.FP32_BULK]:
        P = Problem(ctx, np.zeros(R), K, T=1, R=R, precision=prec, grid_width=w, epssqr=eps)
        P.set_linear_term(b)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=400000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        g = P.get_gamma()
        f, kkt = P.objective()
        print(f"grid prec={prec}: it={r.it} kkt={kkt:.2e} err={np.max(np.abs(g - ref)):.2e}", flush=True)
        assert kkt <= 1e-11 * 1.001
        assert np.max(np.abs(g - ref)) <= 1e-6
        assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)


def test_problem_without_data_is_defined_by_b(ctx, oracle):
    """The SPG-level drop-in: no data vector at all, A = coeff*A0 and b as QPData holds them
    (coeff = epssqr/T with scalef, so epssqr = coeff*T keeps the per-sample KKT scale)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, K, coeff = 9000, 3, 4.0e-3
    b = _rand_b(5, T, K, 1.0 / T).ravel()
    ref, rr = oracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, coeff, maxit=400000,
                               stop_difff=0, stop_kkt=1, kkt_tol=1e-12, kkt_scale=T, nthreads=8,
                               gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    P = Problem(ctx, None, K, T=T, precision=capi.FP64, epssqr=coeff * T)
    P.set_gamma_uniform()
    with pytest.raises(capi.PascError):
        P.gamma_solve(maxit=10)  # no data and no b
    with pytest.raises(capi.PascError):
        P.theta_update()
    P.set_linear_term(b)
    r = P.gamma_solve(maxit=400000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    print(f"no-data problem: it={r.it} kkt={r.kkt:.2e} err={np.max(np.abs(g - ref)):.2e}", flush=True)
    assert r.stop_reason == capi.REASON_KKT
    assert np.max(np.abs(g - ref)) <= 1e-6 and abs(r.fx - rr.fx) <= 1e-8 * abs(rr.fx)
    # the reference's own default stop (|f - f_old| < 1e-9) through the same entry point
    P.set_gamma_uniform()
    r2 = P.gamma_solve(maxit=100000)
    assert r2.stop_reason == capi.REASONS.index("difff") and r2.it > 0
