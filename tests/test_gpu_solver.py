"""GPU parity of the gamma step, theta step and alternation against the FP64 CPU oracle.

The SPG trajectory is path dependent (reduction order), so parity is checked on the converged
minimiser (gamma max-abs, objective relative error, argmax labels) and, for the first
iterations, on the objective sequence of the exact reference iteration.
"""
import numpy as np
import pytest

from helpers import grid_csr, label_agreement

pytestmark = pytest.mark.gpu


def _signal(seed, T, levels, p, sigma):
    from pasc_inference_amd.solver import generate_markov
    return generate_markov(seed, T, levels, p, sigma)


def _oracle_min(oracle, x, theta, T, K, eps, tol=1e-12, graph=None, R=1, xdim=1, heldout=None):
    """FP64 minimiser of the gamma QP: the oracle SPG driven to KKT tol, with the sign-safe <g,d>,
    d'Ad as a sum over edges and g recomputed from gamma (the op-by-op formulas stall near
    KKT 1e-9 on some trajectories)."""
    b = oracle.assemble_b(x, theta, T, R, K, xdim=xdim, heldout=heldout)
    alpha = eps / T
    g0 = np.full(T * R * K, 1.0 / K)
    xs, r = oracle.spg_solve(g0, b, T, R, K, alpha, graph=graph, maxit=400000, stop_difff=0,
                             stop_kkt=1, kkt_tol=tol, kkt_scale=T, nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    assert r.kkt <= tol, "oracle minimiser did not converge"
    return xs, r, b


@pytest.mark.parametrize("K,eps", [(2, 10.0), (3, 100.0), (4, 1.0)])
def test_fp64_solve_matches_oracle_minimiser(ctx, oracle, K, eps):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T = 20000
    levels = [0.0, 1.0, 3.0, 4.5][:K]
    x, _ = _signal(11 + K, T, levels, 1e-3, 0.8)
    theta = np.array(levels)
    ref, rr, b = _oracle_min(oracle, x, theta, T, K, eps)
    print(f"oracle K={K} eps={eps}: it={rr.it} kkt={rr.kkt:.2e} f={rr.fx!r}", flush=True)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P.set_theta(theta)
    P.set_gamma_uniform()
    r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    f, kkt = P.objective()
    print(f"gpu: it={r.it} kkt={r.kkt:.2e} f={r.fx!r} reason={capi.REASONS[r.stop_reason]} "
          f"gamma err={np.max(np.abs(g - ref)):.2e}")
    assert r.stop_reason == 6 and r.kkt <= 1e-11
    assert abs(kkt - r.kkt) <= 1e-12 + 1e-3 * r.kkt
    assert np.max(np.abs(g - ref)) <= 1e-6
    assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)
    assert abs(r.fx - f) <= 1e-10 * abs(f)
    assert label_agreement(g, ref, K) >= 0.9999


def test_fp64_iteration_semantics_follow_reference(ctx, oracle):
    """Default reference settings (stop |df| < 1e-9): the device iteration reproduces the
    oracle's objective sequence closely and stops at (nearly) the same iteration."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, K, eps = 4000, 2, 10.0
    x, _ = _signal(0, T, [1.0, 2.0], 0.002, 0.5)
    theta = np.array([1.0, 2.0])
    b = oracle.assemble_b(x, theta, T, 1, K)
    for maxit in [1, 2, 5, 20]:
        xs, rr = oracle.spg_solve(np.full(T * K, 0.5), b, T, 1, K, eps / T, maxit=maxit,
                                  stop_difff=0)
        P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
        P.set_theta(theta)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=maxit, stop_mask=0)
        g = P.get_gamma()
        print(f"maxit={maxit}: oracle f={rr.fx!r} gpu f={r.fx!r} dg={np.max(np.abs(g - xs)):.2e}")
        assert r.it == maxit
        assert abs(r.fx - rr.fx) <= 1e-12 * abs(rr.fx) + 1e-15
        assert np.max(np.abs(g - xs)) <= 1e-10
    xs, rr = oracle.spg_solve(np.full(T * K, 0.5), b, T, 1, K, eps / T)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P.set_theta(theta)
    P.set_gamma_uniform()
    r = P.gamma_solve()
    print(f"default stop: oracle it={rr.it} f={rr.fx!r}; gpu it={r.it} f={r.fx!r}")
    assert abs(r.it - rr.it) <= max(3, rr.it // 10)
    assert abs(r.fx - rr.fx) <= 1e-8 * abs(rr.fx)


@pytest.mark.parametrize("K", [2, 3, 4])
def test_infeasible_start_and_stored_rows(ctx, oracle, K):
    """The pass stores K - 1 rows (the last is rebuilt from the simplex constraint): from an
    infeasible random gamma_0 (whose last row the first pass must read) the iterates follow
    the oracle's, and the returned gamma has all K rows, non-negative and summing to 1."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, eps = 3001, 10.0
    levels = np.linspace(0.0, 3.0, K)
    x, _ = _signal(30 + K, T, levels, 0.002, 0.5)
    b = oracle.assemble_b(x, levels, T, 1, K)
    g0 = np.random.default_rng(K).random(T * K) * 2.0 - 0.5  # t-major, off the simplex
    for maxit in [1, 2, 7]:
        xs, rr = oracle.spg_solve(g0, b, T, 1, K, eps / T, maxit=maxit, stop_difff=0)
        P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
        P.set_theta(levels)
        P.set_gamma(g0)
        r = P.gamma_solve(maxit=maxit, stop_mask=0)
        g = P.get_gamma()
        P.close()
        assert r.it == maxit
        assert np.max(np.abs(g - xs)) <= 1e-10
        G = g.reshape(T, K)
        assert G.min() >= 0.0 and np.max(np.abs(G.sum(axis=1) - 1.0)) <= 1e-14
        assert abs(r.fx - rr.fx) <= 1e-12 * abs(rr.fx) + 1e-15


@pytest.mark.parametrize("eps", [1.0, 100.0])
def test_fp32_bulk_reaches_fp64_certified_kkt(ctx, oracle, eps):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, K = 50000, 3
    x, _ = _signal(1, T, [0.0, 1.0, 3.0], 1e-3, 1.0)
    theta = np.array([0.0, 1.0, 3.0])
    ref, rr, b = _oracle_min(oracle, x, theta, T, K, eps)
    P = Problem(ctx, x, K, precision=capi.FP32_BULK, epssqr=eps)
    P.set_theta(theta)
    P.set_gamma_uniform()
    # the metric's target (KKT 1e-6), then the parity target (FP64-certified KKT 1e-9)
    # KKT 1e-6 bounds the objective error only to a few 1e-8 (relative) at small epsilon
    for tol, gtol, ftol in [(1e-6, 1e-4, 5e-8), (1e-9, 1e-6, 1e-8)]:
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=tol)
        g = P.get_gamma()
        f, kkt = P.objective()
        err = np.max(np.abs(g - ref))
        print(f"fp32-bulk eps={eps} tol={tol}: it32={r.it_fp32} it64={r.it_fp64} kkt={kkt:.2e} "
              f"gamma err={err:.2e} f rel={abs(f - rr.fx) / abs(rr.fx):.2e}")
        assert r.status == 0 and kkt <= tol
        assert abs(f - rr.fx) <= ftol * abs(rr.fx)
        assert label_agreement(g, ref, K) >= 0.9999
        assert err <= gtol


def test_theta_update_matches_oracle(ctx, oracle):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, K = 30001, 3
    x, _ = _signal(5, T, [0.0, 1.0, 3.0], 1e-3, 1.0)
    rng = np.random.default_rng(3)
    g = oracle.project_simplex(rng.random(T * K), K)
    ref = oracle.theta_update(x, g, T, 1, K)
    P = Problem(ctx, x, K, precision=capi.FP64)
    P.set_gamma(g)
    th = P.theta_update()
    np.testing.assert_allclose(th, ref, rtol=1e-13)
    Pm = Problem(ctx, x, K, precision=capi.FP32_BULK)
    Pm.set_gamma(g)
    th32 = Pm.theta_update()
    np.testing.assert_allclose(th32, ref, rtol=1e-6)


def test_alternation_c1_matches_oracle(ctx, oracle):
    """Full FEM-H1 K-means alternation on the 1-D CPU-runnable configuration."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, SIGNAL_C1 as C1
    x, truth = _signal(C1["seed"], C1["T"], C1["levels"], C1["p_switch"], C1["sigma"])
    T, K = C1["T"], 2
    spg = dict(maxit=100000, stop_difff=0, stop_kkt=1, kkt_tol=1e-11, kkt_scale=T, nthreads=8,
               gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    g_o, th_o, tr_o = oracle.ts_solve(x, np.full(T * K, 0.5), C1["theta0"], T, 1, K, 1,
                                      C1["eps"], ts=dict(maxit=50, eps=1e-10,
                                                         skip_first_theta=1), **spg)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=C1["eps"])
    P.set_theta(C1["theta0"])
    P.set_gamma_uniform()
    tr = P.ts_solve(ts=dict(maxit=50, eps=1e-10, skip_first_theta=1), maxit=100000,
                    stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    th = P.get_theta()
    print(f"oracle: it={tr_o.it} L={tr_o.L!r} theta={th_o}; gpu: it={tr.it} L={tr.L!r} theta={th}")
    np.testing.assert_allclose(th, th_o, atol=1e-7)
    assert np.max(np.abs(g - g_o)) <= 1e-6
    assert abs(tr.L - tr_o.L) <= 1e-8 * abs(tr_o.L)
    assert tr.increased == 0
    assert label_agreement(g, g_o, K) >= 0.9999
    assert np.mean(g.reshape(T, K).argmax(1) == truth) > 0.99


def test_grid_image_matches_oracle(ctx, oracle):
    """4-neighbour grid graph (T = 1), RGB data (xdim = 3, extension of the reference)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_rect_image
    w, h, K = 48, 40, 4
    x, col, lab = generate_rect_image(9, w, h, K, 6, 0.15)
    R = w * h
    theta = col.copy()
    eps = 3.0
    graph = grid_csr(w, h)
    b = oracle.assemble_b(x, theta, 1, R, K, xdim=3)
    xs, rr = oracle.spg_solve(np.full(R * K, 1.0 / K), b, 1, R, K, eps, graph=graph,
                              maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                              kkt_scale=1.0, nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    for prec in [capi.FP64, capi.FP32_BULK]:
        P = Problem(ctx, x, K, T=1, R=R, xdim=3, precision=prec, grid_width=w, epssqr=eps)
        P.set_theta(theta)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        g = P.get_gamma()
        f, kkt = P.objective()
        print(f"grid prec={prec}: it={r.it} kkt={kkt:.2e} f={f!r} oracle f={rr.fx!r} "
              f"err={np.max(np.abs(g - xs)):.2e}")
        assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)
        assert np.max(np.abs(g - xs)) <= 1e-6
        assert label_agreement(g, xs, K) >= 0.999
    # theta update on the image (per-channel weighted means)
    ref_th = oracle.theta_update(x, xs, 1, R, K, xdim=3)
    np.testing.assert_allclose(P.theta_update(), ref_th, atol=1e-5)


def test_batch_matches_individual_solves(ctx, oracle):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Batch
    rng = np.random.default_rng(2)
    xs, Ks, ths = [], [], []
    for i, (T, K) in enumerate([(3000, 2), (10000, 3), (777, 4), (5000, 3)]):
        lv = np.sort(rng.uniform(0, 5, K))
        x, _ = _signal(50 + i, T, lv, 5e-3, 0.6)
        xs.append(x); Ks.append(K); ths.append(lv)
    eps = np.array([0.1, 3.0, 100.0])
    for prec in [capi.FP64, capi.FP32_BULK]:
        B = Batch(ctx, xs, Ks, ths, eps, precision=prec)
        res, t = B.solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-9)
        for s in range(len(xs)):
            T, K = xs[s].size, Ks[s]
            for e in range(eps.size):
                rr = res[s * eps.size + e]
                g = B.get_gamma(s, e)
                ref, ro, b = _oracle_min(oracle, xs[s], ths[s], T, K, eps[e])
                fo = oracle.objective(g, b, T, 1, K, eps[e] / T)
                err = np.max(np.abs(g - ref))
                print(f"prec={prec} s={s} e={e} it={rr.it} ({rr.it_fp32}+{rr.it_fp64}) "
                      f"kkt={rr.kkt:.2e} err={err:.2e}")
                assert rr.status == 0 and rr.stop_reason == 6 and rr.kkt <= 1e-9
                assert abs(fo - ro.fx) <= 1e-8 * abs(ro.fx)
                assert label_agreement(g, ref, K) >= 0.999
        B.close()


def test_autotune_keeps_iterate_and_results(ctx, oracle):
    """The autotuner times every launch shape from the current iterate, restores it, and the
    chosen shape solves to the same minimiser."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, K, eps = 30000, 3, 20.0
    x, _ = _signal(21, T, [0.0, 1.0, 3.0], 1e-3, 0.9)
    theta = np.array([0.0, 1.0, 3.0])
    ref, rr, b = _oracle_min(oracle, x, theta, T, K, eps)
    for prec in [capi.FP64, capi.FP32_BULK]:
        P = Problem(ctx, x, K, precision=prec, epssqr=eps)
        P.set_theta(theta)
        g0 = oracle.project_simplex(np.random.default_rng(1).random(T * K), K)
        P.set_gamma(g0)
        sh = P.autotune()
        print(f"prec={prec}: {sh}")
        m = sh["main"]
        assert m["block"] in (64, 128, 256, 512) and m["prefetch"] in (0, 1) and m["us_per_pass"] > 0
        if prec == capi.FP32_BULK:
            assert m["chunk"] in (4, 8) and sh["polish"]["chunk"] in (2, 4)
            assert sh["polish"]["us_per_pass"] > 0
        np.testing.assert_allclose(P.get_gamma(), g0, atol=1e-7 if prec else 0)
        sh2 = P.autotune()  # cached per context
        assert {k: v for k, v in sh2["main"].items() if k != "us_per_pass"} == \
            {k: v for k, v in m.items() if k != "us_per_pass"}
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-9)
        g = P.get_gamma()
        assert r.kkt <= 1e-9 and np.max(np.abs(g - ref)) <= 1e-6


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
def test_launch_shapes_give_identical_iterates(ctx, prec_name):
    """Prefetching the next tile changes only the load order: with the same block, chunk and
    slot tiling the iterates are bit-identical.  max_slots = 3 (or min_tiles = 4) makes every
    workgroup run several consecutive tiles, so the pipelined loads and the interior / boundary
    tile switch are exercised at a small T.  (Other slot counts change the order of the FP64 partial sums,
    which the BB steps amplify along the trajectory: those runs are not compared mid-solve.)"""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    prec = getattr(capi, prec_name)
    T, K, eps = 50000, 3, 20.0
    x, _ = _signal(31, T, [0.0, 1.0, 3.0], 1e-3, 0.9)
    theta = np.array([0.0, 1.0, 3.0])
    chunks = (4, 8) if prec == capi.FP32_BULK else (2, 4)
    P = Problem(ctx, x, K, precision=prec, epssqr=eps)
    P.set_theta(theta)
    for block in (64, 128, 256):
        for chunk in chunks:
            tilings = [(3, 0), (0, 0), (0, 4)]  # (max_slots, min_tiles)
            outs = {}
            for pf in (0, 1):
                for slots, mt in tilings:
                    P.set_launch_shape(block, chunk, pf, slots, mt)
                    P.set_gamma_uniform()
                    r = P.gamma_solve(maxit=300, stop_mask=0, polish_maxit=50)
                    outs[(pf, slots, mt)] = (r.it, r.fx, P.get_gamma())
            print(block, chunk, {k: v[:2] for k, v in outs.items()})
            # same tiling: prefetch on and off agree bit for bit
            for slots, mt in tilings:
                a, b = outs[(0, slots, mt)], outs[(1, slots, mt)]
                assert a[0] == b[0] and np.array_equal(a[2], b[2]), (slots, mt)
    with pytest.raises(capi.PascError):
        P.set_launch_shape(192, chunks[0])


@pytest.mark.parametrize("prec_name,chunk", [("FP64", 2), ("FP64", 4), ("FP32_BULK", 4), ("FP32_BULK", 8)])
def test_one_wave_workgroups_follow_oracle(ctx, oracle, prec_name, chunk):
    """64-thread workgroups exchange chunk edges by cross-lane shuffles instead of LDS and
    barriers: the iterates follow the oracle's (FP64: to rounding; FP32: to FP32 rounding) and
    the solve reaches the oracle's minimiser."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    prec = getattr(capi, prec_name)
    T, K, eps = 4001, 3, 10.0
    levels = [0.0, 1.0, 3.0]
    x, _ = _signal(41, T, levels, 2e-3, 0.6)
    theta = np.array(levels)
    b = oracle.assemble_b(x, theta, T, 1, K)
    P = Problem(ctx, x, K, precision=prec, epssqr=eps)
    P.set_theta(theta)
    P.set_launch_shape(64, chunk, 1, 0, 0, polish=dict(block=64, chunk=2, prefetch=1))
    if prec == capi.FP64:
        for maxit in (1, 3, 20):
            xs, rr = oracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, eps / T, maxit=maxit,
                                      stop_difff=0)
            P.set_gamma_uniform()
            r = P.gamma_solve(maxit=maxit, stop_mask=0)
            assert r.it == maxit
            assert np.max(np.abs(P.get_gamma() - xs)) <= 1e-10
            assert abs(r.fx - rr.fx) <= 1e-12 * abs(rr.fx)
    ref, _, _ = _oracle_min(oracle, x, theta, T, K, eps)
    P.set_gamma_uniform()
    r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    print(f"{prec_name} chunk {chunk}: it={r.it} kkt={r.kkt:.2e} err={np.max(np.abs(g - ref)):.2e}")
    assert r.stop_reason == capi.REASON_KKT and r.kkt <= 1e-11
    assert np.max(np.abs(g - ref)) <= 1e-6
    P.close()


@pytest.mark.parametrize("block", [64, 256])
def test_many_slots_per_group_follow_oracle(ctx, oracle, block):
    """Up to 8192 partial slots at T = 6e5.  With 64 threads x 2 steps the tiles hold 120 steps
    and every tile is a slot (5000): the fused step's 64 completion groups then hold 78-79 slots,
    more than the one-wave workgroup's 64 threads, so the group reduction loops.  With 256
    threads there are 1191 slots.  The first iterations follow the oracle's to rounding."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    T, K, eps = 600000, 3, 30.0
    levels = [0.0, 1.0, 3.0]
    x, _ = _signal(43, T, levels, 1e-3, 0.7)
    theta = np.array(levels)
    b = oracle.assemble_b(x, theta, T, 1, K)
    P = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P.set_theta(theta)
    P.set_launch_shape(block, 2, 1, 8192, 1)
    for maxit in (1, 3, 20):
        xs, rr = oracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, eps / T, maxit=maxit,
                                  stop_difff=0, nthreads=8)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=maxit, stop_mask=0)
        err = np.max(np.abs(P.get_gamma() - xs))
        print(f"block {block} maxit {maxit}: err {err:.2e} f {r.fx!r} oracle {rr.fx!r}", flush=True)
        assert r.it == maxit
        assert err <= 1e-10
        assert abs(r.fx - rr.fx) <= 1e-12 * abs(rr.fx)
    P.close()


def test_autotune_file_cache(tmp_path):
    """A second context finds the shape the first one tuned in the cache file (no re-tuning)."""
    import json
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Context, Problem
    T, K = 20000, 3
    x, _ = _signal(23, T, [0.0, 1.0, 3.0], 1e-3, 0.9)
    path = str(tmp_path / "tune.jsonl")
    shapes = []
    for _ in range(2):
        with Context(0) as c:
            P = Problem(c, x, K, precision=capi.FP64, epssqr=5.0)
            P.set_theta(np.array([0.0, 1.0, 3.0]))
            P.set_gamma_uniform()
            shapes.append(P.autotune(path))
            P.close()
    lines = [json.loads(l) for l in open(path)]
    assert [l["cached"] for l in lines] == [0, 1]
    assert lines[0]["shape"] == lines[1]["shape"]
    drop = lambda d: {k: v for k, v in d.items() if k != "us_per_pass"}
    assert drop(shapes[0]["main"]) == drop(shapes[1]["main"])


def test_batch_autotune_keeps_results(ctx):
    """The batch autotuner restores gamma0 = 1/K and the tuned batch solves to the same
    minimisers as the untuned one."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Batch
    rng = np.random.default_rng(6)
    xs, Ks, ths = [], [], []
    for i, T in enumerate([3000, 7000, 12000]):
        K = 2 + i
        levels = np.linspace(0.0, 3.0, K)
        x, _ = _signal(40 + i, T, levels, 2e-3, 0.5)
        xs.append(x)
        Ks.append(K)
        ths.append(levels)
    eps = np.array([1.0, 10.0])
    params = dict(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-9)
    B0 = Batch(ctx, xs, Ks, ths, eps, precision=capi.FP32_BULK)
    r0, _ = B0.solve(**params)
    B1 = Batch(ctx, xs, Ks, ths, eps, precision=capi.FP32_BULK)
    sh = B1.autotune()
    print(sh)
    r1, _ = B1.solve(**params)
    for s in range(3):
        for e in range(2):
            j = s * 2 + e
            assert r0[j].kkt <= 1e-9 and r1[j].kkt <= 1e-9
            assert np.max(np.abs(B0.get_gamma(s, e) - B1.get_gamma(s, e))) <= 1e-6


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
def test_tiny_and_wide_problems(ctx, oracle, prec_name):
    """Edge shapes of the 1-D pass: T from 1 to a few hundred steps (one partial slot, boundary
    tiles only) and K from 1 to 8 (K = 1 stores its single row; K >= 2 stores K - 1 rows and
    rebuilds the last): every solve reaches the oracle's minimiser."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    prec = getattr(capi, prec_name)
    rng = np.random.default_rng(12)
    for T in (1, 2, 3, 5, 130):
        for K in (1, 2, 5, 8):
            levels = np.linspace(0.0, 2.0, K)
            x = rng.normal(0.0, 0.7, T) + levels[rng.integers(0, K, T)]
            eps = 3.0
            b = oracle.assemble_b(x, levels, T, 1, K)
            ref, rr = oracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, eps / T,
                                       maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                                       kkt_scale=T, gd_centered=1, gd_certified=1, dad_edges=1,
                                       g_recompute=1, safeguard=1)
            P = Problem(ctx, x, K, precision=prec, epssqr=eps)
            P.set_theta(levels)
            P.set_gamma_uniform()
            r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-10)
            g = P.get_gamma()
            P.close()
            G = g.reshape(T, K)
            assert r.status == 0 and r.kkt <= 1e-10, (T, K, r.kkt)
            assert G.min() >= 0.0 and np.max(np.abs(G.sum(axis=1) - 1.0)) <= 1e-12, (T, K)
            assert np.max(np.abs(g - ref)) <= 1e-6, (T, K, np.max(np.abs(g - ref)))


@pytest.mark.parametrize("case", ["1d_fp64", "1d_fp32", "image", "fem"])
def test_set_data_equals_a_fresh_problem(ctx, case):
    """pasc_set_data replaces the data as TSData::cutdata / scaledata edit the datavector
    (tsdata.cpp:384-478): theta steps and solves afterwards are bitwise those of a problem created
    with the new data (1-D FP64 and FP32-bulk, an image with a held-out mask, a FEM-reduced
    series)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_markov, generate_mask, generate_rect_image
    kw = {}
    if case == "image":
        w, h, K = 90, 40, 3
        x1, col, _ = generate_rect_image(3, w, h, K, 6, 0.2)
        x2, _, _ = generate_rect_image(4, w, h, K, 6, 0.2)
        kw = dict(T=1, xdim=3, grid_width=w, heldout=generate_mask(5, w * h, 0.1),
                  precision=capi.FP64)
        theta = col
    else:
        K = 3
        x1, _ = generate_markov(6, 20000, [0.0, 1.0, 3.0], 1e-3, 0.5)
        x2 = np.clip(x1, 0.0, 2.5)  # cutdata(0, 2.5)
        kw = dict(precision=capi.FP32_BULK if case == "1d_fp32" else capi.FP64)
        if case == "fem":
            kw.update(fem_reduce=0.5, fem_type=capi.FEM_HAT)
        theta = np.array([0.2, 1.1, 2.6])
    outs = []
    for replace in (True, False):
        P = Problem(ctx, x1 if replace else x2, K, epssqr=5.0, **kw)
        if replace:
            P.set_data(x2)
        P.set_theta(theta)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=300, stop_mask=0, polish_maxit=100)
        P.theta_update()
        r2 = P.gamma_solve(maxit=200, stop_mask=0, polish_maxit=100)
        outs.append((P.get_gamma(), P.get_theta(), r.fx, r2.fx))
        P.close()
    print(f"{case}: f {outs[0][2]!r} {outs[1][2]!r}", flush=True)
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2] and outs[0][3] == outs[1][3]
