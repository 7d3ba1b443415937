"""The general graph stencil on the device against the oracle: spatio-temporal graphs of T
steps of R nodes (blockgraphsparse.cpp:48-109; neighbours at t and t+-1), given as a CSR
adjacency or as a 4-neighbour grid (videos: T > 1), and R unconnected time chains."""
import ctypes as C

import numpy as np
import pytest

from devmem import DeviceArray
from helpers import grid_csr, label_agreement, random_graph_csr

pytestmark = pytest.mark.gpu


def _rows(a, n, K, ld):
    h = np.zeros((K, ld))
    h[:, :n] = a.reshape(n, K).T
    return h


@pytest.mark.parametrize("T,R,kind", [(1, 50, "csr"), (7, 50, "csr"), (5, 12 * 9, "grid"),
                                      (4, 33, "none"), (3, 1, "csr0")])
def test_graph_matvec_matches_oracle(ctx, oracle, T, R, kind):
    from pasc_inference_amd import capi
    K = 3
    rng = np.random.default_rng(T * 100 + R)
    n = T * R
    x = rng.random(n * K)
    alpha = 0.61
    ld = (n + 63) // 64 * 64
    dx = DeviceArray(_rows(x, n, K, ld).ravel())
    dy = DeviceArray(np.zeros(K * ld))
    print(f"graph matvec T={T} R={R} {kind}", flush=True)
    if kind in ("csr", "csr0"):
        ptr, idx = random_graph_csr(R, 4.0, 3) if kind == "csr" else \
            (np.zeros(R + 1, dtype=np.int64), np.zeros(0, dtype=np.int32))
        capi.check(capi.lib().pasc_h1_matvec_csr(
            ctx.h, dx.ptr, dy.ptr, T, R, K, ld, alpha,
            ptr.ctypes.data_as(C.POINTER(C.c_int64)), idx.ctypes.data_as(C.POINTER(C.c_int32)),
            capi.DTYPE_F64))
        graph = (ptr, idx)
    elif kind == "grid":
        capi.check(capi.lib().pasc_h1_matvec(ctx.h, dx.ptr, dy.ptr, T, R, K, ld, alpha, 12,
                                             capi.DTYPE_F64))
        graph = grid_csr(12, 9)
    else:
        capi.check(capi.lib().pasc_h1_matvec(ctx.h, dx.ptr, dy.ptr, T, R, K, ld, alpha, 0,
                                             capi.DTYPE_F64))
        graph = (np.zeros(R + 1, dtype=np.int64), np.zeros(0, dtype=np.int32))
    ref = oracle.h1_matvec(x, T, R, K, alpha, graph=graph)
    out = dy.download().reshape(K, ld)[:, :n].T.ravel()
    assert np.max(np.abs(out - ref)) < 1e-13


def _video(seed, w, h, T, K, sigma):
    """Rectangles that move one pixel per frame: T frames of a w x h image, one channel."""
    rng = np.random.default_rng(seed)
    lab = np.zeros((T, h, w), dtype=np.int64)
    for k in range(1, K):
        y0, x0 = rng.integers(0, h // 2), rng.integers(0, w // 2)
        for t in range(T):
            lab[t, y0:y0 + h // 3, (x0 + t) % w:(x0 + t) % w + w // 3] = k
    levels = np.linspace(0.0, 2.0, K)
    x = levels[lab] + rng.normal(0, sigma, size=lab.shape)
    return x.ravel(), lab.ravel(), levels


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
@pytest.mark.parametrize("kind", ["grid", "csr", "none"])
def test_spatiotemporal_solve_matches_oracle(ctx, oracle, prec_name, kind):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    prec = getattr(capi, prec_name)
    K, eps = 3, 2.0
    if kind == "grid":
        w, h, T = 16, 12, 6
        R = w * h
        x, lab, levels = _video(4, w, h, T, K, 0.4)
        graph = grid_csr(w, h)
        P = Problem(ctx, x, K, T=T, grid_width=w, precision=prec, epssqr=eps)
    else:
        R, T = 60, 8
        graph = random_graph_csr(R, 3.0, 9) if kind == "csr" else \
            (np.zeros(R + 1, dtype=np.int64), np.zeros(0, dtype=np.int32))
        rng = np.random.default_rng(5)
        lab = rng.integers(0, K, size=R)[None, :].repeat(T, 0).ravel()
        levels = np.array([0.0, 1.0, 2.0])
        x = levels[lab] + rng.normal(0, 0.4, size=T * R)
        P = Problem(ctx, x, K, T=T, graph=graph if kind == "csr" else None, R=R, precision=prec,
                    epssqr=eps)
    theta = levels.copy()
    b = oracle.assemble_b(x, theta, T, R, K)
    ref, rr = oracle.spg_solve(np.full(T * R * K, 1.0 / K), b, T, R, K, eps / T, graph=graph,
                               maxit=400000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12, kkt_scale=T,
                               nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    assert rr.kkt <= 1e-12
    P.set_theta(theta)
    P.set_gamma_uniform()
    r = P.gamma_solve(maxit=400000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    f, kkt = P.objective()
    print(f"{kind} T={T} R={R} prec={prec_name}: it={r.it} kkt={kkt:.2e} f={f!r} oracle {rr.fx!r} "
          f"err={np.max(np.abs(g - ref)):.2e}", flush=True)
    assert r.stop_reason == capi.REASON_KKT and kkt <= 1e-11 * 1.001
    assert np.max(np.abs(g - ref)) <= 1e-6
    assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)
    assert label_agreement(g, ref, K) >= 0.999
    # theta step on the nodes (weighted means)
    np.testing.assert_allclose(P.theta_update(), oracle.theta_update(x, ref, T, R, K), atol=1e-6)


def test_graph_validation(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x = np.zeros(4 * 3)
    bad = [(np.array([0, 1, 1, 1], dtype=np.int64), np.array([1], dtype=np.int32)),   # not symmetric
           (np.array([0, 1, 2, 2], dtype=np.int64), np.array([0, 0], dtype=np.int32)),  # self loop
           (np.array([0, 1, 2, 3], dtype=np.int64), np.array([1, 0, 7], dtype=np.int32))]  # range
    for g in bad:
        with pytest.raises(capi.PascError):
            Problem(ctx, x, 2, T=4, graph=g)


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
def test_image_pass_matches_general_walk(ctx, prec_name):
    """Images (T = 1 on a 4-neighbour grid) run the fused image pass; the same image given as an
    explicit CSR adjacency (up, left, right, down, the general walk's order) runs the node-wise
    graph kernels.  Per node both form gamma_i, the Laplacian and d_i with the same arithmetic;
    their sums over the nodes run in different orders, so the iterates agree up to rounding:
    after 150 FP64 passes within 1e-8 (1e-10 measured), and at the minimiser within 1e-7 (both precisions; the
    FP32 phase's hand-over point may move by an iteration, so FP32-bulk is compared there
    only)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    prec = getattr(capi, prec_name)
    w, h, K, xdim = 137, 23, 4, 3  # three column tiles, the last one partial
    rng = np.random.default_rng(17)
    lab = rng.integers(0, K, size=(h, w))
    lab[5:15, 8:30] = 1
    col = rng.uniform(0, 1, size=(K, xdim))
    x = (col[lab.ravel()] + rng.normal(0, 0.2, size=(w * h, xdim))).ravel()
    runs = [dict(maxit=150, stop_mask=0, polish_maxit=30),
            dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)]
    outs = []
    for graph in (None, grid_csr(w, h)):
        if graph is None:
            P = Problem(ctx, x, K, T=1, xdim=xdim, grid_width=w, precision=prec, epssqr=3.0)
        else:
            P = Problem(ctx, x, K, T=1, xdim=xdim, graph=graph, precision=prec, epssqr=3.0)
        P.set_theta(col.ravel().copy())
        o = []
        for params in runs:
            P.set_gamma_uniform()
            r = P.gamma_solve(**params)
            o.append((r, P.get_gamma()))
        outs.append(o)
        P.close()
    (ra, ga), (rb, gb) = outs[0][0], outs[1][0]
    print(f"{prec_name} 150 passes: it {ra.it} {rb.it} f {ra.fx!r} {rb.fx!r} "
          f"gamma diff {np.abs(ga - gb).max():.2e}", flush=True)
    if prec == capi.FP64:
        assert ra.it == rb.it == 150
        assert abs(ra.fx - rb.fx) <= 1e-12 * abs(rb.fx)
        assert np.abs(ga - gb).max() <= 1e-8
    (ra, ga), (rb, gb) = outs[0][1], outs[1][1]
    print(f"{prec_name} minimiser: it {ra.it} {rb.it} kkt {ra.kkt:.2e} {rb.kkt:.2e} "
          f"gamma diff {np.abs(ga - gb).max():.2e}", flush=True)
    assert ra.kkt <= 1e-11 and rb.kkt <= 1e-11
    assert abs(ra.fx - rb.fx) <= 1e-12 * abs(rb.fx)
    assert np.abs(ga - gb).max() <= 1e-7


@pytest.mark.parametrize("w,h,rows", [(130, 21, 8), (61, 9, 3), (62, 5, 1), (1, 40, 8),
                                      (200, 1, 4), (70, 33, 64)])
def test_image_pass_tiles_match_oracle(ctx, oracle, w, h, rows):
    """The fused image pass on tile layouts that exercise its edges (a width of one column
    tile, one column past it, a single column, a single row, rows per tile that do not divide
    the height or exceed it), with 10% of the pixels held out, against the oracle's minimiser
    of the same QP on the explicit 4-neighbour graph, FP64 and FP32-bulk."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_mask, generate_rect_image
    K, eps = 3, 2.0
    x, colr, _ = generate_rect_image(w * 1000 + h, w, h, K, 5, 0.2)
    R = w * h
    mask = generate_mask(w + h, R, 0.1)
    b = oracle.assemble_b(x, colr, 1, R, K, xdim=3, heldout=mask)
    xs, rr = oracle.spg_solve(np.full(R * K, 1.0 / K), b, 1, R, K, eps, graph=grid_csr(w, h),
                              maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                              kkt_scale=1.0, nthreads=4, gd_centered=1, gd_certified=1,
                              dad_edges=1, g_recompute=1)
    for prec in (capi.FP64, capi.FP32_BULK):
        P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=prec, heldout=mask,
                    epssqr=eps)
        P.set_theta(colr)
        P.set_launch_shape(64, rows)
        P.set_gamma_uniform()
        print(f"image pass {w}x{h} rows {rows} prec {prec}", flush=True)
        r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        g = P.get_gamma()
        f, kkt = P.objective()
        print(f"  it={r.it} kkt={kkt:.2e} f={f!r} oracle f={rr.fx!r} "
              f"err={np.max(np.abs(g - xs)):.2e}", flush=True)
        assert r.stop_reason == capi.REASON_KKT
        assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)
        assert np.max(np.abs(g - xs)) <= 1e-6
        assert np.abs(g.reshape(R, K).sum(1) - 1.0).max() <= 1e-12
        P.close()


@pytest.mark.parametrize("K,xdim", [(4, 3), (3, 1), (3, 2), (8, 3)])
def test_image_pass_channels_and_heldout_match_oracle(ctx, oracle, K, xdim):
    """The image pass on grayscale, two-channel and RGB data (xdim < K) with 15% of the pixels
    held out (data term 0 there), FP64 and FP32-bulk, against the oracle's minimiser of the same
    QP on the explicit 4-neighbour graph."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_mask, generate_rect_image
    w, h, eps = 75, 44, 2.0
    x3, col3, _ = generate_rect_image(900 + 10 * K + xdim, w, h, K, 8, 0.2)
    R = w * h
    x = x3.reshape(R, 3)[:, :xdim].ravel().copy()
    colr = col3.reshape(K, 3)[:, :xdim].ravel().copy()
    mask = generate_mask(77 + K, R, 0.15)
    b = oracle.assemble_b(x, colr, 1, R, K, xdim=xdim, heldout=mask)
    xs, rr = oracle.spg_solve(np.full(R * K, 1.0 / K), b, 1, R, K, eps, graph=grid_csr(w, h),
                              maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                              kkt_scale=1.0, nthreads=4, gd_centered=1, gd_certified=1,
                              dad_edges=1, g_recompute=1)
    for prec in (capi.FP64, capi.FP32_BULK):
        P = Problem(ctx, x, K, T=1, xdim=xdim, grid_width=w, precision=prec, heldout=mask,
                    epssqr=eps)
        P.set_theta(colr)
        P.set_gamma_uniform()
        print(f"image pass K={K} xdim={xdim} prec {prec}", flush=True)
        r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        g = P.get_gamma()
        f, kkt = P.objective()
        print(f"  it={r.it} kkt={kkt:.2e} f={f!r} oracle f={rr.fx!r} "
              f"err={np.max(np.abs(g - xs)):.2e}", flush=True)
        assert r.stop_reason == capi.REASON_KKT
        assert abs(f - rr.fx) <= 1e-8 * abs(rr.fx)
        assert np.max(np.abs(g - xs)) <= 1e-6
        P.close()


def test_image_pass_rejects_bad_shapes(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x = np.zeros(30 * 3)
    P = Problem(ctx, x, 2, T=1, xdim=3, grid_width=10, precision=capi.FP64, epssqr=1.0)
    for block, chunk in ((128, 8), (64, 0), (64, 257)):
        with pytest.raises(capi.PascError):
            P.set_launch_shape(block, chunk)
    P.close()


def test_image_autotune_keeps_iterate(ctx):
    """The autotuner times the image pass over its tile rows, keeps the iterate, and the tuned
    problem still solves to the target."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_rect_image
    w, h, K = 150, 90, 3
    x, colr, _ = generate_rect_image(31, w, h, K, 7, 0.2)
    for prec in (capi.FP64, capi.FP32_BULK):
        P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=prec, epssqr=2.0)
        P.set_theta(colr)
        P.set_gamma_uniform()
        g0 = P.get_gamma()
        sh = P.autotune()
        print(f"image autotune prec {prec}: {sh}", flush=True)
        for ph in ("main", "polish"):
            assert sh[ph]["block"] == 64 and 6 <= sh[ph]["chunk"] <= 24
        assert np.array_equal(P.get_gamma(), g0)
        r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-10)
        assert r.stop_reason == capi.REASON_KKT
        P.close()


@pytest.mark.parametrize("K", [1, 2, 5, 8])
def test_image_pass_cluster_counts_match_oracle(ctx, oracle, K):
    """The image pass at the smallest and largest cluster counts (K = 1: gamma fixed at 1; K = 8:
    the largest register footprint), RGB data, FP64 and FP32-bulk, against the oracle."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_rect_image
    w, h, eps = 97, 31, 2.0
    x, colr, _ = generate_rect_image(500 + K, w, h, K, 9, 0.2)
    R = w * h
    b = oracle.assemble_b(x, colr, 1, R, K, xdim=3)
    xs, rr = oracle.spg_solve(np.full(R * K, 1.0 / K), b, 1, R, K, eps, graph=grid_csr(w, h),
                              maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                              kkt_scale=1.0, nthreads=4, gd_centered=1, gd_certified=1,
                              dad_edges=1, g_recompute=1)
    for prec in (capi.FP64, capi.FP32_BULK):
        P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=prec, epssqr=eps)
        P.set_theta(colr)
        P.set_gamma_uniform()
        print(f"image pass K={K} prec {prec}", flush=True)
        r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        g = P.get_gamma()
        f, kkt = P.objective()
        print(f"  it={r.it} kkt={kkt:.2e} f={f!r} oracle f={rr.fx!r} "
              f"err={np.max(np.abs(g - xs)):.2e}", flush=True)
        assert r.status == 0
        assert abs(f - rr.fx) <= 1e-8 * max(abs(rr.fx), 1e-12)
        assert np.max(np.abs(g - xs)) <= 1e-6
        P.close()
