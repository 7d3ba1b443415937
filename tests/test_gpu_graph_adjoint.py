"""The adjoint on graph problems (image grid, spatio-temporal CSR graph, RGB data) against
central finite differences of tight FP64 device solves, and the epsilon batch of one problem
(pasc_batch_create_problem) against individual solves -- the multivariate image workflow:
scan epsilon, pick the best held-out error, differentiate there."""
import numpy as np
import pytest

from helpers import label_agreement, random_graph_csr

pytestmark = pytest.mark.gpu

TIGHT = dict(maxit=400000, stop_mask=32, kkt_tol=1e-12)


def _loss(P, x, mask, theta, log_eps, xdim):
    P.set_epssqr(float(np.exp(log_eps)))
    P.set_theta(theta)
    P.set_gamma_uniform()
    r = P.gamma_solve(**TIGHT)
    g = P.get_gamma().reshape(-1, P.K)
    rec = g @ theta.reshape(P.K, xdim)
    res = (x.reshape(-1, xdim) - rec)[mask == 1]
    return float(np.sum(res ** 2)), r


def _check_fd(P, x, mask, theta, eps, xdim):
    from pasc_inference_amd import capi  # noqa: F401
    J, r = _loss(P, x, mask, theta, np.log(eps), xdim)
    loss, dle, dth, cg = P.adjoint(mask, cg_tol=1e-12, cg_maxit=200000)
    print(f"J={J!r} adjoint loss={loss!r} dlogeps={dle!r} cg_it={cg} solve it={r.it}", flush=True)
    assert abs(loss - J) <= 1e-10 * abs(J)
    h = 1e-4
    Jp, _ = _loss(P, x, mask, theta, np.log(eps) + h, xdim)
    Jm, _ = _loss(P, x, mask, theta, np.log(eps) - h, xdim)
    fd = (Jp - Jm) / (2 * h)
    print(f"FD dlogeps={fd!r}", flush=True)
    assert abs(fd - dle) <= 2e-4 * max(abs(fd), 1e-3 * J)
    for i in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fdi = (_loss(P, x, mask, tp, np.log(eps), xdim)[0] - _loss(P, x, mask, tm, np.log(eps), xdim)[0]) / (2 * h)
        print(f"FD dtheta[{i}]={fdi!r} adjoint={dth[i]!r}", flush=True)
        assert abs(fdi - dth[i]) <= 2e-4 * max(abs(fdi), 1e-3 * J)


def test_adjoint_rgb_image_matches_central_differences(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_mask, generate_rect_image
    w, h, K = 40, 32, 3
    x, col, lab = generate_rect_image(11, w, h, K, 5, 0.15)
    mask = generate_mask(12, w * h, 0.1)
    theta = col.ravel().copy()
    P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP64, heldout=mask,
                epssqr=0.5)
    _check_fd(P, x, mask, theta, 0.5, 3)


def test_adjoint_spatiotemporal_graph_matches_central_differences(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_mask
    R, T, K = 40, 10, 2
    rng = np.random.default_rng(3)
    lab = rng.integers(0, K, size=R)[None, :].repeat(T, 0).ravel()
    levels = np.array([0.0, 1.0])
    x = levels[lab] + rng.normal(0, 0.5, size=T * R)
    mask = generate_mask(4, T * R, 0.1)
    P = Problem(ctx, x, K, T=T, graph=random_graph_csr(R, 3.0, 2), precision=capi.FP64,
                heldout=mask, epssqr=1.0)
    _check_fd(P, x, mask, levels.copy(), 1.0, 1)


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
def test_problem_batch_matches_individual_solves(ctx, prec_name):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, ProblemBatch, generate_rect_image
    prec = getattr(capi, prec_name)
    w, h, K = 48, 40, 4
    x, col, lab = generate_rect_image(9, w, h, K, 6, 0.15)
    theta = col.ravel().copy()
    eps = np.logspace(0, 3, 8) / 100.0
    B = ProblemBatch(ctx, x, K, theta, eps, precision=prec, T=1, xdim=3, grid_width=w)
    res, t = B.solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-9, check_every=8,
                     use_graph=1, kkt_every=2)
    print(f"image batch {prec_name}: {t:.3f}s its {[r.it for r in res]}", flush=True)
    for j, e in enumerate(eps):
        assert res[j].status == 0 and res[j].kkt <= 1e-9
        P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP64, epssqr=float(e))
        P.set_theta(theta)
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
        gb, gi = B.get_gamma(j), P.get_gamma()
        err = np.max(np.abs(gb - gi))
        assert err <= 2e-6 * max(1.0, 1e-3 / e), (e, err)
        assert label_agreement(gb, gi, K) >= 0.999
        assert abs(res[j].fx - r.fx) <= 1e-7 * abs(r.fx)


def test_image_scan_then_adjoint_at_best_epsilon(ctx):
    """The multivariate workflow at reduced size: RGB rectangles, K = 4, 10% held-out pixels,
    a batched scan over 8 epsilon values log-spaced in [1, 1e3], the best epsilon by held-out
    error, and the adjoint there checked against central differences in log epsilon."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, ProblemBatch, generate_mask, generate_rect_image
    w = h = 96
    K = 4
    x, col, lab = generate_rect_image(4, w, h, K, 16, 0.15)
    mask = generate_mask(5, w * h, 0.1)
    theta = col.ravel().copy()
    eps = np.logspace(0, 3, 8)
    B = ProblemBatch(ctx, x, K, theta, eps, precision=capi.FP32_BULK, T=1, xdim=3, grid_width=w,
                     heldout=mask)
    res, t = B.solve(maxit=100000, stop_mask=capi.STOP_KKT, kkt_tol=1e-6, check_every=16,
                     use_graph=1, kkt_every=4)
    held = []
    for j in range(eps.size):
        g = B.get_gamma(j).reshape(-1, K)
        r = (x.reshape(-1, 3) - g @ theta.reshape(K, 3))[mask == 1]
        held.append(float(np.sum(r ** 2)))
    best = int(np.argmin(held))
    print(f"scan {t:.3f}s held-out errors {np.round(held, 3)} best eps {eps[best]}", flush=True)
    assert all(r.status == 0 for r in res)
    P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP64, heldout=mask,
                epssqr=float(eps[best]))
    J, r = _loss(P, x, mask, theta, np.log(eps[best]), 3)
    loss, dle, dth, cg = P.adjoint(mask, cg_tol=1e-12, cg_maxit=200000)
    # J is piecewise smooth in log eps (the active face changes where a pixel's gamma_k leaves
    # or reaches the bound); with many pixels near such a change a small step keeps both
    # evaluations on the face the adjoint differentiates
    hh = 1e-6
    Jp, _ = _loss(P, x, mask, theta, np.log(eps[best]) + hh, 3)
    Jm, _ = _loss(P, x, mask, theta, np.log(eps[best]) - hh, 3)
    fd = (Jp - Jm) / (2 * hh)
    print(f"best eps {eps[best]}: J={J!r} dlogeps adjoint {dle!r} FD {fd!r} cg {cg}", flush=True)
    assert abs(loss - J) <= 1e-10 * J
    assert abs(fd - dle) <= 1e-3 * max(abs(fd), 1e-3 * J)
