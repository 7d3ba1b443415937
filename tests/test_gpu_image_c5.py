"""The multivariate graph workload at full size: a 1024 x 1024 RGB image (T = 1048576 pixels) on
the 4-neighbour grid, K = 4, 64 random rectangles, colours uniform in [0,1]^3, noise 0.15 per
channel (seed 4); a batched FP32-bulk scan over 8 epsilon values log-spaced in [1, 1e3] with 10%
of the pixels held out, the best epsilon by held-out error, and the adjoint gradient with
respect to log epsilon there, checked against central differences of FP64 device solves."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_image_1024_scan_and_adjoint(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, ProblemBatch, generate_mask, generate_rect_image
    w = h = 1024
    K = 4
    x, col, lab = generate_rect_image(4, w, h, K, 64, 0.15)
    mask = generate_mask(4, w * h, 0.1)
    theta = col.ravel().copy()
    eps = np.logspace(0, 3, 8)
    print("image 1024^2: batched scan", flush=True)
    B = ProblemBatch(ctx, x, K, theta, eps, precision=capi.FP32_BULK, T=1, xdim=3, grid_width=w,
                     heldout=mask)
    t0 = time.perf_counter()
    res, tdev = B.solve(maxit=50000, stop_mask=capi.STOP_KKT, kkt_tol=1e-6, check_every=32,
                        use_graph=1, kkt_every=4)
    wall = time.perf_counter() - t0
    held, acc = [], []
    for j in range(eps.size):
        g = B.get_gamma(j).reshape(-1, K)
        r = (x.reshape(-1, 3) - g @ theta.reshape(K, 3))[mask == 1]
        held.append(float(np.sum(r ** 2)))
        acc.append(float(np.mean(g.argmax(1) == lab)))
    B.close()
    best = int(np.argmin(held))
    print(f"scan {wall:.2f}s its {[r.it for r in res]} kkt {[f'{r.kkt:.1e}' for r in res]}", flush=True)
    print(f"held-out errors {np.round(held, 2)} label accuracy {np.round(acc, 4)} best eps {eps[best]}",
          flush=True)
    # the smoothest settings (epsilon ~ 1e3 on a 1024^2 grid) need more than the cap of SPG
    # iterations; every instance stays feasible and the scan's choice is certified
    assert all(r.status == 0 for r in res)
    assert res[best].kkt <= 1e-6 and sum(r.kkt <= 1e-6 for r in res) >= 6
    assert acc[best] > 0.99
    P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP64, heldout=mask,
                epssqr=float(eps[best]))
    P.set_theta(theta)
    tight = dict(maxit=400000, stop_mask=capi.STOP_KKT, kkt_tol=1e-11, check_every=32,
                 use_graph=1, kkt_every=4)

    def J(le):
        P.set_epssqr(float(np.exp(le)))
        P.set_gamma_uniform()
        r = P.gamma_solve(**tight)
        g = P.get_gamma().reshape(-1, K)
        res_ = (x.reshape(-1, 3) - g @ theta.reshape(K, 3))[mask == 1]
        return float(np.sum(res_ ** 2)), r

    le = float(np.log(eps[best]))
    J0, r0 = J(le)
    loss, dle, dth, cg = P.adjoint(mask, cg_tol=1e-12, cg_maxit=200000)
    hh = 1e-6
    fd = (J(le + hh)[0] - J(le - hh)[0]) / (2 * hh)
    print(f"best eps {eps[best]}: solve it {r0.it} J {J0!r} adjoint dlogeps {dle!r} FD {fd!r} "
          f"dtheta {np.round(dth, 3)} cg {cg}", flush=True)
    assert abs(loss - J0) <= 1e-9 * J0
    assert abs(fd - dle) <= 1e-3 * max(abs(fd), 1e-3 * J0)
