"""Adjoint gradients of the held-out reconstruction loss w.r.t. log(epsilon) and theta.

Checked against (1) central finite differences of the device solve (h = 1e-4, FP64 solves to a
tight KKT residual) and (2) an FP64 unrolled-SPG reference: forward-mode differentiation through
every iteration of a torch restatement of the SPG loop on a small slice.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(T, seed=3, frac=0.1):
    from pasc_inference_amd.solver import SIGNAL_C4, generate_markov, generate_mask
    c = SIGNAL_C4
    x, _ = generate_markov(seed, T, c["levels"], c["p_switch"], c["sigma"])
    mask = generate_mask(seed, T, frac)
    return x, mask, np.array(c["levels"]), c["eps"]


def _loss(P, x, mask, theta, log_eps, params):
    P.set_epssqr(float(np.exp(log_eps)))
    P.set_theta(theta)
    P.set_gamma_uniform()
    r = P.gamma_solve(**params)
    g = P.get_gamma().reshape(-1, theta.size)
    res = (x - g @ theta)[mask == 1]
    return float(np.sum(res ** 2)), r


@pytest.mark.parametrize("T", [10000, 1000000])
def test_adjoint_matches_central_differences(ctx, T):
    """The differentiation workload (T = 1e6 at full size; levels {0, 1.5, 3}, switch 1e-3,
    noise 0.8, seed 3, 10% held out, epsilon 30) and a T = 1e4 slice of it."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x, mask, theta, eps = _case(T)
    K = theta.size
    P = Problem(ctx, x, K, precision=capi.FP64, heldout=mask, epssqr=eps)
    params = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-12)
    J, r = _loss(P, x, mask, theta, np.log(eps), params)
    loss, dle, dth, cg_it = P.adjoint(mask, cg_tol=1e-12, cg_maxit=100000)
    print(f"J={J!r} adjoint loss={loss!r} dlogeps={dle!r} dtheta={dth} cg_it={cg_it} it={r.it}")
    assert abs(loss - J) <= 1e-10 * abs(J)
    h = 1e-4
    Jp, _ = _loss(P, x, mask, theta, np.log(eps) + h, params)
    Jm, _ = _loss(P, x, mask, theta, np.log(eps) - h, params)
    fd = (Jp - Jm) / (2 * h)
    print(f"FD dlogeps={fd!r}")
    # J is piecewise smooth (the active face changes where a gamma_kt reaches or leaves 0); at
    # T = 1e6 a step of 1e-4 crosses a few such changes, so the bound there is 5e-4
    tol = 1e-4 if T <= 10000 else 5e-4
    assert abs(fd - dle) <= tol * max(abs(fd), 1e-3 * J)
    for k in range(K):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        Jp, _ = _loss(P, x, mask, tp, np.log(eps), params)
        Jm, _ = _loss(P, x, mask, tm, np.log(eps), params)
        fdk = (Jp - Jm) / (2 * h)
        print(f"FD dtheta[{k}]={fdk!r} adjoint={dth[k]!r}")
        assert abs(fdk - dth[k]) <= tol * max(abs(fdk), 1e-3 * J)


def _torch_spg_loss(x, mask, theta, log_eps, T, K, iters):
    """FP64 torch restatement of the SPG loop (spgqpsolver.cpp:140-216) with b built from the
    data and theta (graphh1fem.cpp:258-286), returning the held-out loss after `iters` steps.

    The step lengths alpha_BB and beta are data-dependent scalars whose derivatives multiply the
    vanishing direction d at the fixed point; the tangent recursion treats them as constants
    (otherwise their derivatives, e.g. of dd/dAd with both terms -> 0, swamp the result in
    rounding), which makes the unrolled derivative converge to dgamma*/dp."""
    import torch
    xt = torch.as_tensor(x, dtype=torch.float64)
    m = torch.as_tensor(mask, dtype=torch.float64)
    obs = 1.0 - m
    alpha = torch.exp(log_eps) / T

    def A(v):  # v: [T, K]
        out = torch.zeros_like(v)
        dv = v[1:] - v[:-1]
        out[1:] += 2.0 * dv
        out[:-1] -= 2.0 * dv
        return alpha * out

    def proj(y):
        u, _ = torch.sort(y, dim=1, descending=True)
        css = torch.cumsum(u, 1)
        j = torch.arange(1, K + 1, dtype=torch.float64)
        cond = (u - (css - 1.0) / j) > 0
        rho = cond.sum(1, keepdim=True)
        tau = (torch.gather(css, 1, rho - 1) - 1.0) / rho
        return torch.clamp(y - tau, min=0.0)

    b = -((xt[:, None] - theta[None, :]) ** 2) * obs[:, None] / T
    g0 = torch.full((T, K), 1.0 / K, dtype=torch.float64)
    gam = proj(g0)
    g = A(gam) - b
    f = 0.5 * torch.sum((g - b) * gam)
    fs = [f] * 20
    a = torch.tensor(2.0, dtype=torch.float64)
    for _ in range(iters):
        d = proj(gam - a * g) - gam
        Ad = A(d)
        dd = torch.sum(d * d)
        dAd = torch.sum(d * Ad)
        gd = torch.sum(g * d)
        fmax = torch.max(torch.stack(fs))
        xi = (fmax - f) / dAd
        bb = -gd / dAd
        bh = 0.9 * bb + torch.sqrt(0.81 * bb * bb + 2 * xi)
        beta = torch.clamp(bh, max=1.0).detach()
        gam = gam + beta * d
        g = g + beta * Ad
        f = 0.5 * torch.sum((g - b) * gam)
        fs = fs[1:] + [f]
        a = (dd / dAd).detach()
    res = (xt - gam @ theta) * m
    return torch.sum(res * res)


def test_adjoint_matches_unrolled_spg(ctx):
    """On the first 1e4 steps of the differentiation workload's series and mask (the generators
    are sequential streams, so the T = 1e4 draw of seed 3 is that prefix; the mask too)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    import torch
    T = 10000
    x, mask, theta, eps = _case(T)
    xf, maskf, _, _ = _case(1000000)
    assert np.array_equal(x, xf[:T]) and np.array_equal(mask, maskf[:T])
    K = theta.size
    P = Problem(ctx, x, K, precision=capi.FP64, heldout=mask, epssqr=eps)
    _, r = _loss(P, x, mask, theta, np.log(eps), dict(maxit=200000, stop_mask=capi.STOP_KKT,
                                                      kkt_tol=1e-12))
    loss, dle, dth, _ = P.adjoint(mask, cg_tol=1e-13, cg_maxit=100000)
    iters = max(400, (3 * r.it) // 2)
    le = torch.tensor(np.log(eps), dtype=torch.float64)
    th = torch.tensor(theta, dtype=torch.float64)
    from torch.func import jvp
    f = lambda le_, th_: _torch_spg_loss(x, mask, th_, le_, T, K, iters)  # noqa: E731
    J, dJ_le = jvp(f, (le, th), (torch.tensor(1.0, dtype=torch.float64), torch.zeros(K, dtype=torch.float64)))
    print(f"unrolled ({iters} it): J={J.item()!r} dlogeps={dJ_le.item()!r}; adjoint J={loss!r} dlogeps={dle!r}")
    assert abs(J.item() - loss) <= 1e-8 * loss
    assert abs(dJ_le.item() - dle) <= 1e-5 * max(abs(dle), 1e-3 * loss)
    for k in range(K):
        e = torch.zeros(K, dtype=torch.float64)
        e[k] = 1.0
        _, dJ_k = jvp(f, (le, th), (torch.tensor(0.0, dtype=torch.float64), e))
        print(f"unrolled dtheta[{k}]={dJ_k.item()!r} adjoint={dth[k]!r}")
        assert abs(dJ_k.item() - dth[k]) <= 1e-5 * max(abs(dth[k]), 1e-3 * loss)


def test_autograd_function_wrapper(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    import torch
    from pasc_inference_amd.autograd import HeldoutLoss
    T = 3000
    x, mask, theta, eps = _case(T, seed=7)
    P = Problem(ctx, x, theta.size, precision=capi.FP64, heldout=mask, epssqr=eps)
    params = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-12)
    le = torch.tensor(np.log(eps), dtype=torch.float64, requires_grad=True)
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    J = HeldoutLoss.apply(le, th, P, mask, params)
    J.backward()
    loss, dle, dth, _ = P.adjoint(mask)
    assert abs(le.grad.item() - dle) <= 1e-12 * max(1.0, abs(dle))
    np.testing.assert_allclose(th.grad.numpy(), dth, rtol=1e-12, atol=1e-14)
    # one step of gradient descent on log eps reduces the held-out loss for a small step
    with torch.no_grad():
        le2 = le - 1e-3 * torch.sign(le.grad)
    J2 = HeldoutLoss.apply(le2, th.detach(), P, mask, params)
    assert J2.item() <= J.item() + 1e-12
