"""The batched epsilon scan at its full configuration (the caller it replaces is the sequential
scan of test_signal1D.cpp:268-353): 256 series with T log-uniform in [1e4, 1e6], K in {2, 3, 4},
levels U[0, 5], switch probability U[1e-4, 1e-2], noise U[0.2, 1.5] (seed 2), each scanned over
16 epsilon values log-spaced in [1e-1, 1e3]: 4096 FP32-bulk QP solves in one batched engine.

Checked: every instance ends without error and either certifies the KKT target or stops at the
iteration cap; a stratified subset (one series of every (K, epsilon) cell) is feasible and agrees
with an individual FP64 solve of the same instance (objective, labels, gamma)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAXIT = 20000
KKT = 1e-6


@pytest.mark.timeout(900)
def test_scan_256_series_x_16_eps(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import (SCAN_EPS, Batch, Problem, generate_batch_params,
                                           generate_markov)
    n_series = 256
    T, K, lv, p, s, sd = generate_batch_params(2, n_series)
    xs, ths = [], []
    for i in range(n_series):
        levels = lv[i, :K[i]]
        x, _ = generate_markov(int(sd[i]), int(T[i]), levels, float(p[i]), float(s[i]))
        xs.append(x)
        ths.append(levels.copy())
    eps = SCAN_EPS
    n_eps = eps.size
    print(f"scan: {n_series} series x {n_eps} eps, sum T = {int(T.sum())}", flush=True)
    B = Batch(ctx, xs, K, ths, eps, precision=capi.FP32_BULK)
    res, tdev = B.solve(maxit=MAXIT, polish_maxit=-1, stop_mask=capi.STOP_KKT, kkt_tol=KKT,
                        check_every=64, use_graph=1, kkt_every=4)
    st = np.array([r.status for r in res]).reshape(n_series, n_eps)
    reason = np.array([r.stop_reason for r in res]).reshape(n_series, n_eps)
    kkt = np.array([r.kkt for r in res]).reshape(n_series, n_eps)
    it = np.array([r.it for r in res]).reshape(n_series, n_eps)
    cert = (st == 0) & (kkt <= KKT)
    print(f"batched: {tdev:.1f} s on the device, {int(cert.sum())} of {st.size} certified; "
          f"certified per eps: {cert.sum(0).tolist()}", flush=True)
    assert np.all(st == 0), "an instance ended with an error"
    # an uncertified instance stopped at the shared cap (FP32 + FP64 polish iterations)
    assert np.all(reason[~cert] == capi.REASONS.index("maxit"))
    assert np.all(it[~cert] >= MAXIT - 1) and np.all(it <= MAXIT + 1)
    assert np.all(reason[cert] == capi.REASON_KKT)
    # stratified subset: one series of every (K, epsilon) cell against an FP64 solve
    rng = np.random.default_rng(7)
    checked = 0
    for k in (2, 3, 4):
        ser = np.flatnonzero(K == k)
        for e in range(n_eps):
            si = int(rng.choice(ser))
            gb = B.get_gamma(si, e).reshape(-1, k)
            assert np.abs(gb.sum(1) - 1.0).max() <= 1e-9 and gb.min() >= 0.0
            if not cert[si, e]:
                continue
            P = Problem(ctx, xs[si], k, precision=capi.FP64, epssqr=float(eps[e]))
            P.set_theta(ths[si])
            P.set_gamma_uniform()
            r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=1e-9)
            g1 = P.get_gamma().reshape(-1, k)
            f1, _ = P.objective()
            P.set_gamma(gb.ravel())
            fb, kb = P.objective()
            P.close()
            agree = float(np.mean(gb.argmax(1) == g1.argmax(1)))
            err = float(np.abs(gb - g1).max())
            print(f"  K={k} eps={eps[e]:.3g} series {si} T={xs[si].size}: batch it={it[si, e]} "
                  f"kkt={kb:.2e}; fp64 it={r.it}; f rel={abs(fb - f1) / abs(f1):.2e} "
                  f"gamma err={err:.2e} labels={agree:.6f}", flush=True)
            assert r.stop_reason == capi.REASON_KKT
            assert kb <= KKT * 1.0001
            assert abs(fb - f1) <= 1e-8 * abs(f1)
            assert agree >= 0.999
            checked += 1
    B.close()
    assert checked >= 24
