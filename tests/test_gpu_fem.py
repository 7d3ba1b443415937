"""FEM coarsening on the device against the oracle: the kernel-level reduce / prolongate entry
points (Fem and FemHat, fem.cpp / femhat.cpp) and the FEM-reduced gamma step and alternation of
GraphH1FEMModel (graphh1fem.cpp:115-214,236-301)."""
import numpy as np
import pytest

from devmem import DeviceArray

pytestmark = pytest.mark.gpu


def _rows(a_tmajor, T, K, ld):
    h = np.zeros((K, ld))
    h[:, :T] = a_tmajor.reshape(T, K).T
    return h


@pytest.mark.parametrize("ftype", [1, 2])
@pytest.mark.parametrize("T1,T2", [(23, 7), (1000, 100), (1001, 101), (4097, 333), (50000, 3001)])
def test_fem_kernels_match_oracle(ctx, oracle, ftype, T1, T2):
    from pasc_inference_amd import capi
    K = 3
    rng = np.random.default_rng(T1 + T2)
    fine = rng.random(T1 * K)
    coarse = rng.random(T2 * K)
    ld1, ld2 = (T1 + 63) // 64 * 64, (T2 + 63) // 64 * 64
    print(f"fem kernels type={ftype} T1={T1} T2={T2}", flush=True)
    d_f = DeviceArray(_rows(fine, T1, K, ld1).ravel())
    d_c = DeviceArray(np.zeros(K * ld2))
    capi.check(capi.lib().pasc_fem_reduce(ctx.h, d_f.ptr, d_c.ptr, T1, T2, K, ld1, ld2, ftype))
    red = d_c.download().reshape(K, ld2)[:, :T2].T.ravel()
    ref = oracle.fem_reduce(fine, T1, T2, K, ftype)
    np.testing.assert_allclose(red, ref, rtol=1e-12, atol=1e-12)
    d_c.upload(_rows(coarse, T2, K, ld2).ravel())
    capi.check(capi.lib().pasc_fem_prolongate(ctx.h, d_c.ptr, d_f.ptr, T1, T2, K, ld1, ld2, ftype))
    pro = d_f.download().reshape(K, ld1)[:, :T1].T.ravel()
    refp = oracle.fem_prolongate(coarse, T1, T2, K, ftype)
    np.testing.assert_allclose(pro, refp, rtol=1e-12, atol=1e-12)  # FMA contraction on the device


def test_fem_kernel_arguments_are_checked(ctx):
    from pasc_inference_amd import capi
    d = DeviceArray(np.zeros(64))
    L = capi.lib()
    assert L.pasc_fem_reduce(ctx.h, d.ptr, d.ptr, 10, 1, 1, 64, 64, 2) == capi.PASC_ERR_ARG  # hat needs T2>=2
    assert L.pasc_fem_reduce(ctx.h, d.ptr, d.ptr, 10, 20, 1, 64, 64, 1) == capi.PASC_ERR_ARG  # T2 > T1
    assert L.pasc_fem_reduce(ctx.h, d.ptr, d.ptr, 100, 20, 1, 64, 64, 1) == capi.PASC_ERR_ARG  # ld1 < T1
    assert L.pasc_fem_prolongate(ctx.h, d.ptr, d.ptr, 10, 5, 1, 64, 64, 7) == capi.PASC_ERR_ARG
    assert L.pasc_fem_reduced_T(1000, 0.1) == 100 and L.pasc_fem_reduced_T(1000, 1.0) == 1000


@pytest.mark.parametrize("ftype", [1, 2])
@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
def test_fem_alternation_matches_oracle(ctx, oracle, ftype, prec_name):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_markov
    prec = getattr(capi, prec_name)
    T, K, eps, r = 20000, 2, 10.0, 0.1
    x, truth = generate_markov(0, T, [1.0, 2.0], 0.002, 0.5)
    theta0 = np.array([0.8, 2.2])
    spg_o = dict(maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12, kkt_scale=T * r,
                 nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
    ts = dict(maxit=5, eps=0.0, skip_first_theta=1, fem_reduce=r, fem_type=ftype)
    g_ref, th_ref, res_ref = oracle.ts_solve(x, np.full(T * K, 1.0 / K), theta0, T, 1, K, 1, eps,
                                             ts=ts, **spg_o)
    P = Problem(ctx, x, K, precision=prec, epssqr=eps, fem_reduce=r, fem_type=ftype)
    P.set_theta(theta0)
    P.set_gamma_uniform()
    res = P.ts_solve(ts=dict(maxit=5, eps=0.0, skip_first_theta=1), maxit=200000,
                     stop_mask=capi.STOP_KKT, kkt_tol=1e-11)
    g = P.get_gamma()
    th = P.get_theta()
    print(f"fem type={ftype} prec={prec_name}: it={res.it} L={res.L!r} oracle L={res_ref.L!r} "
          f"theta={th} oracle={th_ref} gamma err={np.max(np.abs(g - g_ref)):.2e}", flush=True)
    assert res.it == 5
    assert abs(res.L - res_ref.L) <= 1e-8 * abs(res_ref.L)
    np.testing.assert_allclose(th, th_ref, atol=1e-7)
    assert np.max(np.abs(g - g_ref)) <= 1e-6
    lab = P.labels()
    assert np.mean(lab == g_ref.reshape(T, K).argmax(1)) >= 0.9999
    assert np.mean(lab == truth) > 0.95
    # abserr on the fine grid against the oracle's
    sol = np.array([1.0, 2.0])[truth]
    assert abs(P.abserr(sol) - oracle.abserr_reconstructed(g_ref, th_ref, sol, T, 1, K)) <= 1e-6 * T
    with pytest.raises(capi.PascError):
        P.set_linear_term(np.zeros(T * K))
