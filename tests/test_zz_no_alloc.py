"""pasc_theta_update_noalloc (the theta step with its partial sums in a static device array)
against pasc_theta_update (graphh1fem.cpp:305-396), bit for bit, on two calls in a row -- the
second on a smaller problem, after the first left its sums in the shared array -- plus the
rejection of a problem beyond the array's size."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _theta_both(P, theta0):
    from pasc_inference_amd.capi import check, lib
    P.set_theta(theta0)
    check(lib().pasc_theta_update(P.h))
    ref = P.get_theta().copy()
    P.set_theta(theta0)
    check(lib().pasc_theta_update_noalloc(P.h))
    return ref, P.get_theta().copy()


def test_noalloc_theta_equals_theta_update(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_markov
    cases = [(200_003, [0.0, 1.0, 3.0], capi.FP64, 30.0),
             (5_001, [0.5, 2.0], capi.FP32_BULK, 3.0)]
    for T, levels, prec, eps in cases:
        print(f"noalloc theta: T={T} K={len(levels)} precision={prec}", flush=True)
        x, _ = generate_markov(17, T, levels, 2e-3, 0.6)
        K = len(levels)
        P = Problem(ctx, x, K, precision=prec, epssqr=eps)
        P.set_theta(levels)
        P.set_gamma_uniform()
        P.gamma_solve(maxit=40, stop_mask=0)  # a non-trivial gamma
        ref, got = _theta_both(P, np.asarray(levels) + 0.1)
        P.close()
        np.testing.assert_array_equal(got, ref)
        assert np.all(np.isfinite(ref)) and np.ptp(ref) > 0.1


def test_noalloc_theta_with_channels_and_fem(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_rect_image
    x, col, _ = generate_rect_image(3, 96, 64, 4, 10, 0.2)
    P = Problem(ctx, x, 4, T=1, xdim=3, grid_width=96, precision=capi.FP64, epssqr=2.0)
    P.set_theta(col.ravel().copy())
    P.set_gamma_uniform()
    P.gamma_solve(maxit=30, stop_mask=0)
    ref, got = _theta_both(P, col.ravel() * 0.9)
    P.close()
    np.testing.assert_array_equal(got, ref)


def test_noalloc_theta_rejects_problems_beyond_the_static_array(ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.capi import PascError, check, lib
    from pasc_inference_amd.solver import Problem
    K, xdim = 8, 3
    nv = K * (1 + xdim)
    T = 256 * (524288 // nv - 1) + 1  # one step beyond the largest served T
    x = np.zeros(T * xdim)
    P = Problem(ctx, x, K, xdim=xdim, precision=capi.FP32_BULK, epssqr=1.0)
    P.set_theta(np.linspace(0, 1, K * xdim))
    P.set_gamma_uniform()
    with pytest.raises(PascError) as ei:
        check(lib().pasc_theta_update_noalloc(P.h))
    assert ei.value.code == 1  # PASC_ERR_ARG
    P.close()
