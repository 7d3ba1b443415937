"""The fused step's hand-over between workgroups leaves every partial-sum location marked empty
for the next launch: repeated solves on one problem, and a solve on a fresh problem, give
bit-identical iterates, for the 1-D pass and the image pass, FP64 and FP32-bulk."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solves(ctx, make, n):
    from pasc_inference_amd import capi
    outs = []
    P = make()
    for _ in range(n):
        P.set_gamma_uniform()
        r = P.gamma_solve(maxit=400, stop_mask=0, polish_maxit=100, check_every=16, use_graph=1)
        outs.append((r.it, r.fx, r.status, P.get_gamma()))
    P.close()
    return outs


@pytest.mark.parametrize("prec_name", ["FP64", "FP32_BULK"])
@pytest.mark.parametrize("kind", ["1d", "image"])
def test_repeated_fused_solves_are_bit_identical(ctx, prec_name, kind):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_markov, generate_rect_image
    prec = getattr(capi, prec_name)
    if kind == "1d":
        x, _ = generate_markov(41, 200000, np.array([0.0, 1.0, 3.0]), 1e-3, 0.8)
        theta = np.array([0.0, 1.0, 3.0])

        def make():
            P = Problem(ctx, x, 3, precision=prec, epssqr=30.0)
            P.set_theta(theta)
            return P
    else:
        w, h = 300, 200
        x, col, _ = generate_rect_image(43, w, h, 4, 12, 0.15)

        def make():
            P = Problem(ctx, x, 4, T=1, xdim=3, grid_width=w, precision=prec, epssqr=5.0)
            P.set_theta(col.ravel().copy())
            return P
    print(f"hand-over {kind} {prec_name}", flush=True)
    a = _solves(ctx, make, 3)
    b = _solves(ctx, make, 1)
    for o in a + b:
        assert o[2] == 0
    for o in a[1:] + b:
        assert o[0] == a[0][0] and o[1] == a[0][1]
        assert np.array_equal(o[3], a[0][3])
