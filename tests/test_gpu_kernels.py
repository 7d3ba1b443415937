"""GPU parity of the kernel-level entry points (projection, H1 matvec) against the oracle.

Device buffers come from hipMalloc (tests/devmem.py); the computation runs in libpasc_hip.so.
"""
import numpy as np
import pytest

from devmem import DeviceArray
from helpers import grid_csr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 8, 16, 40, 257])
@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_projection_matches_oracle(ctx, oracle, K, layout, dtype):
    from pasc_inference_amd import capi
    rng = np.random.default_rng(100 + K)
    nb = 3001
    y = rng.normal(size=(nb, K)) * rng.choice([0.1, 1.0, 10.0], size=(nb, 1))
    y[::7] = np.abs(y[::7]) / np.abs(y[::7]).sum(axis=1, keepdims=True)  # feasible rows
    y[1::11] = 1.0 / K                                                    # ties
    ref = oracle.project_simplex(y.ravel(), K).reshape(nb, K)
    ndt = np.float64 if dtype == "f64" else np.float32
    if layout == 0:
        d = DeviceArray(y.ravel().astype(ndt))
        ld = 0
    else:
        ld = (nb + 63) // 64 * 64
        host = np.zeros((K, ld))
        host[:, :nb] = y.T
        d = DeviceArray(host.ravel().astype(ndt))
    print(f"projection K={K} layout={layout} dtype={dtype}", flush=True)
    capi.check(capi.lib().pasc_project_simplex(
        ctx.h, d.ptr, nb, K, capi.DTYPE_F64 if dtype == "f64" else capi.DTYPE_F32,
        layout, ld))
    out = d.download().astype(np.float64)
    out = out.reshape(nb, K) if layout == 0 else out.reshape(K, ld)[:, :nb].T
    tol = 1e-13 if dtype == "f64" else 2e-6
    scale = np.maximum(1.0, np.abs(y).max(axis=1, keepdims=True))
    assert np.max(np.abs(out - ref) / scale) < tol
    np.testing.assert_allclose(out.sum(axis=1), 1.0, atol=tol * 10 * K)
    assert out.min() >= 0.0


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_time_matvec_matches_oracle(ctx, oracle, dtype):
    from pasc_inference_amd import capi
    rng = np.random.default_rng(7)
    for T in [1, 2, 3, 1000, 4097]:
        K = 3
        x = rng.random((T, K))
        alpha = 0.37
        ref = oracle.h1_matvec(x.ravel(), T, 1, K, alpha).reshape(T, K)
        ld = (T + 63) // 64 * 64
        host = np.zeros((K, ld))
        host[:, :T] = x.T
        ndt = np.float64 if dtype == "f64" else np.float32
        dx = DeviceArray(host.ravel().astype(ndt))
        dy = DeviceArray(np.zeros(host.size, dtype=ndt))
        print(f"time matvec T={T} dtype={dtype}", flush=True)
        capi.check(capi.lib().pasc_h1_matvec(
            ctx.h, dx.ptr, dy.ptr, T, 1, K, ld, alpha, 0,
            capi.DTYPE_F64 if dtype == "f64" else capi.DTYPE_F32))
        out = dy.download().astype(np.float64).reshape(K, ld)[:, :T].T
        tol = 1e-14 if dtype == "f64" else 1e-6
        assert np.max(np.abs(out - ref)) < tol, T


def test_grid_matvec_matches_oracle(ctx, oracle):
    from pasc_inference_amd import capi
    rng = np.random.default_rng(8)
    for (w, h) in [(1, 1), (5, 1), (1, 6), (17, 9), (64, 32)]:
        R, K = w * h, 4
        x = rng.random((R, K))
        alpha = 1.7
        ref = oracle.h1_matvec(x.ravel(), 1, R, K, alpha, graph=grid_csr(w, h)).reshape(R, K)
        ld = (R + 63) // 64 * 64
        host = np.zeros((K, ld))
        host[:, :R] = x.T
        dx = DeviceArray(host.ravel())
        dy = DeviceArray(np.zeros(host.size))
        print(f"grid matvec {w}x{h}", flush=True)
        capi.check(capi.lib().pasc_h1_matvec(ctx.h, dx.ptr, dy.ptr, 1, R, K, ld, alpha, w,
                                             capi.DTYPE_F64))
        out = dy.download().reshape(K, ld)[:, :R].T
        assert np.max(np.abs(out - ref)) < 1e-13, (w, h)
