"""CPU-side checks of the C ABI library: it loads, exports every declared entry point, its host
generators are deterministic and match their published definitions, and it refuses to run
without a device (there is no CPU fallback)."""
import ctypes as C
import math
import re

import numpy as np
import pytest

from pasc_inference_amd import capi


def _declared():
    txt = open(capi.HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", "", txt)
    return sorted(set(re.findall(r"\b(pasc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    # and the binding declares a signature for each of them
    for n in names:
        assert n in capi._SIGS or n in ("pasc_spg_default_params", "pasc_ts_default_params"), n


def _splitmix(seed):
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def test_splitmix64_known_answer():
    # first output of SplitMix64 seeded with 0 (Steele, Lea & Flood 2014; Vigna's reference code)
    assert next(_splitmix(0)) == 0xE220A8397B1DCDAF


def test_markov_generator_matches_its_definition():
    from pasc_inference_amd.solver import generate_markov
    T, levels, p, sigma, seed = 2000, [0.0, 1.0, 3.0], 0.01, 0.7, 42
    x, lab = generate_markov(seed, T, levels, p, sigma)
    g = _splitmix(seed)
    u = lambda: (next(g) >> 11) * 2.0 ** -53  # noqa: E731
    K = len(levels)
    st = min(int(u() * K), K - 1)
    for t in range(T):
        if t > 0 and u() < p:
            st = (st + 1 + min(int(u() * (K - 1)), K - 2)) % K
        u1, u2 = 1.0 - u(), u()
        xv = levels[st] + sigma * math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)
        assert lab[t] == st
        assert abs(x[t] - xv) < 1e-12
    # deterministic
    x2, _ = generate_markov(seed, T, levels, p, sigma)
    np.testing.assert_array_equal(x, x2)


def test_uniform_generator_matches_its_definition_and_the_oracle(oracle):
    """pasc_generate_uniform (the random gamma of the annealing restarts) is (z >> 11) * 2^-53
    of SplitMix64, the same stream the oracle's restarts draw."""
    from pasc_inference_amd.solver import generate_uniform
    u = generate_uniform(13, 1000)
    g = _splitmix(13)
    np.testing.assert_array_equal(u, [(next(g) >> 11) * 2.0 ** -53 for _ in range(1000)])
    np.testing.assert_array_equal(u, oracle.uniform(13, 1000))
    assert u.min() >= 0.0 and u.max() < 1.0


def test_other_generators():
    from pasc_inference_amd.solver import (generate_batch_params, generate_mask,
                                           generate_rect_image)
    m = generate_mask(3, 100000, 0.1)
    assert abs(m.mean() - 0.1) < 0.005
    x, col, lab = generate_rect_image(4, 64, 48, 4, 8, 0.15)
    assert x.shape == (64 * 48 * 3,) and lab.min() >= 0 and lab.max() <= 3
    np.testing.assert_allclose(x.reshape(-1, 3).mean(0), col.reshape(4, 3)[lab].mean(0), atol=0.01)
    T, K, lv, p, s, sd = generate_batch_params(2, 256)
    assert T.min() >= 10 ** 4 and T.max() <= 10 ** 6
    assert set(K.tolist()) <= {2, 3, 4}
    assert np.all((p >= 1e-4) & (p <= 1e-2)) and np.all((s >= 0.2) & (s <= 1.5))
    # log-uniform lengths: about half below 1e5
    assert 0.35 < np.mean(T < 10 ** 5) < 0.65


def test_no_device_means_an_error_not_a_fallback():
    """Without a HIP device the context cannot be created; every solve entry point needs one."""
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a device is present")
    h = C.c_void_p()
    rc = capi.lib().pasc_ctx_create(0, C.byref(h))
    assert rc != capi.PASC_OK
    assert capi.lib().pasc_last_error()
    with pytest.raises(capi.PascError):
        from pasc_inference_amd.solver import Context
        Context(0)


def test_null_arguments_are_rejected():
    L = capi.lib()
    assert L.pasc_problem_create(None, None, None) == capi.PASC_ERR_ARG
    assert L.pasc_gamma_solve(None, None, None) != capi.PASC_OK
    assert L.pasc_project_simplex(None, None, 0, 3, 0, 0, 0) == capi.PASC_ERR_ARG
    assert L.pasc_set_data(None, None) == capi.PASC_ERR_ARG
    p = capi.SpgParams()
    L.pasc_spg_default_params(C.byref(p))
    assert (p.maxit, p.eps, p.m, p.gamma, p.sigma1, p.sigma2, p.alphainit) == \
        (1000, 1e-9, 20, 0.9, 0.0, 1.0, 2.0)  # spgqpsolver.h:19-33
    assert p.stop_mask == capi.STOP_DIFFF
    tp = capi.TsParams()
    L.pasc_ts_default_params(C.byref(tp))
    assert (tp.maxit, tp.eps) == (300, 1e-6)  # tssolver.h:13-14


@pytest.mark.parametrize("name", ["test_signal1d", "test_image"])
def test_cpp_host_driver_is_built_and_fails_loudly_without_device(name):
    import os
    import subprocess
    drv = os.path.join(os.path.dirname(capi.HERE), "pasc_inference_amd", "host", name)
    assert os.path.exists(drv)
    r = subprocess.run([drv], capture_output=True, text=True)
    assert r.returncode == 1 and "test_epssqr has to be set" in r.stdout
    try:
        import torch
        if torch.cuda.device_count() > 0:
            return
    except Exception:
        pass
    r = subprocess.run([drv, "--test_epssqr=1"], capture_output=True, text=True)
    assert r.returncode == 2 and "pasc error" in r.stderr
