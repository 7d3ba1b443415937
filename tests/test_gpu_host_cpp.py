"""The C++ host layer (TSSolver / GraphH1FEMModel / SPGQPSolver over the HipVector backend) run
through its test_signal1d driver on the reference's own data fixture."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "pasc_inference_amd", "host", "test_signal1d")
DATA = os.path.join(HERE, "golden", "reference_data")


def test_cpp_driver_epsilon_scan_on_reference_fixture(oracle, tmp_path):
    from pasc_inference_amd import petsc_io
    assert os.path.exists(DRIVER), "build the host driver first"
    x = os.path.join(DATA, "signal1D_id1_idSigma1.bin")
    sol = os.path.join(DATA, "samplesignal_solution.bin")
    out = str(tmp_path / "res")
    cmd = [DRIVER, f"--test_filename={x}", f"--test_filename_solution={sol}", "--test_K=2",
           "--test_Theta=1", "--test_Theta=2", f"--test_filename_out={out}",
           f"--test_shortinfo_filename={tmp_path / 'short.csv'}",
           "--spgqpsolver_stop_difff=0", "--spgqpsolver_stop_kkt=1", "--spgqpsolver_kkt_tol=1e-11",
           "--spgqpsolver_maxit=100000"]
    for e in [0.1, 1.0, 10.0, 100.0]:
        cmd.append(f"--test_epssqr={e}")
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    m = re.search(r"best_epssqr=([0-9.eE+-]+) abserr=([0-9.eE+-]+)", r.stdout)
    best_eps, abserr = float(m.group(1)), float(m.group(2))
    # the same scan on the oracle: abserr per epsilon, best by abserr (test_signal1D.cpp:346-351)
    xs = petsc_io.read_vec(x)
    truth = petsc_io.read_vec(sol)
    T, K = xs.size, 2
    th = np.array([1.0, 2.0])
    b = oracle.assemble_b(xs, th, T, 1, K)
    errs = {}
    g = np.full(T * K, 0.5)
    for e in [0.1, 1.0, 10.0, 100.0]:
        g, _ = oracle.spg_solve(g, b, T, 1, K, e / T, maxit=100000, stop_difff=0, stop_kkt=1,
                                kkt_tol=1e-12, kkt_scale=T, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1, nthreads=8)
        errs[e] = (oracle.abserr_reconstructed(g, th, truth, T, 1, K), g.copy())
    e_or = min(errs, key=lambda k: errs[k][0])
    assert best_eps == e_or
    assert abs(abserr - errs[e_or][0]) <= 1e-6 * errs[e_or][0] + 1e-6
    # saved files: k-major gamma and the recovered signal (signal1Ddata.cpp:54-131)
    gk = petsc_io.read_vec(out + "_gamma.bin")
    g = petsc_io.kmajor_to_tmajor(gk, K)
    assert np.max(np.abs(g - errs[e_or][1])) <= 1e-6
    rec = petsc_io.read_vec(out + "_recovered.bin")
    np.testing.assert_allclose(rec, g.reshape(T, K) @ th, atol=1e-12)
    gref = petsc_io.read_vec(os.path.join(DATA, "signal1D_gauss_id1_idSigma_1_gamma.bin"))
    assert np.mean(g.reshape(T, K).argmax(1) == gref.reshape(T, K).argmax(1)) >= 0.995
    lines = open(tmp_path / "short.csv").read().strip().splitlines()
    assert len(lines) == 5 and lines[0].startswith("K,epssqr,abserr,")


def test_cpp_driver_fem_reduced_scan(oracle, tmp_path):
    """test_fem_type / test_fem_reduce (test_signal1D.cpp:81-82,218-228): the epsilon scan with
    the gamma QP solved on a 5x coarser grid, against the oracle's FEM alternation."""
    from pasc_inference_amd import petsc_io
    x = os.path.join(DATA, "signal1D_id1_idSigma1.bin")
    sol = os.path.join(DATA, "samplesignal_solution.bin")
    xs, truth = petsc_io.read_vec(x), petsc_io.read_vec(sol)
    T, K, r = xs.size, 2, 0.2
    eps_list = [1.0, 10.0]
    for ref_type, ftype in [(0, oracle.FEM_CONSTANT), (1, oracle.FEM_HAT)]:
        out = str(tmp_path / f"fem{ref_type}")
        cmd = [DRIVER, f"--test_filename={x}", f"--test_filename_solution={sol}", "--test_K=2",
               "--test_Theta=1", "--test_Theta=2", f"--test_filename_out={out}",
               f"--test_shortinfo_filename={out}.csv", f"--test_fem_type={ref_type}",
               f"--test_fem_reduce={r}", "--spgqpsolver_stop_difff=0", "--spgqpsolver_stop_kkt=1",
               "--spgqpsolver_kkt_tol=1e-11", "--spgqpsolver_maxit=200000"]
        cmd += [f"--test_epssqr={e}" for e in eps_list]
        print(" ".join(cmd), flush=True)
        rr = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
        print(rr.stdout[-1500:], rr.stderr[-1000:])
        assert rr.returncode == 0
        rows = [l.split(",") for l in open(out + ".csv").read().strip().splitlines()[1:]]
        abserr_dev = [float(row[2]) for row in rows]
        g = np.full(T * K, 1.0 / K)
        th = np.array([1.0, 2.0])
        for e, a_dev in zip(eps_list, abserr_dev):
            g, _, _ = oracle.ts_solve(xs, g, th, T, 1, K, 1, e,
                                      ts=dict(theta_fixed=1, fem_reduce=r, fem_type=ftype),
                                      maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                                      kkt_scale=T * r, nthreads=8, gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
            a_ref = oracle.abserr_reconstructed(g, th, truth, T, 1, K)
            print(f"fem type {ref_type} eps {e}: abserr device {a_dev} oracle {a_ref}")
            assert abs(a_dev - a_ref) <= 1e-4 * a_ref + 1e-3
        gk = petsc_io.read_vec(out + "_gamma.bin")
        assert gk.size == T * K  # the saved gamma is the fine (prolongated) one
