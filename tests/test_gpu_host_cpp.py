"""The C++ host layer (TSSolver / GraphH1FEMModel / SPGQPSolver / ImageData over the HipVector
backend) run through its test_signal1d and test_image drivers on the reference's own data
fixtures."""
import os
import re
import subprocess

import numpy as np
import pytest
from helpers import grid_csr

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "pasc_inference_amd", "host", "test_signal1d")
IMAGE_DRIVER = os.path.join(ROOT, "pasc_inference_amd", "host", "test_image")
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


def test_cpp_image_driver_on_reference_image(oracle, tmp_path):
    """test_image (the flow of examples/test_image/test_image.cpp) on the reference's own
    250 x 150 noisy text image (examples/test_image/data/usi_text/usi_250_150_02.bin): data cut
    to [0, 1], K = 2, a sorted epsilon list solved in turn from the previous gamma and theta.
    Per epsilon the saved files match the oracle's alternation on the same data and 4-neighbour
    graph: gamma (k-major, permute_TRK inverted) within 1e-6, theta (shortinfo, printed with 6
    digits) within 1e-5, the recovered image sum_k theta_k gamma_k, the original image (cut) on
    the first epsilon only, and one shortinfo line each."""
    from pasc_inference_amd import petsc_io
    assert os.path.exists(IMAGE_DRIVER), "build the host drivers first"
    W, H, K = 250, 150, 2
    R = W * H
    img = os.path.join(DATA, "usi_250_150_02.bin")
    eps_list = [0.3, 0.03, 3.0]  # sorted by the driver
    res = str(tmp_path / "results")
    short = str(tmp_path / "short" / "usi.txt")
    cmd = [IMAGE_DRIVER, f"--test_image_filename={img}", f"--test_width={W}",
           f"--test_height={H}", f"--test_K={K}", f"--test_results_dir={res}",
           "--test_image_out=usi", f"--test_shortinfo_filename={short}", "--test_printstats=1",
           "--spgqpsolver_stop_difff=0", "--spgqpsolver_stop_kkt=1",
           "--spgqpsolver_kkt_tol=1e-11", "--spgqpsolver_maxit=100000"]
    cmd += [f"--test_epssqr={e}" for e in eps_list]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    # printstats of the data as loaded (before cutdata), tsdata.cpp:310-383
    raw = petsc_io.read_vec(img)
    m = re.search(r"- sum:\s+(\S+)\s+- max:\s+(\S+)\s+- min:\s+(\S+)", r.stdout)
    assert m and np.isclose(float(m.group(1)), raw.sum(), rtol=1e-12)
    assert (float(m.group(2)), float(m.group(3))) == (raw.max(), raw.min())
    # the oracle: cutdata(0, 1) (tsdata.cpp:461-478), theta0 spread over the data range, a
    # uniform gamma0 and a gamma step first, each epsilon from the previous gamma and theta
    x = np.clip(petsc_io.read_vec(img), 0.0, 1.0)
    lo, hi = x.min(), x.max()
    th = np.array([lo + (hi - lo) * (k + 0.5) / K for k in range(K)])
    g = np.full(R * K, 1.0 / K)
    graph = grid_csr(W, H)
    for depth, e in enumerate(sorted(eps_list)):
        g, th, rr = oracle.ts_solve(x, g, th, 1, R, K, 1, e, graph=graph,
                                    ts=dict(skip_first_theta=1), maxit=100000, stop_difff=0,
                                    stop_kkt=1, kkt_tol=1e-12, kkt_scale=1.0, nthreads=8,
                                    gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1)
        pre = os.path.join(res, f"usi_epssqr{e:g}")
        gk = petsc_io.read_vec(pre + "_gamma.bin")
        gd = petsc_io.permute_trk(gk, 1, R, K)  # file layout (k-major) -> node-major
        rec = petsc_io.read_vec(pre + "_recovered.bin")
        print(f"eps {e}: oracle it {rr.it} L {rr.L!r} theta {th} gamma err "
              f"{np.abs(gd - g).max():.2e}", flush=True)
        assert np.abs(gd - g).max() <= 1e-6
        np.testing.assert_allclose(rec, gd.reshape(R, K) @ th, atol=1e-7)
        assert os.path.exists(pre + "_original.bin") == (depth == 0)
        if depth == 0:
            np.testing.assert_array_equal(petsc_io.read_vec(pre + "_original.bin"), x)
        thetas = open(short).read().strip().splitlines()[depth + 1].split(",")[5:5 + K]
        # printed with the stream's default 6 significant digits
        np.testing.assert_allclose([float(v) for v in thetas], th, rtol=1e-5)


def test_cpp_image_driver_scaledata_cutgamma_given_theta(oracle, tmp_path):
    """test_image with --test_scaledata (data and theta mapped from [0, 1] to [-1, 1] before each
    solve and back before saving, tsdata.cpp:511-545), --test_cutgamma (gamma set to the
    indicator of its largest component, tsdata.cpp:196-232) and a given theta (never updated).
    As in test_image.cpp:183-236 the given theta is set after the data were scaled, so it is
    taken in the scaled units and unscaled with the data after the solve.  The saved gamma is
    the one-hot argmax of the oracle's gamma QP on the scaled data with that theta, and the
    recovered image uses the unscaled theta."""
    from pasc_inference_amd import petsc_io
    W, H, K, eps = 250, 150, 2, 0.1
    R = W * H
    img = os.path.join(DATA, "usi_250_150_02.bin")
    res = str(tmp_path / "results")
    th = np.array([-0.2, 0.0])  # in the scaled units: 0.4 and 0.5 of the [0, 1] data
    cmd = [IMAGE_DRIVER, f"--test_image_filename={img}", f"--test_width={W}",
           f"--test_height={H}", f"--test_K={K}", f"--test_results_dir={res}",
           "--test_image_out=usi", f"--test_shortinfo_filename={tmp_path / 'short.txt'}",
           "--test_scaledata=1", "--test_cutgamma=1", "--test_Theta=-0.2", "--test_Theta=0",
           f"--test_epssqr={eps}", "--spgqpsolver_stop_difff=0", "--spgqpsolver_stop_kkt=1",
           "--spgqpsolver_kkt_tol=1e-11", "--spgqpsolver_maxit=100000"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    x = np.clip(petsc_io.read_vec(img), 0.0, 1.0)  # cutdata(0, 1) is on by default
    xs = 2.0 * x - 1.0                             # scaledata(-1, 1, 0, 1)
    b = oracle.assemble_b(xs, th, 1, R, K)
    g, _ = oracle.spg_solve(np.full(R * K, 1.0 / K), b, 1, R, K, eps, graph=grid_csr(W, H),
                            maxit=200000, stop_difff=0, stop_kkt=1, kkt_tol=1e-12,
                            kkt_scale=1.0, nthreads=8, gd_centered=1, gd_certified=1,
                            dad_edges=1, g_recompute=1)
    g = g.reshape(R, K)
    # nodes whose two components are within the solves' difference of a tie are left out
    clear = np.abs(g[:, 0] - g[:, 1]) > 1e-6
    onehot = np.eye(K)[g.argmax(1)]
    gd = petsc_io.permute_trk(petsc_io.read_vec(os.path.join(res, f"usi_epssqr{eps:g}_gamma.bin")),
                              1, R, K).reshape(R, K)
    print(f"one-hot nodes {np.mean(np.isin(gd, (0.0, 1.0)).all(1)):.4f}, "
          f"argmax agreement {np.mean((gd == onehot).all(1)[clear]):.6f} on {clear.sum()} nodes",
          flush=True)
    assert np.isin(gd, (0.0, 1.0)).all() and np.all(gd.sum(1) == 1.0)
    assert np.array_equal(gd[clear], onehot[clear])
    rec = petsc_io.read_vec(os.path.join(res, f"usi_epssqr{eps:g}_recovered.bin"))
    np.testing.assert_allclose(rec, gd @ ((th + 1.0) / 2.0), atol=1e-12)  # scaledata(0, 1, -1, 1)
    # the data went to [-1, 1] and back: equal up to the rounding of the two maps
    np.testing.assert_allclose(
        petsc_io.read_vec(os.path.join(res, f"usi_epssqr{eps:g}_original.bin")), x, rtol=0,
        atol=1e-15)
