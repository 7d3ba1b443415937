"""Benchmark of the FEM-H1 K-means gamma step on one MI355X.

Metric: time-to-KKT 1e-6 of one gamma QP at T=1e7, K=3 (levels {0,1,3}, switch 1e-4, noise
sigma 1, seed 1, epsilon 100, theta fixed at the levels, gamma0 = 1/K): the FP64 path and the
FP32-bulk path (FP64 reductions, FP64-certified KKT) are both timed with autotuned launch shapes,
and `value` is the faster of the two (`path` names it); plus the batched epsilon-scan throughput
(256 uneven series x 16 epsilons) against the plain FP64 per-solve path on the same GPU.

A "step" is one complete gamma solve from gamma0 to KKT <= 1e-6 with the data resident in HBM.
Errors are stated against the FP64 CPU oracle's minimiser (KKT <= 1e-10), read from the
committed fixture tests/golden/c2_T1e7_K3_eps100.npz.  A second timed figure solves to the KKT
level at which gamma is within 1e-6 of that minimiser.

The roofline of the dominant kernel (the fused SPG pass of the faster path) divides the bytes
the pass moves per launch by its average launch time (hipEvents around every pass of one solve,
on the engine's stream): HBM traffic measured in this run by two rocprofv3 counter passes
(FETCH_SIZE, WRITE_SIZE, scaled by what the counters read for the copy kernel) when rocprofv3
is available, else the minimum the pass must move, (2 (K-1) T + T) s.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
C2_FIXTURE = os.path.join(ROOT, "tests", "golden", "c2_T1e7_K3_eps100.npz")
COPY_BYTES = 1 << 30


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--T", type=int, default=10_000_000)
    ap.add_argument("--kkt", type=float, default=1e-6)
    ap.add_argument("--kkt-gamma", type=float, default=2e-7,
                    help="KKT level of the second timed figure (gamma within 1e-6 of the "
                         "oracle's minimiser)")
    ap.add_argument("--no-scan", action="store_true")
    ap.add_argument("--no-graph-c5", action="store_true",
                    help="skip the image-graph solve reported beside the headline")
    ap.add_argument("--c5-eps", type=float, default=10.0)
    ap.add_argument("--no-adjoint-c4", action="store_true",
                    help="skip the differentiation workload reported beside the headline")
    ap.add_argument("--scan-series", type=int, default=256)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--plain-per-cell", type=int, default=2,
                    help="instances of every (K, epsilon) cell of the scan timed on the plain "
                         "FP64 per-solve path")
    ap.add_argument("--scan-maxit", type=int, default=20000,
                    help="iteration cap of every scan solve, identical on both paths (FP32 and "
                         "FP64 polish iterations together on the batched path)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=24)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes (roofline on the minimum bytes)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch iterations one by one instead of replaying hipGraphs")
    ap.add_argument("--profile-run", action="store_true",
                    help="a copy kernel, then --passes fused passes of the T=1e7 solve in each "
                         "precision (for rocprofv3 traces and counters)")
    ap.add_argument("--passes", type=int, default=200,
                    help="--profile-run: passes per precision (0: solve to the KKT target)")
    ap.add_argument("--scan-profile", action="store_true",
                    help="run only the batched scan (for rocprofv3 traces of the batched engine)")
    ap.add_argument("--graph-profile", action="store_true",
                    help="run only the image-graph solve (for rocprofv3 traces of the graph kernels)")
    ap.add_argument("--adjoint-profile", action="store_true",
                    help="run only the differentiation workload (for rocprofv3 traces of the adjoint)")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file: launch shapes recorded there are reused, new ones appended")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def log(args, *a):
    if args.verbose:
        print(*a, file=sys.stderr, flush=True)


def c2_workload(args):
    from pasc_inference_amd.solver import SIGNAL_C2, generate_markov
    c = SIGNAL_C2
    x, lab = generate_markov(c["seed"], args.T, c["levels"], c["p_switch"], c["sigma"])
    return x, lab, np.array(c["levels"]), c["eps"]


def _timed_solves(args, P, params, label, steps=None, warmup=None):
    """W untimed warmup solves, then K timed solves from gamma0 = 1/K (wall clock around each,
    device synchronised on both sides); returns the times and the last result."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    for _ in range(warmup):
        P.set_gamma_uniform()
        P.gamma_solve(**params)
    times, res = [], None
    for _ in range(steps):
        P.set_gamma_uniform()
        P.ctx.synchronize()
        t0 = time.perf_counter()
        res = P.gamma_solve(**params)
        P.ctx.synchronize()
        times.append(time.perf_counter() - t0)
        log(args, f"{label}: {times[-1]:.4f}s it={res.it} ({res.it_fp32}+{res.it_fp64}) "
                  f"kkt={res.kkt:.2e}")
    return times, res


def _profiled_pass(P, params, f64):
    """One more solve with hipEvents around every fused pass: average pass time (s) over the
    passes that ran work."""
    P.set_profiling(True)
    P.set_gamma_uniform()
    r = P.gamma_solve(**params)
    P.set_profiling(False)
    if f64:
        return r.pass_time_fp64_s / max(r.passes_fp64, 1), r.passes_fp64
    return r.pass_time_fp32_s / max(r.passes_fp32, 1), r.passes_fp32


def _load_fixture(T):
    if T != 10_000_000 or not os.path.exists(C2_FIXTURE):
        return None
    fx = np.load(C2_FIXTURE)
    lab = np.repeat(fx["label_values"], np.diff(np.r_[fx["label_starts"], T])).astype(np.int32)
    return dict(fx=fx, labels=lab)


def _errors_vs_oracle(P, x, truth, T, K, fix):
    """Errors of the problem's current gamma against the oracle's minimiser (fixture): gamma
    max-abs at the fixture's sample, f relative (explicit FP64 evaluation), argmax agreement on
    every step, the per-cluster sums; plus the KKT residual and the labels against the truth."""
    g = P.get_gamma().reshape(T, K)
    f, kkt = P.objective()
    lab = P.labels()
    out = dict(kkt_fp64=float(kkt), label_vs_truth=float(np.mean(lab == truth)))
    if fix is None:
        out["oracle"] = None
        return out
    fx = fix["fx"]
    f_star = float(fx["f_star"])
    out.update(
        gamma_maxabs=float(np.abs(g[fx["sample_idx"]] - fx["sample_gamma"]).max()),
        f_rel=float(abs(f - f_star) / abs(f_star)),
        label_agreement=float(np.mean(lab == fix["labels"])),
        sum_gamma_maxabs=float(np.abs(g.sum(0) - fx["sum_gamma"]).max()),
        sum_x_gamma_maxabs=float(np.abs(x @ g - fx["sum_x_gamma"]).max()),
    )
    out["meets_bar"] = bool(out["gamma_maxabs"] <= 1e-6 and out["f_rel"] <= 1e-8 and
                            out["label_agreement"] >= 0.9999)
    return out


def run_c2(args, ctx, tune_cache):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x, truth, theta, eps = c2_workload(args)
    T, K = args.T, theta.size
    fix = _load_fixture(T)
    tune = {}
    fast = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt, check_every=32,
                use_graph=0 if args.no_graph else 1, kkt_every=4)
    bar = dict(fast, kkt_tol=args.kkt_gamma)
    bar_steps = max(1, min(args.steps, 5))

    # FP32-bulk path (FP32 storage and bulk arithmetic, FP64 reductions, FP64 polish)
    P32 = Problem(ctx, x, K, precision=capi.FP32_BULK, epssqr=eps)
    P32.set_theta(theta)
    if not args.no_autotune:
        P32.set_gamma_uniform()
        tune["fp32_bulk"] = P32.autotune(tune_cache)
        log(args, f"autotune fp32-bulk: {tune['fp32_bulk']}")
    t32, r32 = _timed_solves(args, P32, fast, "fp32-bulk solve")
    err32 = _errors_vs_oracle(P32, x, truth, T, K, fix)
    pass32, n32 = _profiled_pass(P32, fast, False)
    pass32_polish, _ = _profiled_pass(P32, fast, True)
    tb32, rb32 = _timed_solves(args, P32, bar, "fp32-bulk gamma-bar solve", bar_steps, 0)
    errb32 = _errors_vs_oracle(P32, x, truth, T, K, fix)
    P32.close()

    # plain FP64 per-solve path: default launch shape, host poll every iteration, no graph
    P64 = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P64.set_theta(theta)
    plain = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt)
    tp64, r64 = _timed_solves(args, P64, plain, "plain fp64 solve")
    # FP64 path: autotuned launch shape, graph-captured iteration loop
    if not args.no_autotune:
        P64.set_gamma_uniform()
        tune["fp64"] = P64.autotune(tune_cache)
        log(args, f"autotune fp64: {tune['fp64']}")
    t64g, r64g = _timed_solves(args, P64, fast, "fp64 graph solve")
    err64 = _errors_vs_oracle(P64, x, truth, T, K, fix)
    pass64, n64 = _profiled_pass(P64, fast, True)
    tb64, rb64 = _timed_solves(args, P64, bar, "fp64 gamma-bar solve", bar_steps, 0)
    errb64 = _errors_vs_oracle(P64, x, truth, T, K, fix)
    P64.close()
    return dict(T=T, K=K, x=x, theta=theta, tune=tune, t32=t32, r32=r32, tp64=tp64, r64=r64,
                t64g=t64g, r64g=r64g, pass32=pass32, n32=n32, pass32_polish=pass32_polish,
                pass64=pass64, n64=n64, err32=err32, err64=err64,
                bar={"kkt_tol": args.kkt_gamma,
                     "fp64_graph": {"time_s": float(np.mean(tb64)), "iterations": rb64.it,
                                    "errors_vs_oracle": errb64},
                     "fp32_bulk": {"time_s": float(np.mean(tb32)),
                                   "iterations": [rb32.it_fp32, rb32.it_fp64],
                                   "errors_vs_oracle": errb32}},
                oracle={"fixture": os.path.relpath(C2_FIXTURE, ROOT),
                        "kkt": float(fix["fx"]["kkt"]), "it": int(fix["fx"]["it"]),
                        "f_star": float(fix["fx"]["f_star"])} if fix else None)


# ---- HBM traffic of the pass kernels, measured by rocprofv3 counters in this run --------------
def _counter_rows(d):
    """(kernel name, counter value) of every dispatch in the counter CSV files under d."""
    rows = []
    for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                name = row.get("Kernel_Name") or row.get("Kernel-Name") or ""
                val = row.get("Counter_Value") or row.get("Counter-Value")
                if val is not None:
                    rows.append((name, float(val)))
    return rows


def measure_traffic(args, tune_cache):
    """Two counter passes (FETCH_SIZE, WRITE_SIZE; separate runs, as the counters must not be
    mixed) over `bench.py --profile-run`: a 1 GiB copy kernel and 200 fused passes per
    precision with the shapes this run's autotuner chose.  Bytes per dispatch of each pass
    kernel, each counter scaled by known bytes / counter bytes of the copy kernel."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    out, tmp = {}, tempfile.mkdtemp(prefix="pasc_traffic_")
    env = dict(os.environ, TMPDIR="/tmp")
    try:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, counter.lower())
            cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "c", "--",
                   sys.executable, os.path.join(ROOT, "bench.py"), "--profile-run",
                   "--passes", "200", "--T", str(args.T)]
            if tune_cache:
                cmd += ["--tune-cache", tune_cache]
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True,
                                   timeout=180)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 --pmc {counter} timed out"
            if r.returncode != 0 and "profile-run done" not in r.stdout:
                return None, f"rocprofv3 --pmc {counter} rc={r.returncode}: {r.stderr[-300:]}"
            per = {}
            for name, v in _counter_rows(d):
                per.setdefault(name, []).append(v)
            copy = [v for n, v in per.items() if "pasc_copy_kernel" in n]
            if not copy:
                return None, f"no copy-kernel dispatch in the {counter} output"
            scale = COPY_BYTES / float(np.median(copy[0]))
            out[counter] = {n: float(np.median(v)) * scale for n, v in per.items()
                            if "spg_pass_1d" in n}
            out[counter + "_scale"] = scale
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    kernels = {}
    for n, fb in out["FETCH_SIZE"].items():
        wb = out["WRITE_SIZE"].get(n)
        if wb is not None:
            kernels[n] = {"read_bytes": fb, "write_bytes": wb, "bytes": fb + wb}
    return {"kernels": kernels, "fetch_scale": out["FETCH_SIZE_scale"],
            "write_scale": out["WRITE_SIZE_scale"]}, None


def _pass_kernel_name(s, K, shape):
    pf = "true" if shape.get("prefetch") else "false"
    return f"spg_pass_1d<{s}, {s}, {K}, {shape['chunk']}, {shape['block']}, {pf}>"


def _roofline(kernel, s, n, T, avg, traffic, copy_gbps):
    """The dominant kernel against the HBM roofline: bytes moved per launch (measured HBM
    traffic when the counters ran in this run, else the minimum the pass moves) / average
    launch time."""
    K = n // T
    own = (2.0 * (K - 1 if K >= 2 else K) * T + T) * s
    basis = traffic if traffic else own
    out = {"bound": "hbm", "achieved": basis / avg / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": basis / avg / 1e9 / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
           "bytes_basis": ("HBM traffic per launch measured in this run (rocprofv3 FETCH_SIZE + "
                           "WRITE_SIZE, scaled by the copy kernel)") if traffic else
                          "minimum bytes the pass moves, (2 (K-1) T + T) s",
           "avg_launch_us": avg * 1e6, "own_min_bytes_per_launch": own,
           "own_min_frac": own / avg / 1e9 / HBM_PEAK_GBS, "device_copy_gbps": copy_gbps,
           # the reference iteration moves 6 n s bytes (gamma and g read twice, written once);
           # this pass keeps g and d on chip, so the rate below is reference-equivalent only
           "reference_equivalent_gbps": 6.0 * n * s / avg / 1e9}
    if traffic:
        out["traffic_over_min"] = traffic / own
    return out


def run_graph_c5(args, ctx):
    """The graph stencil at the image configuration: a 1024 x 1024 RGB image on the 4-neighbour
    grid (T = 1 step of R = 1048576 nodes, K = 4, xdim = 3; 64 random rectangles, noise 0.15,
    seed 4), theta at the generating colours, one FP64 gamma solve from gamma0 = 1/K to the
    bench's KKT target, and the same solve with FP32 bulk iterations and an FP64 polish.
    Reported beside the headline, not as it."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_rect_image
    w = h = 1024
    K = 4
    x, col, lab = generate_rect_image(4, w, h, K, 64, 0.15)
    P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP64, epssqr=args.c5_eps)
    P.set_theta(col.ravel().copy())
    params = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt, check_every=32,
                  use_graph=0 if args.no_graph else 1, kkt_every=4)
    times, r = _timed_solves(args, P, params, "graph (image) fp64 solve")
    g = P.get_gamma().reshape(-1, K)
    acc = float(np.mean(g.argmax(1) == lab))
    pass_s, npass = _profiled_pass(P, params, True)
    g64 = P.get_gamma()
    P.close()
    # the same solve with the image's FP32 bulk (FP32 iterations, then an FP64 polish from the
    # FP32 iterate to the same KKT target)
    P32 = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP32_BULK,
                  epssqr=args.c5_eps)
    P32.set_theta(col.ravel().copy())
    times32, r32 = _timed_solves(args, P32, params, "graph (image) fp32-bulk solve")
    g32 = P32.get_gamma()
    P32.set_profiling(True)
    P32.set_gamma_uniform()
    rp = P32.gamma_solve(**params)
    P32.set_profiling(False)
    pass32_s = rp.pass_time_fp32_s / max(rp.passes_fp32, 1)
    polish_s = rp.pass_time_fp64_s / max(rp.passes_fp64, 1)
    P32.close()
    R = w * h
    # the fused image pass reads p_{i-1} (K - 1 rows; the last row follows from the simplex)
    # and the data term q (K rows), and writes gamma_i and p_i (K - 1 rows each), on the passes
    # with beta_{i-1} = 1 (most of them; the others also read gamma_{i-1}):
    # (3 (K - 1) + K) R * 8 bytes per iteration
    own = (3.0 * (K - 1) + K) * R * 8
    return dict(workload=f"gamma QP solve, 1024x1024 RGB image, 4-neighbour grid, K={K}, xdim=3, "
                         f"epsilon {args.c5_eps:g}, theta fixed, gamma0=1/K, KKT<={args.kkt:g}",
                time_s=float(np.mean(times)), iterations=int(r.it), kkt=float(r.kkt),
                us_per_iteration=float(np.mean(times)) / (r.it + 1) * 1e6,
                label_accuracy=acc, pass_us=pass_s * 1e6, own_min_gbps=own / pass_s / 1e9,
                kernels="spg_pass_img (one launch per iteration, step in its last workgroup)",
                fp32_bulk=dict(time_s=float(np.mean(times32)), it_fp32=int(r32.it_fp32),
                               it_fp64=int(r32.it_fp64), kkt=float(r32.kkt),
                               pass_us_fp32=pass32_s * 1e6, pass_us_fp64=polish_s * 1e6,
                               own_min_gbps_fp32=own / 2 / pass32_s / 1e9,
                               gamma_maxabs_vs_fp64=float(np.abs(g32 - g64).max())),
                note="the iteration count depends on the image rows per wave tile (the order "
                     "of the FP64 sums), so pass_us is the figure to compare")


def run_adjoint_c4(args, ctx):
    """The differentiation workload: T = 1e6, K = 3, levels {0, 1.5, 3}, switch 1e-3, noise 0.8,
    seed 3, 10% of the steps held out, epsilon 30, theta at the levels.  Timed: the FP64 gamma
    solve to KKT 1e-12 from gamma0 = 1/K (autotuned launch shape, graph-captured, the residual
    on every 4th iteration)
    and the adjoint (projected CG on the active face to 1e-12; the second of two calls, the
    first also allocates the work vectors), which gives the gradient of the
    held-out squared reconstruction error with respect to log epsilon and theta.  Checked in the
    same run against central differences (h = 1e-4) of that loss over the same solves.
    Reported beside the headline, not as it."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, SIGNAL_C4, generate_markov, generate_mask
    c = SIGNAL_C4
    T = c["T"]
    x, _ = generate_markov(c["seed"], T, c["levels"], c["p_switch"], c["sigma"])
    mask = generate_mask(c["seed"], T, c["heldout"])
    theta = np.array(c["levels"], dtype=np.float64)
    K = theta.size
    P = Problem(ctx, x, K, precision=capi.FP64, heldout=mask, epssqr=c["eps"])
    P.set_theta(theta)
    P.set_gamma_uniform()
    shape = None if args.no_autotune else P.autotune(args.tune_cache)
    params = dict(maxit=400000, stop_mask=capi.STOP_KKT, kkt_tol=1e-12, check_every=32,
                  use_graph=0 if args.no_graph else 1, kkt_every=4)

    def loss(log_eps, th):
        P.set_epssqr(float(np.exp(log_eps)))
        P.set_theta(th)
        P.set_gamma_uniform()
        ctx.synchronize()
        t0 = time.perf_counter()
        r = P.gamma_solve(**params)
        ctx.synchronize()
        t = time.perf_counter() - t0
        g = P.get_gamma().reshape(-1, K)
        res = (x - g @ th)[mask == 1]
        return float(np.sum(res ** 2)), r, t

    le = float(np.log(c["eps"]))
    loss(le, theta)  # warm-up
    J, r, t_solve = loss(le, theta)
    t_adj = []
    for _ in range(2):  # the first call also allocates its work vectors (5 K T doubles)
        ctx.synchronize()
        t0 = time.perf_counter()
        adj_loss, dle, dth, cg_it = P.adjoint(mask, cg_tol=1e-12, cg_maxit=200000)
        ctx.synchronize()
        t_adj.append(time.perf_counter() - t0)
    log(args, f"adjoint c4: solve {t_solve:.4f}s it={r.it} adjoint {t_adj}s cg={cg_it}")
    h = 1e-4
    fd_le = (loss(le + h, theta)[0] - loss(le - h, theta)[0]) / (2 * h)
    fd_th = []
    for k in range(K):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd_th.append((loss(le, tp)[0] - loss(le, tm)[0]) / (2 * h))
    P.close()
    return dict(workload=f"held-out loss gradient, T={T}, K={K}, levels {{0,1.5,3}}, switch 1e-3, "
                         f"sigma 0.8, seed 3, 10% held out, epsilon {c['eps']:g}, FP64",
                launch_shape=shape, solve_s=t_solve, iterations=int(r.it), kkt=float(r.kkt),
                adjoint_s=t_adj[1],
                adjoint_first_call_s=t_adj[0],
                cg_iterations=int(cg_it), loss=J, adjoint_loss=adj_loss,
                dloss_dlogeps=dle, fd_dloss_dlogeps=fd_le,
                dlogeps_rel_diff=abs(dle - fd_le) / max(abs(fd_le), 1e-12),
                dloss_dtheta=[float(v) for v in dth], fd_dloss_dtheta=fd_th,
                dtheta_max_rel_diff=max(abs(a - b) / max(abs(b), 1e-12) for a, b in zip(dth, fd_th)),
                max_scaled_diff=max(abs(a - b) / max(abs(b), 1e-3 * J)
                                    for a, b in zip([dle] + list(dth), [fd_le] + fd_th)),
                note="central differences with h = 1e-4 cross a few changes of the active face "
                     "at this size; the GPU test bounds |FD - adjoint| by "
                     "5e-4 max(|FD|, 1e-3 loss) (max_scaled_diff)")


def run_scan(args, ctx):
    """The epsilon scan: every series x every epsilon, batched FP32-bulk (one engine, graph-
    captured, FP32 and FP64-polish iterations capped together at --scan-maxit) against the
    plain FP64 per-solve path (default launch shape, host poll every iteration, capped at the
    same --scan-maxit) on a stratified sample of the same instances: --plain-per-cell series of
    every (K, epsilon) cell, each cell's plain time and certified share scaled by its instance
    count.  Rates count certified solves only (KKT <= target reached)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import SCAN_EPS, Batch, Problem, generate_batch_params, \
        generate_markov
    T, K, lv, p, s, sd = generate_batch_params(2, args.scan_series)
    xs, ths = [], []
    for i in range(args.scan_series):
        levels = lv[i, :K[i]]
        x, _ = generate_markov(int(sd[i]), int(T[i]), levels, float(p[i]), float(s[i]))
        xs.append(x)
        ths.append(levels.copy())
    eps = SCAN_EPS
    n_eps = eps.size
    n_inst = args.scan_series * n_eps
    params = dict(maxit=args.scan_maxit, polish_maxit=-1, stop_mask=capi.STOP_KKT,
                  kkt_tol=args.kkt, check_every=64, use_graph=0 if args.no_graph else 1,
                  kkt_every=4)
    B = Batch(ctx, xs, K, ths, eps, precision=capi.FP32_BULK)
    if not args.no_autotune:
        log(args, f"autotune batch: {B.autotune(args.tune_cache)}")
    ctx.synchronize()
    t0 = time.perf_counter()
    res, tdev = B.solve(**params)
    ctx.synchronize()
    tb = time.perf_counter() - t0
    cert = np.array([r.status == 0 and r.kkt <= args.kkt for r in res]).reshape(-1, n_eps)
    its = np.array([r.it for r in res])
    B.close()
    log(args, f"batched scan: {tb:.3f}s for {n_inst} solves, {int(cert.sum())} certified, "
              f"it max {its.max()}")
    if args.scan_profile:
        return dict(batched_wall_s=tb, n_inst=n_inst, certified=int(cert.sum()),
                    it_max=int(its.max()), it_mean=float(its.mean()))
    # plain FP64 per-solve path on a stratified sample: --plain-per-cell series per (K, eps)
    rng = np.random.default_rng(12345)
    cells = {}
    for k in sorted(set(K.tolist())):
        ser = np.flatnonzero(K == k)
        for e in range(n_eps):
            pick = rng.choice(ser, size=min(args.plain_per_cell, ser.size), replace=False)
            cells[(k, e)] = dict(n=int(ser.size), pick=[int(q) for q in pick], t=[], ok=[])
    for (k, e), c in cells.items():
        for si in c["pick"]:
            P = Problem(ctx, xs[si], int(K[si]), precision=capi.FP64, epssqr=float(eps[e]))
            P.set_theta(ths[si])
            P.set_gamma_uniform()
            ctx.synchronize()
            t0 = time.perf_counter()
            r = P.gamma_solve(maxit=args.scan_maxit, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt)
            ctx.synchronize()
            c["t"].append(time.perf_counter() - t0)
            c["ok"].append(bool(r.status == 0 and r.kkt <= args.kkt))
            P.close()
            log(args, f"plain fp64 K={k} eps={eps[e]:.3g} series {si}: T={xs[si].size} "
                      f"it={r.it} kkt={r.kkt:.2e} {c['t'][-1]:.3f}s")
    # the whole scan on the plain path, estimated cell by cell
    plain_wall = sum(c["n"] * float(np.mean(c["t"])) for c in cells.values())
    plain_cert = sum(c["n"] * float(np.mean(c["ok"])) for c in cells.values())
    sample_n = sum(len(c["t"]) for c in cells.values())
    batched_rate = float(cert.sum()) / tb
    plain_rate = plain_cert / plain_wall
    per_eps = []
    for e in range(n_eps):
        pc = sum(c["n"] * float(np.mean(c["ok"])) for (k, ee), c in cells.items() if ee == e)
        per_eps.append({"eps": float(eps[e]), "batched_certified": int(cert[:, e].sum()),
                        "plain_certified_est": round(pc, 1), "instances": args.scan_series})
    return dict(certified_solves_per_s={"batched": batched_rate, "plain": plain_rate},
                ratio=batched_rate / plain_rate,
                all_solves_per_s={"batched": n_inst / tb, "plain": n_inst / plain_wall},
                batched_wall_s=tb, plain_wall_est_s=plain_wall, n_inst=n_inst,
                n_series=args.scan_series, eps_values=int(n_eps), maxit_cap=args.scan_maxit,
                certified={"batched": int(cert.sum()), "plain_est": round(plain_cert, 1)},
                plain_sample=sample_n, plain_sample_certified=int(sum(sum(c["ok"]) for c in cells.values())),
                it_max=int(its.max()), it_mean=float(its.mean()), per_eps=per_eps,
                note=("rates count certified solves (KKT <= target) per second; both paths stop "
                      "at the same iteration cap (FP32 and FP64-polish iterations together on "
                      "the batched path); the plain path is timed on a stratified sample, "
                      "every (K, epsilon) cell scaled by its instance count"))


def run_cpu_baseline(args, c2):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    x, theta, T, K = c2["x"], c2["theta"], c2["T"], c2["K"]
    b = pyoracle.assemble_b(x, theta, T, 1, K)
    g0 = np.full(T * K, 1.0 / K)
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    cores = max(1, min(cores, os.cpu_count() or 1))
    out = {}
    for nth in sorted({1, cores}):
        t0 = time.perf_counter()
        pyoracle.spg_solve(g0, b, T, 1, K, 100.0 / T, maxit=args.cpu_iters, stop_difff=0,
                           nthreads=nth)
        dt = time.perf_counter() - t0
        out[nth] = dt / args.cpu_iters
    it_needed = c2["r64"].it
    best = min(out, key=lambda k: out[k])
    # the reference iteration's vectors (6 n s bytes per iteration, s = 8) against a host copy
    # of 2 x 512 MiB (bytes read + written per second)
    alg_bytes = 6.0 * T * K * 8
    stream = {str(nth): pyoracle.host_copy_bandwidth(1 << 29, 3, nth) / 1e9 for nth in sorted({1, cores})}
    return dict(
        value=out[best] * it_needed, unit="s", cores=best, kind="port",
        sample=(f"{args.cpu_iters} op-by-op FP64 SPG iterations of the T={T} K={K} solve on "
                f"{best} host thread(s) ({out[best] * 1e3:.1f} ms/iteration; single thread "
                f"{out[1] * 1e3:.1f} ms/iteration) x {it_needed} iterations to KKT "
                f"{args.kkt:g} (the FP64 device path's count for the same algorithm)"),
        ms_per_iteration={str(k): v * 1e3 for k, v in out.items()},
        algorithmic_gbps={str(k): alg_bytes / v / 1e9 for k, v in out.items()},
        host_copy_gbps=stream,
    )


def profile_run(args):
    """Fixed work for profilers: a 1 GiB copy kernel (3 launches), then --passes fused passes
    of the bench workload per path (FP32-bulk: FP32 phase then FP64 polish; FP64), one host
    poll per pass.  With --passes 0 each path is solved to the bench's KKT target instead.
    The launch shapes are those of the bench (read from --tune-cache, or tuned here)."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Context, Problem
    x, _, theta, eps = c2_workload(args)
    out = {}
    with Context(0) as ctx:
        out["copy_gbps"] = ctx.copy_bandwidth(COPY_BYTES, 3)
        for name, prec in (("fp32", capi.FP32_BULK), ("fp64", capi.FP64)):
            P = Problem(ctx, x, theta.size, precision=prec, epssqr=eps)
            P.set_theta(theta)
            P.set_gamma_uniform()
            # the shapes an earlier bench run recorded in the cache, or tuned here as the bench
            # tunes them
            P.autotune(args.tune_cache)
            P.set_gamma_uniform()
            P.set_profiling(True)
            if args.passes > 0:
                r = P.gamma_solve(maxit=args.passes, polish_maxit=args.passes, stop_mask=0,
                                  check_every=1, kkt_every=4)
            else:
                r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt,
                                  check_every=32, kkt_every=4)
            P.close()
            for ph, t, n in (("fp32", r.pass_time_fp32_s, r.passes_fp32),
                             ("fp64", r.pass_time_fp64_s, r.passes_fp64)):
                if n:
                    out[f"{name}_problem_{ph}_pass_us"] = t / n * 1e6
                    out[f"{name}_problem_{ph}_passes"] = n
    print(json.dumps(out))
    print("profile-run done", flush=True)


def main():
    args = parse()
    if args.profile_run:
        profile_run(args)
        return
    if args.scan_profile:
        from pasc_inference_amd.solver import Context
        with Context(0) as ctx:
            print(json.dumps(run_scan(args, ctx)))
        return
    if args.graph_profile:
        from pasc_inference_amd.solver import Context
        with Context(0) as ctx:
            print(json.dumps(run_graph_c5(args, ctx)))
        return
    if args.adjoint_profile:
        from pasc_inference_amd.solver import Context
        with Context(0) as ctx:
            print(json.dumps(run_adjoint_c4(args, ctx)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(requested --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    from pasc_inference_amd.solver import Context
    tmpdir = tempfile.mkdtemp(prefix="pasc_bench_")
    tune_cache = args.tune_cache or os.path.join(tmpdir, "tune.jsonl")
    try:
        with Context(0) as ctx:
            copy_gbps = ctx.copy_bandwidth(COPY_BYTES, 5)
            log(args, f"device copy kernel: {copy_gbps:.0f} GB/s")
            c2 = run_c2(args, ctx, tune_cache)
            graph = None if args.no_graph_c5 else run_graph_c5(args, ctx)
            adjoint = None if args.no_adjoint_c4 else run_adjoint_c4(args, ctx)
            scan = None if args.no_scan else run_scan(args, ctx)
        traffic, traffic_err = (None, "skipped (--no-traffic)") if args.no_traffic else \
            measure_traffic(args, None if args.no_autotune else tune_cache)
        log(args, f"traffic: {traffic} {traffic_err or ''}")
        cpu = None if args.no_cpu else run_cpu_baseline(args, c2)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    t32, t64g = float(np.mean(c2["t32"])), float(np.mean(c2["t64g"]))
    fp64_best = t64g < t32
    best = t64g if fp64_best else t32
    T, K = c2["T"], c2["K"]
    n = T * K
    r32, r64g = c2["r32"], c2["r64g"]
    tune = c2["tune"]

    def traffic_of(s, shape):
        if not traffic or not shape:
            return None
        want = _pass_kernel_name(s, K, shape)
        for name, v in traffic["kernels"].items():
            if want in name:
                return v["bytes"]
        return None

    sh64 = tune.get("fp64", {}).get("main")
    sh32 = tune.get("fp32_bulk", {}).get("main")
    roof64 = _roofline("spg_pass_1d<double,K=3> (fused SPG iteration, FP64)", 8, n, T,
                       c2["pass64"], traffic_of("double", sh64), copy_gbps)
    roof32 = _roofline("spg_pass_1d<float,K=3> (fused SPG iteration, FP32 phase)", 4, n, T,
                       c2["pass32"], traffic_of("float", sh32), copy_gbps)
    roof, other = (roof64, roof32) if fp64_best else (roof32, roof64)
    line = {
        "metric": "time-to-KKT 1e-6 at T=1e7 K=3; batched eps-scan solves/s vs plain, same GPU",
        "value": best,
        "path": "fp64 (autotuned, graph-captured)" if fp64_best else "fp32-bulk + fp64 polish",
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if fp64_best else "f32 bulk + f64 reductions/polish",
        "data": "synthetic (deterministic SplitMix64 Markov-chain signal generated on the host)",
        "config": {"workload": f"gamma QP solve, T={T}, K={K}, levels {{0,1,3}}, "
                               f"switch 1e-4, sigma 1, seed 1, epsilon 100, theta fixed, "
                               f"gamma0=1/K, KKT<={args.kkt:g}",
                   "T": T, "K": K, "epssqr": 100.0, "kkt_tol": args.kkt},
        "paths_s": {"fp64_graph": t64g, "fp32_bulk": t32, "plain_fp64": float(np.mean(c2["tp64"]))},
        "iterations": {"fp64_graph": r64g.it, "fp32_bulk": [r32.it_fp32, r32.it_fp64],
                       "plain_fp64": c2["r64"].it},
        "speedup_vs_plain_fp64": float(np.mean(c2["tp64"])) / best,
        "autotune": tune,
        "errors_vs_oracle": {"fp64_graph": c2["err64"], "fp32_bulk": c2["err32"],
                             "oracle": c2["oracle"]},
        "time_to_gamma_1e-6": c2["bar"],
        "roofline": roof,
        "roofline_other_path": other,
        "traffic_measurement": traffic if traffic else {"error": traffic_err},
        "fp32_polish_pass_avg_us": c2["pass32_polish"] * 1e6,
        "scan": scan,
        "graph_c5": graph,
        "adjoint_c4": adjoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
