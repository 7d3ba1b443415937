"""Benchmark of the FEM-H1 K-means gamma step on one MI355X.

Metric: time-to-KKT 1e-6 of one gamma QP at T=1e7, K=3 (levels {0,1,3}, switch 1e-4, noise
sigma 1, seed 1, epsilon 100, theta fixed at the levels, gamma0 = 1/K): the FP64 path and the
FP32-bulk path (FP64 reductions, FP64-certified KKT) are both timed with autotuned launch shapes,
and `value` is the faster of the two (`path` names it); plus the batched epsilon-scan throughput
(256 uneven series x 16 epsilons) against the plain FP64 per-solve path on the same GPU.

A "step" is one complete gamma solve from gamma0 to KKT <= 1e-6 with the data resident in HBM.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--T", type=int, default=10_000_000)
    ap.add_argument("--kkt", type=float, default=1e-6)
    ap.add_argument("--no-scan", action="store_true")
    ap.add_argument("--no-graph-c5", action="store_true",
                    help="skip the image-graph solve reported beside the headline")
    ap.add_argument("--c5-eps", type=float, default=10.0)
    ap.add_argument("--scan-series", type=int, default=256)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=32,
                    help="instances of the scan timed on the plain FP64 per-solve path")
    ap.add_argument("--scan-maxit", type=int, default=20000,
                    help="iteration cap of every scan solve, identical on both paths")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=24)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch iterations one by one instead of replaying hipGraphs")
    ap.add_argument("--profile-run", action="store_true",
                    help="only a fixed number of fused passes of the T=1e7 solve in each "
                         "precision plus a copy kernel (for rocprofv3 traces and counters)")
    ap.add_argument("--scan-profile", action="store_true",
                    help="run only the batched scan (for rocprofv3 traces of the batched engine)")
    ap.add_argument("--graph-profile", action="store_true",
                    help="run only the image-graph solve (for rocprofv3 traces of the graph kernels)")
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


def _timed_solves(args, P, params, label):
    """W untimed warmup solves, then K timed solves from gamma0 = 1/K (wall clock around each,
    device synchronised on both sides); returns the times and the last result."""
    for _ in range(args.warmup):
        P.set_gamma_uniform()
        P.gamma_solve(**params)
    times, res = [], None
    for _ in range(args.steps):
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
    """One more solve with hipEvents around every fused pass: average pass time (s)."""
    P.set_profiling(True)
    P.set_gamma_uniform()
    r = P.gamma_solve(**params)
    P.set_profiling(False)
    if f64:
        return r.pass_time_fp64_s / max(r.passes_fp64, 1), r.passes_fp64
    return r.pass_time_fp32_s / max(r.passes_fp32, 1), r.passes_fp32


def _errors(g, f, kkt, gref, fref, truth, T, K):
    lab = g.reshape(T, K).argmax(1)
    labref = gref.reshape(T, K).argmax(1)
    return dict(gamma_maxabs=float(np.max(np.abs(g - gref))), f_rel=float(abs(f - fref) / abs(fref)),
                label_agreement=float(np.mean(lab == labref)),
                label_vs_truth=float(np.mean(lab == truth)), kkt_fp64=float(kkt))


def run_c2(args, ctx):
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem
    x, truth, theta, eps = c2_workload(args)
    T, K = args.T, theta.size
    tune = {}
    fast = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt, check_every=32,
                use_graph=0 if args.no_graph else 1, kkt_every=4)

    # FP32-bulk path (FP32 storage and bulk arithmetic, FP64 reductions, FP64 polish)
    P32 = Problem(ctx, x, K, precision=capi.FP32_BULK, epssqr=eps)
    P32.set_theta(theta)
    if not args.no_autotune:
        P32.set_gamma_uniform()
        tune["fp32_bulk"] = P32.autotune(args.tune_cache)
        log(args, f"autotune fp32-bulk: {tune['fp32_bulk']}")
    t32, r32 = _timed_solves(args, P32, fast, "fp32-bulk solve")
    g32 = P32.get_gamma()
    f32, kkt32 = P32.objective()
    pass32, n32 = _profiled_pass(P32, fast, False)
    pass32_polish, _ = _profiled_pass(P32, fast, True)
    P32.close()

    # plain FP64 per-solve path: default launch shape, host poll every iteration, no graph
    P64 = Problem(ctx, x, K, precision=capi.FP64, epssqr=eps)
    P64.set_theta(theta)
    plain = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt)
    tp64, r64 = _timed_solves(args, P64, plain, "plain fp64 solve")
    # FP64 path: autotuned launch shape, graph-captured iteration loop
    if not args.no_autotune:
        P64.set_gamma_uniform()
        tune["fp64"] = P64.autotune(args.tune_cache)
        log(args, f"autotune fp64: {tune['fp64']}")
    t64g, r64g = _timed_solves(args, P64, fast, "fp64 graph solve")
    g64 = P64.get_gamma()
    f64, kkt64 = P64.objective()
    pass64, n64 = _profiled_pass(P64, fast, True)
    # FP64 reference solve (tight KKT) for the error statement
    P64.set_gamma_uniform()
    rref = P64.gamma_solve(maxit=50000, stop_mask=capi.STOP_KKT, kkt_tol=1e-10, check_every=32,
                           use_graph=0 if args.no_graph else 1)
    log(args, f"fp64 reference solve: it={rref.it} kkt={rref.kkt:.2e}")
    gref = P64.get_gamma()
    fref, kref = P64.objective()
    P64.close()
    err32 = _errors(g32, f32, kkt32, gref, fref, truth, T, K)
    err64 = _errors(g64, f64, kkt64, gref, fref, truth, T, K)
    traffic = {}
    tf = os.path.join(ROOT, "profiles", "traffic_fused_pass_fp32_T1e7_K3.json")
    if os.path.exists(tf):
        try:
            tj = json.load(open(tf))
            traffic = {"fp32": tj.get("bytes_per_launch"), "fp64": tj.get("fp64_bytes_per_launch")}
        except Exception:
            traffic = {}
    return dict(T=T, K=K, x=x, theta=theta, tune=tune, t32=t32, r32=r32, tp64=tp64, r64=r64,
                t64g=t64g, r64g=r64g, pass32=pass32, n32=n32, pass32_polish=pass32_polish,
                pass64=pass64, n64=n64, err32=err32, err64=err64, ref_kkt=float(kref),
                ref_it=int(rref.it), traffic=traffic)


def _roofline(kernel, s, n, T, avg, traffic, copy_gbps):
    """The dominant kernel against the HBM roofline: achieved = algorithmic bytes per launch
    (6*n*s, the vectors the reference iteration moves) / average launch time."""
    alg = 6.0 * n * s
    # K - 1 rows of gamma_{i-1} read and of gamma_i written (the last row is 1 - the others),
    # data read
    K = n // T
    own = (2.0 * (K - 1 if K >= 2 else K) * T + T) * s
    out = {"bound": "hbm", "achieved": alg / avg / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": alg / avg / 1e9 / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel,
           "avg_launch_us": avg * 1e6, "algorithmic_bytes_per_launch": alg,
           "own_min_bytes_per_launch": own,
           "own_min_frac": own / avg / 1e9 / HBM_PEAK_GBS, "device_copy_gbps": copy_gbps}
    if traffic:
        out["traffic_gbps"] = traffic / avg / 1e9
        out["traffic_frac"] = traffic / avg / 1e9 / HBM_PEAK_GBS
    return out


def run_graph_c5(args, ctx):
    """The graph stencil at the image configuration: a 1024 x 1024 RGB image on the 4-neighbour
    grid (T = 1 step of R = 1048576 nodes, K = 4, xdim = 3; 64 random rectangles, noise 0.15,
    seed 4), theta at the generating colours, one FP64 gamma solve from gamma0 = 1/K to the
    bench's KKT target.  Reported beside the headline, not as it."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Problem, generate_rect_image
    w = h = 1024
    K = 4
    x, col, lab = generate_rect_image(4, w, h, K, 64, 0.15)
    P = Problem(ctx, x, K, T=1, xdim=3, grid_width=w, precision=capi.FP64, epssqr=args.c5_eps)
    P.set_theta(col.ravel().copy())
    params = dict(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt, check_every=32,
                  use_graph=1, kkt_every=4)
    times, r = _timed_solves(args, P, params, "graph (image) fp64 solve")
    g = P.get_gamma().reshape(-1, K)
    acc = float(np.mean(g.argmax(1) == lab))
    pass_s, npass = _profiled_pass(P, params, True)
    P.close()
    n = w * h * K
    alg = 6.0 * n * 8
    # the two graph kernels move gamma_{i-1}, d_{i-1} (read), gamma_i, d_i (written) and the
    # data (three channels): (4 n + 3 R) * 8 bytes per iteration
    own = (4.0 * n + 3.0 * w * h) * 8
    return dict(workload=f"gamma QP solve, 1024x1024 RGB image, 4-neighbour grid, K={K}, xdim=3, "
                         f"epsilon {args.c5_eps:g}, theta fixed, gamma0=1/K, KKT<={args.kkt:g}",
                time_s=float(np.mean(times)), iterations=int(r.it), kkt=float(r.kkt),
                label_accuracy=acc, pass_us=pass_s * 1e6,
                algorithmic_gbps=alg / pass_s / 1e9, own_min_gbps=own / pass_s / 1e9,
                kernels="spg_pass_grid_dir + spg_pass_grid_ad (+ spg_step) per iteration")


def run_scan(args, ctx):
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
    n_inst = args.scan_series * eps.size
    params = dict(maxit=args.scan_maxit, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt,
                  check_every=64, use_graph=0 if args.no_graph else 1, kkt_every=4)
    B = Batch(ctx, xs, K, ths, eps, precision=capi.FP32_BULK)
    if not args.no_autotune:
        log(args, f"autotune batch: {B.autotune(args.tune_cache)}")
    ctx.synchronize()
    t0 = time.perf_counter()
    res, tdev = B.solve(**params)
    ctx.synchronize()
    tb = time.perf_counter() - t0
    ok = sum(1 for r in res if r.status == 0 and r.kkt <= args.kkt)
    its = np.array([r.it for r in res])
    B.close()
    log(args, f"batched scan: {tb:.3f}s for {n_inst} solves, {ok} certified, it max {its.max()}")
    if args.scan_profile:
        return dict(batched_wall_s=tb, n_inst=n_inst, certified=ok, it_max=int(its.max()),
                    it_mean=float(its.mean()))
    # plain FP64 per-solve path on a deterministic sample of the same instances
    rng = np.random.default_rng(12345)
    pick = rng.choice(n_inst, size=min(args.plain_sample, n_inst), replace=False)
    tp = 0.0
    plain_ok = 0
    for j in pick:
        si, ei = divmod(int(j), eps.size)
        P = Problem(ctx, xs[si], int(K[si]), precision=capi.FP64, epssqr=float(eps[ei]))
        P.set_theta(ths[si])
        P.set_gamma_uniform()
        ctx.synchronize()
        t0 = time.perf_counter()
        r = P.gamma_solve(maxit=args.scan_maxit, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt)
        ctx.synchronize()
        tp += time.perf_counter() - t0
        plain_ok += int(r.status == 0 and r.kkt <= args.kkt)
        P.close()
        log(args, f"plain fp64 instance {j}: T={xs[si].size} K={K[si]} eps={eps[ei]:.3g} "
                  f"it={r.it} kkt={r.kkt:.2e}")
    # extrapolate the sample to all instances by each instance's share of the batch work
    plain_rate = len(pick) / tp
    batched_rate = n_inst / tb
    log(args, f"plain fp64 sample: {len(pick)} solves in {tp:.3f}s")
    return dict(batched_solves_per_s=batched_rate, plain_solves_per_s=plain_rate,
                ratio=batched_rate / plain_rate, batched_wall_s=tb, n_inst=n_inst,
                n_series=args.scan_series, eps_values=int(eps.size), maxit_cap=args.scan_maxit,
                certified=ok, plain_sample=len(pick), plain_certified=plain_ok,
                it_max=int(its.max()), it_mean=float(its.mean()),
                note=("solves/s = instances / wall time at the same iteration cap on both paths; "
                      "instances still above the KKT target at the cap (near-equal levels make "
                      "the QP almost flat) count as uncertified on both"))


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
    # the same algorithmic-byte convention as the device (6 n s per iteration, s = 8) against a
    # host copy of 2 x 512 MiB (bytes read + written per second)
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
    """Fixed work for profilers: a 1 GiB copy kernel, then one solve of the bench workload per
    path (FP32-bulk with its FP64 polish, then FP64) to the bench's KKT target, so that per-pass
    averages cover a whole trajectory as the bench's own pass timing does.

    With --tune-cache the passes use the launch shapes an earlier bench run recorded there."""
    from pasc_inference_amd import capi
    from pasc_inference_amd.solver import Context, Problem
    x, _, theta, eps = c2_workload(args)
    out = {}
    with Context(0) as ctx:
        out["copy_gbps"] = ctx.copy_bandwidth(1 << 30, 5)
        for name, prec in (("fp32", capi.FP32_BULK), ("fp64", capi.FP64)):
            P = Problem(ctx, x, theta.size, precision=prec, epssqr=eps)
            P.set_theta(theta)
            P.set_gamma_uniform()
            if args.tune_cache:
                P.autotune(args.tune_cache)
            P.set_profiling(True)
            r = P.gamma_solve(maxit=200000, stop_mask=capi.STOP_KKT, kkt_tol=args.kkt,
                              check_every=32, kkt_every=4)
            P.close()
            t = r.pass_time_fp32_s if prec == capi.FP32_BULK else r.pass_time_fp64_s
            n = r.passes_fp32 if prec == capi.FP32_BULK else r.passes_fp64
            out[name + "_pass_us"] = t / max(n, 1) * 1e6
            out[name + "_passes"] = n
    print(json.dumps(out))


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
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(requested --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    from pasc_inference_amd.solver import Context
    with Context(0) as ctx:
        copy_gbps = ctx.copy_bandwidth(1 << 30, 5)
        log(args, f"device copy kernel: {copy_gbps:.0f} GB/s")
        c2 = run_c2(args, ctx)
        graph = None if args.no_graph_c5 else run_graph_c5(args, ctx)
        scan = None if args.no_scan else run_scan(args, ctx)
    cpu = None if args.no_cpu else run_cpu_baseline(args, c2)
    t32, t64g = float(np.mean(c2["t32"])), float(np.mean(c2["t64g"]))
    fp64_best = t64g < t32
    best = t64g if fp64_best else t32
    T, K = c2["T"], c2["K"]
    n = T * K
    r32, r64g = c2["r32"], c2["r64g"]
    if fp64_best:
        solve_bytes = 6.0 * n * 8 * r64g.it
        roof = _roofline("spg_pass_1d<double,K=3> (fused SPG iteration, FP64)", 8, n, T,
                         c2["pass64"], c2["traffic"].get("fp64"), copy_gbps)
        other = _roofline("spg_pass_1d<float,K=3> (fused SPG iteration, FP32 phase)", 4, n, T,
                          c2["pass32"], c2["traffic"].get("fp32"), copy_gbps)
    else:
        solve_bytes = 6.0 * n * (4 * r32.it_fp32 + 8 * r32.it_fp64)
        roof = _roofline("spg_pass_1d<float,K=3> (fused SPG iteration, FP32 phase)", 4, n, T,
                         c2["pass32"], c2["traffic"].get("fp32"), copy_gbps)
        other = _roofline("spg_pass_1d<double,K=3> (fused SPG iteration, FP64)", 8, n, T,
                          c2["pass64"], c2["traffic"].get("fp64"), copy_gbps)
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
        "autotune": c2["tune"],
        # whole solve against the roofline, 6*n*s algorithmic bytes per SPG iteration
        "solve_roofline_frac": solve_bytes / best / (HBM_PEAK_GBS * 1e9),
        "errors_vs_fp64_reference_solve": {"fp64_graph": c2["err64"], "fp32_bulk": c2["err32"],
                                           "ref_kkt": c2["ref_kkt"], "ref_it": c2["ref_it"]},
        "roofline": roof,
        "roofline_other_path": other,
        "fp32_polish_pass_avg_us": c2["pass32_polish"] * 1e6,
        "scan": scan,
        "graph_c5": graph,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
