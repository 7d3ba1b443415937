"""Writes the oracle's minimiser of the single large instance as a compact fixture.

Workload: T = 1e7, K = 3, Markov levels {0, 1, 3}, switch probability 1e-4, N(0, 1) noise,
seed 1 (the library's deterministic generator), epsilon = 100, theta fixed at the levels,
gamma0 = 1/K.  The FP64 CPU oracle (oracle/pasc_oracle.c, SPG-QP of spgqpsolver.cpp:55-367
with the sign-safe dots and g recomputed from gamma, which reach KKT 1e-12 where the op-by-op
formulas stall near 1e-9) is run to the natural KKT residual <= 1e-10 on all host cores.

The fixture (c2_T1e7_K3_eps100.npz) holds what a test at the full size compares against:
  * sha256 of the generated data x (float64 bytes) -- the test regenerates x and checks it;
  * f* (the objective 1/2 g'Ag - b'g, explicit FP64), the KKT residual and the iterations;
  * the argmax labels as run lengths (run starts and their labels);
  * per-cluster sums sum_t gamma_kt and sum_t x_t gamma_kt;
  * gamma (t-major, all K components) at a fixed sample of >= 1e5 time steps covering every
    fractional region: the most fractional step (largest second component) of every block of
    BLOCK consecutive steps that holds a fractional step (some component strictly inside
    (0, 1)), both steps of every label change, and steps drawn uniformly with a fixed seed.

Run in the build container: python tests/golden/make_c2_fixture.py [--threads N]
"""
import argparse
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(HERE, "c2_T1e7_K3_eps100.npz")
FRAC_TOL = 1e-9   # a component in (FRAC_TOL, 1 - FRAC_TOL) makes a step fractional
BLOCK = 100       # at T = 1e7: 1e5 blocks
N_RANDOM = 10_000


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--kkt", type=float, default=1e-10)
    ap.add_argument("--T", type=int, default=10_000_000)
    ap.add_argument("--out", default=OUT)
    a = ap.parse_args()

    import pyoracle
    from pasc_inference_amd.solver import SIGNAL_C2, generate_markov
    c = SIGNAL_C2
    T, K = a.T, len(c["levels"])
    x, _ = generate_markov(c["seed"], T, c["levels"], c["p_switch"], c["sigma"])
    theta = np.array(c["levels"], dtype=np.float64)
    b = pyoracle.assemble_b(x, theta, T, 1, K)
    alpha = c["eps"] / T
    t0 = time.perf_counter()
    g, r = pyoracle.spg_solve(np.full(T * K, 1.0 / K), b, T, 1, K, alpha, maxit=1_000_000,
                              stop_difff=0, stop_kkt=1, kkt_tol=a.kkt, kkt_scale=T,
                              gd_centered=1, gd_certified=1, dad_edges=1, g_recompute=1,
                              safeguard=1, nthreads=a.threads)
    dt = time.perf_counter() - t0
    print(f"oracle: it={r.it} kkt={r.kkt:.3e} fx={r.fx!r} stop={r.stop_reason} {dt:.0f}s",
          flush=True)
    if not r.kkt <= a.kkt:
        raise SystemExit("the oracle did not reach the KKT target")
    f_star = pyoracle.objective(g, b, T, 1, K, alpha)
    G = g.reshape(T, K)
    lab = G.argmax(1).astype(np.int8)
    starts = np.flatnonzero(np.r_[True, lab[1:] != lab[:-1]]).astype(np.int64)
    frac = np.any((G > FRAC_TOL) & (G < 1 - FRAC_TOL), axis=1)
    # the most fractional step of every block that holds one
    second = np.sort(G, axis=1)[:, -2]
    nb = (T + BLOCK - 1) // BLOCK
    sec = np.full(nb * BLOCK, -1.0)
    sec[:T] = np.where(frac, second, -1.0)
    sec = sec.reshape(nb, BLOCK)
    has = sec.max(1) >= 0.0
    pick = set((np.flatnonzero(has) * BLOCK + sec[has].argmax(1)).tolist())
    for s0 in starts[1:]:
        pick.update((int(s0) - 1, int(s0)))
    rng = np.random.default_rng(20261203)
    pick.update(rng.choice(T, size=min(N_RANDOM, T), replace=False).tolist())
    idx = np.array(sorted(pick), dtype=np.int64)
    n_runs = int((np.diff(np.flatnonzero(frac)) > 1).sum() + 1) if frac.any() else 0
    out = dict(
        T=np.int64(T), K=np.int64(K), epssqr=np.float64(c["eps"]), theta=theta,
        x_sha256=np.array(hashlib.sha256(x.tobytes()).hexdigest()),
        f_star=np.float64(f_star), kkt=np.float64(r.kkt), it=np.int64(r.it),
        label_starts=starts, label_values=lab[starts],
        sum_gamma=G.sum(0), sum_x_gamma=x @ G,
        sample_idx=idx, sample_gamma=G[idx].copy(),
        n_fractional=np.int64(frac.sum()), n_fractional_runs=np.int64(n_runs),
        n_fractional_blocks=np.int64(has.sum()),
    )
    np.savez_compressed(a.out, **out)
    print(f"wrote {a.out}: {idx.size} sampled steps, {int(frac.sum())} fractional steps, "
          f"{starts.size} label runs, f*={f_star!r}", flush=True)


if __name__ == "__main__":
    main()
