"""Writes tests/golden/tiny_qp.json: exact minimisers of tiny FEM-H1 gamma QPs.

Each case is min 1/2 g'Ag - b'g over the product of unit simplices, A = alpha*A0 with A0 the
reference's H1 graph-time matrix.  The exact minimiser is found by enumerating the support of
every time step: for each support pattern the equality-constrained QP is solved through its KKT
system, and the pattern whose solution is primal and dual feasible is kept (convex problem, so
this is the global minimiser).  Each result is cross-checked against scipy's SLSQP.  A0 is built
here as the Laplacian of the weighted graph on (t, r) nodes (time edges weight 2, spatial edges
weight 1 between (t, r) and (t', s) for |t - t'| <= 1, s a neighbour of r), which is an
independent construction of the rows assembled in blockgraphsparse.cpp:48-109.

Run: python tests/golden/make_golden.py   (numpy + scipy, CPU only)
"""
import itertools
import json
import os

import numpy as np
from scipy.optimize import minimize

HERE = os.path.dirname(os.path.abspath(__file__))


def laplacian_a0(T, R, nbrs):
    """A0 as the Laplacian of the weighted (t, r) graph; nbrs[r] lists the neighbours of r."""
    N = T * R
    L = np.zeros((N, N))

    def node(t, r):
        return t * R + r

    def edge(i, j, w):
        L[i, i] += w
        L[j, j] += w
        L[i, j] -= w
        L[j, i] -= w

    for t in range(T - 1):
        for r in range(R):
            edge(node(t, r), node(t + 1, r), 2.0)
    seen = set()
    for r in range(R):
        for s in nbrs[r]:
            for t in range(T):
                for t2 in (t - 1, t, t + 1):
                    if 0 <= t2 < T:
                        key = (min((t, r), (t2, s)), max((t, r), (t2, s)))
                        if key in seen:
                            continue
                        seen.add(key)
                        edge(node(t, r), node(t2, s), 1.0)
    return L


def full_matrix(T, R, K, alpha, nbrs):
    """A over gamma in t-major layout (idx = (t*R + r)*K + k)."""
    return alpha * np.kron(laplacian_a0(T, R, nbrs), np.eye(K))


def brute_force(A, b, nb, K):
    n = nb * K
    best = None
    subsets = [s for m in range(1, K + 1) for s in itertools.combinations(range(K), m)]
    for pattern in itertools.product(subsets, repeat=nb):
        S = [t * K + k for t in range(nb) for k in pattern[t]]
        E = np.zeros((nb, len(S)))
        for j, i in enumerate(S):
            E[i // K, j] = 1.0
        M = np.block([[A[np.ix_(S, S)], E.T], [E, np.zeros((nb, nb))]])
        rhs = np.concatenate([b[S], np.ones(nb)])
        sol = np.linalg.lstsq(M, rhs, rcond=None)[0]
        if np.linalg.norm(M @ sol - rhs) > 1e-9:
            continue
        z = np.zeros(n)
        z[S] = sol[: len(S)]
        if z.min() < -1e-12:
            continue
        nu = sol[len(S):]  # A_SS z - b_S + E'nu = 0: g = -nu on the support
        g = A @ z - b
        ok = True
        for i in range(n):
            if i not in S and g[i] + nu[i // K] < -1e-9:  # bound multiplier must be >= 0
                ok = False
                break
        if not ok:
            continue
        f = 0.5 * z @ A @ z - b @ z
        if best is None or f < best[1] - 1e-13:
            best = (np.maximum(z, 0.0), f)
    return best


def slsqp(A, b, nb, K):
    n = nb * K
    cons = [{"type": "eq", "fun": (lambda z, t=t: z[t * K:(t + 1) * K].sum() - 1.0)} for t in range(nb)]
    r = minimize(lambda z: 0.5 * z @ A @ z - b @ z, np.full(n, 1.0 / K),
                 jac=lambda z: A @ z - b, bounds=[(0, None)] * n, constraints=cons,
                 method="SLSQP", options={"ftol": 1e-15, "maxiter": 2000})
    return r.x, r.fun


def main():
    rng = np.random.default_rng(20240607)
    cases = []
    configs = [  # (T, R, K, alpha, graph)
        (1, 1, 1, 0.5, None), (1, 1, 3, 0.5, None), (2, 1, 2, 0.3, None), (3, 1, 2, 1.7, None),
        (4, 1, 2, 0.05, None), (3, 1, 3, 0.8, None), (4, 1, 3, 0.2, None), (5, 1, 2, 3.0, None),
        (1, 4, 2, 0.7, "path"), (1, 4, 3, 0.4, "path"), (2, 2, 2, 0.6, "path"),
        (1, 4, 2, 0.9, "square"),
    ]
    for (T, R, K, alpha, gname) in configs:
        if gname is None:
            nbrs = [[] for _ in range(R)]
        elif gname == "path":
            nbrs = [[s for s in (r - 1, r + 1) if 0 <= s < R] for r in range(R)]
        else:  # 2x2 grid
            nbrs = [[1, 2], [0, 3], [0, 3], [1, 2]]
        for rep in range(2):
            A = full_matrix(T, R, K, alpha, nbrs)
            b = -rng.uniform(0.0, 2.0, T * R * K)
            nb = T * R
            gbest, fbest = brute_force(A, b, nb, K)
            zs, fs = slsqp(A, b, nb, K)
            assert fs >= fbest - 1e-9, (T, R, K, fs, fbest)
            assert abs(fs - fbest) < 1e-7, (T, R, K, fs, fbest)
            ptr = [0]
            idx = []
            for r in range(R):
                idx += nbrs[r]
                ptr.append(len(idx))
            cases.append(dict(T=T, R=R, K=K, alpha=alpha, graph_ptr=ptr, graph_idx=idx,
                              b=b.tolist(), gamma=gbest.tolist(), f=fbest,
                              slsqp_f=fs, A0_dense=(A / alpha).tolist() if T * R * K <= 12 else None))
    with open(os.path.join(HERE, "tiny_qp.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f)
    print(f"wrote {len(cases)} cases")


if __name__ == "__main__":
    main()
