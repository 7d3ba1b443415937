"""Shared helpers of the test-suite (inputs and comparison utilities)."""
import numpy as np


def grid_csr(width: int, height: int):
    """4-neighbour grid adjacency (up, left, right, down) as CSR over row-major nodes."""
    ptr = [0]
    idx = []
    for r in range(height):
        for c in range(width):
            if r > 0:
                idx.append((r - 1) * width + c)
            if c > 0:
                idx.append(r * width + c - 1)
            if c < width - 1:
                idx.append(r * width + c + 1)
            if r < height - 1:
                idx.append((r + 1) * width + c)
            ptr.append(len(idx))
    return np.array(ptr, dtype=np.int64), np.array(idx, dtype=np.int32)


def label_agreement(g1, g2, K):
    a = np.asarray(g1).reshape(-1, K).argmax(axis=1)
    b = np.asarray(g2).reshape(-1, K).argmax(axis=1)
    return float(np.mean(a == b))


def random_graph_csr(R: int, mean_degree: float, seed: int):
    """A random undirected graph on R nodes (symmetric adjacency, no self loops) as CSR."""
    rng = np.random.default_rng(seed)
    m = int(R * mean_degree / 2)
    a = rng.integers(0, R, size=m)
    b = rng.integers(0, R, size=m)
    edges = {(int(min(i, j)), int(max(i, j))) for i, j in zip(a, b) if i != j}
    nbr = [[] for _ in range(R)]
    for i, j in sorted(edges):
        nbr[i].append(j)
        nbr[j].append(i)
    ptr = np.zeros(R + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(v) for v in nbr])
    idx = np.array([j for v in nbr for j in sorted(v)], dtype=np.int32)
    return ptr, idx
