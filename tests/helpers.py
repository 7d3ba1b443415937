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
