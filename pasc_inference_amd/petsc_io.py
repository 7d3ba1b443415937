"""PETSc binary Vec files, the reference's on-disk format for signals and gamma.

A Vec file is big-endian: int32 class id 1211214 (VEC_FILE_CLASSID), int32 length n, then n
float64 values (GeneralVector::load_global / save_binary, PetscViewerBinary).  Gamma files are
k-major ([K][T*R]) on disk, while the solver layout is t-major (idx = t*R*K + r*K + k); the
permutation helpers mirror Decomposition::permute_TRK (decomposition.cpp:265-356) for R = 1.
"""
from __future__ import annotations

import struct

import numpy as np

VEC_FILE_CLASSID = 1211214


def read_vec(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        head = f.read(8)
        if len(head) != 8:
            raise ValueError(f"{path}: truncated header")
        cid, n = struct.unpack(">ii", head)
        if cid != VEC_FILE_CLASSID:
            raise ValueError(f"{path}: class id {cid} is not a PETSc Vec ({VEC_FILE_CLASSID})")
        data = f.read(8 * n)
        if len(data) != 8 * n:
            raise ValueError(f"{path}: expected {n} values")
    return np.frombuffer(data, dtype=">f8").astype(np.float64)


def write_vec(path: str, v) -> None:
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    with open(path, "wb") as f:
        f.write(struct.pack(">ii", VEC_FILE_CLASSID, v.size))
        f.write(v.astype(">f8").tobytes())


def kmajor_to_tmajor(g, K: int) -> np.ndarray:
    """[K][T] on disk -> t-major gamma (idx = t*K + k)."""
    g = np.asarray(g, dtype=np.float64).ravel()
    return g.reshape(K, -1).T.ravel().copy()


def tmajor_to_kmajor(g, K: int) -> np.ndarray:
    g = np.asarray(g, dtype=np.float64).ravel()
    return g.reshape(-1, K).T.ravel().copy()
