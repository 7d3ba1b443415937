"""PETSc binary Vec files, the reference's on-disk format for signals and gamma.

A Vec file is big-endian: int32 class id 1211214 (VEC_FILE_CLASSID), int32 length n, then n
float64 values (GeneralVector::load_global / save_binary, PetscViewerBinary).  Gamma files are
k-major ([K][T*R]) on disk, while the solver layout is t-major (idx = t*R*K + r*K + k).
permute_trblock is Decomposition::permute_TRblocksize (decomposition.cpp:265-356) on one rank,
with permute_TRxdim and permute_TRK (decomposition.cpp:246-263) as its data and gamma cases;
kmajor_to_tmajor / tmajor_to_kmajor are the R = 1 case.
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


def permute_trblock(v, T: int, R: int, bs: int, invert: bool = False) -> np.ndarray:
    """Decomposition::permute_TRblocksize (decomposition.cpp:265-356) on one rank: the original
    layout holds component k of node r at time t at (k*R + r)*T + t ("blockTR"), the permuted
    one at (t*R + r)*bs + k ("TRblock").  invert maps the permuted layout back."""
    v = np.asarray(v, dtype=np.float64).ravel()
    if v.size != T * R * bs:
        raise ValueError(f"permute_trblock: {v.size} values, expected T*R*bs = {T * R * bs}")
    if not invert:
        return v.reshape(bs, R, T).transpose(2, 1, 0).ravel().copy()
    return v.reshape(T, R, bs).transpose(2, 1, 0).ravel().copy()


def permute_trxdim(v, T: int, R: int, xdim: int, invert: bool = False) -> np.ndarray:
    """Decomposition::permute_TRxdim (decomposition.cpp:246-252): data vectors."""
    return permute_trblock(v, T, R, xdim, invert)


def permute_trk(v, T: int, R: int, K: int, invert: bool = False) -> np.ndarray:
    """Decomposition::permute_TRK (decomposition.cpp:255-262): gamma vectors."""
    return permute_trblock(v, T, R, K, invert)
