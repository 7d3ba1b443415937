"""Layout permutations of Decomposition (decomposition.cpp:246-356) in the Python I/O helpers
and the C++ host layer, against a restatement of the reference's index sets."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _reference_permute(orig, T, R, bs):
    """permute_TRblocksize on one rank as written (decomposition.cpp:265-356): the index set of
    the original vector enumerates (t, r, k), that of the new vector (k, r, t), and the values
    are copied from the first subvector to the second in that order."""
    orig_arr, new_arr = [], []
    for t in range(T):
        for i in range(R):
            for k in range(bs):
                orig_arr.append(t * R * bs + i * bs + k)
    for k in range(bs):
        for i in range(R):
            for t in range(T):
                new_arr.append(t * R * bs + i * bs + k)
    new = np.zeros(T * R * bs)
    sub = orig[np.array(orig_arr)]
    new[np.array(new_arr)] = sub
    return new


@pytest.mark.parametrize("T,R,bs", [(5, 1, 3), (4, 3, 2), (1, 7, 4), (6, 2, 1)])
def test_permute_trblock_matches_reference_index_sets(T, R, bs):
    from pasc_inference_amd import petsc_io
    v = np.random.default_rng(T * 100 + R * 10 + bs).random(T * R * bs)
    p = petsc_io.permute_trblock(v, T, R, bs)
    np.testing.assert_array_equal(p, _reference_permute(v, T, R, bs))
    np.testing.assert_array_equal(petsc_io.permute_trblock(p, T, R, bs, invert=True), v)
    np.testing.assert_array_equal(petsc_io.permute_trk(v, T, R, bs), p)
    np.testing.assert_array_equal(petsc_io.permute_trxdim(v, T, R, bs), p)
    if R == 1:
        np.testing.assert_array_equal(p, petsc_io.kmajor_to_tmajor(v, bs))
    with pytest.raises(ValueError):
        petsc_io.permute_trblock(v[:-1], T, R, bs)


def test_cpp_decomposition_matches_python(tmp_path):
    """The host layer's Decomposition<HipVector> (header only) compiled with g++."""
    from pasc_inference_amd import petsc_io
    src = tmp_path / "perm.cpp"
    src.write_text(r'''
#include <cstdio>
#include "pasc/decomposition.h"
using namespace pascinference;
int main() {
    const int64_t T = 4, R = 3;
    const int K = 2, xdim = 3;
    Decomposition<HipVector> d(T, R, K, xdim);
    std::vector<double> g(T * R * K), x(T * R * xdim), gp, xp, gb, xb;
    for (size_t i = 0; i < g.size(); i++) g[i] = (double)i;
    for (size_t i = 0; i < x.size(); i++) x[i] = 100.0 + (double)i;
    d.permute_TRK(g, gp);
    d.permute_TRxdim(x, xp);
    gb.assign(g.size(), -1.0);
    d.permute_TRK(gb, gp, true);
    for (double v : gp) printf("%g ", v);
    printf("\n");
    for (double v : xp) printf("%g ", v);
    printf("\n");
    for (double v : gb) printf("%g ", v);
    printf("\n");
    try { std::vector<double> bad(5), o; d.permute_TRK(bad, o); printf("no throw\n"); }
    catch (const std::invalid_argument &) { printf("throws\n"); }
    return 0;
}
''')
    exe = tmp_path / "perm"
    inc = [f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(ROOT, 'pasc_inference_amd', 'host', 'include')}"]
    r = subprocess.run(["g++", "-std=c++17", "-O1", *inc, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    T, R, K, xdim = 4, 3, 2, 3
    g = np.arange(T * R * K, dtype=float)
    x = 100.0 + np.arange(T * R * xdim, dtype=float)
    np.testing.assert_array_equal(np.array(out[0].split(), float), petsc_io.permute_trk(g, T, R, K))
    np.testing.assert_array_equal(np.array(out[1].split(), float), petsc_io.permute_trxdim(x, T, R, xdim))
    np.testing.assert_array_equal(np.array(out[2].split(), float), g)
    assert out[3] == "throws"
