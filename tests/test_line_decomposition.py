"""The product's line decompositions (tritform's triad, hextform's hexad,
jp_pbfil3.f90) in closed form for the diagonal aspect tensors a namelist
reaches, against the oracle's restatement of the reference's iterations:
same colors, generators and weights; sum_c w_c l_c l_c^T = A; anything
off-diagonal is refused by the product. CPU only: mgbf_coef.cpp is built
with g++ together with a test probe."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "standalone_mgbf_loc_amd", "csrc")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    so = str(tmp_path_factory.mktemp("coef") / "libcoefprobe.so")
    cmd = [gxx, "-std=c++17", "-O1", "-shared", "-fPIC", f"-I{ROOT}/include", f"-I{CSRC}",
           os.path.join(CSRC, "mgbf_coef.cpp"), os.path.join(ROOT, "tests", "coef_probe.cpp"), "-o", so]
    subprocess.run(cmd, check=True, capture_output=True)
    L = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    for f in (L.probe_triad, L.probe_hexad):
        f.argtypes = [dp, ip, dp]
    return L


def _call(f, a, nl, dim):
    a = np.ascontiguousarray(a, dtype=np.float64)
    l = (ctypes.c_int * (nl * dim))()
    w = np.zeros(nl)
    ok = f(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), l,
           w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return bool(ok), np.array(l[:]).reshape(nl, dim), w


def _diag_cases(dim, n=8):
    rng = np.random.default_rng(40 + dim)
    cases = [np.ones(dim), np.full(dim, 1 / np.sqrt(6.125)), np.full(dim, 1 / np.sqrt(0.7))]
    cases += [0.05 + 3 * rng.random(dim) for _ in range(n)]
    return cases


def test_hexad_of_diagonal_tensors_matches_the_reference_iteration(probe):
    for d in _diag_cases(3):
        a = np.concatenate([d, [0.0, 0.0, 0.0]])
        ok, l, w = _call(probe.probe_hexad, a, 7, 3)
        lr, wr, failed = O.hexad(a)
        assert ok and not failed
        assert np.array_equal(l, lr) and np.array_equal(w, wr), (d, l, lr, w, wr)
        A = sum(wc * np.outer(lc, lc) for lc, wc in zip(l, w))
        assert np.allclose(A, np.diag(d), atol=0) and np.all(w >= 0)
        assert np.count_nonzero(np.abs(l).sum(axis=1) == 0) == 1     # one null color


def test_triad_of_diagonal_tensors_matches_the_reference_iteration(probe):
    for d in _diag_cases(2):
        a = np.array([d[0], d[1], 0.0])
        ok, l, w = _call(probe.probe_triad, a, 3, 2)
        lr, wr, failed = O.triad(a)
        assert ok and not failed
        assert np.array_equal(l, lr) and np.array_equal(w, wr)
        A = sum(wc * np.outer(lc, lc) for lc, wc in zip(l, w))
        assert np.allclose(A, np.diag(d), atol=0)


def test_off_diagonal_or_degenerate_tensors_are_refused(probe):
    # the oracle's iterations resolve these onto lattice lines off the axes,
    # which the product does not build: it refuses them
    hex_cases = [[1.0, 1.0, 1.0, 0.3, 0.0, 0.0], [2.0, 1.0, 1.5, 0.0, -0.2, 0.1],
                 [1.0, 0.0, 1.0, 0.0, 0.0, 0.0], [1.0, -1.0, 1.0, 0.0, 0.0, 0.0]]
    for a in hex_cases:
        assert not _call(probe.probe_hexad, a, 7, 3)[0], a
    lr, wr, failed = O.hexad(hex_cases[0])
    assert not failed and np.all(wr >= -1e-14)
    assert any(np.count_nonzero(g) > 1 and wc > 0 for g, wc in zip(lr, wr))
    for a in ([1.0, 1.0, 0.2], [1.0, 2.0, -0.3], [0.0, 1.0, 0.0]):
        assert not _call(probe.probe_triad, a, 3, 2)[0], a
