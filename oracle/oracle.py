"""ctypes/numpy wrapper of the FP64 C oracle (test infrastructure only).

The C code restates the reference algorithm routine by routine; see the
header of mgbf_oracle.c for the reference file:line each routine follows.
Arrays are numpy float64, shape (m, n, K) in C order, i.e. k fastest then x
then y, the memory order of the reference's VALL(k, i, j).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libmgbf_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compiles the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("K", ctypes.c_int),
        ("lm", ctypes.c_int), ("nvcol", ctypes.c_int),
        ("hx", ctypes.c_int), ("hy", ctypes.c_int), ("hz", ctypes.c_int),
        ("p", ctypes.c_int), ("gm", ctypes.c_int),
        ("ampl01", ctypes.c_double), ("ampl02", ctypes.c_double),
        ("p_del", ctypes.c_double), ("p_eps", ctypes.c_double),
        ("l_filt_g1", ctypes.c_int), ("l_vertical_filter", ctypes.c_int),
        ("fast", ctypes.c_int), ("lquart", ctypes.c_int), ("line", ctypes.c_int),
        ("line_tables", ctypes.c_int), ("final_vft", ctypes.c_int),
        ("rad3", ctypes.c_int), ("ampl03", ctypes.c_double), ("lin3", ctypes.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(_Params)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.mgbfo_taps2.argtypes = [P, ctypes.c_int, ip, ip, dp]
        L.mgbfo_taps1.argtypes = [P, ctypes.c_int, ip, dp]
        L.mgbfo_taps_fast.argtypes = [P, ctypes.c_int, ip, dp]
        L.mgbfo_ss1.argtypes = [P]; L.mgbfo_ss1.restype = ctypes.c_double
        L.mgbfo_ss2.argtypes = [P]; L.mgbfo_ss2.restype = ctypes.c_double
        L.mgbfo_radial.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int]
        if hasattr(L, "mgbfo_radial_aspect"):
            L.mgbfo_radial_aspect.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp, dp,
                                              ctypes.c_int]
            L.mgbfo_ss2_aspect.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_vertical.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int]
        L.mgbfo_restrict.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_prolong.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_taps3.argtypes = [P, ctypes.c_int, ip, ip, ip, dp]
        L.mgbfo_hexad.argtypes = [dp, ip, dp]
        L.mgbfo_hexlines.argtypes = [P, ip, ip, ip, dp]
        L.mgbfo_momtab.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_hstform.argtypes = [ctypes.c_int, dp]
        L.mgbfo_blinfil.argtypes = [ctypes.c_int, ctypes.c_double, ip, dp]
        L.mgbfo_triad.argtypes = [dp, ip, dp]
        L.mgbfo_line.argtypes = [P, ctypes.c_int, ip, ip, ip, dp]
        L.mgbfo_restrict2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_prolong2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_pad.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_fold.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.mgbfo_mg_size.argtypes = [P]; L.mgbfo_mg_size.restype = ctypes.c_size_t
        L.mgbfo_wawt.argtypes = [P, dp, ctypes.c_void_p, dp, dp]
        L.mgbfo_wt.argtypes = [P, dp, dp]
        L.mgbfo_w.argtypes = [P, dp, dp]
        L.mgbfo_weight.argtypes = [P, dp, ctypes.c_void_p, dp]
        L.mgbfo_num_threads.restype = ctypes.c_int
        L.mgbfo_filt_to_anal.argtypes = [P, ctypes.POINTER(_Anal), dp, dp]
        L.mgbfo_anal_to_filt.argtypes = [P, ctypes.POINTER(_Anal), dp, dp]
        L.mgbfo_lwq_coef.argtypes = [ctypes.c_int, ctypes.c_int, dp, ip]
        _lib = L
    return _lib


class _Anal(ctypes.Structure):
    _fields_ = [("na", ctypes.c_int), ("ma", ctypes.c_int), ("npx", ctypes.c_int),
                ("npy", ctypes.c_int), ("lm_a", ctypes.c_int), ("scheme", ctypes.c_int),
                ("vmap", ctypes.c_int)]

# horizontal schemes of anal_to_filt / filt_to_anal and the vertical maps
LSQR, LIN, QUAD = 0, 1, 2
VLIN, VLWQ = 0, 1


def _anal_ka(P, lm_a: int) -> int:
    return P.nvcol * lm_a + (P.K - P.nvcol * P.lm)


def filt_to_anal(P, na: int, ma: int, npx: int, npy: int, lm_a: int, vf: np.ndarray,
                 scheme: int = LSQR, vmap: int = VLIN) -> np.ndarray:
    """filt_to_anal: generation-1 field (ny, nx, K) -> analysis field
    (ma, na, Ka) over an npx x npy PE grid. scheme: LSQR, LIN
    (l_lin_horizontal) or QUAD (l_quad_horizontal); vmap: VLIN (the _all2
    routines, or _all with l_lin_vertical) or VLWQ (_all without it)."""
    vf = _f64(vf)
    assert vf.shape == P.gen_shape(0)
    A = _Anal(na, ma, npx, npy, lm_a, scheme, vmap)
    wa = np.empty((ma, na, _anal_ka(P, lm_a)))
    lib().mgbfo_filt_to_anal(ctypes.byref(P.c()), ctypes.byref(A), _dp(vf), _dp(wa))
    return wa


def anal_to_filt(P, na: int, ma: int, npx: int, npy: int, lm_a: int, wa: np.ndarray,
                 scheme: int = LSQR, vmap: int = VLIN) -> np.ndarray:
    """anal_to_filt: the transpose of filt_to_anal."""
    wa = _f64(wa)
    assert wa.shape == (ma, na, _anal_ka(P, lm_a))
    A = _Anal(na, ma, npx, npy, lm_a, scheme, vmap)
    vf = np.empty(P.gen_shape(0))
    lib().mgbfo_anal_to_filt(ctypes.byref(P.c()), ctypes.byref(A), _dp(wa), _dp(vf))
    return vf


def lwq_coef(lm_a: int, lm: int):
    """lwq_vertical_coef: (weights (lm_a, 4), kref (lm_a,), 1-based)."""
    c = np.zeros((lm_a, 4))
    k = np.zeros(lm_a, dtype=np.int32)
    lib().mgbfo_lwq_coef(lm_a, lm, _dp(c), k.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return c, k


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class OracleParams:
    """Problem description on one global domain (names follow mg_parameter.f90)."""
    nx: int
    ny: int
    K: int
    lm: int
    nvcol: int
    hx: int = 6
    hy: int = 6
    hz: int = 3
    p: int = 2
    gm: int = 4
    ampl01: float = 16.0 / 7.0
    ampl02: float = 49.0 / 8.0
    p_del: float = 0.0
    p_eps: float = 0.0
    l_filt_g1: bool = True
    l_vertical_filter: bool = True
    weights: list = field(default_factory=lambda: [1.0, 0.5, 0.25, 0.125])
    fast: bool = False     # filtering_fast_* (procedures 3, 8)
    lquart: bool = False   # quadratic inter-generation transfer (adjoint2 / direct2)
    line: bool = False     # filtering_lin2_* (procedures 2, 7): triad line filters
    line_tables: bool = True   # mgbf_line: the line tables are set (else identity lines)
    final_vft: bool = False    # W A W^T ends with Vf^T (filtering_lin2_ens)
    rad3: bool = False     # filtering_rad3 (procedure 4): 3-D radial filter of the 3-D variables
    ampl03: float = 1.0    # 3-D aspect amplitude (pasp03)
    lin3: bool = False     # filtering_lin3 (procedure 5): hexad line filters of the 3-D variables

    def c(self) -> _Params:
        return _Params(self.nx, self.ny, self.K, self.lm, self.nvcol, self.hx, self.hy,
                       self.hz, self.p, self.gm, self.ampl01, self.ampl02, self.p_del,
                       self.p_eps, int(self.l_filt_g1), int(self.l_vertical_filter), int(self.fast),
                       int(self.lquart), int(self.line), int(self.line_tables), int(self.final_vft),
                       int(self.rad3), float(self.ampl03), int(self.lin3))

    def gen_shape(self, g: int):
        """(m_g, n_g, K) of generation g (0-based)."""
        return (self.ny >> g, self.nx >> g, self.K)

    def wconst(self) -> np.ndarray:
        w = list(self.weights)[: self.gm]
        w = w + [w[-1]] * (self.gm - len(w))
        return np.ascontiguousarray(w, dtype=np.float64)


def taps2(P: OracleParams):
    c = P.c()
    cap = (2 * P.hx + 1) * (2 * P.hy + 1)
    gx = (ctypes.c_int * cap)(); gy = (ctypes.c_int * cap)()
    w = np.zeros(cap)
    nt = lib().mgbfo_taps2(ctypes.byref(c), cap, gx, gy, _dp(w))
    if nt < 0:
        raise ValueError("filter reach exceeds the halo")
    return np.array(gx[:nt]), np.array(gy[:nt]), w[:nt].copy()


def taps1(P: OracleParams):
    c = P.c()
    cap = 2 * P.hz + 1
    gx = (ctypes.c_int * cap)()
    w = np.zeros(cap)
    nt = lib().mgbfo_taps1(ctypes.byref(c), cap, gx, _dp(w))
    if nt < 0:
        raise ValueError("filter reach exceeds the halo")
    return np.array(gx[:nt]), w[:nt].copy()


def taps3(P: OracleParams):
    """(gx, gy, gz, w) of the 3-D filter of procedure 4 in vrbeta3's order."""
    c = P.c()
    cap = (2 * P.hx + 1) * (2 * P.hy + 1) * (2 * P.hz + 1)
    gx = (ctypes.c_int * cap)(); gy = (ctypes.c_int * cap)(); gz = (ctypes.c_int * cap)()
    w = np.zeros(cap)
    nt = lib().mgbfo_taps3(ctypes.byref(c), cap, gx, gy, gz, _dp(w))
    return np.array(gx[:nt]), np.array(gy[:nt]), np.array(gz[:nt]), w[:nt].copy()


def hexad(aspect):
    """hexad: (generators (7, 3) by color 1..7, weights (7,), failure flag)."""
    a = _f64(aspect)
    lh = (ctypes.c_int * 21)()
    w = np.zeros(7)
    ff = lib().mgbfo_hexad(_dp(a), lh, _dp(w))
    return np.array(lh[:]).reshape(7, 3), w, bool(ff)


def hexlines(P: OracleParams):
    """The six hexad lines of the 3-D variables: [(color, (gx, gy, gz), h, fil[0..h])]."""
    c = P.c()
    q = (ctypes.c_int * 6)(); g = (ctypes.c_int * 18)(); h = (ctypes.c_int * 6)()
    fil = np.zeros(6 * 9)
    n = lib().mgbfo_hexlines(ctypes.byref(c), q, g, h, _dp(fil))
    return [(q[l], tuple(g[3 * l: 3 * l + 3]), h[l], fil[9 * l: 9 * l + h[l] + 1].copy())
            for l in range(n)]


def taps_fast(P: OracleParams):
    """1-D factor (offsets, weights) of the fast filter along x."""
    c = P.c()
    cap = 2 * P.hx + 1
    gx = (ctypes.c_int * cap)()
    w = np.zeros(cap)
    nt = lib().mgbfo_taps_fast(ctypes.byref(c), cap, gx, _dp(w))
    if nt < 0:
        raise ValueError("filter reach exceeds the halo")
    return np.array(gx[:nt]), w[:nt].copy()


def momtab(p: int, nh: int):
    """inimomtab: (bnorm, bsprds) for half-spans 1..nh (index 0 unused)."""
    bn = np.zeros(nh + 1); bs = np.zeros(nh + 1)
    if lib().mgbfo_momtab(p, nh, _dp(bn), _dp(bs)):
        raise ValueError("p or nh out of range")
    return bn, bs


def hstform(p: int, spread2: float):
    """hstform: (half-span, failure flag) for a spread**2."""
    v = np.array([float(spread2)])
    ff = lib().mgbfo_hstform(p, _dp(v))
    return float(v[0]), bool(ff)


def blinfil(p: int, hspan: float):
    """blinfil: (h, fil[0..h]) of the dibeta filter, or None on failure."""
    h = ctypes.c_int(0)
    fil = np.zeros(16)
    if lib().mgbfo_blinfil(p, float(hspan), ctypes.byref(h), _dp(fil)):
        return None
    return h.value, fil[: h.value + 1].copy()


def triad(aspect):
    """triad: (generators (3, 2), weights (3,), failure flag)."""
    a = _f64(aspect)
    lt = (ctypes.c_int * 6)()
    w = np.zeros(3)
    ff = lib().mgbfo_triad(_dp(a), lt, _dp(w))
    return np.array(lt[:]).reshape(3, 2), w, bool(ff)


def line(P: OracleParams, c: int):
    """Line c of the domain's line filter: ((dx, dy), h, fil[0..h]); h = -1 for identity."""
    cp = P.c()
    dx, dy, h = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    fil = np.zeros(16)
    lib().mgbfo_line(ctypes.byref(cp), c, ctypes.byref(dx), ctypes.byref(dy), ctypes.byref(h), _dp(fil))
    return (dx.value, dy.value), h.value, fil[: max(h.value, 0) + 1].copy()


def ss1(P: OracleParams) -> float:
    c = P.c()
    return lib().mgbfo_ss1(ctypes.byref(c))


def ss2(P: OracleParams) -> float:
    c = P.c()
    return lib().mgbfo_ss2(ctypes.byref(c))


def _f64(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.float64)


def radial(P: OracleParams, x: np.ndarray, adjoint: bool = False) -> np.ndarray:
    x = _f64(x); m, n, K = x.shape
    y = np.empty_like(x); c = P.c()
    lib().mgbfo_radial(ctypes.byref(c), K, n, m, _dp(x), _dp(y), int(adjoint))
    return y


def radial_aspect(P: OracleParams, asp: np.ndarray, x: np.ndarray, adjoint: bool = False) -> np.ndarray:
    """F (or F^T) of one generation with a per-point aspect tensor field:
    asp (m, n, 3) = (A11, A22, A12) per point, x (m, n, K)."""
    x = _f64(x)
    m, n, K = x.shape
    a = np.ascontiguousarray(asp, dtype=np.float64)
    assert a.shape == (m, n, 3)
    y = np.empty_like(x)
    if lib().mgbfo_radial_aspect(ctypes.byref(P.c()), K, n, m, _dp(a), _dp(x), _dp(y), int(adjoint)) != 0:
        raise ValueError("filter reach exceeds the halo (getlinesum2 stops)")
    return y


def ss2_aspect(P: OracleParams, asp: np.ndarray) -> np.ndarray:
    """getlinesum2's per-point line sums of an aspect field (m, n, 3)."""
    a = np.ascontiguousarray(asp, dtype=np.float64)
    ss = np.empty(a.shape[:2])
    if lib().mgbfo_ss2_aspect(ctypes.byref(P.c()), a.shape[1], a.shape[0], _dp(a), _dp(ss)) != 0:
        raise ValueError("filter reach exceeds the halo (getlinesum2 stops)")
    return ss


def vertical(P: OracleParams, x: np.ndarray, adjoint: bool = False) -> np.ndarray:
    x = _f64(x); m, n, K = x.shape
    assert K == P.K
    y = np.empty_like(x); c = P.c()
    lib().mgbfo_vertical(ctypes.byref(c), n, m, _dp(x), _dp(y), int(adjoint))
    return y


def restrict(x: np.ndarray, quad: bool = False) -> np.ndarray:
    """R (adjoint + bocoT_2d), or with quad R2 (adjoint2, lquart)."""
    x = _f64(x); m, n, K = x.shape
    y = np.empty((m // 2, n // 2, K))
    (lib().mgbfo_restrict2 if quad else lib().mgbfo_restrict)(K, n, m, _dp(x), _dp(y))
    return y


def prolong(xc: np.ndarray, quad: bool = False) -> np.ndarray:
    """P (boco_2d + direct1), or with quad P2 (direct2, lquart)."""
    xc = _f64(xc); mc, nc, K = xc.shape
    y = np.empty((2 * mc, 2 * nc, K))
    (lib().mgbfo_prolong2 if quad else lib().mgbfo_prolong)(K, 2 * nc, 2 * mc, _dp(xc), _dp(y))
    return y


def pad(x: np.ndarray, h: int) -> np.ndarray:
    x = _f64(x); m, n, K = x.shape
    y = np.empty((m + 2 * h, n + 2 * h, K))
    lib().mgbfo_pad(K, n, m, h, _dp(x), _dp(y))
    return y


def fold(xp: np.ndarray, h: int) -> np.ndarray:
    xp = _f64(xp); mp, np_, K = xp.shape
    y = np.empty((mp - 2 * h, np_ - 2 * h, K))
    lib().mgbfo_fold(K, np_ - 2 * h, mp - 2 * h, h, _dp(xp), _dp(y))
    return y


def _wfield_ptrs(P: OracleParams, wfield):
    if wfield is None:
        return None, []
    keep = [None if w is None else _f64(w) for w in wfield]
    arr = (ctypes.c_void_p * P.gm)(*[0 if w is None else w.ctypes.data for w in keep])
    return arr, keep


def mg_split(P: OracleParams, v: np.ndarray):
    """Splits a flat multigrid vector into its generation fields."""
    out, off = [], 0
    for g in range(P.gm):
        s = P.gen_shape(g)
        sz = s[0] * s[1] * s[2]
        out.append(v[off: off + sz].reshape(s))
        off += sz
    return out


def mg_join(fields) -> np.ndarray:
    return np.concatenate([_f64(f).ravel() for f in fields])


def wawt(P: OracleParams, x: np.ndarray, wfield=None) -> np.ndarray:
    x = _f64(x)
    assert x.shape == P.gen_shape(0)
    y = np.empty_like(x); c = P.c(); wc = P.wconst()
    arr, _keep = _wfield_ptrs(P, wfield)
    lib().mgbfo_wawt(ctypes.byref(c), _dp(wc), arr, _dp(x), _dp(y))
    return y


def wt(P: OracleParams, x: np.ndarray) -> np.ndarray:
    x = _f64(x); c = P.c()
    v = np.empty(lib().mgbfo_mg_size(ctypes.byref(c)))
    lib().mgbfo_wt(ctypes.byref(c), _dp(x), _dp(v))
    return v


def w(P: OracleParams, v: np.ndarray) -> np.ndarray:
    v = _f64(v); c = P.c()
    y = np.empty(P.gen_shape(0))
    lib().mgbfo_w(ctypes.byref(c), _dp(v), _dp(y))
    return y


def weight(P: OracleParams, v: np.ndarray, wfield=None) -> np.ndarray:
    v = _f64(v).copy(); c = P.c(); wc = P.wconst()
    arr, _keep = _wfield_ptrs(P, wfield)
    lib().mgbfo_weight(ctypes.byref(c), _dp(wc), arr, _dp(v))
    return v


def num_threads() -> int:
    return lib().mgbfo_num_threads()
