"""Dense-matrix model of the MGBF operator on tiny grids (test infrastructure).

An independent restatement of the operator as explicit matrices, written from
the mathematical definitions rather than from the reference's loop structure:
  - tap set: lattice points strictly inside the ellipse v^T A^{-1} v < 2p+4
    (isotropic: a^2 + b^2 < R^2 with R^2 = ampl02*(2p+4)), weight (1 - r^2)^p,
    normalized to unit sum (jp_pbfil.f90:182-218, 888-928);
  - half-sample mirror for horizontal halos (mg_bocos.f90:225-262, 340-377),
    whole-sample mirror for the vertical column (mg_filtering.f90:2072-2075);
  - interpolation P from the constants of mg_interpolate.f90:76-79 applied on
    the offset grids as in mg_generations.f90:1190-1250; with quad (lquart)
    the 3-point Lagrange interpolation through the coarse points nearest to
    each fine point, its weights computed from the point positions;
  - adjoints are plain matrix transposes.
Flat index of a (m, n, K) C-order field: (j*n + i)*K + k.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

PCOEF = np.array([-9.0, 111.0, 29.0, -3.0]) / 128.0
QCOEF = np.array([-3.0, 29.0, 111.0, -9.0]) / 128.0


def mirror_half(q: int, n: int) -> int:
    if q < 0:
        return -1 - q
    if q >= n:
        return 2 * n - 1 - q
    return q


def mirror_whole(q: int, n: int) -> int:
    if q < 0:
        return -q
    if q >= n:
        return 2 * (n - 1) - q
    return q


def disc_taps(h: int, R2: Fraction, p: int):
    """Taps of the isotropic beta filter with exact rational arithmetic."""
    taps = []
    for b in range(-h, h + 1):
        for a in range(-h, h + 1):
            r2 = Fraction(a * a + b * b)
            if r2 < R2:
                taps.append((a, b, (1 - r2 / R2) ** p))
    s = sum(t[2] for t in taps)
    return [(a, b, float(w / s)) for a, b, w in taps]


def line_taps(h: int, R2: Fraction, p: int):
    taps = []
    for a in range(-h, h + 1):
        r2 = Fraction(a * a)
        if r2 < R2:
            taps.append((a, (1 - r2 / R2) ** p))
    s = sum(t[1] for t in taps)
    return [(a, float(w / s)) for a, w in taps]


def radial_matrix(n: int, m: int, taps) -> np.ndarray:
    """F on an (m, n) plane: gather over the taps from the mirrored field."""
    F = np.zeros((n * m, n * m))
    for j in range(m):
        for i in range(n):
            row = j * n + i
            for a, b, w in taps:
                ii = mirror_half(i + a, n)
                jj = mirror_half(j + b, m)
                F[row, jj * n + ii] += w
    return F


def ball_taps(h: int, hz: int, R2: Fraction, p: int):
    """Taps of the isotropic 3-D beta filter: lattice points strictly inside
    the ball a^2 + b^2 + c^2 < R2, weight (1 - r^2/R2)^p, unit sum."""
    taps = []
    for c in range(-hz, hz + 1):
        for b in range(-h, h + 1):
            for a in range(-h, h + 1):
                r2 = Fraction(a * a + b * b + c * c)
                if r2 < R2:
                    taps.append((a, b, c, (1 - r2 / R2) ** p))
    s = sum(t[3] for t in taps)
    return [(a, b, c, float(w / s)) for a, b, c, w in taps]


def rad3_matrix(n: int, m: int, K: int, lm: int, nv: int, taps3, taps2,
                zero_levels: bool = False) -> np.ndarray:
    """The filter of procedure 4 on an (m, n, K) generation: the first nv*lm
    entries of k are nv columns of lm levels filtered over (i, j, L) with the
    half-sample mirror in i and j and the whole-sample mirror in L (or, with
    zero_levels, zeros beyond the column ends); the rest of k by the 2-D
    filter."""
    N = n * m * K
    F = np.zeros((N, N))
    for j in range(m):
        for i in range(n):
            for k in range(K):
                row = (j * n + i) * K + k
                if k < nv * lm:
                    v, L = divmod(k, lm)
                    for a, b, c, w in taps3:
                        ii, jj = mirror_half(i + a, n), mirror_half(j + b, m)
                        if zero_levels and not 0 <= L + c < lm:
                            continue
                        kk = v * lm + mirror_whole(L + c, lm)
                        F[row, (jj * n + ii) * K + kk] += w
                else:
                    for a, b, w in taps2:
                        ii, jj = mirror_half(i + a, n), mirror_half(j + b, m)
                        F[row, (jj * n + ii) * K + k] += w
    return F


def vertical_matrix(lm: int, taps) -> np.ndarray:
    V = np.zeros((lm, lm))
    for L in range(lm):
        for a, w in taps:
            V[L, mirror_whole(L + a, lm)] += w
    return V


def interp_matrix_1d(nf: int) -> np.ndarray:
    """P along one axis: fine (nf) <- coarse (nf/2), coarse padded by 2 with the
    half-sample mirror. Fine 1-based index i uses coarse iL-1..iL+2, iL=i//2,
    with the q constants for odd i and the p constants for even i."""
    nc = nf // 2
    P = np.zeros((nf, nc))
    for i1 in range(1, nf + 1):
        iL = i1 // 2
        c = QCOEF if i1 % 2 == 1 else PCOEF
        for t in range(4):
            I1 = iL - 1 + t          # 1-based coarse index, may be -1..nc+2
            P[i1 - 1, mirror_half(I1 - 1, nc)] += c[t]
    return P


def lagrange_weights(nodes, x):
    """Lagrange interpolation weights of the nodes at the point x."""
    w = []
    for a in nodes:
        v = 1.0
        for b in nodes:
            if b != a:
                v *= (x - b) / (a - b)
        w.append(v)
    return w


def quad_interp_matrix_1d(nf: int) -> np.ndarray:
    """Quadratic P along one axis: coarse point I sits at 2I + 1/2 in fine
    units (0-based fine i at i). Fine i is interpolated through the coarse
    point u = i // 2 nearest to it and its two neighbours; coarse indices
    outside the domain are the half-sample mirror of those inside."""
    nc = nf // 2
    P = np.zeros((nf, nc))
    for i in range(nf):
        u = i // 2
        nodes = [u - 1, u, u + 1]
        w = lagrange_weights([2.0 * c + 0.5 for c in nodes], float(i))
        for c, wc in zip(nodes, w):
            P[i, mirror_half(c, nc)] += wc
    return P


class DenseMGBF:
    """Dense matrices of every stage for one tiny problem."""

    def __init__(self, nx, ny, K, lm, nvcol, hx, hz, p, gm, R2, R2v,
                 weights, l_filt_g1=True, l_vertical_filter=True, wfield=None, fast=False,
                 quad=False, taps2d=None, final_vft=False, rad3=None, taps3d=None):
        """rad3 = (R2 of the 3-D ball, hz): procedure 4, whose filter couples
        the levels of each 3-D variable and which has no vertical filter.
        taps3d: procedure 5 instead, the given (a, b, c, w) taps over (i, j, L)
        with zeros beyond the column ends."""
        self.nx, self.ny, self.K, self.gm = nx, ny, K, gm
        self.final_vft = final_vft
        if taps2d is not None:
            taps = list(taps2d)
        elif fast:
            # separable: the outer product of two 1-D beta filters whose
            # squared reach is R2 (offsets a with a^2 < R2)
            lt = line_taps(hx, Fraction(R2), p)
            taps = [(a, b, wa * wb) for b, wb in lt for a, wa in lt]
        else:
            taps = disc_taps(hx, Fraction(R2), p)
        ltaps = line_taps(hz, Fraction(R2v), p)
        IK = np.eye(K)
        # vertical: block diagonal over the nvcol columns, identity on the rest
        Vcol = np.eye(K)
        V1 = vertical_matrix(lm, ltaps)
        for c in range(nvcol):
            Vcol[c * lm:(c + 1) * lm, c * lm:(c + 1) * lm] = V1
        self.Vf = np.kron(np.eye(nx * ny), Vcol) if l_vertical_filter else np.eye(nx * ny * K)
        self.F, self.P, self.A = [], [], []
        if rad3 is not None or taps3d is not None:
            self.Vf = np.eye(nx * ny * K)
            t3 = list(taps3d) if taps3d is not None else ball_taps(hx, rad3[1], Fraction(rad3[0]), p)
        for g in range(gm):
            n, m = nx >> g, ny >> g
            if rad3 is not None or taps3d is not None:
                self.F.append(rad3_matrix(n, m, K, lm, nvcol, t3, taps, zero_levels=taps3d is not None))
            else:
                self.F.append(np.kron(radial_matrix(n, m, taps), IK))
            if wfield is not None and wfield[g] is not None:
                self.A.append(np.diag(np.asarray(wfield[g], dtype=np.float64).ravel()))
            else:
                self.A.append(weights[g] * np.eye(n * m * K))
            if g > 0:
                # P: generation g -> generation g-1
                interp = quad_interp_matrix_1d if quad else interp_matrix_1d
                Px = interp(nx >> (g - 1))
                Py = interp(ny >> (g - 1))
                self.P.append(np.kron(np.kron(Py, Px), IK))
        if not l_filt_g1:
            N1 = nx * ny * K
            self.F[0] = np.eye(N1)
            self.A[0] = np.eye(N1)
        self.l_filt_g1 = l_filt_g1

    def chain(self, g):
        """Matrix taking generation g to generation 1 (product of P's)."""
        M = np.eye(self.F[0].shape[0])
        for q in range(g):
            M = M @ self.P[q]
        return M

    def B(self) -> np.ndarray:
        S = 0
        for g in range(self.gm):
            C = self.chain(g)
            S = S + C @ self.F[g] @ self.A[g] @ self.F[g].T @ C.T
        return (self.Vf.T if self.final_vft else self.Vf) @ S @ self.Vf.T

    def Wt(self) -> np.ndarray:
        """Stacked W^T: rows are the generations of the multigrid vector."""
        blocks = []
        for g in range(self.gm):
            C = self.chain(g)
            blocks.append(self.F[g].T @ C.T @ self.Vf.T)
        return np.vstack(blocks)
