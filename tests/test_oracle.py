"""Pins of the FP64 oracle against what the mathematics fixes.

The reference has no tests or fixtures and cannot be compiled here (every one
of its modules has `use mpi`), so each oracle routine is pinned by checks that
a plausible mistake would fail: closed-form lattice counts, exact rational tap
weights, transposition (dot-product) tests between independently written
forward and adjoint codes, conservation laws, reproduction of constants and
linear functions, mirror symmetry of the whole operator, and a brute-force
comparison with an independent dense-matrix model on tiny grids.
"""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
import dense_model as D


def _P(**kw):
    base = dict(nx=32, ny=32, K=6, lm=3, nvcol=2, hx=4, hy=4, hz=2, p=2, gm=3,
                ampl01=Fraction(9, 7), ampl02=Fraction(25, 8))
    base.update(kw)
    base["ampl01"] = float(base["ampl01"])
    base["ampl02"] = float(base["ampl02"])
    return O.OracleParams(**base)


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("h,count2,count1", [(3, 45, 7), (4, 69, 9), (6, 145, 13),
                                             (8, 249, 17), (10, 373, 21)])
def test_tap_counts_match_lattice_count(h, count2, count1):
    # reach = halo: ampl02 = (h+1)^2/(2p+4), ampl01 = (h+1)^2/(2p+3), p = 2
    P = _P(hx=h, hy=h, hz=h, ampl02=Fraction((h + 1) ** 2, 8), ampl01=Fraction((h + 1) ** 2, 7))
    gx, gy, w = O.taps2(P)
    brute = sum(1 for a in range(-h, h + 1) for b in range(-h, h + 1) if a * a + b * b < (h + 1) ** 2)
    assert len(gx) == brute == count2
    g1, w1 = O.taps1(P)
    assert len(g1) == 2 * h + 1 == count1
    assert abs(w.sum() - 1) < 1e-14 and abs(w1.sum() - 1) < 1e-14


@pytest.mark.parametrize("h", [3, 6, 8])
def test_tap_weights_are_the_beta_function(h):
    R2 = Fraction((h + 1) ** 2)
    P = _P(hx=h, hy=h, hz=h, ampl02=R2 / 8, ampl01=R2 / 7)
    gx, gy, w = O.taps2(P)
    exact = {(a, b): wt for a, b, wt in D.disc_taps(h, R2, 2)}
    for a, b, wt in zip(gx, gy, w):
        assert abs(wt - exact[(a, b)]) < 1e-15
    g1, w1 = O.taps1(P)
    exact1 = dict(D.line_taps(h, R2, 2))
    for a, wt in zip(g1, w1):
        assert abs(wt - exact1[a]) < 1e-15
    # the reference's loop order: gy outer (ascending), gx inner (ascending)
    assert list(zip(gy, gx)) == sorted(zip(gy, gx))


def test_aspect_tensor_is_second_moment():
    # rmom2 = 1/sqrt(2p+4) scales the filter so that its second moment per axis
    # approaches the aspect amplitude (continuous limit of the beta function).
    P = _P(hx=10, hy=10, ampl02=Fraction(121, 8))
    gx, gy, w = O.taps2(P)
    m2 = float(np.sum(w * gx.astype(float) ** 2))
    assert abs(m2 / P.ampl02 - 1) < 0.05
    assert abs(np.sum(w * gx * gy)) < 1e-13


def test_anisotropic_aspect_rotates_the_filter():
    # an off-diagonal aspect term makes the taps symmetric only under (a,b)->(-a,-b)
    P = _P(hx=6, hy=6, ampl02=Fraction(3, 1), p_eps=0.3)
    gx, gy, w = O.taps2(P)
    tab = {(a, b): v for a, b, v in zip(gx, gy, w)}
    for (a, b), v in tab.items():
        assert abs(tab[(-a, -b)] - v) < 1e-15
    assert np.sum(w * gx * gy) > 0.1   # positive correlation along the diagonal


def test_reach_beyond_halo_is_refused():
    P = _P(hx=3, hy=3, ampl02=Fraction(25, 8))  # reach 5 > halo 3
    with pytest.raises(ValueError):
        O.taps2(P)


def test_radial_preserves_constants_and_adjoint_conserves_mass():
    P = _P()
    x = np.full((20, 24, 3), 2.5)
    assert np.allclose(O.radial(P, x), 2.5, atol=1e-13)
    z = _rand((20, 24, 3), 1)
    assert abs(O.radial(P, z, adjoint=True).sum() - z.sum()) < 1e-10


def test_radial_adjoint_dot_product():
    P = _P()
    x, y = _rand((20, 24, 3), 2), _rand((20, 24, 3), 3)
    lhs = np.vdot(O.radial(P, x), y)
    rhs = np.vdot(x, O.radial(P, y, adjoint=True))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_radial_is_symmetric_with_half_sample_mirror():
    # gather+pad and scatter+fold give the same matrix for a mirror-symmetric
    # disc, so F^T = F; the GPU path relies on this.
    P = _P()
    x = _rand((20, 24, 3), 4)
    assert np.max(np.abs(O.radial(P, x) - O.radial(P, x, adjoint=True))) < 1e-14


def test_pad_fold_adjoint_and_mirror_values():
    x = _rand((9, 7, 2), 5)
    xp = O.pad(x, 3)
    # half-sample mirror: X(1-t) = X(t), X(n+t) = X(n+1-t) in both directions
    assert np.array_equal(xp[3:-3, 2, :], x[:, 0, :])
    assert np.array_equal(xp[3:-3, 0, :], x[:, 2, :])
    assert np.array_equal(xp[3:-3, -1, :], x[:, -3, :])
    assert np.array_equal(xp[0, 0, :], x[2, 2, :])
    z = _rand(xp.shape, 6)
    assert abs(np.vdot(xp, z) - np.vdot(x, O.fold(z, 3))) < 1e-12


def test_prolong_restrict_adjoint():
    xc, yf = _rand((8, 10, 3), 7), _rand((16, 20, 3), 8)
    lhs = np.vdot(O.prolong(xc), yf)
    rhs = np.vdot(xc, O.restrict(yf))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_prolong_reproduces_constants_and_linear_functions():
    assert np.allclose(O.prolong(np.ones((8, 8, 2))), 1.0, atol=1e-14)
    # coarse point I sits at x = 2I - 1/2 in fine units (0-based fine i at i)
    nc = 12
    X = 2.0 * np.arange(nc) + 0.5
    coarse = np.broadcast_to(X[None, :, None], (nc, nc, 1)).copy()
    fine = O.prolong(coarse)
    xi = np.arange(2 * nc, dtype=float)
    # interior fine points (away from the mirrored 2-halo) reproduce x exactly
    assert np.allclose(fine[:, 4:-4, 0], xi[None, 4:-4], atol=1e-12)


def test_vertical_constants_and_adjoint():
    P = _P(K=7, lm=3, nvcol=2, hz=2, ampl01=Fraction(9, 7))
    x = np.full((4, 5, 7), 1.5)
    y = O.vertical(P, x)
    assert np.allclose(y, 1.5, atol=1e-14)
    a, b = _rand((4, 5, 7), 9), _rand((4, 5, 7), 10)
    lhs = np.vdot(O.vertical(P, a), b)
    rhs = np.vdot(a, O.vertical(P, b, adjoint=True))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # the entries beyond nvcol*lm pass through untouched
    assert np.array_equal(O.vertical(P, a)[..., 6], a[..., 6])
    # whole-sample mirror makes Vf non-symmetric
    assert np.max(np.abs(O.vertical(P, a) - O.vertical(P, a, adjoint=True))) > 1e-3


def _dense_for(P, wfield=None):
    return D.DenseMGBF(P.nx, P.ny, P.K, P.lm, P.nvcol, P.hx, P.hz, P.p, P.gm,
                       Fraction(P.ampl02).limit_denominator(1000) * (2 * P.p + 4),
                       Fraction(P.ampl01).limit_denominator(1000) * (2 * P.p + 3),
                       list(P.wconst()), P.l_filt_g1, P.l_vertical_filter, wfield,
                       quad=P.lquart)


def _oracle_matrix(P, fn, N):
    cols = []
    for e in range(N):
        x = np.zeros(N)
        x[e] = 1.0
        cols.append(fn(x))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("filt_g1,vert", [(True, True), (False, True), (True, False)])
def test_wawt_matches_dense_model(filt_g1, vert):
    P = _P(nx=16, ny=12, K=4, lm=2, nvcol=2, hx=3, hy=3, hz=1, gm=3,
           ampl02=Fraction(16, 8), ampl01=Fraction(4, 7), l_filt_g1=filt_g1,
           l_vertical_filter=vert, weights=[1.0, 0.7, 0.3])
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape)).ravel(), N)
    Bd = _dense_for(P).B()
    assert np.max(np.abs(B - Bd)) < 1e-13 * np.max(np.abs(Bd))
    # symmetric positive semi-definite
    assert np.max(np.abs(B - B.T)) < 1e-13
    assert np.linalg.eigvalsh(0.5 * (B + B.T)).min() > -1e-12


def test_wawt_with_per_point_weights_matches_dense_model():
    P = _P(nx=16, ny=16, K=2, lm=2, nvcol=1, hx=3, hy=3, hz=1, gm=2,
           ampl02=Fraction(16, 8), ampl01=Fraction(4, 7))
    wf = [0.5 + np.random.default_rng(11 + g).random(P.gen_shape(g)) for g in range(P.gm)]
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape), wfield=wf).ravel(), N)
    Bd = _dense_for(P, wf).B()
    assert np.max(np.abs(B - Bd)) < 1e-13 * np.max(np.abs(Bd))


def test_w_and_wt_are_adjoint_and_compose_to_wawt():
    P = _P()
    x = _rand(P.gen_shape(0), 12)
    size = sum(int(np.prod(P.gen_shape(g))) for g in range(P.gm))
    v = _rand(size, 13)
    lhs = np.vdot(O.w(P, v), x)
    rhs = np.vdot(v, O.wt(P, x))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    y1 = O.wawt(P, x)
    y2 = O.w(P, O.weight(P, O.wt(P, x)))
    assert np.max(np.abs(y1 - y2)) < 1e-13 * np.max(np.abs(y1))


def test_wt_matches_dense_model():
    P = _P(nx=16, ny=16, K=2, lm=2, nvcol=1, hx=3, hy=3, hz=1, gm=3,
           ampl02=Fraction(16, 8), ampl01=Fraction(4, 7))
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    Wt = _oracle_matrix(P, lambda x: O.wt(P, x.reshape(shape)), N)
    Wtd = _dense_for(P).Wt()
    assert np.max(np.abs(Wt - Wtd)) < 1e-14


def test_wawt_commutes_with_domain_reflections():
    # the offset grids are mirror symmetric about the domain centre, so B
    # commutes with i -> n+1-i, j -> m+1-j and (square domain) i <-> j
    P = _P(nx=32, ny=32, K=4, lm=2, nvcol=2)
    x = _rand(P.gen_shape(0), 14)
    y = O.wawt(P, x)
    assert np.allclose(O.wawt(P, x[:, ::-1, :].copy()), y[:, ::-1, :], atol=1e-13)
    assert np.allclose(O.wawt(P, x[::-1, :, :].copy()), y[::-1, :, :], atol=1e-13)
    xt = np.ascontiguousarray(x.transpose(1, 0, 2))
    assert np.allclose(O.wawt(P, xt), y.transpose(1, 0, 2), atol=1e-13)


def test_wawt_is_linear():
    P = _P()
    a, b = _rand(P.gen_shape(0), 15), _rand(P.gen_shape(0), 16)
    lhs = O.wawt(P, 2.0 * a - 3.0 * b)
    rhs = 2.0 * O.wawt(P, a) - 3.0 * O.wawt(P, b)
    assert np.max(np.abs(lhs - rhs)) < 1e-12


# ---- analysis <-> filter transfer (anal_to_filt_all2 / filt_to_anal_all2) ----

def _tparams(nx, ny, lm, nvcol, K=None):
    return O.OracleParams(nx=nx, ny=ny, K=K if K is not None else lm * nvcol, lm=lm, nvcol=nvcol,
                          hx=3, hy=3, hz=1, p=2, gm=2, ampl01=4 / 7, ampl02=2.0)


def test_transfer_adjoint_is_the_transpose():
    # <T v, w> = <v, T^T w> with the vertical 2:1 map and 2 PE tiles per axis
    P = _tparams(32, 24, 3, 2)
    rng = np.random.default_rng(11)
    vf = rng.standard_normal(P.gen_shape(0))
    wa = rng.standard_normal((36, 48, 2 * 5))
    a = O.filt_to_anal(P, 48, 36, 2, 2, 5, vf)
    b = O.anal_to_filt(P, 48, 36, 2, 2, 5, wa)
    assert abs(np.sum(a * wa) - np.sum(vf * b)) <= 1e-12 * abs(np.sum(a * wa))


def test_transfer_reproduces_quadratics_inside_and_constants_everywhere():
    # the 4-point weights blend two quadratic interpolants: exact for any
    # quadratic away from the mirrored edges; the half-sample mirror keeps a
    # constant constant up to the edges
    P = _tparams(32, 24, 1, 0, K=1)
    xf = (np.arange(32) + 0.5) * 1.5
    yf = (np.arange(24) + 0.5) * 1.5
    f = lambda x, y: 1 + 0.3 * x - 0.02 * x * x + 0.1 * y + 0.004 * x * y - 0.01 * y * y
    XF, YF = np.meshgrid(xf, yf)
    a = O.filt_to_anal(P, 48, 36, 2, 2, 1, f(XF, YF)[:, :, None])[:, :, 0]
    XA, YA = np.meshgrid(np.arange(48) + 0.5, np.arange(36) + 0.5)
    assert np.max(np.abs(a - f(XA, YA))[4:-4, 4:-4]) < 1e-12
    assert np.max(np.abs(a - f(XA, YA))) > 1e-3          # the edges are mirrored, not extrapolated
    c = O.filt_to_anal(P, 48, 36, 2, 2, 1, np.full(P.gen_shape(0), 2.5))
    assert np.max(np.abs(c - 2.5)) < 1e-13


def test_transfer_is_the_identity_on_matching_grids_and_levels():
    P = _tparams(24, 16, 4, 2)
    x = np.random.default_rng(12).standard_normal(P.gen_shape(0))
    assert np.array_equal(O.filt_to_anal(P, 24, 16, 2, 2, 4, x), x)
    assert np.array_equal(O.anal_to_filt(P, 24, 16, 2, 2, 4, x), x)


def test_vertical_transfer_is_the_two_to_one_linear_map():
    # lm -> 2 lm - 1 levels: odd analysis levels copy, even ones average
    # (l_vertical_direct_spec2)
    P = _tparams(16, 16, 4, 3)
    f = np.random.default_rng(13).standard_normal(P.gen_shape(0))
    a = O.filt_to_anal(P, 16, 16, 1, 1, 7, f)
    for e in range(3):
        col, out = f[:, :, 4 * e: 4 * e + 4], a[:, :, 7 * e: 7 * e + 7]
        assert np.allclose(out[:, :, 0::2], col, atol=1e-15)
        assert np.allclose(out[:, :, 1::2], 0.5 * (col[:, :, :-1] + col[:, :, 1:]), atol=1e-15)


def test_transfer_tiles_join_up_like_one_domain():
    # the per-tile interpolation with halos from the neighbours is the same
    # operator as on one tile covering the domain
    P = _tparams(32, 24, 1, 0, K=3)
    x = np.random.default_rng(14).standard_normal(P.gen_shape(0))
    one = O.filt_to_anal(P, 48, 36, 1, 1, 1, x)
    tiled = O.filt_to_anal(P, 48, 36, 4, 2, 1, x)
    assert np.max(np.abs(one - tiled)) < 1e-12


def _axis_matrix(scheme, nf, na):
    # the 1-D interpolation matrix of one axis (the other axis matching, so
    # exactly the identity), read off unit impulses
    P = _tparams(nf, 4, 1, 0, K=1)
    T = np.zeros((na, nf))
    for i in range(nf):
        v = np.zeros(P.gen_shape(0))
        v[:, i, 0] = 1.0
        a = O.filt_to_anal(P, na, 4, 1, 1, 1, v, scheme=scheme)
        assert np.all(a[1:] == a[0])
        T[:, i] = a[0, :, 0]
    return T


def test_linear_and_quadratic_horizontal_weights_at_ratio_two():
    # analysis points at n - 1/2, filter points at 2 i - 1: the linear
    # weights alternate (3/4, 1/4) and (1/4, 3/4); the quadratic ones through
    # the filter points at 1, 3, 5 evaluated at 1.5 are (21, 14, -3)/32. The
    # first analysis point (at 0.5) sits between the mirrored point at -1 and
    # the point at 1: weights (1/4, 3/4) on the same value.
    L = _axis_matrix(O.LIN, 8, 16)
    assert np.array_equal(L[0], np.eye(8)[0])
    assert np.array_equal(L[1, :2], [0.75, 0.25]) and np.array_equal(L[2, :2], [0.25, 0.75])
    assert np.array_equal(L[15], np.eye(8)[7])
    Q = _axis_matrix(O.QUAD, 8, 16)
    assert np.allclose(Q[1, :3], np.array([21, 14, -3]) / 32, atol=1e-15)
    assert np.allclose(Q.sum(axis=1), 1.0, atol=1e-14)
    # each analysis row has 2 (linear) and at most 3 (quadratic) neighbours
    assert max(np.count_nonzero(r) for r in L) == 2 and max(np.count_nonzero(r) for r in Q) == 3


@pytest.mark.parametrize("scheme,deg", [(O.LIN, 1), (O.QUAD, 2)])
def test_horizontal_schemes_reproduce_their_polynomials_inside(scheme, deg):
    # linear interpolation is exact for bilinear fields, the 3-point Lagrange
    # one for biquadratic fields, away from the mirrored edges; a product
    # term x^deg y^deg checks that the two axes are applied as a product
    P = _tparams(32, 24, 1, 0, K=1)
    xf = (np.arange(32) + 0.5) * 1.5
    yf = (np.arange(24) + 0.5) * 1.5
    f = lambda x, y: 1 + 0.3 * x - 0.2 * y + 0.01 * x ** deg * y ** deg + 0.02 * (x ** deg - y ** deg)
    XF, YF = np.meshgrid(xf, yf)
    a = O.filt_to_anal(P, 48, 36, 2, 2, 1, f(XF, YF)[:, :, None], scheme=scheme)[:, :, 0]
    XA, YA = np.meshgrid(np.arange(48) + 0.5, np.arange(36) + 0.5)
    err = np.abs(a - f(XA, YA))
    assert np.max(err[3:-3, 3:-3]) < 1e-10 * np.max(np.abs(a))
    # one degree more is not reproduced
    g = lambda x, y: x ** (deg + 1)
    b = O.filt_to_anal(P, 48, 36, 2, 2, 1, g(XF, YF)[:, :, None], scheme=scheme)[:, :, 0]
    assert np.max(np.abs(b - g(XA, YA))[3:-3, 3:-3]) > 1e-3
    c = O.filt_to_anal(P, 48, 36, 2, 2, 1, np.full(P.gen_shape(0), 2.5), scheme=scheme)
    assert np.max(np.abs(c - 2.5)) < 1e-13


def _lagrange_blend(y, nodes):
    # the lwq weights rebuilt from Vandermonde solves: the quadratic
    # interpolants through nodes 0..2 and 1..3 (the weights that reproduce
    # 1, y, y^2, in coordinates centred on node 1), blended linearly between
    # nodes 1 and 2
    def quad(ns):
        z = np.asarray(ns) - nodes[1]
        return np.linalg.solve(np.vander(z, 3, increasing=True).T, (y - nodes[1]) ** np.arange(3))
    l, r = quad(nodes[:3]), quad(nodes[1:])
    b = (y - nodes[1]) / (nodes[2] - nodes[1])
    return np.array([(1 - b) * l[0], (1 - b) * l[1] + b * r[0], (1 - b) * l[2] + b * r[1], b * r[2]])


@pytest.mark.parametrize("la,lm", [(65, 33), (13, 5), (10, 4), (7, 3)])
def test_lwq_vertical_coefficients(la, lm):
    # lwq_vertical_coef: analysis level n at height (n-1)(lm-1)/(la-1) on
    # filter heights 0..lm-1; inside, the blend of two quadratic
    # interpolants; at kref = 1 and lm-1 the weight of the mirrored ghost
    # level folds onto its image; the end levels copy
    c, k = O.lwq_coef(la, lm)
    assert k[0] == 1 and k[-1] == lm
    assert np.array_equal(c[0], [0, 1, 0, 0]) and np.array_equal(c[-1], [0, 1, 0, 0])
    for n in range(2, la):
        y = (n - 1) * (lm - 1) / (la - 1)
        kr = int(np.floor(y)) + 1
        assert k[n - 1] == kr
        ref = _lagrange_blend(y, np.arange(kr - 2, kr + 2, dtype=float))
        if kr == 1:
            ref[2] += ref[0]
            ref[0] = 0.0
        if kr == lm - 1:
            ref[1] += ref[3]
            ref[3] = 0.0
        assert np.allclose(c[n - 1], ref, atol=1e-13), n
        assert abs(c[n - 1].sum() - 1.0) < 1e-13


def test_lwq_vertical_map_reproduces_quadratics_and_mirrors_at_the_ends():
    # on matching horizontal grids the transfer is the column map alone:
    # quadratics in the level height are kept wherever the stencil stays
    # inside the column, and quadratics even about an end level are kept at
    # that end (the whole-sample mirror of the fold)
    lm, la, nv = 9, 25, 2
    P = _tparams(8, 8, lm, nv)
    h = np.arange(lm, dtype=float)
    ya = np.arange(la) * (lm - 1) / (la - 1)
    kref = np.floor(ya).astype(int) + 1
    inside = (kref >= 2) & (kref <= lm - 2)
    for q, ok in [(lambda z: 1 + 0.5 * z - 0.07 * z * z, inside),
                  (lambda z: 2 - 0.3 * z * z, kref <= lm - 2),
                  (lambda z: 1 + 0.2 * (z - lm + 1) ** 2, kref >= 2)]:
        f = np.broadcast_to(np.tile(q(h), nv), P.gen_shape(0)).copy()
        a = O.filt_to_anal(P, 8, 8, 1, 1, la, f, vmap=O.VLWQ)
        for e in range(nv):
            col = a[3, 5, la * e: la * (e + 1)]
            assert np.max(np.abs(col - q(ya))[ok]) < 1e-12
    # the end levels copy whatever the column
    g = np.random.default_rng(15).standard_normal(P.gen_shape(0))
    a = O.filt_to_anal(P, 8, 8, 1, 1, la, g, vmap=O.VLWQ)
    assert np.array_equal(a[:, :, 0], g[:, :, 0]) and np.array_equal(a[:, :, la - 1], g[:, :, lm - 1])


def test_linear_vertical_map_with_fewer_than_two_to_one_levels():
    # l_vertical_direct_spec with an odd lm_a < 2 lm - 1: the same
    # odd-copy / even-average rule; filter levels above (lm_a+1)/2 receive
    # nothing. (An even lm_a leaves the top analysis level unwritten in the
    # reference, so it is refused.)
    P = _tparams(8, 8, 5, 1)
    f = np.random.default_rng(16).standard_normal(P.gen_shape(0))
    a = O.filt_to_anal(P, 8, 8, 1, 1, 7, f)
    assert np.array_equal(a[:, :, 0::2], f[:, :, :4])
    assert np.allclose(a[:, :, 1::2], 0.5 * (f[:, :, :3] + f[:, :, 1:4]), atol=1e-15)
    b = O.anal_to_filt(P, 8, 8, 1, 1, 7, a)
    assert np.all(b[:, :, 4] == 0.0)


@pytest.mark.parametrize("scheme", [O.LSQR, O.LIN, O.QUAD])
@pytest.mark.parametrize("vmap,la", [(O.VLIN, 5), (O.VLWQ, 7), (O.VLIN, 3)])
def test_every_transfer_variant_has_its_adjoint(scheme, vmap, la):
    P = _tparams(32, 24, 3, 2)
    rng = np.random.default_rng(17)
    vf = rng.standard_normal(P.gen_shape(0))
    wa = rng.standard_normal((36, 48, 2 * la))
    a = O.filt_to_anal(P, 48, 36, 2, 2, la, vf, scheme=scheme, vmap=vmap)
    b = O.anal_to_filt(P, 48, 36, 2, 2, la, wa, scheme=scheme, vmap=vmap)
    assert abs(np.sum(a * wa) - np.sum(vf * b)) <= 1e-12 * np.sqrt(np.sum(a * a) * np.sum(wa * wa))


# ---- the fast (separable) filter of procedures 3 and 8 ----

def test_fast_taps_take_the_aspect_amplitude_as_the_factor():
    # filtering_fast_*: el = mg_ampl02 itself (no cholaspect), so with
    # mg_ampl02 = 2 and p = 1 the reach is sqrt(2p+3)/2 = 1.118: 3 taps with
    # weights (1 - 4 a^2/5)^1 / (7/5) = (1/7, 5/7, 1/7)
    P = O.OracleParams(nx=16, ny=16, K=1, lm=1, nvcol=0, hx=6, hy=6, hz=3, p=1, gm=2,
                       ampl01=2.0, ampl02=2.0, fast=True)
    gx, w = O.taps_fast(P)
    assert list(gx) == [-1, 0, 1]
    assert np.allclose(w, [1 / 7, 5 / 7, 1 / 7], atol=1e-15)
    P.ampl02 = 0.25   # reach 8.9 > hx: the reference stops
    with pytest.raises(ValueError):
        O.taps_fast(P)


def test_fast_filter_is_separable_and_self_adjoint():
    # on a rank-1 field f(x) g(y) the filter is the product of the 1-D
    # filters with the half-sample mirror; F (bocoy, y gather, bocox, x gather)
    # and F^T (x scatter, bocoTx, y scatter, bocoTy) are the same matrix
    P = O.OracleParams(nx=20, ny=14, K=2, lm=1, nvcol=0, hx=4, hy=4, hz=1, p=2, gm=2,
                       ampl01=4 / 7, ampl02=0.6, fast=True)
    gx, w = O.taps_fast(P)
    assert gx.max() == 4
    rng = np.random.default_rng(15)
    fx, gy = rng.standard_normal(20), rng.standard_normal(14)

    def filt1(v):
        h = int(gx.max())
        vp = np.pad(v, h, mode="symmetric")
        return np.array([sum(wt * vp[i + h + a] for a, wt in zip(gx, w)) for i in range(len(v))])

    x = np.outer(gy, fx)[:, :, None] * np.array([1.0, -2.0])
    y = O.radial(P, x)
    assert np.allclose(y, np.outer(filt1(gy), filt1(fx))[:, :, None] * np.array([1.0, -2.0]), atol=1e-13)
    z = rng.standard_normal((14, 20, 2))
    assert np.allclose(O.radial(P, z), O.radial(P, z, adjoint=True), atol=1e-13)


@pytest.mark.parametrize("filt_g1", [True, False])
def test_fast_wawt_matches_dense_model(filt_g1):
    P = _P(nx=16, ny=12, K=4, lm=2, nvcol=2, hx=3, hy=3, hz=1, gm=3,
           ampl02=Fraction(1, 1), ampl01=Fraction(4, 7), l_filt_g1=filt_g1,
           weights=[1.0, 0.7, 0.3], fast=True)
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape)).ravel(), N)
    Dm = D.DenseMGBF(P.nx, P.ny, P.K, P.lm, P.nvcol, P.hx, P.hz, P.p, P.gm,
                     Fraction(2 * P.p + 3) / Fraction(P.ampl02).limit_denominator(1000) ** 2,
                     Fraction(P.ampl01).limit_denominator(1000) * (2 * P.p + 3),
                     list(P.wconst()), P.l_filt_g1, P.l_vertical_filter, None, fast=True)
    assert np.max(np.abs(B - Dm.B())) < 1e-13



# --------------------------------------------------------------------------
# quadratic inter-generation transfer (lquart: adjoint2 / direct2)
# --------------------------------------------------------------------------

def test_quadratic_prolong_restrict_adjoint():
    xc, yf = _rand((6, 10, 3), 81), _rand((12, 20, 3), 82)
    lhs = np.vdot(O.prolong(xc, quad=True), yf)
    rhs = np.vdot(xc, O.restrict(yf, quad=True))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_quadratic_prolong_reproduces_quadratics_inside_and_constants_everywhere():
    # coarse point I sits at 2I + 1/2 in fine units: a 3-point interpolation
    # is exact for every polynomial of degree 2 away from the 1-point halo
    nc = 10
    X = 2.0 * np.arange(nc) + 0.5

    def q(x, y):
        return 0.3 * x * x - 1.1 * x * y + 0.7 * y * y + 2.0 * x - y + 5.0

    coarse = q(X[None, :], X[:, None])[:, :, None]
    fine = O.prolong(coarse, quad=True)
    xi = np.arange(2 * nc, dtype=float)
    assert np.allclose(fine[2:-2, 2:-2, 0], q(xi[None, 2:-2], xi[2:-2, None]), atol=1e-11)
    # but not for cubics (the stencil has 3 points, not 4)
    cub = (X[None, :] ** 3 + 0 * X[:, None])[:, :, None]
    assert np.abs(O.prolong(cub, quad=True)[2:-2, 2:-2, 0] - xi[None, 2:-2] ** 3).max() > 1e-3
    # the mirror halo keeps constants exact up to the edges
    assert np.allclose(O.prolong(np.ones((5, 7, 2)), quad=True), 1.0, atol=1e-14)


def test_quadratic_prolong_edges_are_the_half_sample_mirror():
    # the half-sample mirror of a coarse field c is the even extension
    # [c reversed, c, c reversed]: prolonging that on a 3x wider domain and
    # keeping its middle third (interior there) gives the edge rows here
    nc, mc = 6, 4
    c = _rand((mc, nc, 2), 83)
    ext = np.concatenate([c[:, ::-1], c, c[:, ::-1]], axis=1)
    ext = np.concatenate([ext[::-1], ext, ext[::-1]], axis=0)
    big = O.prolong(ext, quad=True)
    mid = big[2 * mc:4 * mc, 2 * nc:4 * nc]
    assert np.allclose(O.prolong(c, quad=True), mid, atol=1e-13)


def test_quadratic_transfer_differs_from_the_four_point_one():
    x = _rand((16, 16, 2), 84)
    assert np.abs(O.restrict(x, quad=True) - O.restrict(x)).max() > 1e-3


@pytest.mark.parametrize("filt_g1", [True, False])
def test_lquart_wawt_matches_dense_model(filt_g1):
    P = _P(nx=16, ny=12, K=4, lm=2, nvcol=2, hx=3, hy=3, hz=1, gm=3,
           ampl02=Fraction(16, 8), ampl01=Fraction(4, 7), l_filt_g1=filt_g1,
           weights=[1.0, 0.7, 0.3], lquart=True)
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape)).ravel(), N)
    Bd = _dense_for(P).B()
    assert np.max(np.abs(B - Bd)) < 1e-13 * np.max(np.abs(Bd))
    assert np.max(np.abs(B - B.T)) < 1e-13


def test_lquart_w_and_wt_are_adjoint():
    P = _P(nx=16, ny=16, K=3, lm=3, nvcol=1, gm=3, lquart=True)
    x = _rand(P.gen_shape(0), 85)
    vt = O.wt(P, x)
    v = _rand(vt.shape, 86)
    lhs = np.vdot(O.w(P, v), x)
    rhs = np.vdot(v, vt)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


# --------------------------------------------------------------------------
# triad line filters (filtering_lin2_*, procedures 2 and 7)
# --------------------------------------------------------------------------

def _beta(p, h):
    """Normalized beta filter (1-(i/h)^2)^p, |i| < h, as {offset: weight}."""
    i = np.arange(-h + 1, h)
    w = (1 - (i / h) ** 2) ** p
    return dict(zip(i.tolist(), (w / w.sum()).tolist()))


def _spread2(f):
    return sum(a * a * w for a, w in f.items())


def _dibeta_from_spread(p, s2):
    """The unit-sum filter with spread**2 s2 that mixes the normalized beta
    filters of two consecutive integer half-spans (found by search)."""
    for h in range(1, 7):
        lo, hi = _beta(p, h), _beta(p, h + 1)
        s_lo, s_hi = _spread2(lo), _spread2(hi)
        if s_lo <= s2 <= s_hi:
            lam = (s2 - s_lo) / (s_hi - s_lo)
            return {a: (1 - lam) * lo.get(a, 0.0) + lam * hi.get(a, 0.0) for a in hi}
    raise ValueError("spread out of range")


@pytest.mark.parametrize("p", range(1, 7))
def test_moment_tables_are_the_discrete_beta_moments(p):
    # inimomtab's continuum moments plus Euler-Maclaurin corrections are the
    # exact discrete moments of the beta filter
    bn, bs = O.momtab(p, 6)
    for h in range(1, 7):
        f = _beta(p, h)
        i = np.arange(-h + 1, h)
        w = (1 - (i / h) ** 2) ** p
        assert abs(bn[h] * w.sum() - 1) < 1e-14
        assert abs(bs[h] - _spread2(f)) < 1e-13 * max(1.0, bs[h])


@pytest.mark.parametrize("p,s2", [(2, 0.3), (2, 0.7), (2, 2.0), (1, 1.0), (3, 2.5), (6, 0.1), (4, 1.9)])
def test_dibeta_is_the_beta_mixture_with_the_prescribed_spread(p, s2):
    hs, ff = O.hstform(p, s2)
    assert not ff
    h, fil = O.blinfil(p, hs)
    full = {a: fil[abs(a)] for a in range(-h, h + 1)}
    assert abs(sum(full.values()) - 1) < 1e-14
    assert abs(_spread2(full) - s2) < 1e-13 * max(1.0, s2)
    ref = _dibeta_from_spread(p, s2)
    for a in set(full) | set(ref):
        assert abs(full.get(a, 0.0) - ref.get(a, 0.0)) < 1e-14


@pytest.mark.parametrize("a", [(1.0, 1.0, 0.0), (2.0, 1.0, 0.9), (1.0, 3.0, -1.2),
                               (5.0, 0.3, 1.1), (0.4, 0.4, -0.39), (0.2, 7.0, 1.1)])
def test_triad_resolves_the_aspect_onto_a_superbase(a):
    L, w, ff = O.triad(a)
    assert not ff
    assert (w >= -1e-14).all()
    A = sum(w[c] * np.outer(L[c], L[c]) for c in range(3))
    assert np.allclose([A[0, 0], A[1, 1], A[0, 1]], a, atol=1e-12)
    assert (L.sum(axis=0) == 0).all()
    assert abs(L[0, 0] * L[1, 1] - L[0, 1] * L[1, 0]) == 1


def test_domain_lines_take_the_cholesky_factor_as_their_aspect():
    # init_mg_line runs after cholaspect: the spread**2 of the x and y lines
    # is 1/sqrt(mg_ampl02), and the third line of the superbase is idle
    P = _P(line=True, ampl02=Fraction(1, 4), hx=4, hy=4)
    lines = [O.line(P, c) for c in range(3)]
    assert [ln[0] for ln in lines[:2]] == [(1, 0), (0, 1)]
    for g, h, fil in lines[:2]:
        full = {a: fil[abs(a)] for a in range(-h, h + 1)}
        assert abs(_spread2(full) - 2.0) < 1e-13
    g, h, fil = lines[2]
    assert h == -1 or (fil[0] == 1.0 and not fil[1:].any())
    # without the line tables every line is the identity
    P0 = _P(line=True, line_tables=False)
    x = _rand(P0.gen_shape(0), 91)
    assert np.array_equal(O.radial(P0, x), x)


def test_diagonal_line_gather_and_scatter_are_adjoint():
    # an off-diagonal aspect puts a line on a lattice diagonal; pad + gather
    # and scatter + fold are transposes of each other
    P = _P(line=True, ampl02=Fraction(1, 4), p_eps=0.6, hx=6, hy=6)
    gens = {O.line(P, c)[0] for c in range(3)}
    assert any(dx != 0 and dy != 0 for dx, dy in gens)
    x, y = _rand((12, 14, 2), 92), _rand((12, 14, 2), 93)
    lhs = np.vdot(O.radial(P, x), y)
    rhs = np.vdot(x, O.radial(P, y, adjoint=True))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


@pytest.mark.parametrize("final_vft,filt_g1,tables", [(False, True, True), (True, True, True),
                                                      (True, False, True), (False, True, False)])
def test_line_procedures_match_dense_model(final_vft, filt_g1, tables):
    P = _P(nx=16, ny=12, K=4, lm=2, nvcol=2, hx=4, hy=4, hz=1, gm=3,
           ampl02=Fraction(1, 4), ampl01=Fraction(4, 7), l_filt_g1=filt_g1,
           weights=[1.0, 0.7, 0.3], line=True, line_tables=tables, final_vft=final_vft)
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape)).ravel(), N)
    if tables:
        f = _dibeta_from_spread(P.p, 2.0)           # spread**2 = 1/sqrt(1/4)
        taps = [(a, b, f[a] * f[b]) for b in f for a in f]
    else:
        taps = [(0, 0, 1.0)]
    Dm = D.DenseMGBF(P.nx, P.ny, P.K, P.lm, P.nvcol, P.hx, P.hz, P.p, P.gm, None,
                     Fraction(P.ampl01).limit_denominator(1000) * (2 * P.p + 3),
                     list(P.wconst()), P.l_filt_g1, P.l_vertical_filter, None,
                     taps2d=taps, final_vft=final_vft)
    Bd = Dm.B()
    assert np.max(np.abs(B - Bd)) < 1e-13 * np.max(np.abs(Bd))


# --------------------------------------------------------------------------
# 3-D radial filter (filtering_rad3, procedure 4)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("ampl03,p,count", [(1.0, 2, 93), (Fraction(16, 9), 2, 251), (Fraction(4, 7), 1, 27)])
def test_ball_tap_counts_and_weights(ampl03, p, count):
    # lattice points strictly inside the ball |g|^2 < ampl03 (2p+5), weights
    # (1 - |g|^2/R^2)^p with unit sum, visited z, y, x
    P = _P(rad3=True, ampl03=float(ampl03), p=p, hx=4, hy=4, hz=4)
    gx, gy, gz, w = O.taps3(P)
    assert len(w) == count
    R2 = Fraction(ampl03) * (2 * p + 5)
    ref = [(1 - Fraction(a * a + b * b + c * c) / R2) ** p for a, b, c in zip(gx, gy, gz)]
    assert np.allclose(w, np.array([float(r) for r in ref]) / float(sum(ref)), rtol=1e-13)
    assert list(gz) == sorted(gz)


def test_rad3_filter_adjoint_constants_and_variable_split():
    P = _P(nx=16, ny=12, K=8, lm=3, nvcol=2, hx=3, hy=3, hz=2, gm=2, rad3=True, ampl03=1.0,
           ampl02=Fraction(16, 8))
    x, y = _rand((12, 16, 8), 95), _rand((12, 16, 8), 96)
    lhs = np.vdot(O.radial(P, x), y)
    rhs = np.vdot(x, O.radial(P, y, adjoint=True))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    # constants: preserved by F (the mirrors repeat interior values), not by F^T
    assert np.allclose(O.radial(P, np.ones((12, 16, 8))), 1.0, atol=1e-14)
    # the 2-D variables (k >= nvcol*lm) see only the 2-D filter
    z = np.zeros((12, 16, 8)); z[5, 7, 7] = 1.0
    out = O.radial(P, z)
    assert np.abs(out[..., :6]).max() == 0.0
    P2 = _P(nx=16, ny=12, K=1, lm=3, nvcol=0, hx=3, hy=3, hz=2, ampl02=Fraction(16, 8))
    assert np.allclose(out[..., 7:], O.radial(P2, z[..., 7:]), atol=1e-15)


@pytest.mark.parametrize("lquart", [False, True])
def test_rad3_wawt_matches_dense_model(lquart):
    P = _P(nx=16, ny=12, K=5, lm=3, nvcol=1, hx=3, hy=3, hz=2, gm=2, rad3=True, ampl03=1.0,
           ampl02=Fraction(16, 8), weights=[1.0, 0.6], lquart=lquart, l_vertical_filter=True)
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape)).ravel(), N)
    Dm = D.DenseMGBF(P.nx, P.ny, P.K, P.lm, P.nvcol, P.hx, P.hz, P.p, P.gm,
                     Fraction(P.ampl02).limit_denominator(1000) * (2 * P.p + 4),
                     Fraction(4, 7), list(P.wconst()), True, True, None, quad=lquart,
                     rad3=(Fraction(1) * (2 * P.p + 5), P.hz))
    Bd = Dm.B()
    assert np.max(np.abs(B - Bd)) < 1e-13 * np.max(np.abs(Bd))
    assert np.max(np.abs(B - B.T)) < 1e-13 * np.max(np.abs(B))


# --------------------------------------------------------------------------
# hexad line filters of the 3-D variables (filtering_lin3, procedure 5)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_hexad_resolves_the_aspect_onto_six_lines(seed):
    # sum_c w_c l_c l_c^T reproduces the tensor, w >= 0, one color is null
    rng = np.random.default_rng(200 + seed)
    M = rng.standard_normal((3, 3))
    A = M @ M.T + (0.05 + 0.5 * seed) * np.eye(3)
    a = [A[0, 0], A[1, 1], A[2, 2], A[1, 2], A[2, 0], A[0, 1]]
    L, w, ff = O.hexad(a)
    assert not ff
    assert (w >= -1e-13).all()
    R = sum(w[c] * np.outer(L[c], L[c]) for c in range(7))
    assert np.max(np.abs(R - A)) < 1e-12 * np.max(np.abs(A))
    assert sum(1 for c in range(7) if not L[c].any()) == 1


def test_hexad_lines_of_the_domain_are_the_axes_in_color_order():
    # the diagonal factor 1/sqrt(mg_ampl03) I: x, y, z lines with colors 1, 2, 3
    # and spread**2 1/sqrt(mg_ampl03); colors 4-6 carry idle lines
    P = _P(lin3=True, ampl03=0.25, hx=4, hy=4, hz=4)
    lines = O.hexlines(P)
    assert [(q, g) for q, g, _, _ in lines[:3]] == [(1, (1, 0, 0)), (2, (0, 1, 0)), (3, (0, 0, 1))]
    for q, g, h, fil in lines[:3]:
        full = {a: fil[abs(a)] for a in range(-h, h + 1)}
        assert abs(_spread2(full) - 2.0) < 1e-13
    for q, g, h, fil in lines[3:]:
        assert q in (4, 5, 6) and fil[0] == 1.0 and not fil[1:].any()


def test_lin3_levels_end_with_zeros_and_the_adjoint_is_the_transpose():
    # the level halo is mirrored once, then zeroed by the first line's
    # a = b: the z line (color 3) meets zeros beyond the column ends, and its
    # adjoint drops what it scatters there (bocoT_3d folds x and y only)
    P = _P(nx=16, ny=12, K=9, lm=4, nvcol=2, hx=4, hy=4, hz=3, gm=2, lin3=True, ampl03=0.5,
           ampl02=0.7)
    x, y = _rand((12, 16, 9), 97), _rand((12, 16, 9), 98)
    lhs = np.vdot(O.radial(P, x), y)
    rhs = np.vdot(x, O.radial(P, y, adjoint=True))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
    out = O.radial(P, np.ones((12, 16, 9)))
    fz = _dibeta_from_spread(P.p, 1 / np.sqrt(0.5))
    edge = sum(w for c, w in fz.items() if 0 <= c < 4)     # level 0 keeps taps c >= 0
    assert np.allclose(out[..., 0], edge, atol=1e-14)


@pytest.mark.parametrize("lquart", [False, True])
def test_lin3_wawt_matches_dense_model(lquart):
    P = _P(nx=16, ny=12, K=5, lm=3, nvcol=1, hx=3, hy=3, hz=2, gm=2, lin3=True, ampl03=1.5,
           ampl02=Fraction(3, 4), weights=[1.0, 0.6], lquart=lquart)
    shape = P.gen_shape(0)
    N = int(np.prod(shape))
    B = _oracle_matrix(P, lambda x: O.wawt(P, x.reshape(shape)).ravel(), N)
    f3 = _dibeta_from_spread(P.p, 1 / np.sqrt(1.5))
    f2 = _dibeta_from_spread(P.p, 1 / np.sqrt(0.75))
    taps3 = [(a, b, c, f3[a] * f3[b] * f3[c]) for c in f3 for b in f3 for a in f3]
    taps2 = [(a, b, f2[a] * f2[b]) for b in f2 for a in f2]
    Dm = D.DenseMGBF(P.nx, P.ny, P.K, P.lm, P.nvcol, P.hx, P.hz, P.p, P.gm, None,
                     Fraction(4, 7), list(P.wconst()), True, True, None, quad=lquart,
                     taps2d=taps2, taps3d=taps3)
    Bd = Dm.B()
    assert np.max(np.abs(B - Bd)) < 1e-13 * np.max(np.abs(Bd))
    assert np.max(np.abs(B - B.T)) < 1e-13 * np.max(np.abs(B))


# ---- closed form of the lsqr 4-point weights at ratio 2 ----
# def_offset_coef's constants (mg_interpolate.f90:76-79): an analysis point a
# quarter of the way from filter point i to i+1 reads i-1 .. i+2 with
# p_coef = (-9, 111, 29, -3)/128, three quarters of the way with
# q_coef = (-3, 29, 111, -9)/128; filter points beyond the edges are the
# half-sample mirror images.
P_COEF = np.array([-9, 111, 29, -3]) / 128.0
Q_COEF = np.array([-3, 29, 111, -9]) / 128.0


def _offset_matrix(nf):
    """The 2:1 interpolation of def_offset_coef on nf filter points, read
    from the constants alone: analysis n (0-based, at n + 1/2) with filter
    point i at 2i + 1."""
    T = np.zeros((2 * nf, nf))
    mirror = lambda i: -1 - i if i < 0 else (2 * nf - 1 - i if i >= nf else i)
    for n in range(2 * nf):
        if n % 2:      # at 2u + 1.5: a quarter past filter point u
            u, c = (n - 1) // 2, P_COEF
        else:          # at 2u + 0.5: three quarters past filter point u - 1
            u, c = n // 2, Q_COEF
        base = u - 1 if n % 2 else u - 2
        for t in range(4):
            T[n, mirror(base + t)] += c[t]
    return T


def _lagrange_row(nodes, x):
    w = np.ones(len(nodes))
    for a, xa in enumerate(nodes):
        for b, xb in enumerate(nodes):
            if a != b:
                w[a] *= (x - xb) / (xa - xb)
    return w


def _blend_matrix(nf, shift=0, flip=False):
    """A 4-point blend of the two 3-point quadratic interpolants through the
    filter points around each analysis point (weights t and 1 - t), on the
    mirrored grid; `shift` moves the node window, `flip` swaps the blend
    weights. Every variant reproduces quadratics (an affine blend of exact
    quadratic interpolants), so only the closed form tells them apart."""
    T = np.zeros((2 * nf, nf))
    mirror = lambda i: -1 - i if i < 0 else (2 * nf - 1 - i if i >= nf else i)
    for n in range(2 * nf):
        x = (n + 0.5 - 1.0) / 2.0            # in filter-index units
        i = int(np.floor(x)) + shift
        t = x - np.floor(x)
        nodes = [i - 1, i, i + 1, i + 2]
        left = np.append(_lagrange_row(nodes[:3], x), 0.0)
        right = np.insert(_lagrange_row(nodes[1:], x), 0, 0.0)
        w = (t * left + (1 - t) * right) if flip else ((1 - t) * left + t * right)
        for k, node in enumerate(nodes):
            T[n, mirror(node)] += w[k]
    return T


def test_lsqr_weights_at_ratio_two_are_the_offset_constants():
    # the blend of two quadratic interpolants that lsqr_mg_coef builds
    # (mg_interpolate.f90:129-191) is, at ratio 2, the 4-point offset
    # interpolation of def_offset_coef, edge rows folded by the mirror
    for nf in (8, 12):
        T = _axis_matrix(O.LSQR, nf, 2 * nf)
        assert np.max(np.abs(T - _offset_matrix(nf))) < 1e-14
    # the Lagrange blend written here from the point positions is the same
    # matrix, so the constants and the blend agree independently of the oracle
    assert np.max(np.abs(_blend_matrix(12) - _offset_matrix(12))) < 1e-14


@pytest.mark.parametrize("shift,flip", [(1, False), (-1, False), (0, True)])
def test_closed_form_rejects_a_shifted_window_or_a_flipped_blend(shift, flip):
    # planted mistakes: each variant still reproduces quadratics away from
    # the edges (what the older pins check), but the closed form fails it
    nf = 12
    bad = _blend_matrix(nf, shift=shift, flip=flip)
    xf = np.arange(nf) * 2.0 + 1.0
    xa = np.arange(2 * nf) + 0.5
    f = lambda x: 0.3 + 0.2 * x - 0.01 * x * x
    inside = slice(6, 2 * nf - 6)
    assert np.max(np.abs(bad @ f(xf) - f(xa))[inside]) < 1e-12
    assert np.max(np.abs(bad - _offset_matrix(nf))) > 1e-3


# ---- per-point aspect tensors (vrbeta2 / vrbeta2T with el, ss per point) ----

def aspect_field(n, m, ampl, dmax=0.5, seed=0):
    """A rotated, anisotropic aspect tensor varying over the domain, in the
    reference's parameters: pasp2 = ampl [[1 + p_del, p_eps], [p_eps, 1 - p_del]]
    with p_del = d cos 2t, p_eps = d sin 2t (mg_intstate.f90:383-390);
    returns (m, n, 3) = (A11, A22, A12)."""
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0, 2 * np.pi, 2)
    J, I = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    t = np.pi * (I / n + 0.7 * J / m) + ph[0]
    d = dmax * (0.5 + 0.5 * np.sin(2 * np.pi * I / n + ph[1]) * np.cos(np.pi * J / m))
    pdel, peps = d * np.cos(2 * t), d * np.sin(2 * t)
    return np.stack([ampl * (1 + pdel), ampl * (1 - pdel), ampl * peps], axis=-1)


def _aparams(n, m, K=3, h=4, p=2):
    return O.OracleParams(nx=n, ny=m, K=K, lm=K, nvcol=1, hx=h, hy=h, hz=1, p=p, gm=1,
                          ampl01=4 / 7, ampl02=1.0)


def _dense_aspect(asp, n, m, p):
    """The matrix of F from the definitions alone: output point i takes the
    lattice offsets g with r2 = g^T A_i^{-1} g / (2p+4) < 1, weights
    (1 - r2)^p normalized to sum 1, inputs at the half-sample mirror."""
    mir = lambda q, N: -1 - q if q < 0 else (2 * N - 1 - q if q >= N else q)
    Fm = np.zeros((n * m, n * m))
    for j in range(m):
        for i in range(n):
            a11, a22, a12 = asp[j, i]
            Ai = np.linalg.inv(np.array([[a11, a12], [a12, a22]]))
            taps = []
            for gy in range(-8, 9):
                for gx in range(-8, 9):
                    g = np.array([gx, gy], dtype=float)
                    r2 = g @ Ai @ g / (2 * p + 4)
                    if r2 < 1 - 1e-9:
                        taps.append((gx, gy, (1 - r2) ** p))
            s = sum(w for _, _, w in taps)
            for gx, gy, w in taps:
                Fm[i + n * j, mir(i + gx, n) + n * mir(j + gy, m)] += w / s
    return Fm


def test_aspect_field_uniform_equals_the_domain_aspect():
    # a field equal everywhere to the domain tensor of p_del / p_eps
    # reproduces the pinned single-tensor filter (rotated taps included)
    P = _aparams(20, 16)
    P.p_del, P.p_eps = 0.3, -0.25
    asp = np.broadcast_to([1.0 + 0.3, 1.0 - 0.3, -0.25], (16, 20, 3)).copy()
    x = np.random.default_rng(1).standard_normal((16, 20, 3))
    for adj in (False, True):
        assert np.max(np.abs(O.radial_aspect(P, asp, x, adj) - O.radial(P, x, adj))) < 1e-13


def test_aspect_field_adjoint_constants_and_mass():
    P = _aparams(24, 20)
    asp = aspect_field(24, 20, 1.9, seed=3)
    rng = np.random.default_rng(4)
    x, y = rng.standard_normal((2, 20, 24, 3))
    Fx, Fty = O.radial_aspect(P, asp, x), O.radial_aspect(P, asp, y, adjoint=True)
    assert abs(np.sum(Fx * y) - np.sum(x * Fty)) < 1e-12 * np.sqrt(np.sum(Fx ** 2) * np.sum(y ** 2))
    # per-point normalization: F keeps constants everywhere (the mirror
    # extends a constant), F^T keeps the sum of every k
    c = np.full((20, 24, 3), 2.5)
    assert np.max(np.abs(O.radial_aspect(P, asp, c) - 2.5)) < 1e-13
    assert np.allclose(Fty.sum(axis=(0, 1)), y.sum(axis=(0, 1)), rtol=0, atol=1e-11)
    # the line sums vary over the domain with the tensor
    ss = O.ss2_aspect(P, asp)
    assert ss.max() / ss.min() > 1.1


def test_aspect_field_matches_the_dense_model():
    n, m, p = 14, 12, 2
    P = _aparams(n, m, K=1, p=p)
    asp = aspect_field(n, m, 1.9, seed=5)
    Fm = _dense_aspect(asp, n, m, p)
    I = np.eye(n * m)
    Fo = np.stack([O.radial_aspect(P, asp, I[q].reshape(m, n, 1))[:, :, 0].ravel() for q in range(n * m)], 1)
    Ft = np.stack([O.radial_aspect(P, asp, I[q].reshape(m, n, 1), True)[:, :, 0].ravel() for q in range(n * m)], 1)
    assert np.max(np.abs(Fo - Fm)) < 1e-12
    assert np.max(np.abs(Ft - Fm.T)) < 1e-12
    assert np.max(np.abs(Fm - Fm.T)) > 1e-3      # a varying field is not symmetric


def test_aspect_field_reach_beyond_the_halo_is_refused():
    P = _aparams(12, 12, h=3)
    asp = aspect_field(12, 12, 3.0, seed=6)      # reach up to sqrt(3*1.5*8) = 6 > 3
    with pytest.raises(ValueError, match="halo"):
        O.radial_aspect(P, asp, np.zeros((12, 12, 1)))
