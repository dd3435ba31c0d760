"""Analysis <-> filter transfer on the GPU against the FP64 oracle
(src.mgbf/mg_transfer.f90:58-464): the 4-point, linear and quadratic
horizontal interpolations between a coarser filter grid and the analysis
grid, the linear vertical map of the _all2 routines (odd lm_a <= 2 lm - 1)
and the linear or lwq map of the _all routines, their transposes, and the
driver's whole sequence on an analysis field."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O
from standalone_mgbf_loc_amd import mgbf as M

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")
TOL = 1e-5


def relerr(y, ref):
    y = y.detach().cpu().double().numpy().reshape(ref.shape)
    return float(np.max(np.abs(y - ref)) / max(np.max(np.abs(ref)), 1e-300))


def assert_adjoint(op, a, w, v, b):
    # <F2A v, w> = <v, A2F w> up to the FP32 rounding of a = F2A v and
    # b = A2F w: elementwise errors of about 1e-7 relative, independent
    # across points, move a dot product by about 1e-7 |a * w|_2
    lhs, rhs = op.dot(a, w), op.dot(v, b)
    scale = float((a.double() * w.double()).norm() + (v.double() * b.double()).norm())
    assert abs(lhs - rhs) <= 1e-6 * scale, (lhs, rhs, scale)


def setup(nx, ny, lm, n_ens, na, ma, lm_a, gm=3, h=4, **kw):
    cfg = M.Config.for_grid(nx, ny, lm, n_ens=n_ens, gm=gm, h=h, hz=2, weights=[1.0, 0.5, 0.25],
                            nm0=na, mm0=ma, lm_a=lm_a, **kw)
    op = M.MGBF(cfg)
    d = op.dims
    P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=lm, nvcol=d.km3_all, hx=h, hy=h, hz=2, p=2,
                       gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02, weights=[1.0, 0.5, 0.25])
    return cfg, op, P


# (filter nx, ny, lm, members, analysis na, ma, lm_a): ratios 1.5 and 2, with
# and without the vertical map, one and two PE tiles per axis
CASES = [(32, 24, 4, 2, 48, 36, 7), (32, 32, 6, 1, 64, 64, 6), (48, 32, 3, 3, 72, 48, 5),
         (32, 24, 5, 1, 32, 24, 9)]


@pytest.mark.parametrize("nx,ny,lm,n_ens,na,ma,lm_a", CASES)
def test_transfer_matches_oracle_and_is_its_own_transpose(nx, ny, lm, n_ens, na, ma, lm_a):
    cfg, op, P = setup(nx, ny, lm, n_ens, na, ma, lm_a)
    try:
        d = op.dims
        assert (d.na, d.ma, d.km_a_all) == (na, ma, n_ens * lm_a)
        rng = np.random.default_rng(nx + na + lm_a)
        vf = rng.standard_normal(op.field_shape()).astype(np.float32)
        wa = rng.standard_normal(d.anal_shape()).astype(np.float32)
        a = op.filt_to_anal(torch.from_numpy(vf).to(DEV))
        b = op.anal_to_filt(torch.from_numpy(wa).to(DEV))
        torch.cuda.synchronize()
        args = (P, na, ma, cfg.nxPE, cfg.nyPE, lm_a)
        assert relerr(a, O.filt_to_anal(*args, vf.astype(np.float64))) <= TOL
        assert relerr(b, O.anal_to_filt(*args, wa.astype(np.float64))) <= TOL
        assert_adjoint(op, a, torch.from_numpy(wa).to(DEV), torch.from_numpy(vf).to(DEV), b)
    finally:
        op.close()


def test_driver_sequence_on_an_analysis_field():
    # anal_to_filt_all2 -> filtering_procedure -> filt_to_anal_all2
    # (RBETA_TEST.f90:34-126) on a 1.5x finer analysis grid with 2 lm - 1 levels
    cfg, op, P = setup(32, 24, 4, 2, 48, 36, 7)
    try:
        d = op.dims
        w = np.random.default_rng(21).standard_normal(d.anal_shape()).astype(np.float32)
        y = op.localize_analysis(torch.from_numpy(w).to(DEV))
        torch.cuda.synchronize()
        args = (P, 48, 36, cfg.nxPE, cfg.nyPE, 7)
        ref = O.filt_to_anal(*args, O.wawt(P, O.anal_to_filt(*args, w.astype(np.float64))))
        assert relerr(y, ref) <= TOL
    finally:
        op.close()


def test_matching_grids_make_the_transfer_a_copy():
    cfg, op, P = setup(32, 32, 4, 2, 32, 32, 4)
    try:
        x = torch.randn(op.field_shape(), device=DEV)
        assert torch.equal(op.filt_to_anal(x), x)
        assert torch.equal(op.anal_to_filt(x), x)
    finally:
        op.close()


# horizontal scheme flags, and the routine family with its vertical map
SCHEMES = {O.LSQR: {}, O.LIN: {"l_lin_horizontal": True}, O.QUAD: {"l_quad_horizontal": True}}
FAMILIES = {"all2": (M.OP_F2A, M.OP_A2F, O.VLIN, {}),
            "all_lin": (M.OP_F2A_ALL, M.OP_A2F_ALL, O.VLIN, {"l_new_map": False, "l_lin_vertical": True}),
            "all_lwq": (M.OP_F2A_ALL, M.OP_A2F_ALL, O.VLWQ, {"l_new_map": False})}


@pytest.mark.parametrize("scheme", [O.LSQR, O.LIN, O.QUAD])
@pytest.mark.parametrize("family,grid", [
    ("all2", (32, 24, 4, 2, 48, 36, 7)), ("all2", (32, 32, 5, 1, 64, 64, 7)),
    ("all_lin", (32, 24, 4, 2, 48, 36, 7)), ("all_lwq", (32, 24, 4, 2, 48, 36, 7)),
    ("all_lwq", (32, 32, 5, 1, 64, 64, 12)), ("all_lwq", (48, 32, 3, 3, 72, 48, 3)),
    ("all_lwq", (32, 32, 3, 1, 64, 64, 40))])   # adjoint windows of 40 levels
def test_every_transfer_variant_matches_oracle(scheme, family, grid):
    nx, ny, lm, n_ens, na, ma, lm_a = grid
    f2a, a2f, vmap, flags = FAMILIES[family]
    print(f"transfer scheme={scheme} family={family} grid={grid}", flush=True)
    cfg, op, P = setup(nx, ny, lm, n_ens, na, ma, lm_a, **SCHEMES[scheme], **flags)
    try:
        d = op.dims
        rng = np.random.default_rng(scheme + 10 * lm_a + na)
        vf = rng.standard_normal(op.field_shape()).astype(np.float32)
        wa = rng.standard_normal(d.anal_shape()).astype(np.float32)
        vd, wd = torch.from_numpy(vf).to(DEV), torch.from_numpy(wa).to(DEV)
        a = op.apply(f2a, vd)
        b = op.apply(a2f, wd)
        torch.cuda.synchronize()
        args = (P, na, ma, cfg.nxPE, cfg.nyPE, lm_a)
        assert relerr(a, O.filt_to_anal(*args, vf.astype(np.float64), scheme=scheme, vmap=vmap)) <= TOL
        assert relerr(b, O.anal_to_filt(*args, wa.astype(np.float64), scheme=scheme, vmap=vmap)) <= TOL
        assert_adjoint(op, a, wd, vd, b)
    finally:
        op.close()


def test_driver_sequence_without_the_new_map():
    # l_new_map=.false.: anal_to_filt_all -> filtering_procedure ->
    # filt_to_anal_all with the quadratic horizontal and the lwq vertical map
    cfg, op, P = setup(32, 24, 4, 2, 48, 36, 10, l_new_map=False, l_quad_horizontal=True)
    try:
        w = np.random.default_rng(22).standard_normal(op.dims.anal_shape()).astype(np.float32)
        y = op.localize_analysis(torch.from_numpy(w).to(DEV))
        torch.cuda.synchronize()
        args = (P, 48, 36, cfg.nxPE, cfg.nyPE, 10)
        kw = dict(scheme=O.QUAD, vmap=O.VLWQ)
        ref = O.filt_to_anal(*args, O.wawt(P, O.anal_to_filt(*args, w.astype(np.float64), **kw)), **kw)
        assert relerr(y, ref) <= TOL
    finally:
        op.close()


def test_the_other_routine_family_refuses_what_it_cannot_map():
    # lm_a = 10 is even: the _all2 routines' linear map would leave the top
    # analysis level unwritten, so with l_new_map=.false. the context builds
    # and only the _all routines run
    cfg, op, P = setup(32, 24, 4, 2, 48, 36, 10, l_new_map=False)
    try:
        w = torch.zeros(op.dims.anal_shape(), device=DEV)
        with pytest.raises(M.MGBFError, match="odd lm_a"):
            op.anal_to_filt(w)
        op.anal_to_filt_all(w)
        torch.cuda.synchronize()
    finally:
        op.close()


@pytest.mark.parametrize("scheme,grid", [(O.LSQR, (32, 24, 4, 2, 48, 36, 7)), (O.LIN, (32, 32, 5, 1, 64, 64, 5)),
                                         (O.QUAD, (48, 32, 3, 3, 72, 48, 5)), (O.LSQR, (32, 32, 6, 1, 96, 96, 6))])
def test_tiled_transfer_matches_the_gather(scheme, grid):
    # the 8 x 8 tile kernel (dense tile blocks) against the per-point gather
    # (MGBF_XFER_GATHER=1) and the oracle, both directions; ratios 1.5, 2, 3
    nx, ny, lm, n_ens, na, ma, lm_a = grid
    out = {}
    for gather in (False, True):
        os.environ["MGBF_XFER_GATHER"] = "1" if gather else "0"
        try:
            cfg, op, P = setup(nx, ny, lm, n_ens, na, ma, lm_a, **SCHEMES[scheme])
        finally:
            os.environ.pop("MGBF_XFER_GATHER", None)
        try:
            d = op.dims
            rng = np.random.default_rng(31 + na)
            vf = rng.standard_normal(op.field_shape()).astype(np.float32)
            wa = rng.standard_normal(d.anal_shape()).astype(np.float32)
            a = op.filt_to_anal(torch.from_numpy(vf).to(DEV))
            b = op.anal_to_filt(torch.from_numpy(wa).to(DEV))
            torch.cuda.synchronize()
            args = (P, na, ma, cfg.nxPE, cfg.nyPE, lm_a)
            assert relerr(a, O.filt_to_anal(*args, vf.astype(np.float64), scheme=scheme)) <= TOL
            assert relerr(b, O.anal_to_filt(*args, wa.astype(np.float64), scheme=scheme)) <= TOL
            out[gather] = (a.cpu(), b.cpu())
        finally:
            op.close()
    assert relerr(out[False][0], out[True][0].double().numpy()) <= 1e-6
    assert relerr(out[False][1], out[True][1].double().numpy()) <= 1e-6


@pytest.mark.parametrize("lm,n_ens,lm_a,family", [
    (64, 2, 127, "all2"), (32, 3, 63, "all2"), (64, 2, 127, "all_lin"), (32, 2, 41, "all_lwq"),
    (48, 2, 95, "all2"), (96, 1, 191, "all2")])
def test_fused_vertical_map_matches_the_split_passes(lm, n_ens, lm_a, family):
    # the vertical map inside the tile kernel (windows of <= 4 levels inside
    # 64-level chunks of the filter stack: lm = 64 and 32, several chunks)
    # against the map as a pass of its own (MGBF_XFER_SPLIT=1) and the
    # oracle; lm = 48 and 96 have windows that cross a chunk and keep the
    # split passes either way
    f2a, a2f, vmap, flags = FAMILIES[family]
    nx, ny, na, ma = 32, 24, 48, 36
    print(f"fused vertical map lm={lm} n_ens={n_ens} lm_a={lm_a} family={family}", flush=True)
    out = {}
    for split in (False, True):
        os.environ["MGBF_XFER_SPLIT"] = "1" if split else "0"
        try:
            cfg, op, P = setup(nx, ny, lm, n_ens, na, ma, lm_a, **flags)
        finally:
            os.environ.pop("MGBF_XFER_SPLIT", None)
        try:
            d = op.dims
            rng = np.random.default_rng(lm + lm_a)
            vf = rng.standard_normal(op.field_shape()).astype(np.float32)
            wa = rng.standard_normal(d.anal_shape()).astype(np.float32)
            vd, wd = torch.from_numpy(vf).to(DEV), torch.from_numpy(wa).to(DEV)
            a = op.apply(f2a, vd)
            b = op.apply(a2f, wd)
            torch.cuda.synchronize()
            args = (P, na, ma, cfg.nxPE, cfg.nyPE, lm_a)
            assert relerr(a, O.filt_to_anal(*args, vf.astype(np.float64), vmap=vmap)) <= TOL
            assert relerr(b, O.anal_to_filt(*args, wa.astype(np.float64), vmap=vmap)) <= TOL
            assert_adjoint(op, a, wd, vd, b)
            out[split] = (a.cpu(), b.cpu())
        finally:
            op.close()
    assert relerr(out[False][0], out[True][0].double().numpy()) <= 1e-6
    assert relerr(out[False][1], out[True][1].double().numpy()) <= 1e-6


def seeded_transfer(seed):
    """A transfer configuration drawn from a seed: filter grid 32-64 points
    a side, analysis grid 1-2 times finer per axis (ratios 1, 1.25, 1.5, 2),
    3-8 filter levels and 1-3 members, the routine family with a vertical
    map it accepts (at least lm analysis levels; the linear map: lm_a odd
    and at most 2 lm - 1, or equal to lm; lwq: up to 3 lm - 1), and a
    horizontal scheme."""
    rng = np.random.default_rng(seed)
    nx, ny = (int(v) for v in rng.choice([32, 48, 64], 2))
    na = int(nx * rng.choice([1.0, 1.25, 1.5, 2.0]))
    ma = int(ny * rng.choice([1.0, 1.25, 1.5, 2.0]))
    lm = int(rng.integers(3, 9))
    n_ens = int(rng.integers(1, 4))
    family = str(rng.choice(list(FAMILIES)))
    if family == "all_lwq":
        lm_a = int(rng.integers(lm, 3 * lm))
    else:
        lm_a = int(rng.choice([lm] + [v for v in range(lm, 2 * lm, 1) if v % 2 == 1]))
    scheme = int(rng.choice([O.LSQR, O.LIN, O.QUAD]))
    return (nx, ny, lm, n_ens, na, ma, lm_a), family, scheme


@pytest.mark.parametrize("seed", range(12))
def test_seeded_transfers_match_oracle(seed):
    # drawn grid ratios, level counts, routine families and schemes: F2A and
    # A2F against the oracle and against each other (adjointness)
    grid, family, scheme = seeded_transfer(7000 + seed)
    nx, ny, lm, n_ens, na, ma, lm_a = grid
    f2a, a2f, vmap, flags = FAMILIES[family]
    print(f"seeded transfer {seed}: scheme={scheme} family={family} grid={grid}", flush=True)
    cfg, op, P = setup(nx, ny, lm, n_ens, na, ma, lm_a, **SCHEMES[scheme], **flags)
    try:
        d = op.dims
        rng = np.random.default_rng(8000 + seed)
        vf = rng.standard_normal(op.field_shape()).astype(np.float32)
        wa = rng.standard_normal(d.anal_shape()).astype(np.float32)
        vd, wd = torch.from_numpy(vf).to(DEV), torch.from_numpy(wa).to(DEV)
        a = op.apply(f2a, vd)
        b = op.apply(a2f, wd)
        torch.cuda.synchronize()
        args = (P, na, ma, cfg.nxPE, cfg.nyPE, lm_a)
        assert relerr(a, O.filt_to_anal(*args, vf.astype(np.float64), scheme=scheme, vmap=vmap)) <= TOL
        assert relerr(b, O.anal_to_filt(*args, wa.astype(np.float64), scheme=scheme, vmap=vmap)) <= TOL
        assert_adjoint(op, a, wd, vd, b)
    finally:
        op.close()
