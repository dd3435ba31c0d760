"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerance (FP32 product vs FP64 oracle): normwise relative error
max|y - y64| / max|y64| <= 1e-5 for whole operators, 2e-6 for single stages;
adjoint dot-product mismatch |<Wv,x> - <v,W^T x>| <= 1e-6 * |<Wv,x>|, or
<= 1e-6 * |Wv| |x| (the Cauchy-Schwarz scale) where random x and v make
<Wv,x> cancel down to a small fraction of that scale.
"""
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
TOL_OP = 1e-5
TOL_STAGE = 2e-6


def relerr(y, y64):
    y = y.detach().cpu().double().numpy() if isinstance(y, torch.Tensor) else y
    return float(np.max(np.abs(y - y64)) / max(np.max(np.abs(y64)), 1e-300))


def oparams(cfg: M.Config, d: M.Dims) -> O.OracleParams:
    bckg = cfg.mgbf_proc in (1, 2, 3, 4, 5)
    return O.OracleParams(
        nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=cfg.hx, hy=cfg.hy,
        hz=cfg.hz, p=cfg.p, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
        l_filt_g1=True if bckg else bool(cfg.l_filt_g1),
        l_vertical_filter=bool(cfg.l_vertical_filter), fast=cfg.mgbf_proc in (3, 8),
        lquart=bool(cfg.lquart), line=cfg.mgbf_proc in (2, 7), line_tables=bool(cfg.mgbf_line),
        final_vft=cfg.mgbf_proc == 7, rad3=cfg.mgbf_proc == 4, ampl03=cfg.mg_ampl03,
        lin3=cfg.mgbf_proc == 5,
        weights=cfg.weights()[: d.gm] + [cfg.mg_weig5] * max(0, d.gm - 5))


def field(shape, seed):
    x = np.random.default_rng(seed).standard_normal(shape).astype(np.float32)
    return x


def to_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def sync():
    torch.cuda.synchronize()


# --------------------------------------------------------------------------
# single stages
# --------------------------------------------------------------------------

@pytest.mark.parametrize("h,K,n,gen", [(6, 64, 96, 1), (6, 40, 96, 3), (3, 130, 48, 2),
                                       (8, 64, 128, 1), (10, 24, 160, 1), (4, 70, 64, 1)])
def test_radial_stage_matches_oracle(h, K, n, gen):
    cfg = M.Config.for_grid(n, n, K, gm=3, h=h, hz=min(3, K - 1))
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = field(d.gen_shape(gen), 100 + gen)
        y = op.stage(M.STAGE_F, gen, to_dev(x))
        yt = op.stage(M.STAGE_FT, gen, to_dev(x))
        sync()
        x64 = x.astype(np.float64)
        assert relerr(y, O.radial(P, x64)) <= TOL_STAGE
        assert relerr(yt, O.radial(P, x64, adjoint=True)) <= TOL_STAGE


def test_radial_stage_on_a_rectangle_with_halo_sized_generation():
    # the coarsest generation (12 x 6) is as narrow as the halo allows
    cfg = M.Config(mg_ampl01=9 / 7, mg_ampl02=49 / 8, hx=6, hy=6, hz=2, n_ens=2, lm_a=4, lm=4,
                   gm_max=4, nm0=96, mm0=48, nxPE=8, nyPE=8, im_filt=12, jm_filt=6)
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.gm == 4 and d.gen_shape(4) == (6, 12, 8)
        P = oparams(cfg, d)
        for gen in (1, 4):
            x = field(d.gen_shape(gen), 7 + gen)
            y = op.stage(M.STAGE_F, gen, to_dev(x))
            sync()
            assert relerr(y, O.radial(P, x.astype(np.float64))) <= TOL_STAGE


@pytest.mark.parametrize("lm,n_ens,hz", [(32, 2, 3), (64, 1, 3), (16, 5, 4), (7, 3, 2)])
def test_vertical_stage_matches_oracle(lm, n_ens, hz):
    cfg = M.Config.for_grid(48, 40, lm, n_ens=n_ens, gm=2, h=4, hz=hz)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = field(d.gen_shape(1), 11)
        y = op.stage(M.STAGE_VF, 1, to_dev(x))
        yt = op.stage(M.STAGE_VFT, 1, to_dev(x))
        sync()
        x64 = x.astype(np.float64)
        assert relerr(y, O.vertical(P, x64)) <= TOL_STAGE
        assert relerr(yt, O.vertical(P, x64, adjoint=True)) <= TOL_STAGE


def test_vertical_stage_passes_2d_variables_through():
    # background procedure with km2 = 3 two-dimensional variables after the 3-D ones
    cfg = M.Config.for_grid(32, 32, 8, gm=2, h=3, hz=2, km2=3, km3=2, mgbf_proc=1)
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.km_all == 19
        P = oparams(cfg, d)
        x = field(d.gen_shape(1), 12)
        y = op.stage(M.STAGE_VF, 1, to_dev(x))
        sync()
        ref = O.vertical(P, x.astype(np.float64))
        assert relerr(y, ref) <= TOL_STAGE
        assert np.array_equal(y.cpu().numpy()[..., 16:], x[..., 16:])


@pytest.mark.parametrize("K,n,m", [(64, 96, 96), (33, 64, 48), (128, 32, 16)])
def test_transfer_stages_match_oracle(K, n, m):
    cfg = M.Config(mg_ampl01=16 / 7, mg_ampl02=16 / 8, hx=3, hy=3, hz=3, lm_a=K, lm=K,
                   gm_max=3, nm0=n, mm0=m, nxPE=4, nyPE=4, im_filt=n // 4, jm_filt=m // 4)
    with M.MGBF(cfg) as op:
        d = op.dims
        for gen in (1, 2):
            xf = field(d.gen_shape(gen), 20 + gen)
            xc = field(d.gen_shape(gen + 1), 30 + gen)
            r = op.stage(M.STAGE_R, gen, to_dev(xf))
            p = op.stage(M.STAGE_P, gen, to_dev(xc))
            sync()
            assert relerr(r, O.restrict(xf.astype(np.float64))) <= TOL_STAGE
            assert relerr(p, O.prolong(xc.astype(np.float64))) <= TOL_STAGE


@pytest.mark.parametrize("K,n,m", [(64, 96, 96), (33, 64, 48), (128, 32, 16), (7, 32, 16)])
def test_quadratic_transfer_stages_match_oracle(K, n, m):
    # lquart: direct2 / adjoint2 with the 1-point mirror halo; on the
    # smallest grids every restriction column tile touches an edge
    cfg = M.Config(mg_ampl01=16 / 7, mg_ampl02=16 / 8, hx=3, hy=3, hz=3, lm_a=K, lm=K,
                   gm_max=3, nm0=n, mm0=m, nxPE=4, nyPE=4, im_filt=n // 4, jm_filt=m // 4,
                   lquart=True)
    with M.MGBF(cfg) as op:
        d = op.dims
        for gen in (1, 2):
            xf = field(d.gen_shape(gen), 120 + gen)
            xc = field(d.gen_shape(gen + 1), 130 + gen)
            r = op.stage(M.STAGE_R, gen, to_dev(xf))
            p = op.stage(M.STAGE_P, gen, to_dev(xc))
            sync()
            assert relerr(r, O.restrict(xf.astype(np.float64), quad=True)) <= TOL_STAGE
            assert relerr(p, O.prolong(xc.astype(np.float64), quad=True)) <= TOL_STAGE


# --------------------------------------------------------------------------
# whole operator
# --------------------------------------------------------------------------

def impulses_plus_noise(shape, seed_noise, seed_imp, count=16):
    x = field(shape, seed_noise)
    rng = np.random.default_rng(seed_imp)
    m, n, K = shape
    for _ in range(count):
        x[rng.integers(m), rng.integers(n), rng.integers(K)] += 1.0
    return x


def test_wawt_192x192x32_four_generations_matches_oracle():
    # 192x192 grid, 32 levels, 4 generations, radial half-width 6, vertical 3,
    # p=2, weights [1, .5, .25, .125]; N(0,1) plus 16 unit impulses
    cfg = M.Config.for_grid(192, 192, 32, gm=4, h=6, hz=3, weights=[1, .5, .25, .125])
    with M.MGBF(cfg) as op:
        d = op.dims
        x = impulses_plus_noise(d.gen_shape(1), 1, 2)
        y = op.filtering_procedure(to_dev(x))
        sync()
        ref = O.wawt(oparams(cfg, d), x.astype(np.float64))
        assert relerr(y, ref) <= TOL_OP


@pytest.mark.parametrize("n_ens,lm,gm,h,filt_g1", [(3, 16, 3, 6, True), (2, 20, 4, 4, False),
                                                  (1, 64, 2, 8, True), (2, 24, 4, 3, True)])
def test_wawt_ensemble_matches_oracle(n_ens, lm, gm, h, filt_g1):
    cfg = M.Config.for_grid(96, 64, lm, n_ens=n_ens, gm=gm, h=h, hz=3,
                            weights=[1.0, 0.7, 0.4, 0.2], l_filt_g1=filt_g1)
    with M.MGBF(cfg) as op:
        d = op.dims
        x = field(d.gen_shape(1), 5 + n_ens)
        y = op.filtering_procedure(to_dev(x))
        sync()
        assert relerr(y, O.wawt(oparams(cfg, d), x.astype(np.float64))) <= TOL_OP


@pytest.mark.parametrize("lm,n_ens,hz,exact", [(32, 2, 3, False), (16, 4, 2, False), (64, 2, 4, False),
                                            (32, 2, 3, True), (64, 1, 4, True), (8, 8, 3, True)])
def test_vertical_band_fused_into_the_last_radial_pass(monkeypatch, lm, n_ens, hz, exact):
    # whole columns per wavefront (lm divides 64, K % 64 == 0): the vertical
    # band (Vf Vf^T, half-width 2hz, or Vf alone in the reference's order)
    # rides in the epilogue of the last radial pass, across the lanes
    if exact:
        monkeypatch.setenv("MGBF_EXACT_ORDER", "1")
    cfg = M.Config.for_grid(64, 64, lm, n_ens=n_ens, gm=3, h=4, hz=hz, weights=[1.0, 0.6, 0.3])
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.km_all % 64 == 0 and 64 % lm == 0
        x = field(d.gen_shape(1), 50 + lm + hz)
        y = op.filtering_procedure(to_dev(x))
        sync()
        assert relerr(y, O.wawt(oparams(cfg, d), x.astype(np.float64))) <= TOL_OP


@pytest.mark.parametrize("fuse", ["0", "1", "2", "3"])
def test_every_fusion_level_matches_oracle(monkeypatch, fuse):
    # how much rides in the radial passes (nothing; the prolongation; and the
    # vertical band; and the restriction) changes launches, not results
    monkeypatch.setenv("MGBF_FUSE", fuse)
    cfg = M.Config.for_grid(128, 96, 32, n_ens=2, gm=3, h=6, hz=3, weights=[1.0, 0.6, 0.3])
    with M.MGBF(cfg) as op:
        d = op.dims
        x = field(d.gen_shape(1), 61)
        P = oparams(cfg, d)
        y = op.filtering_procedure(to_dev(x))
        v = op.apply(M.OP_WT, to_dev(x))
        sync()
        assert relerr(y, O.wawt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(v, O.wt(P, x.astype(np.float64))) <= TOL_OP


@pytest.mark.parametrize("case", ["run_namelist", "wide", "filt_g1_off"])
def test_fast_ensemble_procedure_matches_oracle(case):
    # filtering_fast_ens (mgbf_proc = 8): the horizontal filter is the product
    # of 1-D beta filters with el = mg_ampl02 as given. "run_namelist" takes
    # the reference's RUN/mgbf.nml values (mg_ampl01 = mg_ampl02 = 2,
    # p = 1.3 read as 1, hx = 6, hz = 10): a 3-point filter per axis
    kw = dict(n_ens=3, gm=3, h=6, hz=10, p=1, mgbf_proc=8, weights=[1.0, 0.5, 0.25])
    if case == "filt_g1_off":
        kw["l_filt_g1"] = False
    cfg = M.Config.for_grid(96, 64, 16, **kw)
    cfg.mg_ampl01 = 2.0
    cfg.mg_ampl02 = 0.45 if case == "wide" else 2.0
    if case == "wide":
        cfg.p = 2     # reach sqrt(7)/0.45 = 5.9: 11 taps per axis
        cfg.mg_ampl03 = 49 / 9     # keeps the 3-D normalization inside the halo at p = 2
    with M.MGBF(cfg) as op:
        d = op.dims
        x = field(d.gen_shape(1), 71)
        y = op.filtering_procedure(to_dev(x))
        v = op.apply(M.OP_WT, to_dev(x))
        sync()
        P = oparams(cfg, d)
        assert relerr(y, O.wawt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(v, O.wt(P, x.astype(np.float64))) <= TOL_OP


def test_fast_background_procedure_with_2d_variables():
    # filtering_fast_bckg (mgbf_proc = 3): one field, 2-D variables after the
    # 3-D ones, generation 1 always filtered
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=4, hz=2, km2=3, km3=2, mgbf_proc=3)
    cfg.mg_ampl02 = 0.8
    with M.MGBF(cfg) as op:
        d = op.dims
        x = field(d.gen_shape(1), 72)
        y = op.filtering_procedure(to_dev(x))
        sync()
        assert relerr(y, O.wawt(oparams(cfg, d), x.astype(np.float64))) <= TOL_OP


def test_wawt_background_procedure_with_2d_variables():
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=4, hz=2, km2=3, km3=2, mgbf_proc=1)
    with M.MGBF(cfg) as op:
        d = op.dims
        x = field(d.gen_shape(1), 40)
        y = op.filtering_procedure(to_dev(x))
        sync()
        assert relerr(y, O.wawt(oparams(cfg, d), x.astype(np.float64))) <= TOL_OP


def test_wawt_per_point_weights_match_oracle():
    cfg = M.Config.for_grid(64, 64, 16, n_ens=2, gm=3, h=6, hz=3)
    with M.MGBF(cfg) as op:
        d = op.dims
        rng = np.random.default_rng(9)
        wf = [(0.5 + rng.random(d.gen_shape(g))).astype(np.float32) for g in range(1, d.gm + 1)]
        keep = [to_dev(w) for w in wf]
        for g, w in enumerate(keep, start=1):
            op.set_weights(g, field=w)
        x = field(d.gen_shape(1), 41)
        y = op.filtering_procedure(to_dev(x))
        sync()
        ref = O.wawt(oparams(cfg, d), x.astype(np.float64), wfield=[w.astype(np.float64) for w in wf])
        assert relerr(y, ref) <= TOL_OP


def test_wawt_in_place_and_linear_and_symmetric():
    cfg = M.Config.for_grid(128, 128, 32, gm=4, h=6, hz=3)
    with M.MGBF(cfg) as op:
        d = op.dims
        a = to_dev(field(d.gen_shape(1), 50))
        b = to_dev(field(d.gen_shape(1), 51))
        Ba = op.filtering_procedure(a)
        Bb = op.filtering_procedure(b)
        c = (2 * a - 3 * b).contiguous()
        Bc = op.filtering_procedure(c, out=c)      # in place
        sync()
        lin = (Bc - (2 * Ba - 3 * Bb)).abs().max().item() / Bc.abs().max().item()
        assert lin < 1e-5
        lhs, rhs = op.dot(a, Bb), op.dot(Ba, b)
        assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-30) + 1e-9


def test_w_and_wt_match_oracle_and_pass_the_dot_product_test():
    cfg = M.Config.for_grid(96, 96, 24, n_ens=2, gm=3, h=6, hz=3)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = field(d.gen_shape(1), 60)
        v = np.random.default_rng(61).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        assert relerr(wt, O.wt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs = op.dot(w, xd)
        rhs = op.dot(vd, wt)
        assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
        # B = (W A^1/2)(W A^1/2)^T reproduces W A W^T
        s = op.apply(M.OP_BSQRTT, xd)
        y2 = op.apply(M.OP_BSQRT, s)
        y1 = op.filtering_procedure(xd)
        sync()
        assert (y1 - y2).abs().max().item() <= 1e-5 * y1.abs().max().item()


def seeded_configuration(seed):
    """A radial configuration drawn from a seed: background or ensemble
    procedure, 2-4 generations, radial half-width 3-8, vertical 1-4, p = 1
    or 2, 8-64 levels, 1-3 members, either transfer stencil, l_filt_g1 on or
    off, random generation weights; grids of 64-160 points a side (every
    generation at least as wide as the halo)."""
    rng = np.random.default_rng(seed)
    proc = int(rng.choice([1, 6]))
    gm = int(rng.integers(2, 5))
    h = int(rng.integers(3, 9))
    hz = int(rng.integers(1, 5))
    p = int(rng.choice([1, 2]))
    lm = int(rng.choice([8, 16, 32, 64]))
    lm = max(lm, 2 * hz + 2)
    n_ens = 1 if proc == 1 else int(rng.integers(1, 4))
    pe = {2: 2, 3: 4, 4: 8}[gm]
    unit = pe * 2 ** max(0, gm - 1 - int(np.log2(pe)))
    nx = unit * int(np.ceil(64 / unit)) + unit * int(rng.integers(0, 3))
    ny = unit * int(np.ceil(64 / unit)) + unit * int(rng.integers(0, 3))
    weights = [float(w) for w in rng.uniform(0.1, 1.0, 4)]
    kw = dict(mgbf_proc=proc, lquart=bool(rng.integers(0, 2)))
    if proc == 6:
        kw["l_filt_g1"] = bool(rng.integers(0, 2))
    cfg = M.Config.for_grid(nx, ny, lm, n_ens=n_ens, gm=gm, h=h, hz=hz, p=p, weights=weights, **kw)
    return cfg, bool(rng.integers(0, 2))


@pytest.mark.parametrize("seed", range(12))
def test_seeded_configurations_match_oracle(seed):
    # combinations the named cases do not pin one by one, each with the
    # default tiles or the autotuner's: W A W^T, W and W^T against the
    # oracle, and the W / W^T dot-product test
    cfg, tune = seeded_configuration(1000 + seed)
    print(f"seeded configuration {seed}: proc {cfg.mgbf_proc}, {cfg.nm0}x{cfg.mm0}x{cfg.lm} x{cfg.n_ens}, "
          f"gm {cfg.gm_max}, h {cfg.hx}, hz {cfg.hz}, p {cfg.p}, lquart {cfg.lquart}, "
          f"l_filt_g1 {getattr(cfg, 'l_filt_g1', True)}, tuned {tune}", flush=True)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        if tune:
            op.autotune()
        x = field(d.gen_shape(1), 2000 + seed)
        v = np.random.default_rng(3000 + seed).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        y = op.filtering_procedure(xd)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        x64 = x.astype(np.float64)
        assert relerr(y, O.wawt(P, x64)) <= TOL_OP
        assert relerr(wt, O.wt(P, x64)) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs, rhs = op.dot(w, xd), op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale


def seeded_other_procedure(seed):
    """A line, fast or 3-D configuration drawn from a seed: procedures 2, 3,
    7, 8 (half-width 6, mg_ampl02 from the values the line and fast factors
    allow at that halo, triad tables on or off) and 4, 5 (half-width 4,
    3-D variables with or without 2-D ones), 2-3 generations, either
    transfer stencil, random generation weights."""
    rng = np.random.default_rng(seed)
    proc = int(rng.choice([2, 3, 4, 5, 7, 8]))
    gm = int(rng.integers(2, 4))
    nx = 64 + 4 * int(rng.integers(0, 3))
    ny = 64 + 4 * int(rng.integers(0, 3))
    weights = [float(w) for w in rng.uniform(0.1, 1.0, 3)]
    kw = dict(mgbf_proc=proc, lquart=bool(rng.integers(0, 2)), weights=weights)
    if proc in (4, 5):
        lm = int(rng.choice([8, 12, 16]))
        hz = int(rng.choice([2, 3]))
        kw.update(km3=int(rng.integers(1, 4)), km2=int(rng.integers(0, 4)))
        if proc == 5:
            kw["mgbf_line"] = bool(rng.integers(0, 2))
        cfg = M.Config.for_grid(nx, ny, lm, gm=gm, h=4, hz=hz, **kw)
        return cfg
    lm = int(rng.choice([8, 16]))
    hz = int(rng.choice([2, 3]))
    if proc in (2, 3):
        kw.update(km3=int(rng.integers(1, 3)), km2=int(rng.integers(0, 4)))
        n_ens = 1
    else:
        n_ens = int(rng.integers(1, 3))
        kw["l_filt_g1"] = bool(rng.integers(0, 2))
    if proc in (2, 7):
        kw["mgbf_line"] = bool(rng.integers(0, 2))
    cfg = M.Config.for_grid(nx, ny, lm, n_ens=n_ens, gm=gm, h=6, hz=hz, **kw)
    cfg.mg_ampl02 = float(rng.choice([0.5, 0.8, 2.0] if proc in (2, 7) else [0.45, 1.0, 2.0]))
    return cfg


@pytest.mark.parametrize("seed", range(12))
def test_seeded_other_procedures_match_oracle(seed):
    # the line, fast and 3-D procedures on drawn configurations: W A W^T, W
    # and W^T against the oracle, and the W / W^T dot-product test
    cfg = seeded_other_procedure(4000 + seed)
    print(f"seeded procedure {seed}: proc {cfg.mgbf_proc}, {cfg.nm0}x{cfg.mm0}x{cfg.lm} x{cfg.n_ens}, "
          f"km3 {cfg.km3} km2 {cfg.km2}, gm {cfg.gm_max}, hz {cfg.hz}, ampl02 {cfg.mg_ampl02:.2f}, "
          f"line {cfg.mgbf_line}, lquart {cfg.lquart}", flush=True)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = field(d.gen_shape(1), 5000 + seed)
        v = np.random.default_rng(6000 + seed).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        y = op.filtering_procedure(xd)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        x64 = x.astype(np.float64)
        assert relerr(y, O.wawt(P, x64)) <= TOL_OP
        assert relerr(wt, O.wt(P, x64)) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs, rhs = op.dot(w, xd), op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale


@pytest.mark.parametrize("case", ["ens", "ens_band", "ens_filt_g1_off", "bckg", "fast", "fuse0", "tuned"])
def test_lquart_procedures_match_oracle(monkeypatch, case):
    # upsending2_ens / downsending2_ens: the quadratic transfer between
    # generations in every procedure; W and W^T stay adjoint. "tuned":
    # the autotuner's candidates include the matrix-core engine for the
    # plain passes (the quadratic transfer epilogues stay on the vector one)
    if case == "fuse0":
        monkeypatch.setenv("MGBF_FUSE", "0")
    kw = dict(n_ens=2, gm=4, h=6, hz=3, weights=[1.0, 0.6, 0.3, 0.15], lquart=True)
    if case == "ens_filt_g1_off":
        kw["l_filt_g1"] = False
    if case in ("ens_band", "tuned"):
        kw["n_ens"] = 4   # K = 64: the band rides in the last radial pass
    if case == "bckg":
        kw.update(n_ens=1, km2=3, km3=2, mgbf_proc=1)
    if case == "fast":
        kw.update(mgbf_proc=8, p=1)
    cfg = M.Config.for_grid(128, 96, 16, **kw)
    if case == "fast":
        cfg.mg_ampl02 = 1.0
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        if case == "tuned":
            op.autotune()
        x = field(d.gen_shape(1), 140)
        v = np.random.default_rng(141).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        y = op.filtering_procedure(xd)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        assert relerr(y, O.wawt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(wt, O.wt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs = op.dot(w, xd)
        rhs = op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale
        if case in ("ens", "ens_band"):
            # the quadratic restriction rides in the first radial pass of
            # every generation but the last (half-width 6): no restriction
            # launch of its own; with K % 64 == 0 the vertical band rides in
            # the last pass after the quadratic prolongation
            op.timers(True)
            op.filtering_procedure(xd)
            ph = op.read_timers()
            op.timers(False)
            assert ph["upsend"][1] == 0 and ph["hfiltT"][1] == d.gm
            assert ph["vfilt"][1] == (0 if case == "ens_band" else 1)


@pytest.mark.parametrize("case", ["ens", "ens_exact_order", "ens_no_tables", "ens_filt_g1_off",
                                  "ens_lquart", "bckg", "bckg_p1"])
def test_line_procedures_match_oracle(monkeypatch, case):
    # filtering_lin2_ens / _bckg (mgbf_proc = 7 / 2): the triad of dibeta
    # line filters, spread**2 = 1/sqrt(mg_ampl02) along x and y; the
    # ensemble procedure ends with Vf^T as the reference's does. Without
    # mgbf_line the lines are the identity
    if case == "ens_exact_order":
        monkeypatch.setenv("MGBF_EXACT_ORDER", "1")
    kw = dict(n_ens=2, gm=3, h=6, hz=3, mgbf_proc=7, mgbf_line=True, weights=[1.0, 0.5, 0.25])
    if case == "ens_no_tables":
        kw["mgbf_line"] = False
    if case == "ens_filt_g1_off":
        kw["l_filt_g1"] = False
    if case == "ens_lquart":
        kw["lquart"] = True
    if case.startswith("bckg"):
        kw.update(n_ens=1, km2=3, km3=2, mgbf_proc=2)
    if case == "bckg_p1":
        kw["p"] = 1
    cfg = M.Config.for_grid(96, 64, 16, **kw)
    cfg.mg_ampl02 = 0.5    # the smallest the fast factors' normalization allows at hx = 6
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = field(d.gen_shape(1), 150)
        v = np.random.default_rng(151).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        y = op.filtering_procedure(xd)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        assert relerr(y, O.wawt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(wt, O.wt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs = op.dot(w, xd)
        rhs = op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale


@pytest.mark.parametrize("case", ["3d_only", "with_2d", "lquart", "wide"])
def test_rad3_procedure_matches_oracle(case):
    # filtering_rad3 (mgbf_proc = 4): the 3-D variables filtered over
    # (i, j, L) by the beta ball of mg_ampl03 (whole-sample mirror in L), the
    # 2-D variables by the 2-D filter; every generation, no vertical filter
    kw = dict(gm=3, h=4, hz=2, mgbf_proc=4, km3=3, weights=[1.0, 0.5, 0.25])
    if case == "with_2d":
        kw["km2"] = 5
    if case == "lquart":
        kw.update(km2=2, lquart=True)
    cfg = M.Config.for_grid(64, 48, 12, **kw)
    if case == "wide":
        cfg.hz = 4
        cfg.mg_ampl03 = 25 / 9     # |g| < 5: 4 points of reach in i, j and L (485 taps)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        assert d.ntaps3 == len(O.taps3(P)[3])
        x = field(d.gen_shape(1), 160)
        v = np.random.default_rng(161).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        f = op.stage(M.STAGE_F, 1, xd)
        ft = op.stage(M.STAGE_FT, 2, to_dev(field(d.gen_shape(2), 162)))
        y = op.filtering_procedure(xd)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        assert relerr(f, O.radial(P, x.astype(np.float64))) <= TOL_STAGE
        x2 = field(d.gen_shape(2), 162).astype(np.float64)
        assert relerr(ft, O.radial(P, x2, adjoint=True)) <= TOL_STAGE
        assert relerr(y, O.wawt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(wt, O.wt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs = op.dot(w, xd)
        rhs = op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale


@pytest.mark.parametrize("proc", [4, 5])
def test_3d_procedures_weight_generation_1_whatever_l_filt_g1_says(proc):
    # filtering_rad3 / filtering_lin3 weight through weighting_all ->
    # weighting, which scales generation 1 unconditionally
    # (mg_filtering.f90:203, 427; mg_generations.f90:133-155, 973-1009):
    # l_filt_g1 = .false. changes nothing for procedures 4 and 5, neither in
    # W A W^T nor in its halves nor in the weighting A on its own
    kw = dict(gm=3, h=4, hz=3, mgbf_proc=proc, km3=2, km2=2, weights=[0.8, 0.5, 0.25])
    if proc == 5:
        kw["mgbf_line"] = True
    cfgs = []
    for filt_g1 in (True, False):
        c = M.Config.for_grid(64, 48, 16, l_filt_g1=filt_g1, **kw)
        if proc == 5:
            c.mg_ampl02 = 0.7
            c.mg_ampl03 = 0.5
        cfgs.append(c)
    x = field((48, 64, 2 + 2 * 16), 200)
    v = np.random.default_rng(201).standard_normal(
        M.derive(cfgs[0]).mg_size).astype(np.float32)
    res = []
    for cfg in cfgs:
        with M.MGBF(cfg) as op:
            d = op.dims
            xd, vd = to_dev(x), to_dev(v)
            y = op.filtering_procedure(xd)
            a = op.apply(M.OP_A, vd)
            asq = op.apply(M.OP_ASQRT, vd)
            y2 = op.apply(M.OP_BSQRT, op.apply(M.OP_BSQRTT, xd))
            sync()
            assert relerr(y, O.wawt(oparams(cfg, d), x.astype(np.float64))) <= TOL_OP
            assert (y - y2).abs().max().item() <= 1e-5 * y.abs().max().item()
            w = cfg.weights()
            for g, (ag, sg, vg) in enumerate(zip(op.split_mg(a), op.split_mg(asq), op.split_mg(vd))):
                assert torch.allclose(ag, vg * w[g], rtol=1e-6, atol=0), g
                assert torch.allclose(sg, vg * float(np.sqrt(w[g])), rtol=1e-6, atol=0), g
            res.append((y.clone(), a.clone()))
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])


@pytest.mark.parametrize("proc,gm", [(4, 1), (4, 2), (7, 1), (2, 2), (8, 1)])
def test_procedures_with_one_or_two_generations_and_ragged_stack(proc, gm):
    # the smallest hierarchies (gm = 1: no transfer at all) and a stack that
    # fills no whole wavefront (K = 3 + 5*7 = 38 for the background
    # procedures, 2*7 = 14 for the ensembles), with per-point weights on
    # every generation
    kw = dict(gm=gm, h=4, hz=2, mgbf_proc=proc, km3=5, weights=[1.0, 0.5])
    if proc in (7, 8):
        kw.update(n_ens=2, km3=1)
    else:
        kw["km2"] = 3
    if proc in (2, 7):
        kw["mgbf_line"] = True
    cfg = M.Config.for_grid(48, 32, 7, **kw)
    if proc in (2, 7):
        cfg.mg_ampl02 = 0.7
    if proc == 8:
        cfg.mg_ampl02 = 1.0
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.gm == gm
        rng = np.random.default_rng(170 + proc)
        wf = [(0.5 + rng.random(d.gen_shape(g))).astype(np.float32) for g in range(1, gm + 1)]
        keep = [to_dev(w) for w in wf]
        for g, w in enumerate(keep, start=1):
            op.set_weights(g, field=w)
        x = field(d.gen_shape(1), 171)
        y = op.filtering_procedure(to_dev(x))
        sync()
        ref = O.wawt(oparams(cfg, d), x.astype(np.float64), wfield=[w.astype(np.float64) for w in wf])
        assert relerr(y, ref) <= TOL_OP


@pytest.mark.parametrize("case", ["lines", "no_tables", "lquart", "wide"])
def test_lin3_procedure_matches_oracle(case):
    # filtering_lin3 (mgbf_proc = 5, mgbf_line): hexad lines along x, y and
    # the levels (colors 1, 2, 3; zeros beyond the column ends), the 2-D
    # variables through the triad lines; every generation, no vertical filter
    kw = dict(gm=3, h=4, hz=3, mgbf_proc=5, km3=3, km2=2, mgbf_line=True, weights=[1.0, 0.5, 0.25])
    if case == "no_tables":
        kw["mgbf_line"] = False
    if case == "lquart":
        kw["lquart"] = True
    cfg = M.Config.for_grid(64, 48, 12, **kw)
    cfg.mg_ampl02 = 0.7
    cfg.mg_ampl03 = 0.25 if case == "wide" else 0.5
    if case == "wide":
        cfg.hz = 4          # level reach 4 at spread**2 2
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        if case != "no_tables":
            assert d.ntaps3 == int(np.prod([2 * h + 1 for _, _, h, _ in O.hexlines(P)[:3]]))
        x = field(d.gen_shape(1), 180)
        v = np.random.default_rng(181).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = to_dev(x), to_dev(v)
        f = op.stage(M.STAGE_F, 1, xd)
        ft = op.stage(M.STAGE_FT, 1, xd)
        y = op.filtering_procedure(xd)
        wt = op.apply(M.OP_WT, xd)
        w = op.apply(M.OP_W, vd)
        sync()
        assert relerr(f, O.radial(P, x.astype(np.float64))) <= TOL_STAGE
        assert relerr(ft, O.radial(P, x.astype(np.float64), adjoint=True)) <= TOL_STAGE
        assert relerr(y, O.wawt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(wt, O.wt(P, x.astype(np.float64))) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
        lhs = op.dot(w, xd)
        rhs = op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale


@pytest.mark.parametrize("proc,lm,km3,km2,hz,case", [
    (4, 16, 3, 0, 2, ""),        # K3 = 48: a partial wavefront of whole columns
    (4, 16, 4, 5, 3, ""),        # the 2-D variables in the next wavefront (gather)
    (4, 32, 3, 2, 1, ""),        # a wavefront shared by 3-D and 2-D variables
    (4, 8, 2, 0, 3, ""),         # lm = 2 hz + 2: the four corrected lanes adjacent
    (4, 64, 1, 0, 4, ""),        # one column per wavefront, half-width 4
    (4, 16, 4, 0, 3, ""),        # the restriction fused into the transposed pass (H = 3)
    (4, 16, 3, 2, 2, "lquart"),
    (5, 16, 3, 2, 3, ""),
    (5, 32, 2, 0, 4, "wide"),
    (5, 16, 3, 2, 3, "no_tables"),
    (4, 16, 3, 2, 3, "h2"),      # 2-D variables at half-width 2, run at 3 with zero taps
    (4, 32, 1, 3, 2, "h2"),      # ... beside a 3-D ball of half-width 2
])
def test_streaming_3d_filter_matches_oracle(proc, lm, km3, km2, hz, case):
    # the streaming 3-D filter (columns of lm | 64 levels across the lanes of
    # a wavefront, level pairs by lane shifts) against the FP64 oracle, and
    # against the gather it replaces (MGBF_BALL_GATHER=1), with per-point
    # weights on generation 1
    h = 2 if case == "h2" else 4
    kw = dict(gm=3, h=h, hz=hz, mgbf_proc=proc, km3=km3, km2=km2, weights=[1.0, 0.5, 0.25])
    if proc == 5:
        kw["mgbf_line"] = case != "no_tables"
    if case == "lquart":
        kw["lquart"] = True
    cfg = M.Config.for_grid(64, 48, lm, **kw)
    if proc == 5:
        cfg.mg_ampl02 = 0.7
        cfg.mg_ampl03 = 0.25 if case == "wide" else 0.5
    print(f"streaming 3-D filter proc={proc} lm={lm} km3={km3} km2={km2} hz={hz} {case}", flush=True)
    outs = {}
    for gather in (False, True):
        os.environ["MGBF_BALL_GATHER"] = "1" if gather else "0"
        try:
            op = M.MGBF(cfg)
        finally:
            os.environ.pop("MGBF_BALL_GATHER", None)
        with op:
            d = op.dims
            P = oparams(cfg, d)
            rng = np.random.default_rng(190 + lm + hz)
            wf = (0.5 + rng.random(d.gen_shape(1))).astype(np.float32)
            wd = to_dev(wf)
            op.set_weights(1, field=wd)
            x = field(d.gen_shape(1), 191)
            v = np.random.default_rng(192).standard_normal(d.mg_size).astype(np.float32)
            xd, vd = to_dev(x), to_dev(v)
            f = op.stage(M.STAGE_F, 1, xd)
            ft = op.stage(M.STAGE_FT, 2, to_dev(field(d.gen_shape(2), 193)))
            y = op.filtering_procedure(xd)
            wt = op.apply(M.OP_WT, xd)
            w = op.apply(M.OP_W, vd)
            sync()
            wfl = [wf.astype(np.float64), None, None]
            assert relerr(f, O.radial(P, x.astype(np.float64))) <= TOL_STAGE
            x2 = field(d.gen_shape(2), 193).astype(np.float64)
            assert relerr(ft, O.radial(P, x2, adjoint=True)) <= TOL_STAGE
            assert relerr(y, O.wawt(P, x.astype(np.float64), wfield=wfl)) <= TOL_OP
            assert relerr(wt, O.wt(P, x.astype(np.float64))) <= TOL_OP
            assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP
            lhs, rhs = op.dot(w, xd), op.dot(vd, wt)
            scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
            assert abs(lhs - rhs) <= 1e-6 * scale
            outs[gather] = y.clone()
            if not gather and km2 > 0 and case in ("", "lquart", "h2"):
                # the 2-D variables take the restriction, the weighting and
                # the prolongation in their own radial pass beside the 3-D one
                # (procedure 5's line filter at half-width 3, zero taps beyond);
                # with lquart the 3-D pass takes the quadratic restriction and
                # the prolongation stays a pass of its own
                op.timers(True)
                op.filtering_procedure(xd)
                ph = op.read_timers()
                op.timers(False)
                assert ph["upsend"][1] == 0
                assert (ph["dnsend"][1] > 0) == (case == "lquart")
    # the two engines sum in different orders: equal to rounding, not bitwise
    # (without line tables the filter is the identity, exact in both)
    if case != "no_tables":
        assert not torch.equal(outs[False], outs[True])
    assert relerr(outs[False], outs[True].double().cpu().numpy()) <= 1e-6


def test_dot_is_fp64_accumulated():
    cfg = M.Config.for_grid(64, 64, 8, gm=2, h=3, hz=2)
    with M.MGBF(cfg) as op:
        n = 3_000_001
        a = torch.full((n,), 1.0 + 2.0 ** -20, device=DEV)
        b = torch.ones(n, device=DEV)
        r = op.dot(a, b)
        assert abs(r - n * (1.0 + 2.0 ** -20)) < 1e-6


def test_autotuned_tiles_keep_parity():
    # tiles change how work is split over wavefronts; each instantiation may
    # contract multiply-adds differently, so results agree to rounding, and
    # every engine is held to the oracle bar
    cfg = M.Config.for_grid(128, 96, 32, n_ens=2, gm=3, h=6, hz=3)
    with M.MGBF(cfg) as op:
        d = op.dims
        xh = field(d.gen_shape(1), 70)
        x = to_dev(xh)
        y0 = op.filtering_procedure(x)
        tiles = op.autotune()
        y1 = op.filtering_procedure(x)
        sync()
        assert all((ti in (2, 4, 16) or (key == (1, 2) and ti == 0)) and tj >= 1 for key, (ti, tj) in tiles.items())
        assert (y1 - y0).abs().max().item() <= 1e-6 * y0.abs().max().item() or \
            any(ti == 16 for ti, _ in tiles.values())
        assert relerr(y1, O.wawt(oparams(cfg, d), xh.astype(np.float64))) <= TOL_OP
        # pinned in a second context, the same tiles give the same bits
        with M.MGBF(cfg) as op2:
            op2.set_tiles(tiles)
            assert op2.tiles() == tiles
            y2 = op2.filtering_procedure(x)
            sync()
            assert torch.equal(y1, y2)


@pytest.mark.parametrize("engine", [2, 4, 16])
def test_odd_tile_heights_match_oracle(engine):
    # the autotuner also tries tile heights fitted to the chip's wave count,
    # odd ones included for the passes without the restriction: tiles then
    # start at odd rows, where the prolongation's coarse window and the
    # vertical band must still line up
    cfg = M.Config.for_grid(96, 80, 32, n_ens=2, gm=3, h=6, hz=3, weights=[1.0, 0.6, 0.3])
    heights = {(1, 0): 13, (1, 1): 11, (1, 2): 9, (1, 3): 14, (2, 0): 7, (2, 1): 5, (2, 3): 10,
               (3, 0): 3}
    print(f"odd tile heights, engine {engine}", flush=True)
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.km_all == 64
        P = oparams(cfg, d)
        tiles = {}
        for (g, kind), tj in heights.items():
            ti = engine
            if kind == 3 and ti == 2:
                ti = 4
            if ti == 16 and d.gen_shape(g)[1] % 16:
                ti = 4
            tiles[(g, kind)] = (ti, tj)
        op.set_tiles(tiles)
        got = op.tiles()
        for key, val in tiles.items():
            assert got[key] == val
        x = field(d.gen_shape(1), 171)
        v = np.random.default_rng(172).standard_normal(d.mg_size).astype(np.float32)
        y = op.filtering_procedure(to_dev(x))
        wt = op.apply(M.OP_WT, to_dev(x))
        w = op.apply(M.OP_W, to_dev(v))
        sync()
        x64 = x.astype(np.float64)
        assert relerr(y, O.wawt(P, x64)) <= TOL_OP
        assert relerr(wt, O.wt(P, x64)) <= TOL_OP
        assert relerr(w, O.w(P, v.astype(np.float64))) <= TOL_OP


@pytest.mark.parametrize("h", [6, 8])
def test_vertical_band_as_a_pass_of_its_own_matches_fused(h):
    # a kind-2 tile with TI = 0 (what the autotuner picks where the kind-1
    # pass and a separate band beat the fused pass) runs the band as a
    # vertical pass after the last radial pass; the result is the fused one's
    # to rounding and the oracle's
    cfg = M.Config.for_grid(96, 64, 32, n_ens=2, gm=3, h=h, hz=4, weights=[1.0, 0.6, 0.3])
    print(f"band as a pass of its own, h={h}", flush=True)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = to_dev(field(d.gen_shape(1), 181))
        op.timers(True)
        y_fused = op.filtering_procedure(x)
        fused = op.read_timers()
        op.timers(False)
        tj = op.tiles()[(1, 2)][1]
        op.set_tiles({(1, 2): (0, tj)})
        assert op.tiles()[(1, 2)] == (0, tj)
        op.timers(True)
        y_split = op.filtering_procedure(x)
        split = op.read_timers()
        op.timers(False)
        sync()
        assert fused["vfilt"][1] == 0 and split["vfilt"][1] == 1
        assert relerr(y_split, y_fused.double().cpu().numpy()) <= 1e-6
        assert relerr(y_split, O.wawt(P, x.cpu().numpy().astype(np.float64))) <= TOL_OP
        # a kind-2 tile of an engine brings the fused pass back
        op.set_tiles({(1, 2): (4, tj)})
        assert op.tiles()[(1, 2)] == (4, tj)
        y_back = op.filtering_procedure(x)
        sync()
        assert relerr(y_back, y_fused.double().cpu().numpy()) <= 1e-6


@pytest.mark.parametrize("h,K,n,m,gm,scale", [(6, 64, 96, 96, 3, 1.0), (4, 64, 80, 48, 3, 1e-6),
                                              (8, 128, 64, 64, 2, 1e4), (3, 192, 64, 48, 2, 1.0),
                                              (6, 128, 192, 96, 4, 1.0), (5, 64, 128, 64, 3, 1e-30)])
def test_matrix_core_engine_matches_oracle(monkeypatch, h, K, n, m, gm, scale):
    # every eligible radial pass on v_mfma_f32_16x16x32_f16 (MGBF_MFMA=1):
    # the plain pass, the restriction, the prolongation and the vertical band
    # epilogues; input scales from 1e-30 to 1e4 through the per-k exponents
    cfg = M.Config.for_grid(n, m, K, gm=gm, h=h, hz=min(3, K - 1))
    x = (field((m, n, K), 90 + h) * scale).astype(np.float32)
    print(f"matrix cores h={h} K={K} {n}x{m} gm={gm} scale={scale}", flush=True)
    with M.MGBF(cfg) as op:
        y_vec = op.filtering_procedure(to_dev(x))
        sync()
    monkeypatch.setenv("MGBF_MFMA", "1")
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        t = op.tiles()
        assert t[(1, 0)][0] == 16 and t[(1, 3)][0] == 16 and t[(1, 1)][0] == 16
        y = op.filtering_procedure(to_dev(x))
        v = op.apply(M.OP_WT, to_dev(x))
        vv = np.random.default_rng(91).standard_normal(d.mg_size).astype(np.float32) * scale
        w = op.apply(M.OP_W, to_dev(vv))
        sync()
        x64 = x.astype(np.float64)
        assert relerr(y, O.wawt(P, x64)) <= TOL_OP
        assert relerr(v, O.wt(P, x64)) <= TOL_OP
        assert relerr(w, O.w(P, vv.astype(np.float64))) <= TOL_OP
        assert not torch.equal(y, y_vec)   # the matrix-core path ran


@pytest.mark.parametrize("case", ["members_apart", "rows_grow", "wfield", "fast", "band64", "lm32", "lquart",
                                  "lquart_lm32"])
def test_matrix_core_engine_scales_and_epilogues(monkeypatch, case):
    # members whose magnitudes differ by 1e10 in one stack (each member's k
    # keep their own exponent; every member is held to the bar on its own),
    # rows that grow by 2^24 down the field (the exponents are lowered and the
    # ring rescaled on the way), per-point weights, the separable fast filter,
    # the vertical band fused into the last pass with whole 64-level columns
    # and with two 32-level columns per chunk, and the quadratic transfer
    # (lquart) in the restriction, prolongation and band epilogues
    monkeypatch.setenv("MGBF_MFMA", "1")
    lm, n_ens, kw = 64, 2, {}
    if case == "fast":
        kw = dict(mgbf_proc=8, p=2)
    if case.startswith("lquart"):
        kw = dict(lquart=True)
    if case in ("lm32", "lquart_lm32"):
        lm, n_ens = 32, 4
    cfg = M.Config.for_grid(96, 64, lm, n_ens=n_ens, gm=3, h=6, hz=3, weights=[1.0, 0.5, 0.25], **kw)
    if case == "fast":
        cfg.mg_ampl02 = 0.45       # reach sqrt(7)/0.45 = 5.9: 11 taps per axis (half-width 5)
    print(f"matrix cores: {case}", flush=True)
    with M.MGBF(cfg) as op:
        d = op.dims
        P = oparams(cfg, d)
        x = field(d.gen_shape(1), 95)
        if case == "members_apart":
            x = x.reshape(64, 96, n_ens, lm)
            x[:, :, 0, :] *= 1e-6
            x[:, :, 1, :] *= 1e4
            x = x.reshape(d.gen_shape(1))
        if case == "rows_grow":
            x = (x * np.exp2(np.arange(64) * 24.0 / 63)[:, None, None]).astype(np.float32)
        wf = None
        if case == "wfield":
            rng = np.random.default_rng(96)
            wf = [(0.5 + rng.random(d.gen_shape(g))).astype(np.float32) for g in (1, 2, 3)]
            keep = [to_dev(a) for a in wf]
            for g, a in enumerate(keep, start=1):
                op.set_weights(g, field=a)
        t = op.tiles()
        assert all(t[(1, kind)][0] == 16 for kind in (0, 1, 3))
        y = op.filtering_procedure(to_dev(x))
        if case in ("band64", "lm32", "members_apart", "rows_grow", "lquart", "lquart_lm32"):
            # the band rode in the last radial pass (no vertical launch)
            op.timers(True)
            op.filtering_procedure(to_dev(x))
            ph = op.read_timers()
            op.timers(False)
            assert ph["vfilt"][1] == 0 and t[(1, 2)][0] == 16
        sync()
        ref = O.wawt(P, x.astype(np.float64),
                     wfield=[a.astype(np.float64) for a in wf] if wf is not None else None)
        if case == "members_apart":
            yv = y.cpu().double().numpy().reshape(64, 96, n_ens, lm)
            rv = ref.reshape(64, 96, n_ens, lm)
            for mbr in range(n_ens):
                assert relerr(yv[:, :, mbr], rv[:, :, mbr]) <= TOL_OP, mbr
        else:
            assert relerr(y, ref) <= TOL_OP
        if case == "rows_grow":
            # the small rows keep their accuracy relative to their own size
            yr = y.cpu().double().numpy()[:8]
            assert np.max(np.abs(yr - ref[:8])) <= 1e-5 * np.max(np.abs(ref[:8]))
        if case == "wfield":
            # the weighted passes ran on the matrix cores with a scale pass
            # after them: the vector kernel, which reads the field in its
            # epilogue, sums in another order
            op.set_tiles({(g, 3): (4, t[(g, 3)][1]) for g in (1, 2)})
            y_vec = op.filtering_procedure(to_dev(x))
            sync()
            assert not torch.equal(y, y_vec)
            assert relerr(y_vec, ref) <= TOL_OP


def box3(a):
    """3x3 box average with the half-sample mirror at the edges, (m, n, K)."""
    p = np.pad(a, ((1, 1), (1, 1), (0, 0)), mode="symmetric")
    return sum(p[1 + dj: p.shape[0] - 1 + dj, 1 + di: p.shape[1] - 1 + di]
               for dj in (-1, 0, 1) for di in (-1, 0, 1)) / 9.0


def test_single_field_768x768x64_five_generations_full_size():
    # one field 768x768x64, 5 generations, radial half-width 8, vertical 4,
    # p=2, weights 2^-g (g = 0..4); seeded N(0,1) smoothed by a 3x3 box
    cfg = M.Config.for_grid(768, 768, 64, gm=5, h=8, hz=4, weights=[2.0 ** -g for g in range(5)])
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.gm == 5 and d.ntaps2 == 249 and d.ntaps1 == 9
        x = box3(np.random.default_rng(3).standard_normal(d.gen_shape(1))).astype(np.float32)
        y = op.filtering_procedure(to_dev(x))
        sync()
        ref = O.wawt(oparams(cfg, d), x.astype(np.float64))
        assert relerr(y, ref) <= TOL_OP
        # with the autotuner's tiles (on this field: matrix-core passes with
        # tile heights fitted to one round of waves, and the vertical band
        # as a pass of its own where that was faster)
        print(f"single field: autotuned tiles {op.autotune()}", flush=True)
        y2 = op.filtering_procedure(to_dev(x))
        sync()
        assert relerr(y2, ref) <= TOL_OP


def test_batched_ensemble_80x768x768x64_full_size():
    # the benchmark workload: 80 members x 768x768x64 stacked in k, 4
    # generations, radial half-width 6, vertical half-width 3, p = 2, weights
    # [1, 1/2, 1/4, 1/8]; member m ~ N(0,1) with seed 100+m. Members 0, 39
    # and 79 against the FP64 oracle, every member against the plain
    # per-member context, and the W / W^T dot-product test at full size
    # (filtering_rad2_ens, mg_filtering.f90:1000-1147)
    M_, n, lm, gm = 80, 768, 64, 4
    w = [1.0, 0.5, 0.25, 0.125]
    cfg = M.Config.for_grid(n, n, lm, n_ens=M_, gm=gm, h=6, hz=3, p=2, weights=w)
    print("batched ensemble 80 x 768x768x64: generating members", flush=True)
    x = torch.empty((n, n, M_ * lm), dtype=torch.float32, device=DEV)
    xv = x.view(n, n, M_, lm)
    g = torch.Generator(device="cuda")
    for m in range(M_):
        g.manual_seed(100 + m)
        xv[:, :, m, :] = torch.randn((n, n, lm), generator=g, device=DEV, dtype=torch.float32)
    with M.MGBF(cfg) as op:
        d = op.dims
        assert d.gm == gm and d.km_all == M_ * lm and d.ntaps2 == 145 and d.ntaps1 == 7
        op.autotune()
        y = op.filtering_procedure(x)
        sync()
        yv = y.view(n, n, M_, lm)
        cfg1 = M.Config.for_grid(n, n, lm, n_ens=1, gm=gm, h=6, hz=3, p=2, weights=w)
        P = O.OracleParams(nx=n, ny=n, K=lm, lm=lm, nvcol=1, hx=6, hy=6, hz=3, p=2, gm=gm,
                           ampl01=cfg1.mg_ampl01, ampl02=cfg1.mg_ampl02, weights=w)
        for m in (0, 39, 79):
            print(f"batched ensemble: member {m} against the FP64 oracle", flush=True)
            ref = O.wawt(P, xv[:, :, m, :].double().cpu().numpy())
            err = relerr(yv[:, :, m, :], ref)
            assert err <= TOL_OP, (m, err)
        print("batched ensemble: every member against the plain context", flush=True)
        with M.MGBF(cfg1) as op1:
            worst = 0.0
            for m in range(M_):
                ym = op1.filtering_procedure(xv[:, :, m, :].contiguous())
                ref = yv[:, :, m, :]
                worst = max(worst, ((ym - ref).abs().max() / ref.abs().max()).item())
            assert worst <= 1e-6, worst
        print("batched ensemble: W / W^T dot-product test", flush=True)
        gv = torch.Generator(device="cuda").manual_seed(7)
        v = torch.randn((d.mg_size,), generator=gv, device=DEV, dtype=torch.float32)
        wv = op.apply(M.OP_W, v)
        lhs = op.dot(wv, x)
        del wv, y
        wt = op.apply(M.OP_WT, x)
        rhs = op.dot(v, wt)
        scale = float(torch.linalg.vector_norm(v.double())) * float(torch.linalg.vector_norm(wt.double()))
        assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-3 * scale), (lhs, rhs, scale)
        del v, wt
    del x
    torch.cuda.empty_cache()
