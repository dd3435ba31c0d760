"""Per-point aspect tensors (the pasp2(:,:,i,j) of def_mg_weights built from
p_del and p_eps fields, mg_intstate.f90:383-390, read per point by vrbeta2 /
vrbeta2T, jp_pbfil.f90:888-928, 1011-1050): the HIP path against the FP64
oracle's per-point restatement, on rotated anisotropic tensors varying over
the domain (and so line sums ss varying per point).

Tolerance as in test_gpu_parity.py: normwise relative error 2e-6 for single
stages, 1e-5 for operators; dot-product mismatch 1e-6 of the Cauchy-Schwarz
scale."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O
from standalone_mgbf_loc_amd import mgbf as M
from test_oracle import aspect_field

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")


def relerr(y, y64):
    y = y.detach().cpu().double().numpy() if isinstance(y, torch.Tensor) else y
    return float(np.max(np.abs(y - y64)) / max(np.max(np.abs(y64)), 1e-300))


def _setup(n_ens=2, lm=32, gm=3, h=6):
    cfg = M.Config.for_grid(96, 64, lm, n_ens=n_ens, gm=gm, h=h, hz=3, weights=[1.0, 0.6, 0.3])
    # reach sqrt(A (2p+4)) <= h + 1 for the largest axis ampl (1 + 0.5)
    ampl = (h + 1) ** 2 / 8.0 / 1.55
    return cfg, ampl


def _oracle(cfg, d):
    return O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=cfg.hx, hy=cfg.hy,
                          hz=cfg.hz, p=cfg.p, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                          weights=cfg.weights()[: d.gm])


def _wawt_composed(P, asps, x):
    """W A W^T from the oracle's stages: Vf^T, the restriction chain, per
    generation F^T, A, F (per-point aspects where given), the prolongation
    up-sweep, Vf (filtering_rad2_ens, mg_filtering.f90:1000-1147)."""
    gm = P.gm
    u = [O.vertical(P, x, adjoint=True)]
    for g in range(1, gm):
        u.append(O.restrict(u[-1]))
    z = []
    for g in range(gm):
        filt = (lambda v, adj: O.radial_aspect(P, asps[g], v, adj)) if asps[g] is not None \
            else (lambda v, adj: O.radial(P, v, adj))
        z.append(filt(P.weights[g] * filt(u[g], True), False))
    s = z[-1]
    for g in range(gm - 2, -1, -1):
        s = z[g] + O.prolong(s)
    return O.vertical(P, s)


def test_aspect_stages_match_oracle():
    cfg, ampl = _setup()
    with M.MGBF(cfg) as op:
        d = op.dims
        P = _oracle(cfg, d)
        for gen in (1, 2):
            ny, nx, _ = d.gen_shape(gen)
            asp = aspect_field(nx, ny, ampl, seed=10 + gen)
            op.set_aspect(gen, asp)
            x = np.random.default_rng(20 + gen).standard_normal(d.gen_shape(gen)).astype(np.float32)
            xd = torch.from_numpy(x).to(DEV)
            f = op.stage(M.STAGE_F, gen, xd)
            ft = op.stage(M.STAGE_FT, gen, xd)
            torch.cuda.synchronize()
            x64 = x.astype(np.float64)
            assert relerr(f, O.radial_aspect(P, asp, x64)) <= 2e-6
            assert relerr(ft, O.radial_aspect(P, asp, x64, adjoint=True)) <= 2e-6
            # F^T is no longer F
            assert (f - ft).abs().max().item() > 1e-3 * f.abs().max().item()


@pytest.mark.parametrize("which", ["g1", "g12"])
def test_aspect_operator_matches_composed_oracle(which):
    cfg, ampl = _setup()
    with M.MGBF(cfg) as op:
        d = op.dims
        P = _oracle(cfg, d)
        asps = [None] * d.gm
        for gen in ((1,) if which == "g1" else (1, 2)):
            ny, nx, _ = d.gen_shape(gen)
            asps[gen - 1] = aspect_field(nx, ny, ampl, seed=30 + gen)
            op.set_aspect(gen, asps[gen - 1])
        x = np.random.default_rng(31).standard_normal(d.gen_shape(1)).astype(np.float32)
        v = np.random.default_rng(32).standard_normal(d.mg_size).astype(np.float32)
        xd, vd = torch.from_numpy(x).to(DEV), torch.from_numpy(v).to(DEV)
        y = op.filtering_procedure(xd)
        w = op.apply(M.OP_W, vd)
        wt = op.apply(M.OP_WT, xd)
        y2 = op.apply(M.OP_BSQRT, op.apply(M.OP_BSQRTT, xd))
        torch.cuda.synchronize()
        assert relerr(y, _wawt_composed(P, asps, x.astype(np.float64))) <= 1e-5
        assert (y - y2).abs().max().item() <= 1e-5 * y.abs().max().item()
        lhs, rhs = op.dot(w, xd), op.dot(vd, wt)
        scale = float(torch.linalg.vector_norm(w.double())) * float(torch.linalg.vector_norm(xd.double()))
        assert abs(lhs - rhs) <= 1e-6 * scale
        # B stays symmetric: <x, B z> = <B x, z>
        z = torch.from_numpy(np.random.default_rng(33).standard_normal(d.gen_shape(1)).astype(np.float32)).to(DEV)
        Bz = op.filtering_procedure(z)
        assert abs(op.dot(xd, Bz) - op.dot(y, z)) <= 1e-6 * float(torch.linalg.vector_norm(y.double())) * \
            float(torch.linalg.vector_norm(z.double()))
        # back to the namelist's tensor: the operator of a fresh context
        for gen in (1, 2):
            op.set_aspect(gen, None)
        y0 = op.filtering_procedure(xd)
        with M.MGBF(cfg) as op2:
            yf = op2.filtering_procedure(xd)
        torch.cuda.synchronize()
        assert torch.equal(y0, yf)


def test_uniform_aspect_field_matches_the_domain_tensor():
    # a field equal everywhere to the namelist's tensor reproduces the
    # radial kernel's result to rounding
    cfg, _ = _setup()
    with M.MGBF(cfg) as op:
        d = op.dims
        x = torch.from_numpy(np.random.default_rng(40).standard_normal(d.gen_shape(1)).astype(np.float32)).to(DEV)
        y_iso = op.filtering_procedure(x)
        for gen in range(1, d.gm + 1):
            ny, nx, _ = d.gen_shape(gen)
            op.set_aspect(gen, np.broadcast_to([cfg.mg_ampl02, cfg.mg_ampl02, 0.0], (ny, nx, 3)))
        y_asp = op.filtering_procedure(x)
        torch.cuda.synchronize()
        assert (y_asp - y_iso).abs().max().item() <= 2e-6 * y_iso.abs().max().item()


def test_aspect_refusals():
    cfg, _ = _setup(gm=2)
    with M.MGBF(cfg) as op:
        ny, nx, _ = op.dims.gen_shape(1)
        with pytest.raises(M.MGBFError, match="reach"):
            op.set_aspect(1, np.broadcast_to([40.0, 40.0, 0.0], (ny, nx, 3)))
        with pytest.raises(M.MGBFError, match="positive definite"):
            op.set_aspect(1, np.broadcast_to([1.0, 1.0, 2.0], (ny, nx, 3)))
    fast = M.Config.for_grid(96, 64, 16, n_ens=2, gm=2, h=6, hz=3, mgbf_proc=8, mg_ampl02=0.45)
    with M.MGBF(fast) as op:
        ny, nx, _ = op.dims.gen_shape(1)
        with pytest.raises(M.MGBFError, match="radial procedures"):
            op.set_aspect(1, np.broadcast_to([1.0, 1.0, 0.0], (ny, nx, 3)))
