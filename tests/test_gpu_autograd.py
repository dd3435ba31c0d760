"""Differentiation through the MGBF operator (torch.autograd over the C ABI).

Loss J = 1/2 ||W A_w W^T (a * x) - y||^2 with learnable generation weights w
and amplitude field a. J is quadratic in w and in a, so central differences
of the FP64 oracle (h = 1e-3) are exact up to rounding: the GPU gradients
must match them to 1e-4 relative. At the full size (16 members x 512x512x64,
5 generations) the check is homogeneity: B_w z is linear in w and in a, so
sum_g w_g dJ/dw_g = <r, B_w z> = sum_i a_i dJ/da_i.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O
from standalone_mgbf_loc_amd import mgbf as M
from standalone_mgbf_loc_amd.autograd import localize, loss_and_grads

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")


def smooth(a, passes=2):
    """3x3 box smoothing with the half-sample mirror, on (m, n, K) arrays."""
    for _ in range(passes):
        p = np.pad(a, ((1, 1), (1, 1), (0, 0)), mode="symmetric")
        a = sum(p[1 + dj: p.shape[0] - 1 + dj, 1 + di: p.shape[1] - 1 + di]
                for dj in (-1, 0, 1) for di in (-1, 0, 1)) / 9.0
    return a


def box3_passes(t, passes):
    """3x3 box smoothing of a (m, n, K) device tensor with the half-sample
    mirror (edge replication for a 1-point pad), `passes` times."""
    F = torch.nn.functional
    v = t.permute(2, 0, 1).unsqueeze(0)          # (1, K, m, n)
    for _ in range(passes):
        v = F.avg_pool2d(F.pad(v, (1, 1, 1, 1), mode="replicate"), 3, stride=1)
    return v.squeeze(0).permute(1, 2, 0).contiguous()


def test_box3_passes_matches_numpy_smoothing():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((12, 10, 3))
    got = box3_passes(torch.from_numpy(a).to(DEV), 2).cpu().numpy()
    assert np.allclose(got, smooth(a, 2), atol=1e-12)


def inputs(shape, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(shape).astype(np.float32)
    y = rng.standard_normal(shape).astype(np.float32)
    a = (1.0 + 0.2 * smooth(rng.standard_normal(shape))).astype(np.float32)
    return x, y, a


def oracle_J(P, a, x, y, w):
    P.weights = list(w)
    r = O.wawt(P, a.astype(np.float64) * x.astype(np.float64)) - y.astype(np.float64)
    return 0.5 * float(np.sum(r * r))


@pytest.mark.parametrize("filt_g1", [True, False])
def test_gradients_match_fp64_central_differences(filt_g1):
    cfg = M.Config.for_grid(48, 48, 8, n_ens=2, gm=3, h=4, hz=2, l_filt_g1=filt_g1)
    w0 = [0.9, 0.6, 0.3]
    with M.MGBF(cfg) as op:
        d = op.dims
        x, y, a = inputs(d.gen_shape(1), 3)
        ws = [torch.tensor(v, dtype=torch.float64, device=DEV) for v in w0]
        J, ga, gw = loss_and_grads(op, torch.from_numpy(a).to(DEV), torch.from_numpy(x).to(DEV),
                                   torch.from_numpy(y).to(DEV), ws)
        P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=4, hy=4,
                           hz=2, p=2, gm=3, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                           l_filt_g1=filt_g1)
        J64 = oracle_J(P, a, x, y, w0)
        assert abs(J - J64) <= 1e-5 * abs(J64)
        h = 1e-3
        for g in range(3):
            wp, wm = list(w0), list(w0)
            wp[g] += h
            wm[g] -= h
            fd = (oracle_J(P, a, x, y, wp) - oracle_J(P, a, x, y, wm)) / (2 * h)
            got = gw[g].item()
            if g == 0 and not filt_g1:
                assert got == 0.0 and abs(fd) < 1e-9 * abs(J64)
            else:
                assert abs(got - fd) <= 1e-4 * abs(fd), (g, got, fd)
        rng = np.random.default_rng(7)
        gan = ga.cpu().double().numpy()
        for _ in range(4):
            idx = tuple(int(rng.integers(s)) for s in a.shape)
            ap, am = a.astype(np.float64).copy(), a.astype(np.float64).copy()
            ap[idx] += h
            am[idx] -= h
            fd = (oracle_J(P, ap, x, y, w0) - oracle_J(P, am, x, y, w0)) / (2 * h)
            assert abs(gan[idx] - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(gan).max()), (idx, gan[idx], fd)


@pytest.mark.parametrize("case", ["rad3", "fast", "lquart", "line_bckg", "rad3_stream", "lin3_stream",
                                  "rad3_filt_g1_off", "lin3_filt_g1_off"])
def test_weight_gradients_of_the_other_procedures(case):
    # procedures 4 (3-D radial), 8 (fast), 6 with lquart and 2 (line), and
    # procedures 4 and 5 on the streaming 3-D filter (columns of 16 levels):
    # the weight gradients <(W^T g)_g, (W^T z)_g> against FP64 central
    # differences. Procedures 4 and 5 weight generation 1 whatever
    # l_filt_g1 says (weighting_all, mg_filtering.f90:203, 427), so with
    # l_filt_g1 = .false. dJ/dw_1 is still the central difference
    kw = dict(gm=3, h=4, hz=2)
    if case.endswith("_filt_g1_off"):
        kw["l_filt_g1"] = False
        case = case[: -len("_filt_g1_off")] + "_stream"
    lm = 6
    if case == "rad3":
        kw.update(mgbf_proc=4, km3=2, km2=1)
    elif case == "rad3_stream":
        kw.update(mgbf_proc=4, km3=2)
        lm = 16
    elif case == "lin3_stream":
        kw.update(mgbf_proc=5, km3=2, mgbf_line=True, hz=3)
        lm = 16
    elif case == "fast":
        kw.update(mgbf_proc=8, n_ens=2)
    elif case == "lquart":
        kw.update(n_ens=2, lquart=True)
    else:
        kw.update(mgbf_proc=2, km3=2, km2=1, mgbf_line=True)
    cfg = M.Config.for_grid(48, 48, lm, **kw)
    if case == "fast":
        cfg.mg_ampl02 = 1.0
    if case in ("line_bckg", "lin3_stream"):
        cfg.mg_ampl02 = 0.7
    if case == "lin3_stream":
        cfg.mg_ampl03 = 0.5
    w0 = [0.9, 0.6, 0.3]
    with M.MGBF(cfg) as op:
        d = op.dims
        x, y, a = inputs(d.gen_shape(1), 11)
        ws = [torch.tensor(v, dtype=torch.float64, device=DEV) for v in w0]
        J, ga, gw = loss_and_grads(op, torch.from_numpy(a).to(DEV), torch.from_numpy(x).to(DEV),
                                   torch.from_numpy(y).to(DEV), ws)
        P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=4, hy=4,
                           hz=cfg.hz, p=2, gm=3, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                           fast=case == "fast", lquart=case == "lquart", line=case == "line_bckg",
                           rad3=case in ("rad3", "rad3_stream"), lin3=case == "lin3_stream",
                           ampl03=cfg.mg_ampl03)
        J64 = oracle_J(P, a, x, y, w0)
        assert abs(J - J64) <= 1e-5 * abs(J64)
        h = 1e-3
        for g in range(3):
            wp, wm = list(w0), list(w0)
            wp[g] += h
            wm[g] -= h
            fd = (oracle_J(P, a, x, y, wp) - oracle_J(P, a, x, y, wm)) / (2 * h)
            assert abs(gw[g].item() - fd) <= 1e-4 * abs(fd), (g, gw[g].item(), fd)


def test_ensemble_line_procedure_is_refused_by_autograd():
    # filtering_lin2_ens ends with Vf^T: its W A W^T is not symmetric
    cfg = M.Config.for_grid(32, 32, 4, n_ens=2, gm=2, h=3, hz=1, mgbf_proc=7)
    with M.MGBF(cfg) as op:
        z = torch.zeros(op.dims.gen_shape(1), device=DEV)
        w = [torch.tensor(1.0, dtype=torch.float64, device=DEV, requires_grad=True)] * 2
        with pytest.raises(ValueError, match="not symmetric"):
            localize(op, z, w)


def test_per_point_weight_field_gradient():
    cfg = M.Config.for_grid(32, 32, 4, n_ens=2, gm=2, h=3, hz=1)
    with M.MGBF(cfg) as op:
        d = op.dims
        x, y, _ = inputs(d.gen_shape(1), 5)
        rng = np.random.default_rng(6)
        f = [(0.5 + rng.random(d.gen_shape(g))).astype(np.float64) for g in (1, 2)]
        ws = [torch.from_numpy(v).to(DEV).requires_grad_(True) for v in f]
        xd, yd = torch.from_numpy(x).to(DEV), torch.from_numpy(y).to(DEV)
        r = localize(op, xd, ws) - yd
        (0.5 * (r.double() ** 2).sum()).backward()
        P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=3, hy=3,
                           hz=1, p=2, gm=2, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02)

        def J(fields):
            rr = O.wawt(P, x.astype(np.float64), wfield=fields) - y
            return 0.5 * float(np.sum(rr * rr))

        h = 1e-3
        for g in (0, 1):
            for _ in range(3):
                idx = tuple(int(rng.integers(s)) for s in f[g].shape)
                fp = [v.copy() for v in f]
                fm = [v.copy() for v in f]
                fp[g][idx] += h
                fm[g][idx] -= h
                fd = (J(fp) - J(fm)) / (2 * h)
                got = ws[g].grad.cpu().numpy()[idx]
                scale = np.abs(ws[g].grad.cpu().numpy()).max()
                assert abs(got - fd) <= 1e-4 * max(abs(fd), 1e-2 * scale), (g, idx, got, fd)


def test_full_size_gradients_are_homogeneous():
    # 16 members x 512x512x64, 5 generations, learnable weights [1, .5, .25, .125, .0625];
    # homogeneity and central differences of the weights and of the amplitude field
    cfg = M.Config.for_grid(512, 512, 64, n_ens=16, gm=5, h=6, hz=3)
    with M.MGBF(cfg) as op:
        d = op.dims
        gen = torch.Generator(device="cuda").manual_seed(11)
        x = torch.randn(d.gen_shape(1), generator=gen, device=DEV)
        y = torch.randn(d.gen_shape(1), generator=gen, device=DEV)
        # amplitude field a = 1 + 0.2 N(0,1) smoothed by two 3x3 box passes
        # with the half-sample mirror (for a 1-point pad, edge replication)
        noise = torch.randn(d.gen_shape(1), generator=gen, device=DEV)
        a = 1.0 + 0.2 * box3_passes(noise, 2)
        del noise
        w0 = [1.0, 0.5, 0.25, 0.125, 0.0625]
        ws = [torch.tensor(v, dtype=torch.float64, device=DEV) for v in w0]
        J, ga, gw = loss_and_grads(op, a, x, y, ws)
        z = (a * x).contiguous()
        for g, v in enumerate(w0, start=1):
            op.set_weights(g, value=v)
        Bz = op.filtering_procedure(z)
        r = (Bz - y).contiguous()
        rBz = op.dot(r, Bz)
        lhs_w = sum(v * g.item() for v, g in zip(w0, gw))
        lhs_a = op.dot(a.contiguous(), ga.to(torch.float32).contiguous())
        assert abs(lhs_w - rBz) <= 1e-5 * abs(rBz)
        assert abs(lhs_a - rBz) <= 1e-4 * abs(rBz)

        # central differences at full size (h = 1e-3): J is quadratic in
        # every w_g and in a, so J(+h) - J(-h) = 2h dJ exactly; J is
        # accumulated in FP64 over the FP32 operator
        def J_of(wv, av):
            for g, v in enumerate(wv, start=1):
                op.set_weights(g, value=v)
            rr = op.filtering_procedure((av * x).contiguous()) - y
            return 0.5 * float((rr.double() ** 2).sum())

        # the steps are taken as the library stores them (FP32), so the
        # difference quotient divides by the step actually taken
        h = 1e-3
        f32 = lambda v: float(np.float32(v))
        for g in range(5):
            wp, wm = list(w0), list(w0)
            wp[g], wm[g] = f32(w0[g] + h), f32(w0[g] - h)
            fd = (J_of(wp, a) - J_of(wm, a)) / (wp[g] - wm[g])
            assert abs(gw[g].item() - fd) <= 1e-4 * abs(fd), (g, gw[g].item(), fd)
        # the amplitude gradient along a smooth random direction; J is
        # quadratic in a, so a step of 0.05 has no truncation error and keeps
        # the difference far above the FP32 noise of J (~1e8) itself
        da = box3_passes(torch.randn(d.gen_shape(1), generator=gen, device=DEV), 2)
        ha = 0.05
        ap, am = a + ha * da, a - ha * da
        fd = (J_of(w0, ap) - J_of(w0, am))
        got = float((ga.double() * (ap.double() - am.double())).sum())
        assert abs(got - fd) <= 1e-4 * abs(fd), (got, fd)
        del x, y, a, z, Bz, r, ga, da, ap, am
