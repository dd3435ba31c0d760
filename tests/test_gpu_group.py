"""Uneven batch: regional domains of different sizes, generations and filter
widths applied as one hipGraph (mgbf_group_*), against the same contexts
called one by one and against the FP64 oracle. The graph is replayed
several times with eager work in between and every replay is checked."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import oracle as O
from standalone_mgbf_loc_amd import mgbf as M

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")

# (nx, ny, levels, generations, radial half-width) of each domain
DOMAINS = [(96, 96, 16, 4, 6), (128, 64, 32, 3, 8), (48, 48, 8, 2, 4), (160, 96, 24, 4, 10)]


def _cfg(nx, ny, lm, gm, h):
    return M.Config.for_grid(nx, ny, lm, gm=gm, h=h, hz=3, weights=[1.0, 0.6, 0.3, 0.15])


def _oracle(cfg, d, x):
    P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=cfg.hx,
                       hy=cfg.hy, hz=cfg.hz, p=2, gm=d.gm, ampl01=cfg.mg_ampl01,
                       ampl02=cfg.mg_ampl02, weights=[1.0, 0.6, 0.3, 0.15][: d.gm])
    return O.wawt(P, x.astype(np.float64))


def test_group_graph_matches_eager_and_oracle():
    cfgs = [_cfg(*dm) for dm in DOMAINS]
    ops = [M.MGBF(c) for c in cfgs]
    try:
        xs_h = [np.random.default_rng(300 + i).standard_normal(op.field_shape()).astype(np.float32)
                for i, op in enumerate(ops)]
        xs = [torch.from_numpy(x).to(DEV) for x in xs_h]
        eager = [op.filtering_procedure(x) for op, x in zip(ops, xs)]
        torch.cuda.synchronize()
        for cfg, op, x, y in zip(cfgs, ops, xs_h, eager):
            ref = _oracle(cfg, op.dims, x)
            assert np.max(np.abs(y.cpu().double().numpy() - ref)) / np.max(np.abs(ref)) <= 1e-5
        outs = [torch.empty_like(x) for x in xs]
        with M.Group(ops) as grp:
            for rep in range(3):
                for o in outs:
                    o.fill_(float("nan"))
                grp.apply(M.OP_WAWT, xs, outs)
                torch.cuda.synchronize()
                for y, o in zip(eager, outs):
                    assert torch.equal(y, o), rep
                # eager work on the same contexts between replays
                ops[rep % len(ops)].filtering_procedure(xs[rep % len(ops)])
                torch.cuda.synchronize()
            # new buffers force a new capture
            outs2 = [torch.empty_like(x) for x in xs]
            grp.apply(M.OP_WAWT, xs, outs2)
            torch.cuda.synchronize()
            for y, o in zip(eager, outs2):
                assert torch.equal(y, o)
    finally:
        for op in ops:
            op.close()


def test_group_recaptures_after_weights_or_tiles_change():
    # the captured kernels read the weights and tiles of the moment of
    # capture: after set_weights / set_tiles on one member the next replay
    # must be captured again, and equal the eager calls with the new state
    cfgs = [_cfg(*dm) for dm in DOMAINS[:3]]
    ops = [M.MGBF(c) for c in cfgs]
    try:
        xs = [torch.from_numpy(np.random.default_rng(340 + i).standard_normal(op.field_shape())
                               .astype(np.float32)).to(DEV) for i, op in enumerate(ops)]
        outs = [torch.empty_like(x) for x in xs]
        with M.Group(ops) as grp:
            grp.apply(M.OP_WAWT, xs, outs)
            torch.cuda.synchronize()
            before = [o.clone() for o in outs]
            gen = torch.Generator(device="cuda").manual_seed(341)
            wf = 0.5 + torch.rand(ops[0].dims.gen_shape(2), generator=gen, device=DEV)
            ops[0].set_weights(2, field=wf)
            ops[1].set_weights(1, value=0.3)
            grp.apply(M.OP_WAWT, xs, outs)
            torch.cuda.synchronize()
            eager = [op.filtering_procedure(x) for op, x in zip(ops, xs)]
            torch.cuda.synchronize()
            for i, (y, o) in enumerate(zip(eager, outs)):
                assert torch.equal(y, o), i
            assert not torch.equal(before[0], outs[0]) and not torch.equal(before[1], outs[1])
            assert torch.equal(before[2], outs[2])
            # a new tile on one member: captured again, same bits as eager
            t = ops[2].tiles()
            t[(1, 0)] = (2, 16)
            ops[2].set_tiles(t)
            grp.apply(M.OP_WAWT, xs, outs)
            torch.cuda.synchronize()
            y2 = ops[2].filtering_procedure(xs[2])
            torch.cuda.synchronize()
            assert torch.equal(y2, outs[2])
    finally:
        for op in ops:
            op.close()


def test_group_rejects_a_context_twice():
    op = M.MGBF(_cfg(*DOMAINS[2]))
    try:
        with pytest.raises(M.MGBFError):
            M.Group([op, op])
    finally:
        op.close()


# 24 regional domains drawn from (n, levels, generations): the full sizes of
# the uneven-batch workload, radial half-width drawn from {6, 8, 10}
UNEVEN_SIZES = [(384, 32, 4), (512, 64, 5), (768, 64, 5), (1024, 127, 6)]


def _uneven_domains(seed=24, count=24):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n, lm, gm = UNEVEN_SIZES[rng.integers(len(UNEVEN_SIZES))]
        out.append((n, lm, gm, int(rng.choice([6, 8, 10]))))
    return out


def test_uneven_batch_full_size_with_amplitude_fields():
    """24 domains of 384^2x32 .. 1024^2x127 in one graph, each with tiles from
    the autotuner; every generation carries a per-point amplitude field
    0.5 + U(0,1) in its a_diff (mg_intstate.f90:161-164). Every replay equals
    the eager calls bit for bit; the first domain of each of the three
    smaller sizes is checked against the FP64 oracle; the largest domain is
    checked for symmetry <x,By> = <Bx,y>."""
    doms = _uneven_domains()
    assert {d[0] for d in doms} == {384, 512, 768, 1024}
    weights = [1.0, 0.5, 0.25, 0.125, 0.0625]
    cfgs = [M.Config.for_grid(n, n, lm, gm=gm, h=h, hz=3, weights=weights) for n, lm, gm, h in doms]
    ops = [M.MGBF(c) for c in cfgs]
    try:
        gen = torch.Generator(device="cuda")
        xs, amps = [], []
        for i, op in enumerate(ops):
            print(f"uneven batch: domain {i} {doms[i]}: inputs and autotune", flush=True)
            gen.manual_seed(500 + i)
            xs.append(torch.randn(op.field_shape(), generator=gen, device=DEV))
            amps.append([0.5 + torch.rand(op.dims.gen_shape(g), generator=gen, device=DEV)
                         for g in range(1, op.dims.gm + 1)])
            for g, a in enumerate(amps[-1], start=1):
                op.set_weights(g, field=a)
            op.autotune()
        eager = [op.filtering_procedure(x) for op, x in zip(ops, xs)]
        torch.cuda.synchronize()
        outs = [torch.empty_like(x) for x in xs]
        with M.Group(ops) as grp:
            for rep in range(2):
                for o in outs:
                    o.fill_(float("nan"))
                grp.apply(M.OP_WAWT, xs, outs)
                torch.cuda.synchronize()
                for i, (y, o) in enumerate(zip(eager, outs)):
                    assert torch.equal(y, o), (rep, i, doms[i])
                ops[rep].filtering_procedure(xs[rep])
                torch.cuda.synchronize()
        for size in (384, 512, 768):
            i = next(i for i, d in enumerate(doms) if d[0] == size)
            print(f"uneven batch: domain {i} {doms[i]} against the FP64 oracle", flush=True)
            cfg, op = cfgs[i], ops[i]
            d = op.dims
            P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=cfg.hx,
                               hy=cfg.hy, hz=cfg.hz, p=2, gm=d.gm, ampl01=cfg.mg_ampl01,
                               ampl02=cfg.mg_ampl02, weights=weights[: d.gm])
            wf = [a.cpu().double().numpy() for a in amps[i]]
            ref = O.wawt(P, xs[i].cpu().double().numpy(), wfield=wf)
            err = np.max(np.abs(eager[i].cpu().double().numpy() - ref)) / np.max(np.abs(ref))
            assert err <= 1e-5, (doms[i], err)
        i = next(i for i, d in enumerate(doms) if d[0] == 1024)
        op = ops[i]
        y2 = torch.randn(op.field_shape(), generator=gen, device=DEV)
        By2 = op.filtering_procedure(y2)
        lhs, rhs = op.dot(xs[i], By2), op.dot(eager[i], y2)
        assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    finally:
        for op in ops:
            op.close()


def test_group_of_other_procedures_replays_like_eager():
    # the streaming 3-D filter (procedures 4 and 5), the separable fast and
    # line filters and the quadratic transfer captured in one graph: every
    # replay equals the eager calls bit for bit
    cfgs = [M.Config.for_grid(64, 48, 16, gm=3, h=4, hz=2, mgbf_proc=4, km3=2),
            M.Config.for_grid(64, 48, 16, gm=3, h=4, hz=3, mgbf_proc=5, km3=2, mgbf_line=True,
                              mg_ampl02=0.7, mg_ampl03=0.5),
            M.Config.for_grid(64, 64, 16, n_ens=2, gm=3, h=6, hz=3, mgbf_proc=8, mg_ampl02=0.4),
            M.Config.for_grid(64, 64, 16, n_ens=2, gm=3, h=6, hz=2, mgbf_proc=7, mgbf_line=True,
                              mg_ampl02=0.5),
            M.Config.for_grid(96, 64, 16, n_ens=2, gm=3, h=4, hz=2, lquart=True)]
    ops = [M.MGBF(c) for c in cfgs]
    try:
        xs = [torch.from_numpy(np.random.default_rng(320 + i).standard_normal(op.field_shape())
                               .astype(np.float32)).to(DEV) for i, op in enumerate(ops)]
        eager = [op.filtering_procedure(x) for op, x in zip(ops, xs)]
        torch.cuda.synchronize()
        outs = [torch.empty_like(x) for x in xs]
        with M.Group(ops) as grp:
            for rep in range(3):
                for o in outs:
                    o.fill_(float("nan"))
                grp.apply(M.OP_WAWT, xs, outs)
                torch.cuda.synchronize()
                for y, o in zip(eager, outs):
                    assert torch.equal(y, o), rep
                ops[rep % len(ops)].filtering_procedure(xs[rep % len(ops)])
                torch.cuda.synchronize()
    finally:
        for op in ops:
            op.close()
