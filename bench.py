#!/usr/bin/env python3
"""Benchmark of the MGBF localization operator W A W^T on one MI355X.

Workload: batched ensemble localization, FP32, 80 members x 768 x 768 x 64
levels, 4 generations, radial half-width 6, vertical half-width 3, p = 2,
generation weights [1, 1/2, 1/4, 1/8]; member m is an independent seeded
N(0,1) field (seed 100+m). One step = one application of W A W^T to the whole
ensemble (members stacked in k, as the reference stacks them in VALL), with
inputs resident in HBM. The plain form (one member at a time) is timed beside
it. The value is gridpoint.members per second of the batched form.

Also reported: the maximum relative error of member 0 against an FP64 run
(max|y-y64|/max|y64|), the roofline of the dominant kernel (the radial
filter) from device-event timers around its launches, a device copy
bandwidth, and the FP64 CPU restatement of the algorithm timed on member 0
on this host's cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HBM_PEAK_GBPS = 8000.0
METRIC = "gridpoint·members/sec for W·Wᵀ apply (batched vs plain GPU) at rel err ≤1e-5"


def bytes_per_unit(gm: int) -> float:
    """Two-sweep minimum HBM traffic of W A W^T per generation-1 unit (FP32)."""
    S = sum(4.0 ** -(g - 1) for g in range(1, gm + 1))
    return 4.0 * (4 * S + 2 * (S - 1))


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "radial_80x768x768x64_hbm_traffic.md")


def profiled_traffic(M, n, lm, gm):
    """HBM bytes per radial launch measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE on this workload, calibrated against a copy kernel, and the
    tiles of that profile run; recorded in the profile summary as
    `traffic_bytes_per_launch: <workload> <bytes>` and
    `traffic_tiles: <workload> <tiles as JSON>`. (None, None) without it."""
    key = f"{M}x{n}x{n}x{lm}_gm{gm}"
    nbytes, tiles = None, None
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split(None, 2)
                if len(parts) == 3 and parts[1] == key:
                    if parts[0] == "traffic_bytes_per_launch:":
                        nbytes = float(parts[2])
                    elif parts[0] == "traffic_tiles:":
                        tiles = json.loads(parts[2])
    except (OSError, ValueError):
        pass
    return nbytes, tiles


def make_members(torch, M, n, lm, seed0=100):
    x = torch.empty((n, n, M * lm), dtype=torch.float32, device="cuda")
    v = x.view(n, n, M, lm)
    g = torch.Generator(device="cuda")
    for m in range(M):
        g.manual_seed(seed0 + m)
        v[:, :, m, :] = torch.randn((n, n, lm), generator=g, device="cuda", dtype=torch.float32)
    return x


def stream_copy_gbps(torch, nbytes=4 << 30, reps=5):
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    return 2 * nbytes / t / 1e9


# The other filtering procedures, the quadratic inter-generation transfer
# and the analysis <-> filter transfer, each on a workload of its own:
#   ensemble rows (procedures 7, 8 and lquart with 6): the headline workload,
#     80 members x 768x768x64, 4 generations;
#   background rows (procedures 1-5): one field of 8 three-dimensional
#     variables x 64 levels (K = 512) on 1536x1536, 4 generations;
#   transfer: anal_to_filt_all2 then filt_to_anal_all2 between the
#     768x768x64 filter grid and a 1152x1152 analysis grid with 127 levels,
#     16 members;
#   group: 24 regional domains of uneven size in one hipGraph;
#   single: one field 768x768x64, 5 generations, radial half-width 8;
#   diff: the loss J = 1/2 |W A_w W^T (a x) - y|^2 and its gradients with
#     respect to the generation weights and the amplitude field, 16 members
#     x 512x512x64, 5 generations.
ROWS = ("proc1", "proc2", "proc3", "proc4", "proc5", "proc7", "proc8", "lquart", "transfer", "group", "single",
        "diff")
ROW_NAMES = {"proc1": "filtering_rad2_bckg", "proc2": "filtering_lin2_bckg (mgbf_line)",
             "proc3": "filtering_fast_bckg", "proc4": "filtering_rad3", "proc5": "filtering_lin3 (mgbf_line)",
             "proc7": "filtering_lin2_ens (mgbf_line)", "proc8": "filtering_fast_ens",
             "lquart": "filtering_rad2_ens with the quadratic transfer (lquart)",
             "transfer": "anal_to_filt_all2 + filt_to_anal_all2 (4-point horizontal, 2:1 vertical)",
             "group": "uneven batch of regional domains in one hipGraph",
             "single": "one field (filtering_rad2_ens, one member)",
             "diff": "differentiation: J = 1/2 |W A_w W^T (a x) - y|^2, its gradients w.r.t. the generation "
                     "weights w and the amplitude field a"}


def _oracle_params(O, cfg, d):
    """The FP64 oracle's description of a configuration."""
    bckg = cfg.mgbf_proc in (1, 2, 3, 4, 5)
    return O.OracleParams(
        nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=cfg.hx, hy=cfg.hy, hz=cfg.hz,
        p=cfg.p, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
        l_filt_g1=True if bckg else bool(cfg.l_filt_g1), l_vertical_filter=bool(cfg.l_vertical_filter),
        fast=cfg.mgbf_proc in (3, 8), lquart=bool(cfg.lquart), line=cfg.mgbf_proc in (2, 7),
        line_tables=bool(cfg.mgbf_line), final_vft=cfg.mgbf_proc == 7, rad3=cfg.mgbf_proc == 4,
        ampl03=cfg.mg_ampl03, lin3=cfg.mgbf_proc == 5, weights=cfg.weights()[: d.gm])


def box3(t, passes):
    """3x3 box smoothing of a (m, n, K) device field with the half-sample
    mirror (edge replication for a 1-point pad), `passes` times."""
    import torch
    F = torch.nn.functional
    v = t.permute(2, 0, 1).unsqueeze(0)
    for _ in range(passes):
        v = F.avg_pool2d(F.pad(v, (1, 1, 1, 1), mode="replicate"), 3, stride=1)
    return v.squeeze(0).permute(1, 2, 0).contiguous()


def row_cpu_baseline(MG, cfg, row):
    """The FP64 oracle on a bounded sample of a row's workload: one member (or
    one 3-D variable) of the same filter, on 768x768 (384x384 for the 3-D
    radial procedure), timed on this host's cores."""
    import numpy as np
    from oracle import oracle as O
    n = {"proc4": 384, "diff": 512}.get(row, 768)
    kw = {k: getattr(cfg, k) for k in ("mgbf_proc", "mgbf_line", "lquart", "mg_ampl01", "mg_ampl02",
                                       "mg_ampl03", "hx", "hy", "hz", "p", "gm_max")}
    if row == "transfer":
        cs = MG.Config.for_grid(n, n, cfg.lm, n_ens=1, gm=cfg.gm_max, weights=cfg.weights(),
                                nm0=cfg.nm0 * n // (cfg.nxPE * cfg.im_filt),
                                mm0=cfg.mm0 * n // (cfg.nyPE * cfg.jm_filt), lm_a=cfg.lm_a)
    else:
        cs = MG.Config.for_grid(n, n, cfg.lm, n_ens=1, gm=cfg.gm_max, weights=cfg.weights(),
                                km3=1 if cfg.mgbf_proc <= 5 else cfg.km3)
        for k, v in kw.items():
            setattr(cs, k, v)
    d = MG.derive(cs)
    P = _oracle_params(O, cs, d)
    x = np.random.default_rng(1).standard_normal(P.gen_shape(0))
    t0 = time.perf_counter()
    if row == "transfer":
        args = (P, d.na, d.ma, cs.nxPE, cs.nyPE, cs.lm_a)
        w = O.filt_to_anal(*args, x)
        O.anal_to_filt(*args, w)
        units = d.na * d.ma * d.km_a_all
        what = f"one member: {n}x{n}x{cs.lm} <-> {d.na}x{d.ma}x{cs.lm_a}"
    elif row == "diff":
        # the loss and its gradients on one member: B z, B r (dJ/dz), and
        # W^T z, W^T r for the weight gradients
        rng = np.random.default_rng(2)
        a = 1.0 + 0.2 * rng.standard_normal(x.shape)
        yv = rng.standard_normal(x.shape)
        z = a * x
        r = O.wawt(P, z) - yv
        gz = O.wawt(P, r)
        vz, vr = O.mg_split(P, O.wt(P, z)), O.mg_split(P, O.wt(P, r))
        grad_w = [float(np.sum(u * v)) for u, v in zip(vz, vr)]
        grad_a = x * gz
        del grad_w, grad_a
        units = d.nx * d.ny * d.km_all
        what = f"one member on {n}x{n}x{cs.lm}: loss and gradients"
    else:
        O.wawt(P, x)
        units = d.nx * d.ny * d.km_all
        what = f"one {'variable' if cfg.mgbf_proc <= 5 else 'member'} on {n}x{n}x{cs.lm}"
    tc = time.perf_counter() - t0
    return {"value": units / tc, "unit": "gridpoint·members/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{what}, FP64 C restatement with OpenMP, {tc:.2f} s"}


def pin_or_tune(args, op, log, profile_tiles=None):
    """Radial tiles: pinned from --tiles; else, unless --tune, the tiles of
    the traffic profile of this workload (found by the autotuner in that
    session, so that the profiled traffic is this run's); else autotuned
    unless --no-tune (then the defaults, reported as None)."""
    pin = args.tiles
    if pin is None and profile_tiles is not None and not args.tune and not args.no_tune:
        pin = json.dumps(profile_tiles)
        log("[bench] tiles of the traffic profile")
    if pin:
        pinned = {(int(k[1:k.index("k")]), int(k[k.index("k") + 1:])): tuple(v)
                  for k, v in json.loads(pin).items()}
        op.set_tiles(pinned)
        tiles = op.tiles()
        log(f"[bench] pinned tiles: {tiles}")
        return tiles
    if args.no_tune:
        return None
    t0 = time.perf_counter()
    tiles = op.autotune()
    log(f"[bench] autotune {time.perf_counter() - t0:.1f} s: {tiles}")
    return tiles


# The uneven batch: the regional domains of the workload drawn as the GPU
# test draws them (seed 24): sizes from {384^2 x 32, 512^2 x 64, 768^2 x 64,
# 1024^2 x 127} with 4, 5, 5, 6 generations, radial half-width from {6, 8,
# 10}; per-domain N(0,1) inputs and per-point amplitude fields 0.5 + U(0,1)
# on every generation (a_diff_f / a_diff_h); tiles from the autotuner.
UNEVEN_SIZES = [(384, 32, 4), (512, 64, 5), (768, 64, 5), (1024, 127, 6)]


def uneven_domains(seed=24, count=24):
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n, lm, gm = UNEVEN_SIZES[rng.integers(len(UNEVEN_SIZES))]
        out.append((n, lm, gm, int(rng.choice([6, 8, 10]))))
    return out


def run_group(args, torch, MG):
    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    weights = [1.0, 0.5, 0.25, 0.125, 0.0625]
    doms = uneven_domains()
    log(f"[bench] group: {len(doms)} domains {doms}")
    ops, xs, outs, amps = [], [], [], []
    gen = torch.Generator(device="cuda")
    # --tiles: a list of per-domain tile dicts (the "tiles" of an earlier
    # group line), pinned instead of autotuning
    pinned = json.loads(args.tiles) if args.tiles else None
    if pinned is not None and len(pinned) != len(doms):
        raise SystemExit(f"--tiles: {len(pinned)} tile sets for {len(doms)} domains")
    t0 = time.perf_counter()
    for i, (n, lm, gm, h) in enumerate(doms):
        op = MG.MGBF(MG.Config.for_grid(n, n, lm, gm=gm, h=h, hz=3, weights=weights))
        gen.manual_seed(500 + i)
        xs.append(torch.randn(op.field_shape(), generator=gen, device="cuda"))
        outs.append(torch.empty_like(xs[-1]))
        amps.append([0.5 + torch.rand(op.dims.gen_shape(g), generator=gen, device="cuda")
                     for g in range(1, op.dims.gm + 1)])
        for g, a in enumerate(amps[-1], start=1):
            op.set_weights(g, field=a)
        if pinned is not None:
            op.set_tiles({(int(k[1:k.index("k")]), int(k[k.index("k") + 1:])): tuple(v)
                          for k, v in pinned[i].items()})
        elif not args.no_tune:
            op.autotune()
        ops.append(op)
    log(f"[bench] group: contexts and {'pinned tiles' if pinned is not None else 'autotune'} "
        f"{time.perf_counter() - t0:.1f} s")
    units = sum(op.dims.nx * op.dims.ny * op.dims.km_all for op in ops)
    # algorithmic bytes: the two-sweep minimum of each domain plus one read
    # of every generation's weight field
    abytes = 0.0
    for op in ops:
        d = op.dims
        S = sum(4.0 ** -(g - 1) for g in range(1, d.gm + 1))
        abytes += d.nx * d.ny * d.km_all * (bytes_per_unit(d.gm) + 4.0 * S)
    grp = MG.Group(ops)
    for _ in range(args.warmup):
        grp.apply(MG.OP_WAWT, xs, outs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        grp.apply(MG.OP_WAWT, xs, outs)
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - t0) * 1e3 / args.steps
    # the same applications eagerly, one after the other on one stream
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for op, x, y in zip(ops, xs, outs):
            op.filtering_procedure(x, out=y)
    torch.cuda.synchronize()
    ms_eager = (time.perf_counter() - t0) * 1e3 / args.steps
    value = units / (ms_graph / 1e3)
    achieved = abytes / (ms_graph / 1e3) / 1e9
    log(f"[bench] group: graph {ms_graph:.3f} ms, eager {ms_eager:.3f} ms, {value:.4e} units/s")
    out = {
        "metric": METRIC, "row": "group", "value": value, "unit": "gridpoint·members/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_graph, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: per-domain N(0,1) (seed 500+i), amplitude fields 0.5+U(0,1)",
        "config": {"workload": f"uneven batch: {len(doms)} regional domains (n, levels, generations, "
                               f"half-width) {doms}, per-point amplitude fields on every generation, "
                               f"{'pinned' if pinned is not None else 'autotuned'} tiles, one hipGraph"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                     "kernel": "the whole graph (two-sweep minimum bytes of every domain plus one "
                               "read of every weight field, over the graph's time)"},
        "cpu_baseline": None,
        "eager_ms_per_step": round(ms_eager, 4),
        "graph_over_eager": round(ms_eager / ms_graph, 3),
        "tiles": [{f"g{g}k{k}": list(v) for (g, k), v in op.tiles().items()} for op in ops],
    }
    grp.close()
    for op in ops:
        op.close()
    print(json.dumps(out), flush=True)


def run_row(args, torch, MG):
    row = args.row
    weights = [1.0, 0.5, 0.25, 0.125, 0.0625]
    gm = args.gm

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    tiles = None
    if row == "transfer":
        M, n, lm, na, la = 16, 768, 64, 1152, 127
        cfg = MG.Config.for_grid(n, n, lm, n_ens=M, gm=gm, h=6, hz=3, p=2, weights=weights,
                                 nm0=na, mm0=na, lm_a=la)
        workload = (f"{ROW_NAMES[row]}: {M} members, filter grid {n}x{n}x{lm}, "
                    f"analysis grid {na}x{na}x{la}")
    elif row == "single":
        # one field 768x768x64, 5 generations, radial half-width 8, vertical
        # 4, p = 2, weights 2^-g; seeded N(0,1) smoothed by a 3x3 box
        M, n, lm = 1, 768, 64
        cfg = MG.Config.for_grid(n, n, lm, gm=5, h=8, hz=4, p=2, weights=[2.0 ** -g for g in range(5)])
        workload = f"{ROW_NAMES[row]}: {n}x{n}x{lm}, 5 generations, h=8, hz=4, p=2, weights 2^-g"
    elif row == "diff":
        # 16 members x 512x512x64, 5 generations, learnable weights
        # [1, 1/2, 1/4, 1/8, 1/16], amplitude field a = 1 + 0.2 N(0,1) smoothed
        M, n, lm = 16, 512, 64
        cfg = MG.Config.for_grid(n, n, lm, n_ens=M, gm=5, h=6, hz=3, p=2, weights=weights)
        workload = f"{ROW_NAMES[row]}: {M} members x {n}x{n}x{lm}, 5 generations, h=6, hz=3, p=2"
    elif row in ("proc7", "proc8", "lquart"):
        M, n, lm = 80, 768, 64
        kw = {"proc7": dict(mgbf_proc=7, mgbf_line=True), "proc8": dict(mgbf_proc=8),
              "lquart": dict(lquart=True)}[row]
        cfg = MG.Config.for_grid(n, n, lm, n_ens=M, gm=gm, h=6, hz=3, p=2, weights=weights, **kw)
        workload = f"{ROW_NAMES[row]}: {M} members x {n}x{n}x{lm}, {gm} generations, h=6, hz=3, p=2"
    else:
        proc = int(row[4:])
        M, n, lm = 1, 1536, 64
        cfg = MG.Config.for_grid(n, n, lm, n_ens=1, gm=gm, h=6, hz=3, p=2, weights=weights,
                                 mgbf_proc=proc, km3=8, km2=args.km2, mgbf_line=proc in (2, 5))
        extra = f" plus {args.km2} 2-D variables" if args.km2 else ""
        workload = (f"{ROW_NAMES[row]}: one field of 8 variables x {lm} levels{extra} on {n}x{n}, "
                    f"{gm} generations, h=6, hz=3, p=2")
    if row in ("proc2", "proc3", "proc7", "proc8"):
        # the line and fast filters take mg_ampl02 as their factor: 0.4 is
        # about the smallest value whose fast factor still fits the halo of 6
        # that def_mg_weights checks for every procedure, so the widest
        # filters (line 7x7, fast 13x13)
        cfg.mg_ampl02 = 0.4
        workload += ", mg_ampl02=0.4"
    log(f"[bench] row {row}: {workload}")
    op = MG.MGBF(cfg)
    d = op.dims
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    data = "synthetic: N(0,1), seed 7"
    if row == "transfer":
        w = torch.randn(d.anal_shape(), generator=g, device="cuda", dtype=torch.float32)
        v = torch.empty(op.field_shape(), dtype=torch.float32, device="cuda")
        w2 = torch.empty_like(w)
        units = d.na * d.ma * d.km_a_all
        # algorithmic bytes of the pair: each map reads its input and writes
        # its output once
        pair_bytes = 2 * 4.0 * (d.na * d.ma * d.km_a_all + d.nx * d.ny * d.km_all)

        def step():
            op.anal_to_filt(w, out=v)
            op.filt_to_anal(v, out=w2)
    elif row == "diff":
        from standalone_mgbf_loc_amd.autograd import loss_and_grads
        x = torch.randn(op.field_shape(), generator=g, device="cuda", dtype=torch.float32)
        yv = torch.randn(op.field_shape(), generator=g, device="cuda", dtype=torch.float32)
        a = 1.0 + 0.2 * box3(torch.randn(op.field_shape(), generator=g, device="cuda"), 2)
        ws = [torch.tensor(v, dtype=torch.float64, device="cuda") for v in weights]
        units = d.nx * d.ny * d.km_all
        tiles = pin_or_tune(args, op, log)
        data = ("synthetic: x, y ~ N(0,1), a = 1 + 0.2 N(0,1) smoothed by two 3x3 box passes, seed 7; "
                "weights [1, 1/2, 1/4, 1/8, 1/16]")

        def step():
            loss_and_grads(op, a, x, yv, ws)
    else:
        x = torch.randn(op.field_shape(), generator=g, device="cuda", dtype=torch.float32)
        if row == "single":
            x = box3(x, 1)
            data = "synthetic: N(0,1) smoothed by a 3x3 box, seed 7"
        y = torch.empty_like(x)
        units = d.nx * d.ny * d.km_all
        tiles = pin_or_tune(args, op, log)

        def step():
            op.filtering_procedure(x, out=y)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    op.timers(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        step()
    e1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = op.read_timers()
    op.timers(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = units / (ms_step / 1e3)
    if row == "transfer":
        # events on the launch stream around the timed steps: nothing else runs there
        k_ms = e0.elapsed_time(e1)
        if os.environ.get("MGBF_XFER_SPLIT") == "1":
            per_step = 4
            kernel = "xfer_tile_kernel + kwin_kernel (A2F and F2A, the vertical map a pass of its own)"
        else:
            per_step = 2
            kernel = ("xfer_tile_kernel with the vertical map fused (A2F: the map first, F2A: the map "
                      "last; one launch each)")
        k_n, k_bytes = per_step * args.steps, pair_bytes * args.steps
        avg = k_ms / k_n
        achieved = k_bytes / (k_ms / 1e3) / 1e9
    else:
        k_ms = phases["hfiltT"][0] + phases["hfilt"][0]
        k_n = phases["hfiltT"][1] + phases["hfilt"][1]
        k_bytes = phases["hfiltT"][2] + phases["hfilt"][2]
        kernel = ("ball3_kernel (3-D filter, F and F^T on every generation)" if row in ("proc4", "proc5")
                  else "radial filter passes (F and F^T on every generation; radial_mx_kernel where "
                       "the tiles say 16, radial_kernel elsewhere)")
        avg = k_ms / max(k_n, 1)
        achieved = k_bytes / (k_ms / 1e3) / 1e9
    phase_out = {k: {"ms": round(v_[0] / args.steps, 4), "launches": v_[1] // args.steps}
                 for k, v_ in phases.items()}
    log(f"[bench] {ms_step:.3f} ms/step, {value:.4e} units/s, phases {phase_out}")
    out = {
        "metric": METRIC, "row": row, "value": value, "unit": "gridpoint·members/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": data,
        "config": {"workload": workload},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": kernel,
                     "launches": k_n,
                     "avg_launch_ms": round(avg, 4),
                     "algorithmic_bytes_per_launch": k_bytes / max(k_n, 1)},
        "cpu_baseline": None,
        "phases": phase_out,
        "tiles": {f"g{g}k{k}": list(v) for (g, k), v in tiles.items()} if tiles else None,
    }
    op.close()
    if not args.no_cpu:
        out["cpu_baseline"] = row_cpu_baseline(MG, cfg, row)
        log(f"[bench] cpu: {out['cpu_baseline']['sample']}")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--members", type=int, default=80)
    ap.add_argument("--n", type=int, default=768)
    ap.add_argument("--levels", type=int, default=64)
    ap.add_argument("--gm", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default radial tiles")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the radial tiles instead of pinning those of the traffic profile")
    ap.add_argument("--tiles", default=None,
                    help='pin radial tiles instead of autotuning, JSON as printed under "tiles", '
                         'e.g. \'{"g1k0": [4, 384], ...}\' (for --row group a list of such, one per domain)')
    ap.add_argument("--no-cpu", action="store_true", help="skip the FP64 CPU leg (error check and baseline)")
    ap.add_argument("--km2", type=int, default=0,
                    help="2-D variables after the 3-D ones in the background rows proc1..proc5")
    ap.add_argument("--row", choices=ROWS, default=None,
                    help="measure one of the other procedures or the analysis transfer instead "
                         "(one JSON line with the same keys; no plain or CPU leg)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(the operator would shard by member with no exchange between GPUs)"}))
        sys.exit(2)

    import torch
    from standalone_mgbf_loc_amd import mgbf as MG

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)

    if args.row == "group":
        run_group(args, torch, MG)
        return
    if args.row:
        run_row(args, torch, MG)
        return

    M, n, lm, gm = args.members, args.n, args.levels, args.gm
    weights = [1.0, 0.5, 0.25, 0.125, 0.0625]
    cfg = MG.Config.for_grid(n, n, lm, n_ens=M, gm=gm, h=6, hz=3, p=2, weights=weights)
    units = M * n * n * lm

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    log(f"[bench] workload {M} x {n}x{n}x{lm}, gm={gm}; generating members")
    x = make_members(torch, M, n, lm)
    y = torch.empty_like(x)
    op = MG.MGBF(cfg)
    assert op.dims.gm == gm
    traffic_bytes, traffic_tiles = profiled_traffic(M, n, lm, gm)
    tiles = pin_or_tune(args, op, log, traffic_tiles)

    # ---- batched: one application over all members ----
    for _ in range(args.warmup):
        op.filtering_procedure(x, out=y)
    torch.cuda.synchronize()
    # the library's per-phase device timers (HIP events on the launch stream
    # around every kernel of a phase) run through the timed steps
    op.timers(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        op.filtering_procedure(x, out=y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    phases = op.read_timers()
    g1 = op.read_timers_gen(1)
    op.timers(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = units / (ms_step / 1e3)
    log(f"[bench] batched: {ms_step:.3f} ms/step, {value:.4e} units/s")
    r_ms = phases["hfiltT"][0] + phases["hfilt"][0]
    r_n = phases["hfiltT"][1] + phases["hfilt"][1]
    r_bytes = phases["hfiltT"][2] + phases["hfilt"][2]
    achieved = r_bytes / (r_ms / 1e3) / 1e9
    phase_out = {k: {"ms": round(v[0] / args.steps, 4), "launches": v[1] // args.steps,
                     "GBps": round(v[2] / (v[0] / 1e3) / 1e9, 1) if v[0] > 0 else None}
                 for k, v in phases.items()}
    log(f"[bench] phases: {phase_out}")

    # ---- plain: one member at a time ----
    plain = None
    if not args.no_plain:
        cfg1 = MG.Config.for_grid(n, n, lm, n_ens=1, gm=gm, h=6, hz=3, p=2, weights=weights)
        op1 = MG.MGBF(cfg1)
        if not args.no_tune:
            op1.autotune()
        xp = x.view(n, n, M, lm).permute(2, 0, 1, 3).contiguous()
        yp = torch.empty_like(xp)
        for m in range(min(M, 2)):
            op1.filtering_procedure(xp[m], out=yp[m])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            for m in range(M):
                op1.filtering_procedure(xp[m], out=yp[m])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        p_ms = (t1 - t0) * 1e3 / args.plain_steps
        diff = (yp.permute(1, 2, 0, 3).reshape(n, n, M * lm) - y).abs().max().item()
        plain = {"value": units / (p_ms / 1e3), "ms_per_step": round(p_ms, 3),
                 "max_abs_diff_vs_batched": diff}
        log(f"[bench] plain: {p_ms:.3f} ms/step ({plain['value']:.4e} units/s), diff {diff:.3e}")
        op1.close()
        del xp, yp

    gbps_copy = stream_copy_gbps(torch)
    log(f"[bench] device copy: {gbps_copy:.1f} GB/s")

    # ---- FP64 CPU leg on member 0: error check and CPU baseline ----
    relerr = None
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        P = O.OracleParams(nx=n, ny=n, K=lm, lm=lm, nvcol=1, hx=6, hy=6, hz=3, p=2, gm=gm,
                           ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02, weights=weights[:gm])
        x0 = x.view(n, n, M, lm)[:, :, 0, :].double().cpu().numpy()
        y0 = y.view(n, n, M, lm)[:, :, 0, :].double().cpu().numpy()
        t0 = time.perf_counter()
        ref = O.wawt(P, x0)
        tc = time.perf_counter() - t0
        import numpy as np
        relerr = float(np.max(np.abs(y0 - ref)) / np.max(np.abs(ref)))
        cpu = {"value": n * n * lm / tc, "unit": "gridpoint·members/s", "cores": O.num_threads(),
               "kind": "port",
               "sample": f"member 0 of the workload ({n}x{n}x{lm}, gm={gm}) by the FP64 C "
                         f"restatement with OpenMP, {tc:.2f} s"}
        log(f"[bench] cpu: {tc:.2f} s, rel err {relerr:.3e}")

    # the traffic figure comes from a profile run of this workload; say which
    # tiles it ran and whether this run's tiles are the same
    run_tiles = {f"g{g}k{k}": list(v) for (g, k), v in tiles.items()} if tiles else None
    traffic_source = None
    if traffic_bytes is not None:
        same = traffic_tiles is not None and run_tiles is not None and \
            all(traffic_tiles.get(k) == v for k, v in run_tiles.items())
        traffic_source = (f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, calibrated, from "
                          f"{os.path.relpath(TRAFFIC_PROFILE, os.path.dirname(os.path.abspath(__file__)))} "
                          f"(a profile run with tiles {json.dumps(traffic_tiles)}); this run's tiles "
                          f"{'are the same' if same else 'differ'}")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "gridpoint·members/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: member m ~ N(0,1), seed 100+m",
        "config": {"workload": f"batched ensemble localization W A W^T: {M} members x {n}x{n}x{lm}, "
                               f"{gm} generations, radial half-width 6, vertical half-width 3, p=2, "
                               f"weights [1,1/2,1/4,1/8]",
                   "members": M, "nx": n, "ny": n, "levels": lm, "gm": gm},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4),
                     "traffic": traffic_bytes,
                     "traffic_source": traffic_source,
                     "kernel": "radial filter passes (F and F^T on every generation; matrix-core "
                               "engine radial_mx_kernel where the tiles say 16, radial_kernel elsewhere)",
                     "launches": r_n, "avg_launch_ms": round(r_ms / max(r_n, 1), 4),
                     "algorithmic_bytes_per_launch": r_bytes / max(r_n, 1),
                     # generation 1's two radial passes on their own (the same
                     # event timers, launches on generation 1 only)
                     "gen1_passes": {
                         name: {"avg_launch_ms": round(g1[ph][0] / max(g1[ph][1], 1), 4),
                                "algorithmic_bytes_per_launch": g1[ph][2] / max(g1[ph][1], 1),
                                "achieved": round(g1[ph][2] / (g1[ph][0] / 1e3) / 1e9, 1) if g1[ph][0] else None,
                                "frac": round(g1[ph][2] / (g1[ph][0] / 1e3) / 1e9 / HBM_PEAK_GBPS, 4)
                                if g1[ph][0] else None}
                         for name, ph in (("first (F^T, A, restriction)", "hfiltT"),
                                          ("second (F, prolongation; the vertical band a pass of its own)"
                                           if tiles and tiles.get((1, 2), (1, 1))[0] == 0 else
                                           "second (F, prolongation, vertical band)", "hfilt"))}},
        "cpu_baseline": cpu,
        "max_rel_err_vs_fp64": relerr,
        "plain": plain,
        "batched_over_plain": round(value / plain["value"], 3) if plain else None,
        "operator_hbm_frac": round(bytes_per_unit(gm) * value / 1e9 / HBM_PEAK_GBPS, 4),
        "device_copy_GBps": round(gbps_copy, 1),
        "phases": phase_out,
        "tiles": {f"g{g}k{k}": list(v) for (g, k), v in tiles.items()} if tiles else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
