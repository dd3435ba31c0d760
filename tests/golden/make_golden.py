"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Each fixture holds seeded inputs, the configuration and the FP64 outputs of
W A W^T, W^T and the single stages at small sizes; the xfer_* fixtures hold
both directions of the analysis <-> filter transfer. They are regression
vectors: the oracle itself is pinned by tests/test_oracle.py. The json file
holds the oracle's parameters, the namelist values the library is built
from ("cfg", on top of Config.for_grid) and, for the stages, whether the
generation transfer is the quadratic one ("quad").
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

# name: grid, stack and filter of the case; the seed is the case's position
CASES = {
    "ens_32x32_gm3_h3": dict(nx=32, ny=32, n_ens=2, lm=2, gm=3, h=3, hz=1, weights=[1.0, 0.6, 0.3]),
    "ens_24x24_gm2_h6_nofilt1": dict(nx=24, ny=24, n_ens=2, lm=2, gm=2, h=6, hz=1, weights=[1.0, 0.5],
                                     l_filt_g1=False),
    "single_32x24_gm3_h4": dict(nx=32, ny=24, n_ens=1, lm=3, gm=3, h=4, hz=2, weights=[1.0, 0.5, 0.25]),
    # filtering_rad2_ens with the quadratic transfer (lquart)
    "lquart_ens_32x32_gm3_h3": dict(nx=32, ny=32, n_ens=2, lm=4, gm=3, h=3, hz=1,
                                    weights=[1.0, 0.6, 0.3], lquart=True),
    # filtering_lin2_ens (mgbf_proc 7, mgbf_line): triad line filters, Vf^T at the end
    "line_ens_32x24_gm3_h6": dict(nx=32, ny=24, n_ens=2, lm=4, gm=3, h=6, hz=2,
                                  weights=[1.0, 0.5, 0.25], proc=7, ampl02=0.5),
    # filtering_lin2_bckg (mgbf_proc 2) with 2-D variables
    "line_bckg_32x32_gm3_km2": dict(nx=32, ny=32, n_ens=1, lm=3, km3=2, km2=2, gm=3, h=4, hz=2,
                                    weights=[1.0, 0.5, 0.25], proc=2, ampl02=0.75),
    # filtering_rad3 (mgbf_proc 4): 3-D variables over (i, j, L), 2-D variables
    "rad3_32x24_gm3_km2": dict(nx=32, ny=24, n_ens=1, lm=5, km3=2, km2=3, gm=3, h=4, hz=2,
                               weights=[1.0, 0.5, 0.25], proc=4),
    # filtering_lin3 (mgbf_proc 5, mgbf_line): hexad lines of the 3-D variables
    "lin3_32x24_gm3_km2": dict(nx=32, ny=24, n_ens=1, lm=6, km3=2, km2=2, gm=3, h=4, hz=3,
                               weights=[1.0, 0.5, 0.25], proc=5, ampl02=0.7, ampl03=0.5),
}


def config(nx, ny, n_ens, lm, gm, h, hz, weights, l_filt_g1=True, lquart=False, proc=6,
           km3=1, km2=0, ampl02=None, ampl03=None):
    ampl01 = (hz + 1) ** 2 / 7.0
    a2 = (h + 1) ** 2 / 8.0 if ampl02 is None else ampl02
    a3 = (min(h, hz) + 1) ** 2 / 9.0 if ampl03 is None else ampl03
    bckg = proc in (1, 2, 3, 4, 5)
    c = dict(nx=nx, ny=ny, K=n_ens * (km2 + lm * km3), lm=lm, nvcol=n_ens * km3, hx=h, hy=h, hz=hz,
             p=2, gm=gm, ampl01=ampl01, ampl02=a2, weights=weights,
             l_filt_g1=True if bckg else l_filt_g1, n_ens=n_ens, lquart=lquart,
             line=proc in (2, 7), line_tables=True, final_vft=proc == 7, rad3=proc == 4, ampl03=a3,
             lin3=proc == 5)
    cfg = dict(mgbf_proc=proc, lquart=lquart, mgbf_line=proc in (2, 5, 7), km2=km2, km3=km3,
               l_filt_g1=l_filt_g1, mg_ampl02=a2, mg_ampl03=a3)
    return c, cfg


# analysis <-> filter transfer: filter grid (nx, ny) with n_ens members of lm
# levels, analysis grid (na, ma) with lm_a levels, over a pe x pe PE grid;
# horizontal scheme and vertical map as in oracle.filt_to_anal, and the
# namelist values that select them
XFER_CASES = {
    "xfer_lin_all2_32x24_to_48x36": dict(nx=32, ny=24, n_ens=2, lm=4, na=48, ma=36, lm_a=7, pe=4,
                                         scheme=O.LIN, vmap=O.VLIN, cfg=dict(l_lin_horizontal=True)),
    "xfer_quad_lwq_32x32_to_64x64": dict(nx=32, ny=32, n_ens=1, lm=5, na=64, ma=64, lm_a=12, pe=4,
                                         scheme=O.QUAD, vmap=O.VLWQ,
                                         cfg=dict(l_quad_horizontal=True, l_new_map=False)),
    "xfer_lsqr_lwq_48x32_to_72x48": dict(nx=48, ny=32, n_ens=3, lm=4, na=72, ma=48, lm_a=10, pe=4,
                                         scheme=O.LSQR, vmap=O.VLWQ, cfg=dict(l_new_map=False)),
}


def transfer_main():
    for seed, (name, a) in enumerate(XFER_CASES.items()):
        K = a["n_ens"] * a["lm"]
        P = O.OracleParams(nx=a["nx"], ny=a["ny"], K=K, lm=a["lm"], nvcol=a["n_ens"], hx=3, hy=3, hz=1,
                           p=2, gm=3)
        rng = np.random.default_rng(2000 + seed)
        x = rng.standard_normal(P.gen_shape(0)).astype(np.float32).astype(np.float64)
        w = rng.standard_normal((a["ma"], a["na"], a["n_ens"] * a["lm_a"])).astype(np.float32).astype(np.float64)
        args = (P, a["na"], a["ma"], a["pe"], a["pe"], a["lm_a"])
        kw = dict(scheme=a["scheme"], vmap=a["vmap"])
        out = dict(x=x, w=w, f2a=O.filt_to_anal(*args, x, **kw), a2f=O.anal_to_filt(*args, w, **kw))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(a, f, indent=1)
        print(name, {k: v.shape for k, v in out.items()})


def main():
    transfer_main()
    for seed, (name, args) in enumerate(CASES.items()):
        c, cfg = config(**args)
        P = O.OracleParams(**{k: v for k, v in c.items() if k != "n_ens"})
        rng = np.random.default_rng(1000 + seed)
        x = rng.standard_normal(P.gen_shape(0)).astype(np.float32).astype(np.float64)
        x1 = rng.standard_normal(P.gen_shape(1)).astype(np.float32).astype(np.float64)
        q = bool(c["lquart"])
        out = dict(x=x, x_coarse=x1, wawt=O.wawt(P, x), wt=O.wt(P, x),
                   radial=O.radial(P, x), vertical=O.vertical(P, x),
                   vertical_t=O.vertical(P, x, adjoint=True),
                   restrict=O.restrict(x, quad=q), prolong=O.prolong(x1, quad=q))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(dict(c, cfg=cfg), f, indent=1)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
