"""Writes the golden fixtures of tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Each fixture holds seeded inputs, the configuration and the FP64 outputs of
W A W^T, W^T and the single stages at small sizes. They are regression
vectors: the oracle itself is pinned by tests/test_oracle.py.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402

CASES = {
    # name: (nx, ny, n_ens, lm, gm, h, hz, weights, l_filt_g1)
    "ens_32x32_gm3_h3": (32, 32, 2, 2, 3, 3, 1, [1.0, 0.6, 0.3], True),
    "ens_24x24_gm2_h6_nofilt1": (24, 24, 2, 2, 2, 6, 1, [1.0, 0.5], False),
    "single_32x24_gm3_h4": (32, 24, 1, 3, 3, 4, 2, [1.0, 0.5, 0.25], True),
}


def config(nx, ny, n_ens, lm, gm, h, hz, weights, filt_g1):
    return dict(nx=nx, ny=ny, K=n_ens * lm, lm=lm, nvcol=n_ens, hx=h, hy=h, hz=hz, p=2, gm=gm,
                ampl01=(hz + 1) ** 2 / 7.0, ampl02=(h + 1) ** 2 / 8.0, weights=weights,
                l_filt_g1=filt_g1, n_ens=n_ens)


def main():
    for seed, (name, args) in enumerate(CASES.items()):
        c = config(*args)
        P = O.OracleParams(**{k: v for k, v in c.items() if k != "n_ens"})
        rng = np.random.default_rng(1000 + seed)
        x = rng.standard_normal(P.gen_shape(0)).astype(np.float32).astype(np.float64)
        x1 = rng.standard_normal(P.gen_shape(1)).astype(np.float32).astype(np.float64)
        out = dict(x=x, x_coarse=x1, wawt=O.wawt(P, x), wt=O.wt(P, x),
                   radial=O.radial(P, x), vertical=O.vertical(P, x),
                   vertical_t=O.vertical(P, x, adjoint=True),
                   restrict=O.restrict(x), prolong=O.prolong(x1))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(c, f, indent=1)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
