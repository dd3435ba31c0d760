"""Golden fixtures (tests/golden/, written by make_golden.py from the FP64
oracle): the oracle must keep reproducing them (CPU), and the HIP path must
match them within the FP32 bars (GPU)."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_ALL = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))
CASES = [c for c in _ALL if not c.startswith("xfer_")]
XFER = [c for c in _ALL if c.startswith("xfer_")]


def _load(name):
    c = json.load(open(os.path.join(HERE, name + ".json")))
    d = np.load(os.path.join(HERE, name + ".npz"))
    return c, {k: d[k] for k in d.files}


def _oparams(c):
    return O.OracleParams(**{k: v for k, v in c.items() if k not in ("n_ens", "cfg")})


def test_fixtures_present():
    # radial (ensemble, no generation-1 filter, single field), lquart, line
    # procedures 7 and 2, and the 3-D procedures 4 and 5; the transfer with
    # the linear, quadratic and 4-point schemes and both vertical maps
    assert len(CASES) == 8
    assert len(XFER) == 3


def _xfer_params(c):
    P = O.OracleParams(nx=c["nx"], ny=c["ny"], K=c["n_ens"] * c["lm"], lm=c["lm"], nvcol=c["n_ens"], hx=3,
                       hy=3, hz=1, p=2, gm=3)
    return P, (P, c["na"], c["ma"], c["pe"], c["pe"], c["lm_a"]), dict(scheme=c["scheme"], vmap=c["vmap"])


@pytest.mark.parametrize("name", XFER)
def test_oracle_reproduces_transfer_golden(name):
    c, g = _load(name)
    _, args, kw = _xfer_params(c)
    for key, val in (("f2a", O.filt_to_anal(*args, g["x"], **kw)), ("a2f", O.anal_to_filt(*args, g["w"], **kw))):
        assert np.max(np.abs(val - g[key])) <= 1e-12 * np.max(np.abs(g[key])), key


@pytest.mark.gpu
@pytest.mark.parametrize("name", XFER)
def test_hip_transfer_matches_golden(name):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from standalone_mgbf_loc_amd import mgbf as M
    c, g = _load(name)
    cfg = M.Config.for_grid(c["nx"], c["ny"], c["lm"], n_ens=c["n_ens"], gm=3, h=3, hz=1, nm0=c["na"],
                            mm0=c["ma"], lm_a=c["lm_a"], **c["cfg"])
    assert cfg.nxPE == c["pe"] and cfg.nyPE == c["pe"]
    new_map = cfg.l_new_map
    with M.MGBF(cfg) as op:
        x = torch.from_numpy(g["x"].astype(np.float32)).to("cuda:0")
        w = torch.from_numpy(g["w"].astype(np.float32)).to("cuda:0")
        got = {"f2a": op.apply(M.OP_F2A if new_map else M.OP_F2A_ALL, x),
               "a2f": op.apply(M.OP_A2F if new_map else M.OP_A2F_ALL, w)}
        torch.cuda.synchronize()
        for key, y in got.items():
            ref = g[key]
            err = np.max(np.abs(y.cpu().double().numpy().reshape(ref.shape) - ref)) / np.max(np.abs(ref))
            assert err <= 1e-5, (key, err)


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    c, g = _load(name)
    P = _oparams(c)
    x = g["x"]
    q = bool(c.get("lquart", False))
    for key, val in (("wawt", O.wawt(P, x)), ("wt", O.wt(P, x)), ("radial", O.radial(P, x)),
                     ("vertical", O.vertical(P, x)), ("vertical_t", O.vertical(P, x, adjoint=True)),
                     ("restrict", O.restrict(x, quad=q)), ("prolong", O.prolong(g["x_coarse"], quad=q))):
        assert np.max(np.abs(val - g[key])) <= 1e-12 * np.max(np.abs(g[key])), key


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_path_matches_golden(name):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from standalone_mgbf_loc_amd import mgbf as M
    c, g = _load(name)
    n_ens, lm = c["n_ens"], c["lm"]
    over = dict(c.get("cfg", {"l_filt_g1": c["l_filt_g1"]}))
    cfg = M.Config.for_grid(c["nx"], c["ny"], lm, n_ens=n_ens, gm=c["gm"], h=c["hx"], hz=c["hz"],
                            weights=c["weights"], **over)
    with M.MGBF(cfg) as op:
        x = torch.from_numpy(g["x"].astype(np.float32)).to("cuda:0")
        xc = torch.from_numpy(g["x_coarse"].astype(np.float32)).to("cuda:0")
        got = {"wawt": op.filtering_procedure(x), "wt": op.apply(M.OP_WT, x),
               "radial": op.stage(M.STAGE_F, 1, x), "vertical": op.stage(M.STAGE_VF, 1, x),
               "vertical_t": op.stage(M.STAGE_VFT, 1, x), "restrict": op.stage(M.STAGE_R, 1, x),
               "prolong": op.stage(M.STAGE_P, 1, xc)}
        torch.cuda.synchronize()
        for key, y in got.items():
            ref = g[key]
            err = np.max(np.abs(y.cpu().double().numpy().reshape(ref.shape) - ref)) / np.max(np.abs(ref))
            assert err <= 1e-5, (key, err)
