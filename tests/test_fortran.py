"""The Fortran host layer: namelist parsing, the bind(C) record layout, and
(on a GPU) the RBETA_TEST-style driver against the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from standalone_mgbf_loc_amd import mgbf as M
from standalone_mgbf_loc_amd.namelist import read_namelist, write_namelist

LIB = os.path.join(os.path.dirname(M.__file__), "lib")
CHECK = os.path.join(LIB, "mgbf_nml_check")
DRIVER = os.path.join(LIB, "mgbf_driver")


def _need(path):
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (python -m standalone_mgbf_loc_amd.fortran_build)")


@pytest.mark.parametrize("cfg", [
    M.Config.for_grid(192, 192, 32, gm=4),
    M.Config.for_grid(96, 64, 20, n_ens=3, gm=3, h=4, hz=2, weights=[1, .7, .4]),
    M.Config.for_grid(64, 64, 8, gm=3, h=4, hz=2, km2=3, km3=2, mgbf_proc=1),
])
def test_fortran_namelist_and_record_layout_match_c(cfg, tmp_path):
    _need(CHECK)
    nml = tmp_path / "mgbeta.nml"
    write_namelist(cfg, str(nml))
    out = subprocess.run([CHECK, str(nml)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = [int(v) for v in out.stdout.split("dims", 1)[1].split()]
    d = M.derive(cfg)
    assert vals == [d.gm, d.nx, d.ny, d.km, d.km_all, d.km3_all, d.ntaps2, d.ntaps1, d.mg_size]


def test_fortran_reports_configuration_errors(tmp_path):
    _need(CHECK)
    cfg = M.Config.for_grid(192, 192, 32)
    cfg.mg_ampl02 = 10.0   # reach beyond hx = 6
    write_namelist(cfg, str(tmp_path / "bad.nml"))
    out = subprocess.run([CHECK, str(tmp_path / "bad.nml")], capture_output=True, text=True, timeout=60)
    assert out.returncode != 0 and "too large" in out.stdout


def test_python_namelist_roundtrip_and_integer_p(tmp_path):
    cfg = M.Config.for_grid(96, 96, 16, n_ens=2, gm=3)
    write_namelist(cfg, str(tmp_path / "a.nml"))
    assert read_namelist(str(tmp_path / "a.nml")) == cfg
    txt = open(tmp_path / "a.nml").read().replace("p=2,", "p=1.3,")
    open(tmp_path / "b.nml", "w").write(txt)
    with pytest.raises(ValueError):
        read_namelist(str(tmp_path / "b.nml"))


@pytest.mark.gpu
@pytest.mark.parametrize("lquart", [False, True])
def test_fortran_driver_matches_oracle(tmp_path, lquart):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle import oracle as O
    _need(DRIVER)
    cfg = M.Config.for_grid(96, 96, 16, n_ens=2, gm=3, h=6, hz=3, weights=[1.0, 0.5, 0.25],
                            lquart=lquart)
    d = M.derive(cfg)
    write_namelist(cfg, str(tmp_path / "mgbeta.nml"))
    # WORKA(km_a_all, nm0, mm0) in Fortran order == C array (mm0, nm0, km_a_all)
    x = np.random.default_rng(5).standard_normal(d.gen_shape(1))
    x.astype("<f8").tofile(tmp_path / "in.bin")
    out = subprocess.run([DRIVER, "mgbeta.nml", "in.bin", "out.bin"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(d.gen_shape(1))
    P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=6, hy=6,
                       hz=3, p=2, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                       weights=[1.0, 0.5, 0.25], lquart=lquart)
    # the driver rounds the input to FP32 at the device boundary
    ref = O.wawt(P, x.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) <= 1e-5


@pytest.mark.gpu
def test_fortran_driver_on_a_finer_analysis_grid_matches_oracle(tmp_path):
    # nm0 = 1.5 nxPE im_filt and lm_a = 2 lm - 1: anal_to_filt_all2 and
    # filt_to_anal_all2 interpolate (mg_transfer.f90:154-203)
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle import oracle as O
    _need(DRIVER)
    cfg = M.Config.for_grid(64, 48, 8, n_ens=2, gm=3, h=4, hz=2, weights=[1.0, 0.5, 0.25],
                            nm0=96, mm0=72, lm_a=15)
    d = M.derive(cfg)
    write_namelist(cfg, str(tmp_path / "mgbeta.nml"))
    x = np.random.default_rng(6).standard_normal(d.anal_shape())
    x.astype("<f8").tofile(tmp_path / "in.bin")
    out = subprocess.run([DRIVER, "mgbeta.nml", "in.bin", "out.bin"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(d.anal_shape())
    P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=4, hy=4,
                       hz=2, p=2, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                       weights=[1.0, 0.5, 0.25])
    args = (P, 96, 72, cfg.nxPE, cfg.nyPE, 15)
    ref = O.filt_to_anal(*args, O.wawt(P, O.anal_to_filt(*args, x.astype(np.float32).astype(np.float64))))
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) <= 1e-5


@pytest.mark.gpu
def test_fortran_driver_without_the_new_map_matches_oracle(tmp_path):
    # l_new_map=.false.: the driver calls anal_to_filt_all / filt_to_anal_all
    # (RBETA_TEST.f90:48-52, 92-96), here with the linear horizontal
    # interpolation and the lwq vertical map onto 12 analysis levels
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle import oracle as O
    _need(DRIVER)
    cfg = M.Config.for_grid(64, 48, 8, n_ens=2, gm=3, h=4, hz=2, weights=[1.0, 0.5, 0.25],
                            nm0=96, mm0=72, lm_a=12, l_new_map=False, l_lin_horizontal=True)
    d = M.derive(cfg)
    write_namelist(cfg, str(tmp_path / "mgbeta.nml"))
    x = np.random.default_rng(7).standard_normal(d.anal_shape())
    x.astype("<f8").tofile(tmp_path / "in.bin")
    out = subprocess.run([DRIVER, "mgbeta.nml", "in.bin", "out.bin"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(d.anal_shape())
    P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=4, hy=4,
                       hz=2, p=2, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                       weights=[1.0, 0.5, 0.25])
    args = (P, 96, 72, cfg.nxPE, cfg.nyPE, 12)
    kw = dict(scheme=O.LIN, vmap=O.VLWQ)
    xin = x.astype(np.float32).astype(np.float64)
    ref = O.filt_to_anal(*args, O.wawt(P, O.anal_to_filt(*args, xin, **kw)), **kw)
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("proc", [4, 7])
def test_fortran_driver_runs_the_3d_and_line_procedures(tmp_path, proc):
    # filtering_procedure(4) (filtering_rad3) and (7) (filtering_lin2_ens
    # with mgbf_line) through the Fortran module surface
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from oracle import oracle as O
    _need(DRIVER)
    kw = dict(gm=3, h=4, hz=2, weights=[1.0, 0.5, 0.25], mgbf_proc=proc)
    if proc == 4:
        kw.update(km3=2, km2=2)
    else:
        kw.update(n_ens=2, mgbf_line=True)
    cfg = M.Config.for_grid(64, 64, 8, **kw)
    if proc == 7:
        cfg.mg_ampl02 = 0.7
    d = M.derive(cfg)
    write_namelist(cfg, str(tmp_path / "mgbeta.nml"))
    x = np.random.default_rng(6 + proc).standard_normal(d.gen_shape(1))
    x.astype("<f8").tofile(tmp_path / "in.bin")
    out = subprocess.run([DRIVER, "mgbeta.nml", "in.bin", "out.bin"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    y = np.fromfile(tmp_path / "out.bin", dtype="<f8").reshape(d.gen_shape(1))
    P = O.OracleParams(nx=d.nx, ny=d.ny, K=d.km_all, lm=cfg.lm, nvcol=d.km3_all, hx=4, hy=4,
                       hz=2, p=2, gm=d.gm, ampl01=cfg.mg_ampl01, ampl02=cfg.mg_ampl02,
                       weights=[1.0, 0.5, 0.25], line=proc == 7, final_vft=proc == 7,
                       rad3=proc == 4, ampl03=cfg.mg_ampl03)
    ref = O.wawt(P, x.astype(np.float32).astype(np.float64))
    assert np.max(np.abs(y - ref)) / np.max(np.abs(ref)) <= 1e-5
