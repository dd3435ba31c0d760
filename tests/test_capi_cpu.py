"""C ABI checks that need no GPU: the library loads, exports every entry
point include/mgbf.h declares, and derives/validates configurations the way
init_mg_parameter does (src.mgbf/mg_parameter.f90:179-635)."""
import ctypes
import dataclasses
import os
import re

import numpy as np
import pytest

from standalone_mgbf_loc_amd import mgbf as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "mgbf.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mgbf_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = M.load_library()
    names = _declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(M.EXPORTED_SYMBOLS)


def test_version_and_default_config():
    L = M.load_library()
    assert L.mgbf_version().decode().startswith("mgbf")
    c = M._CConfig()
    assert L.mgbf_config_default(ctypes.byref(c)) == 0
    assert c.mgbf_proc == 6 and c.hx == 6 and c.hz == 3 and c.p == 2


def test_derive_config1_needs_8x8_pes_for_4_generations():
    # 192x192x32 with 4 generations; 2x2 PEs would cap gm at 2 (def_maxgen)
    cfg = M.Config.for_grid(192, 192, 32, gm=4)
    d = M.derive(cfg)
    assert (d.gm, d.nx, d.ny, d.km_all) == (4, 192, 192, 32)
    assert d.ntaps2 == 145 and d.ntaps1 == 7
    assert d.mg_size == 32 * (192 * 192 + 96 * 96 + 48 * 48 + 24 * 24)
    cfg2 = M.Config.for_grid(192, 192, 32, gm=4)
    cfg2.nxPE = cfg2.nyPE = 2
    cfg2.im_filt = cfg2.jm_filt = 96
    assert M.derive(cfg2).gm == 2


@pytest.mark.parametrize("pe,gm", [(1, 2), (2, 2), (4, 3), (5, 4), (8, 4), (9, 5), (16, 5), (17, 6), (32, 6)])
def test_def_maxgen_table(pe, gm):
    assert M._def_maxgen(pe, pe) == gm


def test_ensemble_stacks_members_in_k():
    cfg = M.Config.for_grid(768, 768, 64, n_ens=80, gm=4)
    d = M.derive(cfg)
    assert d.km_all == 80 * 64 and d.km3_all == 80
    assert d.field_size == 80 * 64 * 768 * 768


def _err(cfg):
    with pytest.raises(M.MGBFError) as e:
        M.derive(cfg)
    return e.value


def test_reach_beyond_halo_is_refused():
    cfg = M.Config.for_grid(192, 192, 32)
    cfg.mg_ampl02 = 64 / 8.0  # reach 8 > hx 6
    assert "too large" in str(_err(cfg))


def test_nm0_divisibility_is_checked():
    cfg = M.Config.for_grid(192, 192, 32)
    cfg.nm0 = 190
    assert "not divisible" in str(_err(cfg))


@pytest.mark.parametrize("key,val", [("mgbf_proc", 0), ("mgbf_proc", 9)])
def test_options_not_built_are_refused(key, val):
    cfg = M.Config.for_grid(192, 192, 32)
    setattr(cfg, key, val)
    assert _err(cfg).status == -4


def test_l_loc_is_accepted_and_changes_no_dimension():
    """l_loc only sets up the MPI tables and weight arrays of the uncalled
    filtering_rad2_z_loc_g3/g4 (mg_filtering.f90:1510-1830), so the operator
    is the one without it."""
    cfg = M.Config.for_grid(192, 192, 32)
    base = M.derive(cfg)
    cfg.l_loc = True
    cfg.l_loc_vertical = True
    d = M.derive(cfg)
    assert dataclasses.astuple(d) == dataclasses.astuple(base)


def test_ensemble_vertical_filter_with_2d_variables_is_refused():
    cfg = M.Config.for_grid(192, 192, 32, n_ens=2)
    cfg.km2 = 4
    assert _err(cfg).status == -2
    cfg.mgbf_proc = 1
    cfg.n_ens = 1
    assert M.derive(cfg).km_all == 36


def test_init_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(M.MGBFError):
        M.MGBF(M.Config.for_grid(192, 192, 32))


@pytest.mark.parametrize("proc", [2, 7])
@pytest.mark.parametrize("ampl02,p", [(0.5, 2), (2.0, 2), (0.6, 1), (1.0, 3)])
def test_line_procedure_taps_follow_the_oracle_lines(proc, ampl02, p):
    # the library's line filter is the product of the x and y lines of the
    # oracle's init_mg_line restatement (the third line idle); without
    # mgbf_line it is the identity
    from oracle import oracle as O
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=6, hz=2, mgbf_proc=proc, p=p)
    cfg.mg_ampl02 = ampl02
    assert M.derive(cfg).ntaps2 == 1
    cfg.mgbf_line = True
    P = O.OracleParams(nx=64, ny=64, K=8, lm=8, nvcol=1, hx=6, hy=6, hz=2, p=p, ampl02=ampl02,
                       line=True)
    hx = O.line(P, 0)[1]
    hy = O.line(P, 1)[1]
    assert O.line(P, 0)[0] == (1, 0) and O.line(P, 1)[0] == (0, 1)
    assert M.derive(cfg).ntaps2 == (2 * max(hx, 0) + 1) * (2 * max(hy, 0) + 1)


def test_line_procedure_needs_exponent_in_the_moment_table():
    # every filter fits its halo at p = 7; only the moment table refuses it
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=6, hz=3, mgbf_proc=7, p=7, mgbf_line=True,
                            mg_ampl01=0.5, mg_ampl02=1.0, mg_ampl03=0.5)
    assert "inimomtab" in str(_err(cfg))
    cfg.mgbf_line = False
    assert M.derive(cfg).ntaps2 == 1


@pytest.mark.parametrize("key,val,msg", [("mg_ampl02", 0.3, "getlinesum1"), ("mg_ampl03", 3.0, "getlinesum3"),
                                         ("hz", 1, "getlinesum")])
def test_every_filter_of_the_namelist_must_fit_its_halo(key, val, msg):
    # def_mg_weights normalizes the fast factors, the vertical, the 2-D and
    # the 3-D filters whatever the procedure; each stops beyond its halo
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=6, hz=3)
    assert M.derive(cfg).gm == 3
    setattr(cfg, key, val)
    assert msg in str(_err(cfg))


@pytest.mark.parametrize("ampl03,p,hz", [(1.0, 2, 2), (25 / 9, 2, 4), (4 / 7, 1, 2)])
def test_rad3_procedure_taps_follow_the_oracle(ampl03, p, hz):
    from oracle import oracle as O
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=4, hz=hz, mgbf_proc=4, km3=2, p=p)
    cfg.mg_ampl03 = ampl03
    P = O.OracleParams(nx=64, ny=64, K=16, lm=8, nvcol=2, hx=4, hy=4, hz=hz, p=p, rad3=True,
                       ampl03=ampl03)
    gx, gy, gz, w = O.taps3(P)
    d = M.derive(cfg)
    assert d.ntaps3 == len(w) and d.km_all == 16


def test_rad3_procedure_needs_more_levels_than_hz():
    cfg = M.Config.for_grid(64, 64, 2, gm=3, h=4, hz=2, mgbf_proc=4, km3=2, l_vertical_filter=False)
    assert "lm must exceed hz" in str(_err(cfg))


@pytest.mark.parametrize("ampl03,hz", [(0.5, 3), (0.25, 4), (1.0, 2)])
def test_lin3_procedure_taps_follow_the_oracle_hexad(ampl03, hz):
    # the product of the x, y and level lines of the oracle's hextform
    from oracle import oracle as O
    cfg = M.Config.for_grid(64, 64, 8, gm=3, h=4, hz=hz, mgbf_proc=5, km3=2, mgbf_line=True)
    cfg.mg_ampl02 = 0.7
    cfg.mg_ampl03 = ampl03
    P = O.OracleParams(nx=64, ny=64, K=16, lm=8, nvcol=2, hx=4, hy=4, hz=hz, p=2, lin3=True,
                       ampl03=ampl03, ampl02=0.7)
    lines = O.hexlines(P)
    assert [g for _, g, _, _ in lines[:3]] == [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    assert M.derive(cfg).ntaps3 == int(np.prod([2 * h + 1 for _, _, h, _ in lines[:3]]))
    cfg.mgbf_line = False
    assert M.derive(cfg).ntaps3 == 1


@pytest.mark.parametrize("kw,msg", [
    (dict(lm_a=10), "odd lm_a"),                                       # _all2: even lm_a
    (dict(lm_a=9), "odd lm_a"),                                        # _all2: lm_a > 2 lm - 1
    (dict(lm_a=3), "lm_a=3 < lm=4"),
    (dict(l_new_map=False, km2=2, lm_a=4, mgbf_proc=1), "S2C_ens"),    # _all: 3-D variables only
    (dict(l_new_map=False, l_lin_vertical=True, lm_a=10), "odd lm_a"),
])
def test_analysis_transfer_families_refuse_what_they_cannot_map(kw, msg):
    cfg = M.Config.for_grid(32, 24, 4, gm=3, h=4, hz=2, nm0=48, mm0=36, **kw)
    assert msg in str(_err(cfg))


def test_lwq_map_needs_three_levels_and_takes_any_lm_a_above_lm():
    cfg = M.Config.for_grid(32, 24, 2, gm=3, h=4, hz=1, nm0=48, mm0=36, lm_a=3, l_new_map=False)
    assert "lm >= 3" in str(_err(cfg))
    for lm_a in (4, 10, 64):
        cfg = M.Config.for_grid(32, 24, 3, gm=3, h=4, hz=2, nm0=48, mm0=36, lm_a=lm_a, l_new_map=False)
        assert M.derive(cfg).km_a_all == lm_a


def test_wrapper_sizes_every_operation_and_stage():
    # the C ABI takes bare pointers: the wrapper checks every buffer's
    # element count against the operation before a call
    torch = pytest.importorskip("torch")
    cfg = M.Config.for_grid(64, 48, 8, n_ens=2, gm=3, h=4, hz=2, nm0=96, mm0=72)
    d = M.derive(cfg)
    fld, mg, an = d.gen_shape(1), (d.mg_size,), d.anal_shape()
    assert an == (72, 96, 16) and mg == (16 * (64 * 48 + 32 * 24 + 16 * 12),)
    expect = {M.OP_WAWT: (fld, fld), M.OP_WT: (fld, mg), M.OP_W: (mg, fld), M.OP_A: (mg, mg),
              M.OP_ASQRT: (mg, mg), M.OP_BSQRT: (mg, fld), M.OP_BSQRTT: (fld, mg),
              M.OP_A2F: (an, fld), M.OP_F2A: (fld, an), M.OP_A2F_ALL: (an, fld),
              M.OP_F2A_ALL: (fld, an)}
    for op, shapes in expect.items():
        assert M.op_shapes(d, op) == shapes
    with pytest.raises(ValueError):
        M.op_shapes(d, 99)
    assert M.stage_shapes(d, M.STAGE_R, 2) == (d.gen_shape(2), d.gen_shape(3))
    assert M.stage_shapes(d, M.STAGE_P, 1) == (d.gen_shape(2), d.gen_shape(1))
    assert M.stage_shapes(d, M.STAGE_F, 3) == (d.gen_shape(3), d.gen_shape(3))
    for bad in ((M.STAGE_R, 3), (M.STAGE_P, 3), (M.STAGE_VF, 2), (M.STAGE_F, 0), (M.STAGE_F, 4)):
        with pytest.raises(ValueError):
            M.stage_shapes(d, *bad)
    # a field-sized vector where a multigrid vector belongs, and the reverse
    x_fld = torch.empty(fld)
    x_mg = torch.empty(mg)
    M.check_sizes("W", x_mg, x_fld, *M.op_shapes(d, M.OP_W))
    with pytest.raises(ValueError, match="input has"):
        M.check_sizes("W", x_fld, x_fld, *M.op_shapes(d, M.OP_W))
    with pytest.raises(ValueError, match="output has"):
        M.check_sizes("WT", x_fld, x_fld, *M.op_shapes(d, M.OP_WT))
    M.check_sizes("WAWT", x_fld, None, *M.op_shapes(d, M.OP_WAWT))
