"""The library's device timers: the per-phase accumulators of mg_timers
(src.mgbf/mg_timers.f90:17-60) and their split by generation."""
import pytest

torch = pytest.importorskip("torch")

from standalone_mgbf_loc_amd import mgbf as M

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")


@pytest.mark.parametrize("fuse", ["3", "0"])
def test_timers_split_by_generation_add_up(monkeypatch, fuse):
    # every launch belongs to one generation: the per-generation launches and
    # bytes of a phase add up to the phase's exactly, and the times to its
    # time; at full fusion generation 1 has one first and one second radial
    # pass per application, and generation gm never restricts or prolongs
    monkeypatch.setenv("MGBF_FUSE", fuse)
    cfg = M.Config.for_grid(96, 96, 32, n_ens=2, gm=3, h=4, hz=2)
    reps = 2
    with M.MGBF(cfg) as op:
        gm = op.dims.gm
        x = torch.randn(op.field_shape(), device=DEV)
        op.filtering_procedure(x)
        torch.cuda.synchronize()
        op.timers(True)
        for _ in range(reps):
            op.filtering_procedure(x)
        torch.cuda.synchronize()
        tot = op.read_timers()
        gens = [op.read_timers_gen(g) for g in range(1, gm + 1)]
        op.timers(False)
        with pytest.raises(M.MGBFError):
            op.read_timers_gen(gm + 1)
    for name in M.TIMER_NAMES:
        ms, n, nbytes = tot[name]
        assert sum(g[name][1] for g in gens) == n, name
        assert sum(g[name][2] for g in gens) == pytest.approx(nbytes, rel=1e-12), name
        assert sum(g[name][0] for g in gens) == pytest.approx(ms, rel=1e-6, abs=1e-6), name
    assert tot["hfilt"][1] > 0 and tot["hfiltT"][1] > 0
    if fuse == "3":
        g1 = gens[0]
        assert g1["hfiltT"][1] == reps and g1["hfilt"][1] == reps
    assert gens[-1]["upsend"][1] == 0 and gens[-1]["dnsend"][1] == 0
