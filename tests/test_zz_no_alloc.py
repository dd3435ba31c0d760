"""mgbf_apply_noalloc: the variant of mgbf_apply whose workspaces live in the
library's static device scratch. Through the C ABI it must give mgbf_apply's
results bit for bit, on two calls in a row of which the second is smaller,
allocate no device memory, and refuse a context whose workspace exceeds the
scratch."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from standalone_mgbf_loc_amd import mgbf as M

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")


def _rand(shape, seed):
    return torch.from_numpy(np.random.default_rng(seed).standard_normal(shape).astype(np.float32)).to(DEV)


def test_noalloc_matches_apply_on_two_calls_and_allocates_nothing():
    big = M.Config.for_grid(256, 256, 32, n_ens=2, gm=4, h=6, hz=3, weights=[1.0, 0.5, 0.25, 0.125])
    small = M.Config.for_grid(96, 64, 16, n_ens=1, gm=3, h=4, hz=2, weights=[1.0, 0.6, 0.3])
    print("noalloc: contexts 256x256x64 gm 4 and 96x64x16 gm 3", flush=True)
    # the runtime loads each code object of the library on the first launch
    # of one of its kernels (once per process and device): contexts of the
    # same shapes run first, so that the measured calls meet loaded code
    with M.MGBF(big) as W1, M.MGBF(small) as W2:
        for w_ in (W1, W2):
            xw = torch.zeros(w_.field_shape(), device=DEV)
            w_.apply_noalloc(M.OP_WAWT, xw, torch.empty_like(xw))
            w_.apply_noalloc(M.OP_WT, xw, torch.empty((w_.dims.mg_size,), device=DEV))
            w_.apply_noalloc(M.OP_W, torch.zeros((w_.dims.mg_size,), device=DEV), torch.empty_like(xw))
        torch.cuda.synchronize()
    with M.MGBF(big) as A, M.MGBF(small) as B:
        assert 2 * A.dims.mg_size + A.dims.gm * A.dims.km_all <= M.NOALLOC_SCRATCH_FLOATS
        xa, xb = _rand(A.field_shape(), 1), _rand(B.field_shape(), 2)
        va = _rand((A.dims.mg_size,), 3)
        ya, wta, wa = torch.empty_like(xa), torch.empty_like(va), torch.empty_like(xa)
        yb = torch.empty_like(xb)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        # first call: the larger context; second call: the smaller one
        A.apply_noalloc(M.OP_WAWT, xa, ya)
        A.apply_noalloc(M.OP_WT, xa, wta)
        A.apply_noalloc(M.OP_W, va, wa)
        B.apply_noalloc(M.OP_WAWT, xb, yb)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 == free0, (free0, free1)
        # the allocating path on the same contexts: workspaces appear
        ref_ya = A.apply(M.OP_WAWT, xa)
        ref_wta = A.apply(M.OP_WT, xa)
        ref_wa = A.apply(M.OP_W, va)
        ref_yb = B.apply(M.OP_WAWT, xb)
        torch.cuda.synchronize()
        free2 = torch.cuda.mem_get_info()[0]
        assert free0 - free2 >= 2 * 4 * A.dims.mg_size, (free0, free2)
        assert torch.equal(ya, ref_ya) and torch.equal(wta, ref_wta) and torch.equal(wa, ref_wa)
        assert torch.equal(yb, ref_yb)
        # again after the allocating calls, smaller context first: the
        # scratch keeps the last call's contents, which must not matter
        yb2, ya2 = torch.empty_like(xb), torch.empty_like(xa)
        B.apply_noalloc(M.OP_WAWT, xb, yb2)
        A.apply_noalloc(M.OP_WAWT, xa, ya2)
        torch.cuda.synchronize()
        assert torch.equal(yb2, ref_yb) and torch.equal(ya2, ref_ya)


def test_noalloc_refuses_a_workspace_beyond_the_scratch():
    cfg = M.Config.for_grid(768, 768, 64, n_ens=1, gm=4, h=6, hz=3)
    with M.MGBF(cfg) as op:
        assert 2 * op.dims.mg_size > M.NOALLOC_SCRATCH_FLOATS
        x = torch.zeros(op.field_shape(), device=DEV)
        y = torch.empty_like(x)
        with pytest.raises(M.MGBFError, match="exceeds the static scratch"):
            op.apply_noalloc(M.OP_WAWT, x, y)
