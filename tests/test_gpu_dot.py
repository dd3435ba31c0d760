"""The library's FP64-accumulated inner product (mgbf_dot), which the
adjoint tests and the weight gradients rest on: against an FP64 sum on the
host, for lengths that are not multiples of the 16-byte load width and for
arrays that are not 16-byte aligned, and the same bits on every call."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from standalone_mgbf_loc_amd import mgbf as M

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no HIP device", allow_module_level=True)

DEV = torch.device("cuda:0")


@pytest.mark.parametrize("n,off", [(1, 0), (3, 1), (1000, 0), (1001, 3), (262144 * 4 + 5, 0),
                                   (262144 * 4 + 5, 2), (7_000_003, 1)])
def test_dot_matches_fp64_and_repeats_bit_for_bit(n, off):
    cfg = M.Config.for_grid(32, 32, 4, n_ens=1, gm=2, h=3, hz=1)
    rng = np.random.default_rng(n + off)
    a = rng.standard_normal(n + off).astype(np.float32)
    b = rng.standard_normal(n + off).astype(np.float32)
    with M.MGBF(cfg) as op:
        ad, bd = torch.from_numpy(a).to(DEV), torch.from_numpy(b).to(DEV)
        # the views start `off` floats into their allocations
        av, bv = ad[off:], bd[off:]
        got = op.dot(av, bv)
        again = op.dot(av, bv)
    ref = float(np.dot(a[off:].astype(np.float64), b[off:].astype(np.float64)))
    # every product is exact in FP64; only the order of the sum differs
    assert abs(got - ref) <= 1e-12 * float(np.abs(a[off:].astype(np.float64) * b[off:]).sum())
    assert got == again
