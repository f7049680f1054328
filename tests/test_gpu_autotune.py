"""The tile-shape autotuner of the residual kernels: every candidate (I, J, K) shape gives
the default shape's residual bit for bit (laminar Roe and k-omega, ragged blocks), the
autotuned cycles equal the default ones bit for bit, and the tuner records a shape per
block size."""
import numpy as np
import pytest

from finflo_amd.cases import make_block
from finflo_amd.layout import default_params

pytestmark = pytest.mark.gpu

NSHAPE = 8


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_every_shape_gives_the_same_residual(hip_device, iturb):
    from finflo_amd.ffhip import Context
    prm = default_params(iturb=iturb)
    for dims in ((37, 11, 9), (70, 20, 18)):
        print(f"case tile shapes iturb={iturb} dims={dims}", flush=True)
        blk = make_block(*dims, prm, distort=0.1, seed=5, amp=0.03, kind="all",
                         turb=(1.0, 80.0) if iturb else None)
        ref = None
        for sh in [-1] + list(range(NSHAPE)):
            b = blk.copy()
            ctx = Context(prm)
            ctx.tile_shape(sh)
            ctx.register(b)
            ctx.upload()
            ctx.statim()
            ctx.turbdi()
            r = ctx.residi()
            ctx.download()
            ctx.close()
            if ref is None:
                ref, rr = b.F.copy(), r
            else:
                np.testing.assert_array_equal(b.F, ref, err_msg=f"shape {sh}")
                np.testing.assert_array_equal(r, rr)


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_autotuned_cycles(hip_device, iturb):
    from finflo_amd.ffhip import Context
    print(f"case autotuned cycles iturb={iturb}", flush=True)
    prm = default_params(iturb=iturb)
    blk = make_block(96, 48, 32, prm, distort=0.05, seed=9, amp=0.02, kind="all",
                     turb=(1.0, 80.0) if iturb else None)
    a, b = blk.copy(), blk.copy()
    ca, cb = Context(prm), Context(prm)
    cb.autotune(True)
    for c, x in ((ca, a), (cb, b)):
        c.register(x)
        c.upload()
        c.cycle(ncycles=3)
        c.download()
    np.testing.assert_array_equal(a.F, b.F)
    tuned = cb.tuned_shape()
    print("tuned shape", tuned, flush=True)
    assert tuned is not None and tuned[1] > 0
    assert ca.tuned_shape() is None
    ca.close()
    cb.close()
