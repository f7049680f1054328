"""The mixed-precision form (FF_PREC_MIXED: the staged LU-SGS records -- diagonal,
split-eigenvalue coefficients, face bases, SYI/SYM1I scalars -- stored in FP32; residual,
norms, increments, update and multigrid corrections in FP64) against the FP64 path and the
FP64 oracle.  The residual of a given state is the FP64 one bit for bit (the mixed form
changes only the implicit operator); the iterates follow the FP64 ones to the FP32 rounding
of that operator, and the error of every solved field after several cycles is stated."""
import numpy as np
import pytest

from finflo_amd.cases import config2, make_block
from finflo_amd.layout import default_params
from oracle.oracle_py import OracleHierarchy
from tests.parity import make_walls

pytestmark = pytest.mark.gpu

SOLVED = ["TEMP", "U", "V", "W", "PDIFF", "RK", "REPS"]


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_mixed_follows_fp64(hip_device, iturb):
    from finflo_amd.ffhip import Context
    print(f"case mixed precision iturb={iturb}", flush=True)
    prm = default_params(iturb=iturb)
    blk = make_block(40, 24, 16, prm, distort=0.08, seed=7, amp=0.03, kind="all",
                     turb=(1.0, 80.0) if iturb else None)
    init = blk.copy()
    a, b = blk.copy(), blk.copy()
    wp, pa = make_walls(a, prm, [(2, 1, 40, 1, 16, 0)], 9)
    pb = [p.copy() for p in pa]
    ca, cb = Context(prm), Context(prm)
    cb.precision("mixed")
    for c, x, pw in ((ca, a, pa), (cb, b, pb)):
        c.register(x)
        c.upload()
        c.wall_register(wp, pw)
    r_a = ca.cycle(ncycles=1)
    r_b = cb.cycle(ncycles=1)
    # the first residual is computed from the same state by the same FP64 kernels
    np.testing.assert_array_equal(r_a, r_b)
    ca.cycle(ncycles=4)
    cb.cycle(ncycles=4)
    ca.download()
    cb.download()
    names = SOLVED if iturb else SOLVED[:5]
    for f in names:
        change = np.abs(a.interior(f) - init.interior(f)).max()
        err = np.abs(b.interior(f) - a.interior(f)).max()
        print(f"{f}: change over 5 cycles {change:.3e}, mixed - FP64 {err:.3e} "
              f"({err / change:.2e} of the change)", flush=True)
        assert err <= 1e-4 * change, f
        assert np.isfinite(b.interior(f)).all()
    ca.close()
    cb.close()


def test_config2_mixed_error_against_the_oracle(hip_device):
    """The 192x96x96 flat plate, 3-level V-cycles with the wall, in the mixed form, against
    the FP64 oracle: the error of every solved field after 3 V-cycles, relative to that
    field's largest change from the initial state."""
    from finflo_amd.ffhip import Context
    print("case config 2 mixed precision, 3 V-cycles", flush=True)
    prm, blocks, walls = config2(seed=0)
    init = blocks[0].copy()
    host = [b.copy() for b in blocks]
    wh = [(wp, [p.copy() for p in ps]) for wp, ps in walls]
    ctx = Context(prm)
    ctx.precision("mixed")
    ctx.register_levels(blocks)
    oh = OracleHierarchy(prm, host)
    for m in range(3):
        ctx.upload(level=m + 1)
        ctx.wall_register(*walls[m], level=m + 1)
        oh.levels[m].set_walls(*wh[m])
    for _ in range(3):
        r1 = oh.vcycle()
    r2 = ctx.vcycle(3, ncycles=3)
    ctx.download()
    for f in SOLVED:
        change = np.abs(host[0].interior(f) - init.interior(f)).max()
        err = np.abs(blocks[0].interior(f) - host[0].interior(f)).max()
        print(f"config 2 {f}: change {change:.3e}, mixed - FP64 oracle {err:.3e} "
              f"({err / change:.2e} of the change)", flush=True)
        assert err <= 1e-4 * change, f
    print("RESL2 sums, mixed / FP64 oracle:", r2[:5] / r1[:5], flush=True)
    np.testing.assert_allclose(r2[:5], r1[:5], rtol=1e-4)
    ctx.close()
