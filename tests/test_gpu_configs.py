"""The benchmark configurations at their own workloads on the HIP path, through the C-ABI,
against the oracle:
* the plain method's 48x24x24 laminar block with its J-low wall: 200 LU-SGS iterations,
  Roe and van Leer, bit for bit (fields, wall patch arrays, force sums; RESL2 sums to
  rounding, whose reduction order differs);
* the 192x96x96 flat plate with k-omega SST, its wall on all three multigrid levels: two
  V-cycles at 1e-12 of each field's largest magnitude (the device tanh/pow round
  differently in the last bit);
* the 256x128x128 EARSM block over the 3-D bump: STATIM and TURBDI with the exact EARSM1
  at 1e-12, and the learned closure's BF16-MFMA inference on every one of its 4.19M cells
  against the BF16-emulating oracle."""
import numpy as np
import pytest

from finflo_amd.cases import config1, config2, config4
from oracle.oracle_py import OracleBlock, OracleHierarchy
from tests.parity import PRIM, SST_PRIM, assert_fields_equal, assert_patches_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("iflux", [1, 0], ids=["roe", "vanleer"])
def test_config1_200_iterations(hip_device, iflux):
    from finflo_amd.ffhip import Context
    print(f"case config 1 iflux={iflux}: 48x24x24, 200 iterations", flush=True)
    prm, g, (wp, pg) = config1(iflux=iflux, seed=0)
    a = g.copy()
    pa = [p.copy() for p in pg]
    ctx = Context(prm)
    ctx.register(g)
    ctx.upload()
    ctx.wall_register(wp, pg)
    ctx.wall_sync(0)
    oa = OracleBlock(prm, a)
    oa.set_walls(wp, pa)
    hist = []
    for it in range(200):
        hist.append(oa.cycle())
    r_gpu = ctx.cycle(ncycles=200)
    ctx.download()
    ctx.wall_sync(1)
    print("density RESL2 sum: first", hist[0][0], "last", hist[-1][0], flush=True)
    assert np.all(np.isfinite(a.F))
    bad = [n for n in PRIM + ["P", "RO", "E", "C", "DRO", "DM", "DN", "DW", "DE", "DTL"]
           if not np.array_equal(g[n], a[n])]
    assert not bad, f"fields differ after 200 iterations: {bad}"
    np.testing.assert_array_equal(g.F, a.F)
    assert_patches_equal(pg, pa, what="config 1 wall")
    np.testing.assert_array_equal(ctx.wall_forces(0), oa.forces(0))
    np.testing.assert_allclose(r_gpu[:5], hist[-1][:5], rtol=1e-13)
    ctx.close()


def test_config2_two_vcycles(hip_device):
    from finflo_amd.ffhip import Context
    print("case config 2: 192x96x96 SST, 3 levels, wall, 2 V-cycles", flush=True)
    prm, blocks, walls = config2(seed=0)
    host = [b.copy() for b in blocks]
    wh = [(wp, [p.copy() for p in ps]) for wp, ps in walls]
    ctx = Context(prm)
    ctx.register_levels(blocks)
    oh = OracleHierarchy(prm, host)
    for m in range(3):
        ctx.upload(level=m + 1)
        ctx.wall_register(*walls[m], level=m + 1)
        ctx.wall_sync(0, level=m + 1)
        oh.levels[m].set_walls(*wh[m])
    for _ in range(2):
        r1 = oh.vcycle()
    r2 = ctx.vcycle(3, ncycles=2)
    for m in range(3):
        ctx.download(level=m + 1)
        ctx.wall_sync(1, level=m + 1)
        assert np.all(np.isfinite(host[m].F))
        assert_fields_equal(blocks[m], host[m], SST_PRIM + ["P", "RO", "EPS2", "VIST"],
                            interior=False, rtol=1e-12, what=f"config 2 level {m + 1}")
    assert_patches_equal(walls[0][1], wh[0][1], rtol=1e-12, what="config 2 wall")
    f0 = oh.levels[0].forces(0)
    np.testing.assert_allclose(ctx.wall_forces(0), f0, rtol=1e-10, atol=1e-10 * np.abs(f0).max())
    np.testing.assert_allclose(r2, r1, rtol=1e-10)
    ctx.close()


def test_config4_exact_earsm_and_closure_inference(hip_device):
    import torch
    from finflo_amd.ffhip import Context
    from finflo_amd.mlp import init_params
    from oracle import mlp_oracle as M
    print("case config 4: 256x128x128 EARSM over the bump", flush=True)
    prm, g, _ = config4(seed=0)
    a = g.copy()
    ctx = Context(prm)
    ctx.register(g)
    ctx.upload()
    ctx.save_old()
    ctx.statim()
    ctx.turbdi()
    ctx.download()
    oa = OracleBlock(prm, a)
    oa.save_old()
    oa.state()
    oa.turb()
    names = ["P", "RO", "C", "VIS", "EPS2", "VIST", "TTS", "STRAIN", "OHMI",
             "S11", "S12", "S13", "S22", "S23", "S33", "W12", "W13", "W23",
             "B11", "B12", "B13", "B22", "B23", "B33"]
    assert_fields_equal(g, a, names, rtol=1e-12, what="config 4 STATIM + TURBDI")
    # the closure's inference on every cell's invariants against the emulating oracle
    x, y = ctx.earsm_samples()
    n = len(x)
    assert n == 256 * 128 * 128
    ctx.mlp_fit_norm(x, y)
    ctx.mlp_set(init_params(3))
    p, norm = ctx.mlp_get()
    yp = ctx.mlp_infer(x).cpu().numpy()
    xh = x.cpu().numpy()
    worst, nbad = 0.0, 0
    for s in range(0, n, 1 << 19):
        ref = M.predict(p, norm, xh[s:s + (1 << 19)], emulate=True)
        err = np.abs(yp[s:s + (1 << 19)] - ref) / np.abs(ref)
        worst = max(worst, float(err.max()))
        nbad += int(np.count_nonzero(err > 2e-6))
    print(f"closure inference on {n} cells: max relative error {worst:.3e}, "
          f"{nbad} cells above 2e-6", flush=True)
    # the FP32 accumulation order and asinhf differ from the oracle's by ~1e-7 relative, so
    # an activation that lies that close to a BF16 rounding boundary (spacing 2^-8) rounds to
    # the other neighbour: ~3e-5 per activation, ~133 activations per cell, ~0.4 % of cells
    # (measured 16,859 of 4,194,304 on MI355X), each moved by about one BF16 ulp through the
    # network (max 1.2e-2 measured).  Held to: at most 1 % of cells above 2e-6, none above
    # 5e-2; every other cell within 2e-6.
    assert nbad <= 1e-2 * n
    assert worst < 5e-2
    ctx.close()
