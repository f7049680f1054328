"""The boundary routines on the HIP path (ff_bc_register, ff_state with REFLEC, ff_reflec,
ff_ghostext through the C-ABI) against the oracle: REFLEC bit for bit on every patch kind,
GHOSTEXT bit for bit, whole iterations and V-cycles with REFLEC in every STATIM (laminar bit
for bit, k-omega at 1e-12 of each field's largest magnitude), and the captured V-cycle."""
import numpy as np
import pytest

from finflo_amd.cases import make_hierarchy
from finflo_amd.layout import FIELDS, BcPatch, default_params
from oracle.oracle_py import OracleBlock, OracleHierarchy
from tests.parity import PRIM, SST_PRIM, assert_fields_equal, make_case, make_walls
from tests.test_oracle_bc import CASES, make_bcs

pytestmark = pytest.mark.gpu


def ctx_for(prm, blk, bcs=None, level=1, block=1):
    from finflo_amd.ffhip import Context
    ctx = Context(prm)
    ctx.register(blk, level=level, block=block)
    ctx.upload(level=level, block=block)
    if bcs is not None:
        ctx.bc_register(bcs, level=level, block=block)
    return ctx


@pytest.mark.parametrize("name,dims,pover,ckw,specs", CASES, ids=[c[0] for c in CASES])
def test_reflec_in_statim(hip_device, name, dims, pover, ckw, specs):
    print(f"case REFLEC {name}", flush=True)
    prm, g = make_case(dims, pover, ckw)
    a = g.copy()
    bcs = make_bcs(prm, specs, 7)
    ctx = ctx_for(prm, g, bcs)
    ctx.statim()
    ctx.download()
    oa = OracleBlock(prm, a)
    oa.set_bcs(bcs)
    oa.state()
    bad = [f for i, f in enumerate(FIELDS) if not np.array_equal(g.F[i], a.F[i])]
    assert not bad, f"{name}: fields differ: {bad}"
    # REFLEC alone, on a changed interior
    rng = np.random.default_rng(3)
    for f in ("U", "V", "W", "TEMP", "P"):
        g[f] = g[f] * (1 + 0.01 * rng.uniform(-1, 1, g.ntot))
    a.F[:] = g.F
    ctx.upload()
    ctx.reflec()
    ctx.download()
    oa.reflec()
    np.testing.assert_array_equal(g.F, a.F)
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6])
def test_ghostext(hip_device, iturb):
    prm, g = make_case((10, 8, 6), dict(iturb=iturb),
                       dict(kind="all", amp=0.05, seed=5, turb=(1.0, 80.0) if iturb else None))
    a = g.copy()
    bcs = [BcPatch(2, 5, 1, 10, 1, 6), BcPatch(2, 4, 2, 7, 1, 6), BcPatch(2, 3, 1, 10, 1, 8),
           BcPatch(2, 1, 1, 8, 2, 5)]
    ctx = ctx_for(prm, g, bcs)
    ctx.ghostext(alpha=0.05, beta=0.02)
    ctx.download()
    oa = OracleBlock(prm, a)
    oa.set_bcs(bcs)
    oa.ghostext(0.05, 0.02)
    np.testing.assert_array_equal(g.F, a.F)
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6])
def test_cycles_with_reflec(hip_device, iturb):
    print(f"case cycles with REFLEC iturb={iturb}", flush=True)
    prm, g = make_case((24, 16, 12), dict(iturb=iturb),
                       dict(kind="all", amp=0.03, distort=0.08, seed=213,
                            turb=(1.0, 80.0) if iturb else None))
    a = g.copy()
    wp, pg = make_walls(g, prm, [(2, 1, 24, 1, 12, 0)], 19)
    pa = [p.copy() for p in pg]
    bcs = make_bcs(prm, [(8, 2, 1, 24, 1, 12, 0, 1, False), (7, 1, 1, 16, 1, 12, 0, 0, False),
                         (8, 5, 3, 20, 1, 12, 1, 1, True)], 23)
    ctx = ctx_for(prm, g, bcs)
    ctx.wall_register(wp, pg)
    oa = OracleBlock(prm, a)
    oa.set_walls(wp, pa)
    oa.set_bcs(bcs)
    for _ in range(10):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=10)
    ctx.download()
    assert_fields_equal(g, a, SST_PRIM if iturb else PRIM, interior=False,
                        rtol=1e-12 if iturb else 0.0, what="cycles with REFLEC")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if iturb else 1e-13)
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6])
def test_vcycles_with_reflec_and_graph(hip_device, iturb):
    """3-level V-cycles with walls and REFLEC patches on every level, eager and replayed from
    the captured graph, against the oracle's V-cycles."""
    from finflo_amd.ffhip import Context
    print(f"case V-cycles with REFLEC iturb={iturb}", flush=True)
    prm = default_params(iturb=iturb, cfll=3.0)
    blocks = make_hierarchy(16, 12, 12, 3, prm, distort=0.1, kind="all", amp=0.03, seed=89,
                            turb=(1.0, 150.0) if iturb else None)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = Context(prm)
    ctx.register_levels(blocks)
    keep = []
    for m, b in enumerate(blocks):
        ctx.upload(level=m + 1)
        wp, pw = make_walls(b, prm, [(2, 1, b.imax, 1, b.kmax, 0)], 51 + m)
        ph = [p.copy() for p in pw]
        bcs = make_bcs(prm, [(8, 2, 1, b.imax, 1, b.kmax, 0, 1, False),
                             (7, 3, 1, b.imax, 1, b.jmax, 0, 0, False),
                             (7, 4, 1, b.jmax, 1, b.kmax, 0, 0, False)], 61 + m)
        keep.append((pw, ph, bcs))
        ctx.wall_register(wp, pw, level=m + 1)
        ctx.bc_register(bcs, level=m + 1)
        oh.levels[m].set_walls(wp, ph)
        oh.levels[m].set_bcs(bcs)
    prim = SST_PRIM if iturb else PRIM
    for rep, (ne, ng) in enumerate(((2, 0), (0, 2), (1, 1))):
        for _ in range(ne + ng):
            oh.vcycle()
        if ne:
            ctx.vcycle(3, ncycles=ne)
        if ng:
            ctx.vcycle_graph(3, ncycles=ng)
        for m in range(3):
            ctx.download(level=m + 1)
            assert_fields_equal(blocks[m], host[m], prim, interior=False,
                                rtol=1e-12 if iturb else 0.0, what=f"rep {rep} level {m + 1}")
    ctx.close()
