"""INLET/OUTLET on the HIP path (ff_bc_register with INL and OUT patches, ff_inlout,
ff_bc_cycle, ff_bc_read and the level-1 residual that starts with the update) against the
oracle, whose INLET/OUTLET is pinned against the reference's own in
tests/test_oracle_inlout.py:
* the update alone, bit for bit in every field and in the BOUNP/BOUNPD that INLET 4
  writes: every supported INLOUT type, every face direction, outflow and inflow on one
  face, ICYCLE 0, 1 and 2, laminar and k-omega SST;
* whole iterations with an inlet, an outlet and symmetry planes (laminar bit for bit,
  k-omega at 1e-12 of each field's largest magnitude: the device tanh and pow round
  differently in the last bit), so the first cycle's ICYCLE = 1 branch and the count on
  the device are exercised;
* 2-level V-cycles, eager and replayed from the captured graph (the count advances inside
  the graph);
* registrations the path does not support fail with the library's error codes."""
import numpy as np
import pytest

from finflo_amd.cases import make_hierarchy
from finflo_amd.layout import FIELDS, BcPatch, default_params
from oracle.oracle_py import OracleBlock, OracleHierarchy
from tests.parity import PRIM, SST_PRIM, assert_fields_equal, make_case
from tests.test_oracle_inlout import PATCH_SETS, _case, _prepared, boun_arrays, make_patches

pytestmark = pytest.mark.gpu


def ctx_for(prm, blk, bcs, level=1):
    from finflo_amd.ffhip import Context
    ctx = Context(prm)
    ctx.register(blk, level=level)
    ctx.upload(level=level)
    ctx.bc_register(bcs, level=level)
    return ctx


@pytest.mark.parametrize("iturb", [0, 6])
@pytest.mark.parametrize("icycle", [0, 1, 2])
@pytest.mark.parametrize("pset", list(PATCH_SETS))
def test_inlout_update_bit_exact(hip_device, pset, icycle, iturb):
    print(f"case INLET/OUTLET {pset} icycle {icycle} iturb {iturb}", flush=True)
    prm, a = _prepared(iturb, 301)
    g = a.copy()
    pa = make_patches(prm, PATCH_SETS[pset], 11)
    pg = [bp.copy() for bp in pa]
    ctx = ctx_for(prm, g, pg)
    ctx.bc_cycle(icycle)
    ctx.inlout()
    ctx.download()
    oa = OracleBlock(prm, a)
    oa.set_bcs(pa)
    oa.bc_cycle(icycle)
    oa.inlout()
    bad = [f for i, f in enumerate(FIELDS) if not np.array_equal(g.F[i], a.F[i])]
    assert not bad, f"{pset}: fields differ: {bad}"
    for l, bp in enumerate(pa):
        bpp, bpd = ctx.bc_read(l, bp.size)
        np.testing.assert_array_equal(bpp, bp.boun["bounp"])
        np.testing.assert_array_equal(bpd, bp.boun["bounpd"])
    ctx.close()


def _duct_patches(prm, dims, seed):
    """A subsonic inlet on I-low, a pressure outlet on I-high (some faces see inflow),
    symmetry planes on the J and K faces."""
    i, j, k = dims
    pa = make_patches(prm, [(3, 1, 1, j, 1, k, 4, False), (5, 4, 1, j, 1, k, 2, True)], seed)
    sym = [BcPatch(7, f, 1, n1, 1, n2, viscous=0)
           for f, n1, n2 in ((2, i, k), (5, i, k), (3, i, j), (6, i, j))]
    return pa, sym


@pytest.mark.parametrize("iturb", [0, 6])
def test_cycles_with_inlet_and_outlet(hip_device, iturb):
    print(f"case cycles with INLET/OUTLET iturb={iturb}", flush=True)
    dims = (24, 16, 12)
    prm, g = make_case(dims, dict(iturb=iturb),
                       dict(kind="all", amp=0.03, distort=0.08, seed=223,
                            turb=(1.0, 80.0) if iturb else None))
    a = g.copy()
    pa, sym = _duct_patches(prm, dims, 71)
    pg = [bp.copy() for bp in pa]
    ctx = ctx_for(prm, g, pg + sym)
    oa = OracleBlock(prm, a)
    oa.set_bcs(pa + sym)
    for _ in range(10):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=10)
    ctx.download()
    assert_fields_equal(g, a, SST_PRIM if iturb else PRIM, interior=False,
                        rtol=1e-12 if iturb else 0.0, what="cycles with INLET/OUTLET")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if iturb else 1e-13)
    bpp, _ = ctx.bc_read(0, pa[0].size)
    np.testing.assert_allclose(bpp, pa[0].boun["bounp"], rtol=1e-12 if iturb else 0.0)
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6])
def test_vcycles_with_inlet_outlet_and_graph(hip_device, iturb):
    """2-level V-cycles with the inlet, outlet and symmetry patches on both levels (INLET/
    OUTLET act on level 1 only, REFLEC's inlet pressures on both), eager and replayed from
    the captured graph, against the oracle's V-cycles."""
    from finflo_amd.ffhip import Context
    print(f"case V-cycles with INLET/OUTLET iturb={iturb}", flush=True)
    prm = default_params(iturb=iturb, cfll=3.0)
    blocks = make_hierarchy(16, 12, 12, 2, prm, distort=0.1, kind="all", amp=0.03, seed=97,
                            turb=(1.0, 150.0) if iturb else None)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = Context(prm)
    ctx.register_levels(blocks)
    keep = []
    for m, b in enumerate(blocks):
        ctx.upload(level=m + 1)
        pa, sym = _duct_patches(prm, (b.imax, b.jmax, b.kmax), 81 + 2 * m)
        pg = [bp.copy() for bp in pa]
        keep.append((pa, pg, sym))
        ctx.bc_register(pg + sym, level=m + 1)
        oh.levels[m].set_bcs(pa + sym)
    prim = SST_PRIM if iturb else PRIM
    for rep, (ne, ng) in enumerate(((2, 0), (0, 2), (1, 1))):
        for _ in range(ne + ng):
            oh.vcycle()
        if ne:
            ctx.vcycle(2, ncycles=ne)
        if ng:
            ctx.vcycle_graph(2, ncycles=ng)
        for m in range(2):
            ctx.download(level=m + 1)
            assert_fields_equal(blocks[m], host[m], prim, interior=False,
                                rtol=1e-12 if iturb else 0.0, what=f"rep {rep} level {m + 1}")
    ctx.close()


def test_unsupported_registrations(hip_device):
    from finflo_amd.ffhip import Context, FFError
    print("case INLET/OUTLET registrations", flush=True)
    prm, g = _case(0, 309)
    ctx = Context(prm)
    ctx.register(g)
    ctx.upload()
    bad_type = BcPatch(5, 4, 1, 7, 1, 6, inlout=1)
    bad_type.boun.update(boun_arrays(prm, bad_type, 3))
    missing = BcPatch(3, 1, 1, 7, 1, 6, inlout=4)
    outside = BcPatch(3, 1, 1, 8, 1, 6, inlout=4)
    outside.boun.update(boun_arrays(prm, outside, 5))
    for bp, code in ((bad_type, -5), (missing, -1), (outside, -1)):      # FF_E_UNSUP, FF_E_ARG
        with pytest.raises(FFError) as e:
            ctx.bc_register([bp])
        assert f"({code})" in str(e.value), (bp.inlout, str(e.value))
    with pytest.raises(FFError):
        ctx.bc_cycle(2)                   # no INL/OUT patch registered
    ctx.close()
    prm2, g2 = _case(6, 311)
    prm2.istres = 1
    ctx = Context(prm2)
    ctx.register(g2)
    ctx.upload()
    ok = BcPatch(5, 4, 1, 7, 1, 6, inlout=4)
    with pytest.raises(FFError) as e:
        ctx.bc_register([ok])
    assert "(-5)" in str(e.value)
    ctx.close()
