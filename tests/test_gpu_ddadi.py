"""The DDADI implicit stage (ff_params.ddadi = 1: BLKS%LUSGS = 0, imps3.f:321-373) on the HIP
path, through the C-ABI, against the oracle (whose DDADI is pinned against the reference's
own PRED/DIAG/CORR(KE) in tests/test_oracle_ddadi.py):
* the implicit stage from the oracle's state, residual and increments: bit for bit in DTL
  and every increment field, laminar and k-omega SST, on blocks whose line counts are
  ragged against the 128-thread workgroups;
* whole cycles and a 2-level V-cycle: laminar bit for bit, SST at 1e-12 of each field's
  largest magnitude (the device tanh and pow round differently in the last bit);
* a captured V-cycle graph with DDADI replays the eager V-cycles bit for bit."""
import numpy as np
import pytest

from oracle.oracle_py import OracleBlock, OracleHierarchy
from tests.parity import (CASES, PRIM, RES, SST_CASES, SST_PRIM, SST_RES, assert_fields_equal,
                          make_case)

pytestmark = pytest.mark.gpu

SST_RTOL = 1e-12
DD_CASES = [CASES[1], CASES[3], CASES[6], SST_CASES[0], SST_CASES[2]]


def _ctx(prm, blk):
    from finflo_amd.ffhip import Context
    ctx = Context(prm)
    ctx.register(blk)
    ctx.upload()
    return ctx


@pytest.mark.parametrize("name,dims,pover,ckw", DD_CASES, ids=[c[0] for c in DD_CASES])
def test_ddadi_stage_bit_exact(hip_device, name, dims, pover, ckw):
    print(f"case DDADI stage {name} dims {dims}", flush=True)
    prm, blk = make_case(dims, dict(pover, ddadi=1), ckw)
    a = blk.copy()
    oa = OracleBlock(prm, a)
    oa.save_old(); oa.state(); oa.turb(); oa.residual()
    g = a.copy()
    ctx = _ctx(prm, g)
    oa.implicit(); ctx.impsdi(); ctx.download()
    kw = prm.iturb == 6
    assert_fields_equal(g, a, ["DTL"] + (SST_RES[:7] if kw else RES), interior=False,
                        what="IMPS with DDADI")
    ctx.close()


@pytest.mark.parametrize("name,dims,pover,ckw", DD_CASES[:1] + DD_CASES[3:4],
                         ids=[c[0] for c in DD_CASES[:1] + DD_CASES[3:4]])
def test_ddadi_cycles(hip_device, name, dims, pover, ckw):
    print(f"case DDADI cycles {name} dims {dims}", flush=True)
    prm, blk = make_case(dims, dict(pover, ddadi=1), ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g)
    for _ in range(4):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=4)
    ctx.download()
    kw = prm.iturb == 6
    assert_fields_equal(g, a, SST_PRIM if kw else PRIM, interior=False,
                        rtol=SST_RTOL if kw else 0.0, what="4 DDADI cycles")
    np.testing.assert_allclose(r2[:5], r1[:5], rtol=1e-10 if kw else 1e-13)
    ctx.close()


@pytest.mark.parametrize("dims", [(37, 21, 18), (16, 40, 9)], ids=["37x21x18", "16x40x9"])
def test_ddadi_large_ragged(hip_device, dims):
    """Blocks with more lines than one workgroup in every direction and ragged counts."""
    print(f"case DDADI large {dims}", flush=True)
    prm, blk = make_case(dims, dict(iturb=6, ddadi=1),
                         dict(kind="all", amp=0.03, distort=0.1, seed=47, turb=(1.0, 200.0)))
    a = blk.copy()
    oa = OracleBlock(prm, a)
    oa.save_old(); oa.state(); oa.turb(); oa.residual()
    g = a.copy()
    ctx = _ctx(prm, g)
    oa.implicit(); ctx.impsdi(); ctx.download()
    assert_fields_equal(g, a, ["DTL"] + SST_RES[:7], interior=False, what=f"DDADI {dims}")
    ctx.close()


def test_ddadi_vcycle_and_graph(hip_device):
    """A 2-level V-cycle with DDADI on both levels against the oracle's, then the captured
    V-cycle graph against the eager V-cycles (bit for bit)."""
    from finflo_amd.cases import make_hierarchy
    from finflo_amd.ffhip import Context
    from finflo_amd.layout import default_params
    print("case DDADI 2-level V-cycle 16x12x12", flush=True)
    prm = default_params(ddadi=1, cfll=3.0)
    blocks = make_hierarchy(16, 12, 12, 2, prm, distort=0.1, kind="all", amp=0.03, seed=3)
    ob = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, ob)
    ge = [b.copy() for b in blocks]
    gg = [b.copy() for b in blocks]
    ce, cg = Context(prm), Context(prm)
    for c, bl in ((ce, ge), (cg, gg)):
        c.register_levels(bl)
        for m in (1, 2):
            c.upload(level=m)
    for _ in range(2):
        oh.vcycle()
    ce.vcycle(2, ncycles=2)
    ce.download(level=1)
    assert_fields_equal(ge[0], ob[0], PRIM, interior=False, what="DDADI V-cycles")
    cg.vcycle(2, ncycles=1)
    cg.vcycle_graph(2, ncycles=1)
    cg.download(level=1)
    assert_fields_equal(gg[0], ge[0], PRIM, interior=False, what="DDADI graph replay")
    ce.close()
    cg.close()
