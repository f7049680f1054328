"""Multigrid on the HIP path (libffhip.so through its C-ABI) against the oracle: the
transfers TRANSF/TRANKE, FORCFU/ADDRES and INTERP/INTERT one by one, the coarse-level
stages, and whole V-cycles.  Laminar results are bit for bit; k-omega results are held to
1e-12 of each field's largest magnitude where a tanh (FUN1KO, EPS2KO) lies upstream, since the
device and host math libraries round tanh differently in the last bit."""
import numpy as np
import pytest

from finflo_amd.cases import make_hierarchy
from finflo_amd.layout import default_params
from oracle.oracle_py import OracleBlock, OracleHierarchy
from tests.parity import (PRIM, RES, SST_PRIM, SST_RES, assert_fields_equal,
                          assert_patches_equal, make_walls)

pytestmark = pytest.mark.gpu

RTOL = 1e-12
FORC = ["ROP2H", "RMP2H", "RNP2H", "RWP2H", "EP2H"]


def hierarchy(iturb, nlev=2, dims=(16, 8, 8), seed=71, fullns=0):
    prm = default_params(iturb=iturb, cfll=3.0, fullns=fullns)
    blocks = make_hierarchy(*dims, nlev, prm, distort=0.1, kind="all", amp=0.03, seed=seed,
                            turb=(1.0, 150.0) if iturb else None)
    return prm, blocks


def context(prm, blocks):
    from finflo_amd.ffhip import Context
    ctx = Context(prm)
    ctx.register_levels(blocks)
    for m in range(len(blocks)):
        ctx.upload(level=m + 1)
    return ctx


@pytest.mark.parametrize("dims", [(16, 8, 8), (16, 8, 1)], ids=["3d", "planar"])
@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_transfers(hip_device, iturb, dims):
    print(f"case multigrid transfers iturb={iturb} dims {dims}", flush=True)
    prm, blocks = hierarchy(iturb, dims=dims)
    rng = np.random.default_rng(5)
    for name in ("DRO", "DM", "DN", "DW", "DE", "DRK", "DEPS"):
        blocks[0][name] = rng.uniform(-1, 1, blocks[0].ntot)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = context(prm, blocks)
    # restriction
    oh.restrict(1); ctx.restrict(1); ctx.download(level=2)
    names = ["TEMP", "U", "V", "W", "PDIFF", "EPS2", "VIST", "TOLD", "UOLD", "VOLD", "WOLD",
             "POLD"] + FORC + (["RK", "REPS", "RKOLD", "EPSOLD", "RKP2H", "EPSP2H"] if iturb else [])
    assert_fields_equal(blocks[1], host[1], names, interior=False, what="TRANSF")
    # FORCFU / ADDRES
    for name in ("DRO", "DM", "DN", "DW", "DE", "DRK", "DEPS"):
        host[1][name] = rng.uniform(-1, 1, host[1].ntot)
    blocks[1].F[:] = host[1].F
    ctx.upload(level=2)
    oc = oh.levels[1]
    for phase in (0, 1):
        oc.forcing(phase); ctx.forcing(2, phase); ctx.download(level=2)
        assert_fields_equal(blocks[1], host[1], FORC + RES, interior=False, what=f"forcing {phase}")
    # prolongation of a coarse correction
    for name in ("TEMP", "U", "V", "W", "PDIFF") + (("RK", "REPS") if iturb else ()):
        host[1][name] = host[1][name] * (1.0 + 0.01 * rng.uniform(-1, 1, host[1].ntot))
    blocks[1].F[:] = host[1].F
    ctx.upload(level=2)
    oh.prolong(2); ctx.prolong(2); ctx.download(level=1)
    prim = SST_PRIM if iturb else PRIM
    assert_fields_equal(blocks[0], host[0], prim, interior=False, what="INTERP")
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_coarse_level_stages(hip_device, iturb):
    """Level 2: first-order residual (and RESIKE's own gradients), CFLL, coarse limiters."""
    print(f"case coarse stages iturb={iturb}", flush=True)
    prm, blocks = hierarchy(iturb)
    c = blocks[1]
    if iturb:
        rng = np.random.default_rng(9)
        c["EPS2"] = 1.0 + rng.uniform(5.0, 50.0, c.ntot)
    a = c.copy()
    oa = OracleBlock(prm, a, level=2)
    ctx = context(prm, blocks)
    ctx.upload(level=2)
    oa.save_old(); ctx.save_old(level=2)
    oa.state(); ctx.statim(level=2)
    if iturb:
        ctx.download(level=2)
        a["VIST"] = (a["EPS2"] - 1.0) * a["VIS"]
        c["VIST"] = a["VIST"]
        ctx.upload(level=2, fields=["VIST"])
    oa.residual(m2=2); ctx.residi(level=2, m2=2); ctx.download(level=2)
    res = SST_RES if iturb else RES
    assert_fields_equal(c, a, res, rtol=RTOL if iturb else 0.0, what="coarse residual")
    c.F[:] = a.F
    ctx.upload(level=2)
    oa.implicit(); ctx.impsdi(level=2); ctx.download(level=2)
    assert_fields_equal(c, a, ["DTL"] + res[:7], what="coarse implicit")
    oa.update(); ctx.fin(level=2); ctx.download(level=2)
    prim = SST_PRIM if iturb else PRIM
    assert_fields_equal(c, a, prim, interior=False, what="coarse update")
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("nlev", [2, 3])
@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycles(hip_device, iturb, nlev, lusgs):
    print(f"case V-cycles iturb={iturb} nlev={nlev} LU-SGS {lusgs}", flush=True)
    prm, blocks = hierarchy(iturb, nlev=nlev, dims=(16, 8, 12), seed=73)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = context(prm, blocks)
    ctx.lusgs_mode(lusgs)
    for _ in range(3):
        r1 = oh.vcycle()
    r2 = ctx.vcycle(nlev, ncycles=3)
    for m in range(nlev):
        ctx.download(level=m + 1)
    prim = SST_PRIM if iturb else PRIM
    for m in range(nlev):
        assert_fields_equal(blocks[m], host[m], prim, interior=False,
                            rtol=RTOL if iturb else 0.0, what=f"V-cycle level {m + 1}")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if iturb else 1e-13)
    ctx.close()


@pytest.mark.parametrize("nlev", [2, 3])
@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycles_planar(hip_device, iturb, nlev):
    """A block one cell thick in K (KMAX = 1 on every level): TRANSF's planar path."""
    print(f"case planar V-cycles iturb={iturb} nlev={nlev}", flush=True)
    prm, blocks = hierarchy(iturb, nlev=nlev, dims=(24, 16, 1), seed=75)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = context(prm, blocks)
    for _ in range(3):
        r1 = oh.vcycle()
    r2 = ctx.vcycle(nlev, ncycles=3)
    for m in range(nlev):
        ctx.download(level=m + 1)
    prim = SST_PRIM if iturb else PRIM
    for m in range(nlev):
        assert_fields_equal(blocks[m], host[m], prim, interior=False,
                            rtol=RTOL if iturb else 0.0, what=f"planar V-cycle level {m + 1}")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if iturb else 1e-13)
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycles_fullns(hip_device, iturb):
    """FULLNS V-cycles: FLUXBL (and FLKEBL) with the TEMDER gradients on level 1, the
    thin-layer fluxes on the coarse levels."""
    print(f"case FULLNS V-cycles iturb={iturb} nlev=3", flush=True)
    prm, blocks = hierarchy(iturb, nlev=3, dims=(16, 8, 12), seed=75, fullns=1)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = context(prm, blocks)
    for _ in range(3):
        r1 = oh.vcycle()
    r2 = ctx.vcycle(3, ncycles=3)
    for m in range(3):
        ctx.download(level=m + 1)
    prim = SST_PRIM if iturb else PRIM
    for m in range(3):
        assert_fields_equal(blocks[m], host[m], prim, interior=False,
                            rtol=RTOL if iturb else 0.0, what=f"FULLNS V-cycle level {m + 1}")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if iturb else 1e-13)
    ctx.close()


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycles_walls(hip_device, iturb):
    """V-cycles with a solid wall on the J-low face of every level (and a partial K-high
    wall): BOTSKE on all levels, the force sums on level 1."""
    print(f"case V-cycles with walls iturb={iturb} nlev=3", flush=True)
    prm, blocks = hierarchy(iturb, nlev=3, dims=(16, 12, 12), seed=77)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = context(prm, blocks)
    pw_h, pw_g = [], []
    for m, b in enumerate(blocks):
        specs = [(2, 1, b.imax, 1, b.kmax, 0), (6, 2, b.imax - 1, 1, b.jmax, 1)]
        wp, patches = make_walls(b, prm, specs, 31 + m)
        pw_h.append([p.copy() for p in patches])
        pw_g.append([p.copy() for p in patches])
        oh.levels[m].set_walls(wp, pw_h[m])
        ctx.wall_register(wp, pw_g[m], level=m + 1)
    for _ in range(3):
        r1 = oh.vcycle()
    r2 = ctx.vcycle(3, ncycles=3)
    for m in range(3):
        ctx.download(level=m + 1)
        ctx.wall_sync(1, level=m + 1)
    prim = SST_PRIM if iturb else PRIM
    for m in range(3):
        assert_fields_equal(blocks[m], host[m], prim, interior=False,
                            rtol=RTOL if iturb else 0.0, what=f"V-cycle with walls level {m + 1}")
    assert_patches_equal(pw_g[0], pw_h[0], rtol=RTOL if iturb else 0.0, what="level-1 patches")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if iturb else 1e-13)
    for l in range(2):
        # with the SST the state differs from the oracle's by the device tanh after a few
        # cycles; sums whose terms cancel (QY, QZ) are held to the scale of the largest sum
        f0 = oh.levels[0].forces(l)
        np.testing.assert_allclose(ctx.wall_forces(l), f0, rtol=1e-10 if iturb else 0,
                                   atol=1e-10 * np.abs(f0).max() if iturb else 0)
    ctx.close()


@pytest.mark.parametrize("nlev", [1, 3])
@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycle_graph_replays(hip_device, iturb, nlev):
    """The captured V-cycle against eager V-cycles of a twin context, bit for bit in every
    field of every level after each replay, with eager cycles, downloads and an option change
    (which captures again) between the replays."""
    print(f"case V-cycle graph iturb={iturb} nlev={nlev}", flush=True)
    prm, blocks = hierarchy(iturb, nlev=nlev, dims=(16, 8, 12), seed=79)
    twin = [b.copy() for b in blocks]
    cg, ce = context(prm, blocks), context(prm, twin)

    def check(what):
        for m in range(nlev):
            cg.download(level=m + 1)
            ce.download(level=m + 1)
            np.testing.assert_array_equal(blocks[m].F, twin[m].F, err_msg=f"{what} level {m + 1}")

    for rep in range(3):
        ce.vcycle(nlev)
        cg.vcycle_graph(nlev)
        check(f"replay {rep}")
    # eager work on the same context between replays
    ce.vcycle(nlev, ncycles=2)
    cg.vcycle(nlev, ncycles=2)
    check("eager cycles")
    ce.vcycle(nlev, ncycles=2)
    cg.vcycle_graph(nlev, ncycles=2)
    check("two replays in one call")
    for mode in ("tiles", "staged1", "staged"):
        ce.lusgs_mode(mode)
        cg.lusgs_mode(mode)
        for rep in range(2):
            ce.vcycle(nlev)
            cg.vcycle_graph(nlev)
            check(f"LU-SGS {mode} replay {rep}")
    cg.close()
    ce.close()


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycle_graph_fullns(hip_device, iturb):
    """FULLNS V-cycles (TEMDER gradients and FLUXBL/FLKEBL on level 1) replayed from the
    graph, bit for bit the eager ones of a twin context, with eager cycles between."""
    print(f"case V-cycle graph FULLNS iturb={iturb} nlev=3", flush=True)
    prm, blocks = hierarchy(iturb, nlev=3, dims=(16, 8, 12), seed=83, fullns=1)
    twin = [b.copy() for b in blocks]
    cg, ce = context(prm, blocks), context(prm, twin)
    for rep in range(3):
        ce.vcycle(3, ncycles=rep + 1)
        cg.vcycle_graph(3, ncycles=rep + 1)
        for m in range(3):
            cg.download(level=m + 1)
            ce.download(level=m + 1)
            np.testing.assert_array_equal(blocks[m].F, twin[m].F, err_msg=f"replay {rep} level {m + 1}")
        ce.vcycle(3)
        cg.vcycle(3)
    cg.close()
    ce.close()


@pytest.mark.parametrize("iturb", [0, 6], ids=["laminar", "sst"])
def test_vcycle_graph_walls(hip_device, iturb):
    """V-cycles with solid walls on every level replayed from the captured graph (the force
    pass forked onto its side stream and joined inside the graph) against the oracle's
    V-cycles: every solved field of every level, the level-1 patch arrays and force sums,
    with eager V-cycles of the same context between the replays."""
    print(f"case V-cycle graph with walls iturb={iturb} nlev=3", flush=True)
    prm, blocks = hierarchy(iturb, nlev=3, dims=(16, 12, 12), seed=87)
    host = [b.copy() for b in blocks]
    oh = OracleHierarchy(prm, host)
    ctx = context(prm, blocks)
    pw_h, pw_g = [], []
    for m, b in enumerate(blocks):
        specs = [(2, 1, b.imax, 1, b.kmax, 0), (6, 2, b.imax - 1, 1, b.jmax, 1)]
        wp, patches = make_walls(b, prm, specs, 41 + m)
        pw_h.append([p.copy() for p in patches])
        pw_g.append([p.copy() for p in patches])
        oh.levels[m].set_walls(wp, pw_h[m])
        ctx.wall_register(wp, pw_g[m], level=m + 1)
    prim = SST_PRIM if iturb else PRIM
    tol = RTOL if iturb else 0.0
    for rep, (ng, ne) in enumerate(((2, 0), (1, 1), (3, 0))):
        for _ in range(ng + ne):
            oh.vcycle()
        ctx.vcycle_graph(3, ncycles=ng)
        if ne:
            ctx.vcycle(3, ncycles=ne)
        for m in range(3):
            ctx.download(level=m + 1)
            ctx.wall_sync(1, level=m + 1)
            assert_fields_equal(blocks[m], host[m], prim, interior=False, rtol=tol,
                                what=f"graph replay {rep} level {m + 1}")
        assert_patches_equal(pw_g[0], pw_h[0], rtol=tol, what=f"graph replay {rep} patches")
        for l in range(2):
            f0 = oh.levels[0].forces(l)
            np.testing.assert_allclose(ctx.wall_forces(l), f0, rtol=1e-10 if iturb else 0,
                                       atol=1e-10 * np.abs(f0).max() if iturb else 0)
    ctx.close()
