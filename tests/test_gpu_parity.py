"""HIP path (libffhip.so, called through its C-ABI) against the oracle.

Laminar: bit-exact FP64.  k-omega SST: the SST blending functions call tanh (FUN1KO,
EPS2KO), which the device math library and the host C library round differently in the last
bit, so SST fields are compared to 1e-12 of each field's largest magnitude (the FP64
parity bar), with everything upstream of the first tanh still compared bit for bit."""
import numpy as np
import pytest

from oracle.oracle_py import OracleBlock
from tests.parity import (BIJ, CASES, EARSM_CASES, FULLNS_CASES, GRADKW, GRADT, PRIM, RES,
                          SST_CASES, SST_PRIM, SST_RES, SST_TURB, STATE, WALL_CASES,
                          assert_fields_equal, assert_patches_equal, make_case,
                          make_earsm_case, make_fullns_case, make_wall_case)

pytestmark = pytest.mark.gpu

SST_RTOL = 1e-12


def _ctx(prm, blk, lusgs="tiles"):
    from finflo_amd.ffhip import Context
    ctx = Context(prm)
    ctx.lusgs_mode(lusgs)
    ctx.register(blk)
    ctx.upload()
    return ctx


@pytest.mark.parametrize("name,dims,pover,ckw", CASES, ids=[c[0] for c in CASES])
def test_stages(hip_device, name, dims, pover, ckw):
    print(f"case {name} dims {dims}", flush=True)
    prm, blk = make_case(dims, pover, ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g)
    oa.save_old(); ctx.save_old()
    oa.state(); ctx.statim(); ctx.download()
    assert_fields_equal(g, a, STATE, interior=False, what="STATIM")
    oa.turb(); ctx.turbdi()
    r1 = oa.residual(); r2 = ctx.residi(); ctx.download()
    assert_fields_equal(g, a, RES, what="RESIDU")
    np.testing.assert_allclose(r2[:5], r1[:5], rtol=1e-13)
    oa.implicit(); ctx.impsdi(); ctx.download()
    assert_fields_equal(g, a, ["DTL"] + RES, what="IMPS")
    oa.update(); ctx.fin(); ctx.download()
    assert_fields_equal(g, a, PRIM, interior=False, what="FIN")
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "planes", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("name,dims,pover,ckw", CASES[:4], ids=[c[0] for c in CASES[:4]])
def test_cycles(hip_device, name, dims, pover, ckw, lusgs):
    print(f"case {name} dims {dims} 5 cycles, LU-SGS {lusgs}", flush=True)
    prm, blk = make_case(dims, pover, ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g, lusgs)
    for _ in range(5):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=5)
    ctx.download()
    assert_fields_equal(g, a, PRIM, interior=False, what="5 cycles")
    np.testing.assert_allclose(r2[:5], r1[:5], rtol=1e-13)
    ctx.close()


# blocks spanning several 8x8 (J,K) tiles and 16x4 (I,J) tiles of the LU-SGS wavefronts,
# with ragged last tiles and extents that are multiples of the tile edges
MULTI_TILE = [(40, 24, 20), (20, 19, 13), (9, 17, 26), (33, 8, 16)]


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("dims", MULTI_TILE, ids=["x".join(map(str, d)) for d in MULTI_TILE])
def test_larger_block_cycle(hip_device, dims, lusgs):
    """Blocks with more cells than one workgroup per row/plane/tile."""
    print(f"case large {dims} LU-SGS {lusgs}", flush=True)
    prm, blk = make_case(dims, {}, dict(kind="all", amp=0.03, distort=0.1, seed=21))
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g, lusgs)
    for _ in range(2):
        oa.cycle()
    ctx.cycle(ncycles=2)
    ctx.download()
    assert_fields_equal(g, a, PRIM, interior=False, what=f"{dims}")
    ctx.close()


def test_implicit_stage_tiles_vs_planes(hip_device):
    """The three LU-SGS forms give the same increments bit for bit, stage by stage."""
    print("case tiles-vs-planes-vs-kmarch-vs-staged 26x21x18", flush=True)
    prm, blk = make_case((26, 21, 18), {}, dict(kind="all", amp=0.04, distort=0.1, seed=31))
    out = {}
    for mode in ("tiles", "planes", "kmarch", "staged", "staged1"):
        g = blk.copy()
        ctx = _ctx(prm, g, mode)
        ctx.save_old(); ctx.statim(); ctx.turbdi(); ctx.residi(); ctx.impsdi()
        ctx.download()
        out[mode] = g
        ctx.close()
    assert_fields_equal(out["tiles"], out["planes"], ["DTL"] + RES, what="tiles vs planes")
    assert_fields_equal(out["kmarch"], out["planes"], ["DTL"] + RES, what="kmarch vs planes")
    assert_fields_equal(out["staged"], out["planes"], ["DTL"] + RES, what="staged vs planes")
    assert_fields_equal(out["staged1"], out["planes"], ["DTL"] + RES, what="staged1 vs planes")


@pytest.mark.parametrize("dims", [(70, 9, 17), (5, 3, 2), (130, 6, 9)], ids=["70x9x17", "5x3x2", "130x6x9"])
def test_residual_tiles_vs_cells_vs_oracle(hip_device, dims):
    """The face-once tile residual, the per-cell residual and the oracle agree bit for bit
    on blocks with ragged I-, J- and K-tiles."""
    print(f"case residual forms {dims}", flush=True)
    prm, blk = make_case(dims, {}, dict(kind="all", amp=0.05, distort=0.12, seed=41))
    a = blk.copy()
    oa = OracleBlock(prm, a)
    oa.save_old(); oa.state(); oa.turb(); r0 = oa.residual()
    for mode in ("tiles", "cells"):
        g = blk.copy()
        from finflo_amd.ffhip import Context
        ctx = Context(prm)
        ctx.residual_mode(mode)
        ctx.register(g)
        ctx.upload()
        ctx.save_old(); ctx.statim(); ctx.turbdi(); r = ctx.residi()
        ctx.download()
        assert_fields_equal(g, a, RES, what=f"RESIDU ({mode})")
        np.testing.assert_allclose(r[:5], r0[:5], rtol=1e-13)
        ctx.close()


@pytest.mark.parametrize("name,dims,pover,ckw", SST_CASES, ids=[c[0] for c in SST_CASES])
def test_sst_stages(hip_device, name, dims, pover, ckw):
    """k-omega SST stage by stage; each GPU stage starts from the oracle's state of the
    previous stage, so a stage's error is its own."""
    print(f"case {name} dims {dims}", flush=True)
    prm, blk = make_case(dims, pover, ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g)
    oa.save_old(); ctx.save_old()
    oa.state(); ctx.statim(); ctx.download()
    assert_fields_equal(g, a, STATE, interior=False, what="STATIM")
    oa.turb(); ctx.turbdi(); ctx.download()
    # gradients, vorticity and strain use no tanh: bit for bit
    assert_fields_equal(g, a, SST_TURB[:11], what="DERFUN")
    assert_fields_equal(g, a, ["TTS"], interior=False, what="TTIMES")
    assert_fields_equal(g, a, ["EPS2", "VIST"], rtol=SST_RTOL, what="EPS2KO")
    g.F[:] = a.F
    ctx.upload()
    r1 = oa.residual(); r2 = ctx.residi(); ctx.download()
    assert_fields_equal(g, a, ["RK", "REPS"], interior=False, what="CROSKO rounding")
    assert_fields_equal(g, a, SST_RES, rtol=SST_RTOL, what="RESIKE")
    np.testing.assert_allclose(r2, r1, rtol=1e-11)
    g.F[:] = a.F
    ctx.upload()
    oa.implicit(); ctx.impsdi(); ctx.download()
    assert_fields_equal(g, a, ["DTL"], what="TIME3")
    assert_fields_equal(g, a, SST_RES[:7], what="IMPS")
    g.F[:] = a.F
    ctx.upload()
    oa.update(); ctx.fin(); ctx.download()
    assert_fields_equal(g, a, SST_PRIM, interior=False, what="FIN")
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("name,dims,pover,ckw", SST_CASES[:2], ids=[c[0] for c in SST_CASES[:2]])
def test_sst_cycles(hip_device, name, dims, pover, ckw, lusgs):
    print(f"case {name} dims {dims} 5 SST cycles, LU-SGS {lusgs}", flush=True)
    prm, blk = make_case(dims, pover, ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g, lusgs)
    for _ in range(5):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=5)
    ctx.download()
    assert_fields_equal(g, a, SST_PRIM, interior=False, rtol=SST_RTOL, what="5 SST cycles")
    np.testing.assert_allclose(r2, r1, rtol=1e-10)
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("dims", [(40, 24, 20), (21, 19, 13)], ids=["40x24x20", "21x19x13"])
def test_sst_multi_tile(hip_device, dims, lusgs):
    """SST blocks spanning several LU-SGS tiles and residual tiles, with ragged last tiles."""
    print(f"case SST large {dims} LU-SGS {lusgs}", flush=True)
    prm, blk = make_case(dims, dict(iturb=6), dict(kind="all", amp=0.03, distort=0.1, seed=43,
                                                   turb=(1.0, 200.0)))
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g, lusgs)
    for _ in range(2):
        oa.cycle()
    ctx.cycle(ncycles=2)
    ctx.download()
    assert_fields_equal(g, a, SST_PRIM, interior=False, rtol=SST_RTOL, what=f"SST {dims}")
    ctx.close()


@pytest.mark.parametrize("name,dims,pover,ckw", EARSM_CASES, ids=[c[0] for c in EARSM_CASES])
def test_earsm_stages(hip_device, name, dims, pover, ckw):
    """SST + Wallin-Johansson EARSM stage by stage; each GPU stage starts from the oracle's
    state of the previous stage.  EARSM1's cube roots (pow), ACOS and COS, like the SST's
    tanh, round differently on the device in the last bit: compared to 1e-12."""
    print(f"case {name} dims {dims}", flush=True)
    prm, blk = make_earsm_case(dims, pover, ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g)
    oa.save_old(); ctx.save_old()
    oa.state(); ctx.statim()
    oa.turb(); ctx.turbdi(); ctx.download()
    assert_fields_equal(g, a, BIJ, interior=False, rtol=SST_RTOL, what="EARSM1")
    g.F[:] = a.F
    ctx.upload()
    r1 = oa.residual(); r2 = ctx.residi(); ctx.download()
    assert_fields_equal(g, a, SST_RES, rtol=SST_RTOL, what="RESIKE with PEEKOO and ANISFLUX")
    np.testing.assert_allclose(r2, r1, rtol=1e-11)
    g.F[:] = a.F
    ctx.upload()
    oa.implicit(); ctx.impsdi(); ctx.download()
    assert_fields_equal(g, a, SST_RES[:7], what="IMPS")
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("name,dims,pover,ckw", EARSM_CASES[:2], ids=[c[0] for c in EARSM_CASES[:2]])
def test_earsm_cycles(hip_device, name, dims, pover, ckw, lusgs):
    print(f"case {name} dims {dims} 5 EARSM cycles, LU-SGS {lusgs}", flush=True)
    prm, blk = make_earsm_case(dims, pover, ckw)
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g, lusgs)
    for _ in range(5):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=5)
    ctx.download()
    assert_fields_equal(g, a, SST_PRIM + BIJ, interior=False, rtol=SST_RTOL, what="5 EARSM cycles")
    np.testing.assert_allclose(r2, r1, rtol=1e-10)
    ctx.close()


@pytest.mark.parametrize("name,dims,pover,ckw", FULLNS_CASES, ids=[c[0] for c in FULLNS_CASES])
def test_fullns_stages(hip_device, name, dims, pover, ckw):
    """FULLNS on level 1: the DERFUN and TEMDER gradients (first ghost layer included) bit
    for bit; the FLUXBL residual bit for bit (laminar) or, with FLKEBL and the SST's tanh
    upstream, to 1e-12; then IMPS and FIN.  Each GPU stage starts from the oracle's state."""
    print(f"case {name} dims {dims}", flush=True)
    prm, blk = make_fullns_case(dims, pover, ckw)
    kw = prm.iturb == 6
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g)
    oa.save_old(); ctx.save_old()
    oa.state(); ctx.statim()
    oa.turb(); ctx.turbdi(); ctx.download()
    g.F[:] = a.F
    ctx.upload()
    r1 = oa.residual(); r2 = ctx.residi(); ctx.download()
    assert_fields_equal(g, a, SST_TURB[:10] + GRADT + (GRADKW if kw else []), interior=False,
                        what="DERFUN+TEMDER")
    assert_fields_equal(g, a, SST_RES if kw else RES, rtol=SST_RTOL if kw else 0.0,
                        what="residual with FLUXBL")
    np.testing.assert_allclose(r2, r1, rtol=1e-11 if kw else 1e-13)
    g.F[:] = a.F
    ctx.upload()
    oa.implicit(); ctx.impsdi(); ctx.download()
    assert_fields_equal(g, a, ["DTL"] + (SST_RES[:7] if kw else RES), what="IMPS")
    oa.update(); ctx.fin(); ctx.download()
    assert_fields_equal(g, a, SST_PRIM if kw else PRIM, interior=False, what="FIN")
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("name,dims,pover,ckw", [FULLNS_CASES[i] for i in (0, 2, 3, 5)],
                         ids=[FULLNS_CASES[i][0] for i in (0, 2, 3, 5)])
def test_fullns_cycles(hip_device, name, dims, pover, ckw, lusgs):
    print(f"case {name} dims {dims} 5 FULLNS cycles, LU-SGS {lusgs}", flush=True)
    prm, blk = make_fullns_case(dims, pover, ckw)
    kw = prm.iturb == 6
    a, g = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    ctx = _ctx(prm, g, lusgs)
    for _ in range(5):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=5)
    ctx.download()
    assert_fields_equal(g, a, SST_PRIM if kw else PRIM, interior=False,
                        rtol=SST_RTOL if kw else 0.0, what="5 FULLNS cycles")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if kw else 1e-13)
    ctx.close()


@pytest.mark.parametrize("dims", [(70, 9, 17), (5, 3, 2), (130, 6, 9), (40, 24, 1)],
                         ids=["70x9x17", "5x3x2", "130x6x9", "40x24x1"])
def test_fullns_residual_tiles_vs_cells_vs_oracle(hip_device, dims):
    """The FULLNS residual in its tile and per-cell forms against the oracle on ragged
    tiles, including a single K-plane (TEMDER then leaves out the K ghost layer)."""
    print(f"case FULLNS residual forms {dims}", flush=True)
    prm, blk = make_case(dims, dict(fullns=1, rmult=2000.0),
                         dict(kind="all", amp=0.05, distort=0.12, seed=45))
    a = blk.copy()
    oa = OracleBlock(prm, a)
    oa.save_old(); oa.state(); oa.turb(); r0 = oa.residual()
    for mode in ("tiles", "cells"):
        g = blk.copy()
        from finflo_amd.ffhip import Context
        ctx = Context(prm)
        ctx.residual_mode(mode)
        ctx.register(g)
        ctx.upload()
        ctx.save_old(); ctx.statim(); ctx.turbdi(); r = ctx.residi()
        ctx.download()
        assert_fields_equal(g, a, GRADT + SST_TURB[:10], interior=False, what=f"gradients ({mode})")
        assert_fields_equal(g, a, RES, what=f"RESIDU FLUXBL ({mode})")
        np.testing.assert_allclose(r[:5], r0[:5], rtol=1e-13)
        ctx.close()


@pytest.mark.parametrize("name,dims,pover,ckw,specs,grav", WALL_CASES, ids=[c[0] for c in WALL_CASES])
def test_walls_stages(hip_device, name, dims, pover, ckw, specs, grav):
    """BOTSKE at solid-wall patches on the GPU against the oracle: the residual with the wall
    fluxes, the patch arrays and the force sums (side stream), then IMPS and FIN.  Laminar is
    bit for bit.  With the SST the residual carries FUN1KO's tanh (1e-12); the wall quantities
    are bit for bit except F3EPS, whose surface omega takes a device pow."""
    print(f"case {name} dims {dims} walls {specs}", flush=True)
    prm, blk, wp, patches = make_wall_case(dims, pover, ckw, specs, grav)
    kw = prm.iturb == 6
    tol = SST_RTOL if kw else 0.0
    a, g = blk.copy(), blk.copy()
    pa, pg = [p.copy() for p in patches], [p.copy() for p in patches]
    oa = OracleBlock(prm, a)
    oa.set_walls(wp, pa)
    ctx = _ctx(prm, g)
    ctx.wall_register(wp, pg)
    oa.save_old(); ctx.save_old()
    oa.state(); ctx.statim()
    oa.turb(); ctx.turbdi(); ctx.download()
    g.F[:] = a.F
    ctx.upload()
    r1 = oa.residual(); r2 = ctx.residi(); ctx.download(); ctx.wall_sync(1)
    assert_fields_equal(g, a, (SST_RES + ["PTUR"]) if kw else RES, rtol=tol,
                        what="residual with BOTSKE")
    np.testing.assert_allclose(r2, r1, rtol=1e-11 if kw else 1e-13)
    assert_patches_equal(pg, pa, what="patch arrays")
    for l in range(len(patches)):
        f1, f2 = oa.forces(l), ctx.wall_forces(l)
        np.testing.assert_array_equal(f2, f1, err_msg=f"forces of patch {l}")
    g.F[:] = a.F
    ctx.upload()
    oa.implicit(); ctx.impsdi(); ctx.download()
    assert_fields_equal(g, a, ["DTL"] + (SST_RES[:7] if kw else RES), what="IMPS")
    oa.update(); ctx.fin(); ctx.download()
    assert_fields_equal(g, a, SST_PRIM if kw else PRIM, interior=False, what="FIN")
    ctx.close()


@pytest.mark.parametrize("lusgs", ["tiles", "kmarch", "staged", "staged1"])
@pytest.mark.parametrize("name,dims,pover,ckw,specs,grav", [WALL_CASES[i] for i in (1, 2)],
                         ids=[WALL_CASES[i][0] for i in (1, 2)])
def test_walls_cycles(hip_device, name, dims, pover, ckw, specs, grav, lusgs):
    """Five cycles with wall patches: the residual form with the wall fluxes, the per-cell
    form agreeing with it, and the force sums of the last residual."""
    print(f"case {name} dims {dims} 5 cycles with walls, LU-SGS {lusgs}", flush=True)
    prm, blk, wp, patches = make_wall_case(dims, pover, ckw, specs, grav)
    kw = prm.iturb == 6
    a, g = blk.copy(), blk.copy()
    pa, pg = [p.copy() for p in patches], [p.copy() for p in patches]
    oa = OracleBlock(prm, a)
    oa.set_walls(wp, pa)
    ctx = _ctx(prm, g, lusgs)
    ctx.wall_register(wp, pg)
    for _ in range(5):
        r1 = oa.cycle()
    r2 = ctx.cycle(ncycles=5)
    ctx.download(); ctx.wall_sync(1)
    assert_fields_equal(g, a, SST_PRIM if kw else PRIM, interior=False,
                        rtol=SST_RTOL if kw else 0.0, what="5 cycles with walls")
    np.testing.assert_allclose(r2, r1, rtol=1e-10 if kw else 1e-13)
    assert_patches_equal(pg, pa, rtol=SST_RTOL if kw else 0.0, what="patch arrays")
    for l in range(len(patches)):
        f0 = oa.forces(l)
        np.testing.assert_allclose(ctx.wall_forces(l), f0, rtol=1e-10 if kw else 0,
                                   atol=1e-10 * np.abs(f0).max() if kw else 0)
    ctx.close()


def test_walls_residual_cells_form(hip_device):
    """The per-cell residual form takes the wall fluxes like the tile form."""
    name, dims, pover, ckw, specs, grav = WALL_CASES[1]
    prm, blk, wp, patches = make_wall_case(dims, pover, ckw, specs, grav)
    a = blk.copy()
    pa = [p.copy() for p in patches]
    oa = OracleBlock(prm, a)
    oa.set_walls(wp, pa)
    oa.save_old(); oa.state(); oa.turb(); oa.residual()
    from finflo_amd.ffhip import Context
    g = blk.copy()
    pg = [p.copy() for p in patches]
    ctx = Context(prm)
    ctx.residual_mode("cells")
    ctx.register(g)
    ctx.upload()
    ctx.wall_register(wp, pg)
    ctx.save_old(); ctx.statim(); ctx.turbdi(); ctx.residi()
    ctx.download(); ctx.wall_sync(1)
    assert_fields_equal(g, a, RES, what="per-cell RESIDU with walls")
    assert_patches_equal(pg, pa, what="patch arrays (cells form)")
    ctx.close()


def test_wall_register_rejects(hip_device):
    """Patches outside the face, overlapping patches and unsupported heat options are refused
    with an error, before any launch."""
    from finflo_amd.ffhip import Context, FFError
    from tests.parity import make_walls
    prm, blk = make_case((8, 6, 6), {}, dict(kind="all", amp=0.02, seed=3))
    ctx = Context(prm)
    ctx.register(blk)
    for specs in ([(2, 1, 9, 1, 6, 0)], [(2, 1, 8, 1, 6, 4)],
                  [(2, 1, 4, 1, 6, 0), (2, 4, 8, 1, 6, 0)], [(7, 1, 2, 1, 2, 0)]):
        wp, patches = make_walls(blk, prm, specs, 5)
        with pytest.raises(FFError):
            ctx.wall_register(wp, patches)
    ctx.close()
