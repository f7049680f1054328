"""The C restatement (oracle/ff_oracle.c) against the reference's own Fortran routines
(oracle/_ref, built from the unmodified sources by oracle/Makefile): bit-exact per stage
and over several cycles."""
import numpy as np
import pytest

from oracle.oracle_py import OracleBlock, RefBlock, have_ref
from tests.parity import (BIJ, CASES, EARSM_CASES, FULLNS_CASES, GRADKW, GRADT, PRIM, RES,
                          SST_CASES, SST_PRIM, SST_RES, SST_TURB, STATE, WALL_CASES,
                          assert_patches_equal, make_earsm_case, make_fullns_case,
                          make_wall_case, assert_fields_equal, make_case)

pytestmark = pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (needs /root/reference)")


@pytest.mark.parametrize("name,dims,pover,ckw", CASES, ids=[c[0] for c in CASES])
def test_stages_bit_exact(name, dims, pover, ckw):
    prm, blk = make_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    oa.save_old(); rb.save_old()
    oa.state(); rb.state()
    assert_fields_equal(a, b, STATE, interior=False, what="STATIM")
    oa.turb(); rb.turb()
    r1, r2 = oa.residual(), rb.residual()
    assert_fields_equal(a, b, RES, what="RESIDU")
    assert np.array_equal(r1[:5], r2[:5]), "RESL2 sums differ"
    oa.implicit(); rb.implicit()
    assert_fields_equal(a, b, ["DTL"] + RES, what="IMPS")
    oa.update(); rb.update()
    assert_fields_equal(a, b, PRIM, interior=False, what="FIN")


@pytest.mark.parametrize("name,dims,pover,ckw", CASES[:3], ids=[c[0] for c in CASES[:3]])
def test_cycles_bit_exact(name, dims, pover, ckw):
    prm, blk = make_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    for _ in range(4):
        r1, r2 = oa.cycle(), rb.cycle()
        assert np.array_equal(r1[:5], r2[:5])
    assert_fields_equal(a, b, PRIM, interior=False, what="4 cycles")


@pytest.mark.parametrize("name,dims,pover,ckw", SST_CASES, ids=[c[0] for c in SST_CASES])
def test_sst_stages_bit_exact(name, dims, pover, ckw):
    """k-omega SST: TURBEQ, RESIKE, IMPS with PRKEGS/DIKEGS/COKEGS and FIN, stage by stage.
    CROSKO rounds RK and REPS on every cell, so they are compared after RESIDI too."""
    prm, blk = make_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    oa.save_old(); rb.save_old()
    oa.state(); rb.state()
    assert_fields_equal(a, b, STATE, interior=False, what="STATIM")
    oa.turb(); rb.turb()
    assert_fields_equal(a, b, SST_TURB, what="TURBEQ")
    assert_fields_equal(a, b, ["TTS"], interior=False, what="TTIMES")
    r1, r2 = oa.residual(), rb.residual()
    assert_fields_equal(a, b, ["RK", "REPS"], interior=False, what="CROSKO")
    assert_fields_equal(a, b, SST_RES, what="RESIKE")
    assert np.array_equal(r1, r2), "RESL2 sums differ"
    oa.implicit(); rb.implicit()
    assert_fields_equal(a, b, ["DTL"] + SST_RES[:7], what="IMPS")
    oa.update(); rb.update()
    assert_fields_equal(a, b, SST_PRIM, interior=False, what="FIN")


@pytest.mark.parametrize("name,dims,pover,ckw", SST_CASES[:2], ids=[c[0] for c in SST_CASES[:2]])
def test_sst_cycles_bit_exact(name, dims, pover, ckw):
    prm, blk = make_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    for _ in range(4):
        r1, r2 = oa.cycle(), rb.cycle()
        assert np.array_equal(r1, r2)
    assert_fields_equal(a, b, SST_PRIM, interior=False, what="4 cycles")


# coarse multigrid levels: RESIDI with INTI = 55 and, for SST, RESIKE's own DERFUN (M2 >= 2,
# interior + EXTEND); IMPSDI with CFLL; FIN with the coarse-level limiters of CYCLES
COARSE_CASES = [
    ("lam_coarse", (6, 8, 4), dict(cfll=3.0), dict(kind="all", amp=0.05, distort=0.12, seed=51)),
    ("sst_coarse", (8, 6, 6), dict(iturb=6, cfll=3.0),
     dict(kind="all", amp=0.05, distort=0.12, seed=53, turb=(1.0, 60.0))),
    ("sst_coarse_wall", (6, 6, 4), dict(iturb=6, re=1e14, cfll=2.0),
     dict(kind="velocity", amp=0.03, jratio=1.3, seed=57, turb=(0.5, 300.0))),
]


@pytest.mark.parametrize("level", [2, 3])
@pytest.mark.parametrize("name,dims,pover,ckw", COARSE_CASES, ids=[c[0] for c in COARSE_CASES])
def test_coarse_level_stages_bit_exact(name, dims, pover, ckw, level):
    prm, blk = make_case(dims, pover, ckw)
    if prm.iturb == 6:
        # on a coarse level EPS2 and VIST are the restricted fine values
        rng = np.random.default_rng(61)
        blk["EPS2"] = 1.0 + rng.uniform(5.0, 50.0, blk.ntot)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a, level=level), RefBlock(prm, b, level=level)
    oa.save_old(); rb.save_old()
    oa.state(); rb.state()
    if prm.iturb == 6:
        a["VIST"] = (a["EPS2"] - 1.0) * a["VIS"]
        b["VIST"] = a["VIST"]
    oa.residual(m2=level); rb.residual(m2=level)
    if prm.iturb == 6:
        assert_fields_equal(a, b, SST_TURB[:10], interior=False, what="DERFUN+EXTEND")
        assert_fields_equal(a, b, SST_RES, what="RESIKE coarse")
    else:
        assert_fields_equal(a, b, RES, what="RESIDU coarse")
    oa.implicit(); rb.implicit()
    nres = SST_RES[:7] if prm.iturb == 6 else RES
    assert_fields_equal(a, b, ["DTL"] + nres, what="IMPS coarse")
    oa.update(); rb.update()
    prim = SST_PRIM if prm.iturb == 6 else PRIM
    assert_fields_equal(a, b, prim, interior=False, what="FIN coarse")


@pytest.mark.parametrize("name,dims,pover,ckw", SST_CASES[:2], ids=[c[0] for c in SST_CASES[:2]])
def test_sst_fine_level_second_residual(name, dims, pover, ckw):
    """The level-1 residual the V-cycle re-evaluates after the update (RESIDI(1,N,2)):
    RESIKE forms the velocity gradients itself, first ghost layer included."""
    prm, blk = make_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    for o in (oa, rb):
        o.save_old(); o.state(); o.turb(); o.residual(); o.implicit(); o.update(); o.state()
    r1, r2 = oa.residual(m2=2), rb.residual(m2=2)
    assert_fields_equal(a, b, SST_TURB[:11], interior=False, what="DERFUN level 1")
    assert_fields_equal(a, b, SST_RES, what="RESIKE M2=2")
    assert np.array_equal(r1, r2)


@pytest.mark.parametrize("name,dims,pover,ckw", EARSM_CASES, ids=[c[0] for c in EARSM_CASES])
def test_earsm_stages_bit_exact(name, dims, pover, ckw):
    """SST with the Wallin-Johansson EARSM (ISTRES = 1): EARSM1 in TURBEQ, PEEKOO and
    ANISFLUX in RESIKE, then IMPS and FIN, stage by stage against the reference's Fortran."""
    prm, blk = make_earsm_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    oa.save_old(); rb.save_old()
    oa.state(); rb.state()
    oa.turb(); rb.turb()
    assert_fields_equal(a, b, SST_TURB, what="TURBEQ")
    assert_fields_equal(a, b, BIJ, interior=False, what="EARSM1")
    assert np.abs(a.interior("B12")).max() > 0, "EARSM1 wrote no anisotropy"
    r1, r2 = oa.residual(), rb.residual()
    assert_fields_equal(a, b, SST_RES, what="RESIKE with PEEKOO and ANISFLUX")
    assert np.array_equal(r1, r2), "RESL2 sums differ"
    oa.implicit(); rb.implicit()
    assert_fields_equal(a, b, ["DTL"] + SST_RES[:7], what="IMPS")
    oa.update(); rb.update()
    assert_fields_equal(a, b, SST_PRIM, interior=False, what="FIN")


@pytest.mark.parametrize("name,dims,pover,ckw", EARSM_CASES[:2], ids=[c[0] for c in EARSM_CASES[:2]])
def test_earsm_cycles_bit_exact(name, dims, pover, ckw):
    prm, blk = make_earsm_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    for _ in range(4):
        r1, r2 = oa.cycle(), rb.cycle()
        assert np.array_equal(r1, r2)
    assert_fields_equal(a, b, SST_PRIM + BIJ, interior=False, what="4 EARSM cycles")


@pytest.mark.parametrize("name,dims,pover,ckw", FULLNS_CASES, ids=[c[0] for c in FULLNS_CASES])
def test_fullns_stages_bit_exact(name, dims, pover, ckw):
    """FULLNS on level 1: DERFUN (laminar: inside RESIDU) and TEMDER with the first ghost
    layer, FLUXBL in place of FLUXDO and, for the SST, FLKEBL in place of FLUDKE."""
    prm, blk = make_fullns_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    kw = prm.iturb == 6
    for o in (oa, rb):
        o.save_old(); o.state(); o.turb()
    r1, r2 = oa.residual(), rb.residual()
    assert_fields_equal(a, b, SST_TURB[:10] + GRADT + (GRADKW if kw else []), interior=False,
                        what="DERFUN+TEMDER")
    assert max(np.abs(a.interior(g)).max() for g in GRADT + (GRADKW if kw else [])) > 0, \
        "TEMDER wrote no gradient"
    assert_fields_equal(a, b, SST_RES if kw else RES, what="RESIDU/RESIKE with FLUXBL")
    assert np.array_equal(r1, r2), "RESL2 sums differ"
    oa.implicit(); rb.implicit()
    assert_fields_equal(a, b, ["DTL"] + (SST_RES[:7] if kw else RES), what="IMPS")
    oa.update(); rb.update()
    assert_fields_equal(a, b, SST_PRIM if kw else PRIM, interior=False, what="FIN")


@pytest.mark.parametrize("name,dims,pover,ckw", FULLNS_CASES[::3], ids=[c[0] for c in FULLNS_CASES[::3]])
def test_fullns_cycles_bit_exact(name, dims, pover, ckw):
    prm, blk = make_fullns_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    for _ in range(4):
        r1, r2 = oa.cycle(), rb.cycle()
        assert np.array_equal(r1, r2)
    prim = SST_PRIM if prm.iturb == 6 else PRIM
    assert_fields_equal(a, b, prim, interior=False, what="4 FULLNS cycles")


@pytest.mark.parametrize("name,dims,pover,ckw", [FULLNS_CASES[0], FULLNS_CASES[3]],
                         ids=[FULLNS_CASES[0][0], FULLNS_CASES[3][0]])
def test_fullns_coarse_level_is_thin_layer(name, dims, pover, ckw):
    """On a coarse level FULLNS leaves the thin-layer FLUXDO/FLUDKE in place (FLUXBL and
    FLKEBL run on M = 1 only); the reference still forms the gradients there, which no
    coarse-level flux reads."""
    prm, blk = make_fullns_case(dims, pover, ckw)
    if prm.iturb == 6:
        rng = np.random.default_rng(61)
        blk["EPS2"] = 1.0 + rng.uniform(5.0, 50.0, blk.ntot)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a, level=2), RefBlock(prm, b, level=2)
    for o in (oa, rb):
        o.save_old(); o.state()
    if prm.iturb == 6:
        a["VIST"] = (a["EPS2"] - 1.0) * a["VIS"]
        b["VIST"] = a["VIST"]
    oa.residual(m2=2); rb.residual(m2=2)
    assert_fields_equal(a, b, SST_RES if prm.iturb == 6 else RES, what="coarse residual")


@pytest.mark.parametrize("name,dims,pover,ckw,specs,grav", WALL_CASES, ids=[c[0] for c in WALL_CASES])
def test_walls_stages_bit_exact(name, dims, pover, ckw, specs, grav):
    """BOTSKE on solid-wall patches: the wall fluxes inside RESIDU/RESIKE, the patch arrays
    (Cp, T, heat fluxes, shear stresses, surface forces), PTUR next to the wall (TURCOR),
    the force sums, then IMPS and FIN."""
    prm, blk, wp, patches = make_wall_case(dims, pover, ckw, specs, grav)
    kw = prm.iturb == 6
    a, b = blk.copy(), blk.copy()
    pa, pb = [p.copy() for p in patches], [p.copy() for p in patches]
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    oa.set_walls(wp, pa); rb.set_walls(wp, pb)
    for o in (oa, rb):
        o.save_old(); o.state(); o.turb()
    r1, r2 = oa.residual(), rb.residual()
    assert_fields_equal(a, b, (SST_RES if kw else RES), what="residual with BOTSKE")
    assert np.array_equal(r1, r2), "RESL2 sums differ"
    assert_patches_equal(pa, pb, what="patch arrays")
    for l in range(len(patches)):
        f1, f2 = oa.forces(l), rb.forces(l)
        assert np.array_equal(f1, f2), f"forces of patch {l}: {f1 - f2}"
        assert np.abs(f1[:3]).max() > 0, "no force"
    oa.implicit(); rb.implicit()
    assert_fields_equal(a, b, ["DTL"] + (SST_RES[:7] if kw else RES), what="IMPS")
    oa.update(); rb.update()
    assert_fields_equal(a, b, SST_PRIM if kw else PRIM, interior=False, what="FIN")


@pytest.mark.parametrize("name,dims,pover,ckw,specs,grav", [WALL_CASES[i] for i in (1, 2)],
                         ids=[WALL_CASES[i][0] for i in (1, 2)])
def test_walls_cycles_and_coarse_level(name, dims, pover, ckw, specs, grav):
    """Four whole cycles with walls, and a coarse-level residual (no forces there)."""
    prm, blk, wp, patches = make_wall_case(dims, pover, ckw, specs, grav)
    kw = prm.iturb == 6
    a, b = blk.copy(), blk.copy()
    pa, pb = [p.copy() for p in patches], [p.copy() for p in patches]
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    oa.set_walls(wp, pa); rb.set_walls(wp, pb)
    for _ in range(4):
        r1, r2 = oa.cycle(), rb.cycle()
        assert np.array_equal(r1, r2)
    assert_fields_equal(a, b, SST_PRIM if kw else PRIM, interior=False, what="4 cycles with walls")
    assert_patches_equal(pa, pb, what="patch arrays after 4 cycles")
    c, d = blk.copy(), blk.copy()
    if kw:
        rng = np.random.default_rng(61)
        c["EPS2"] = 1.0 + rng.uniform(5.0, 50.0, c.ntot)
        d.F[:] = c.F
    qa, qb = [p.copy() for p in patches], [p.copy() for p in patches]
    oc, rc = OracleBlock(prm, c, level=2), RefBlock(prm, d, level=2)
    oc.set_walls(wp, qa); rc.set_walls(wp, qb)
    for o in (oc, rc):
        o.save_old(); o.state()
    if kw:
        c["VIST"] = (c["EPS2"] - 1.0) * c["VIS"]
        d["VIST"] = c["VIST"]
    oc.residual(m2=2); rc.residual(m2=2)
    assert_fields_equal(c, d, SST_RES if kw else RES, what="coarse residual with walls")
    assert_patches_equal(qa, qb, what="coarse patch arrays")
