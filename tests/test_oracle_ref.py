"""The C restatement (oracle/ff_oracle.c) against the reference's own Fortran routines
(oracle/_ref, built from the unmodified sources by oracle/Makefile): bit-exact per stage
and over several cycles."""
import numpy as np
import pytest

from oracle.oracle_py import OracleBlock, RefBlock, have_ref
from tests.parity import (CASES, PRIM, RES, SST_CASES, SST_PRIM, SST_RES, SST_TURB, STATE,
                          assert_fields_equal, make_case)

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
