"""The DDADI implicit stage (LUSGS = 0, imps3.f:321-373) of the C restatement against the
reference's own PRED/DIAG/CORR and PREDKE/DIAGKE/CORRKE (imps4.f, compiled into
oracle/_ref), bit for bit in every entry of the increment fields, ghost entries included:
* every direction (K, I, J) and phase (PRED, DIAG, CORR) alone, on seeded increments and
  time steps, so that each sweep's strides are checked against the reference's sorted
  arrays (the driver's own JKI and KIJ copies; multi.f's SORTZZ does not build here);
* whole implicit stages (TIME3, LOCDT(K), IMSOKO, COPR(KE), the three directions, DIVV12,
  MULV12) and whole cycles, laminar and k-omega SST.
Plus what the mathematics fixes: every sweep is linear in the increments it starts from
(scaling them scales the result), and a block of one cell per line (KMAX = 1 after the
sort) takes PRED's and CORR's single cell to the diagonal scaling alone."""
import numpy as np
import pytest

from oracle.oracle_py import OracleBlock, RefBlock, have_ref
from tests.parity import CASES, PRIM, RES, SST_CASES, SST_PRIM, SST_RES, assert_fields_equal, make_case

INC5 = ["DRO", "DM", "DN", "DW", "DE"]
INC7 = INC5 + ["DRK", "DEPS"]

DD_CASES = [CASES[1], CASES[3], CASES[6], SST_CASES[0], SST_CASES[2]]


def _seeded(dims, pover, ckw, seed):
    """A case after STATIM and TURBDI, with seeded increments on every entry and positive
    local time steps."""
    prm, blk = make_case(dims, pover, ckw)
    o = OracleBlock(prm, blk)
    o.save_old()
    o.state()
    o.turb()
    rng = np.random.default_rng(seed)
    for name in INC7:
        blk[name] = rng.normal(0.0, 1e-3, blk.ntot)
    blk["DTL"] = rng.uniform(0.5e-4, 2e-4, blk.ntot)
    return prm, blk


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("name,dims,pover,ckw", DD_CASES, ids=[c[0] for c in DD_CASES])
def test_every_direction_and_phase_bit_exact(name, dims, pover, ckw):
    prm, blk = _seeded(dims, pover, ckw, seed=5)
    names = INC7 if prm.iturb == 6 else INC5
    for direction in range(3):
        for phase in range(3):
            a, b = blk.copy(), blk.copy()
            OracleBlock(prm, a).ddadi_dir(direction, phase)
            RefBlock(prm, b).ddadi_dir(direction, phase)
            assert_fields_equal(a, b, names, interior=False,
                                what=f"direction {direction} phase {phase}")
            # the sweep changed the increments
            assert not np.array_equal(a.interior("DRO"), blk.interior("DRO"))


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("name,dims,pover,ckw", DD_CASES, ids=[c[0] for c in DD_CASES])
def test_implicit_stage_and_cycles_bit_exact(name, dims, pover, ckw):
    pover = dict(pover, ddadi=1)
    prm, blk = make_case(dims, pover, ckw)
    a, b = blk.copy(), blk.copy()
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    kw = prm.iturb == 6
    oa.save_old(); rb.save_old()
    oa.state(); rb.state()
    oa.turb(); rb.turb()
    oa.residual(); rb.residual()
    oa.implicit(); rb.implicit()
    assert_fields_equal(a, b, ["DTL"] + (SST_RES if kw else RES), interior=False, what="IMPS DDADI")
    oa.update(); rb.update()
    assert_fields_equal(a, b, SST_PRIM if kw else PRIM, interior=False, what="FIN")
    for _ in range(3):
        r1, r2 = oa.cycle(), rb.cycle()
        assert np.array_equal(r1[:5], r2[:5])
    assert_fields_equal(a, b, SST_PRIM if kw else PRIM, interior=False, what="3 DDADI cycles")


@pytest.mark.parametrize("name,dims,pover,ckw", DD_CASES[:2] + DD_CASES[3:4],
                         ids=[c[0] for c in DD_CASES[:2] + DD_CASES[3:4]])
def test_sweeps_are_linear(name, dims, pover, ckw):
    """Each of PRED, DIAG, CORR maps the increments linearly (its coefficients depend only on
    the state): 4 x the increments give 4 x the result, bit for bit (a power of two)."""
    prm, blk = _seeded(dims, pover, ckw, seed=9)
    names = INC7 if prm.iturb == 6 else INC5
    for direction in range(3):
        for phase in range(3):
            a, b = blk.copy(), blk.copy()
            for n in names:
                b[n] = 4.0 * b[n]
            OracleBlock(prm, a).ddadi_dir(direction, phase)
            OracleBlock(prm, b).ddadi_dir(direction, phase)
            for n in names:
                np.testing.assert_array_equal(4.0 * a.interior(n), b.interior(n),
                                              err_msg=f"{n} direction {direction} phase {phase}")


def test_single_cell_lines_are_pointwise():
    """A block one cell deep in K: every line of the K sweep holds one cell, so PRED and CORR
    see no neighbour flux and act on each cell alone: a zero increment field stays zero, and
    doubling one cell's increments doubles that cell's result and changes no other entry."""
    prm, blk = _seeded((5, 4, 1), {}, dict(kind="all", amp=0.05, distort=0.1, seed=3), seed=13)
    for phase in (0, 2):
        z = blk.copy()
        for n in INC5:
            z[n] = np.zeros(z.ntot)
        OracleBlock(prm, z).ddadi_dir(0, phase)
        for n in INC5:
            assert not np.any(z[n]), (phase, n)
        b, c = blk.copy(), blk.copy()
        idx = c.cell(3, 2, 1)
        for n in INC5:
            c[n][idx] *= 2.0
        OracleBlock(prm, b).ddadi_dir(0, phase)
        OracleBlock(prm, c).ddadi_dir(0, phase)
        for n in INC5:
            d = c[n] - b[n]
            assert d[idx] == b[n][idx] and b[n][idx] != 0.0, (phase, n)
            d[idx] = 0.0
            assert not np.any(d), (phase, n)
