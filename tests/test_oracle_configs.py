"""The plain method's CPU-runnable configuration at its own workload: the 48x24x24 laminar
block with its no-slip J-low wall, M = 0.3, Re = 1e5, CFL 5, seeded +/-1 % density
perturbation, 200 LU-SGS iterations with Roe and with van Leer fluxes -- the C restatement
against the reference's own Fortran (oracle/_ref), bit for bit in every field, the RESL2
sums of every iteration, the wall patch arrays and the force sums."""
import numpy as np
import pytest

from finflo_amd.cases import config1
from oracle.oracle_py import OracleBlock, RefBlock, have_ref
from tests.parity import assert_patches_equal

pytestmark = pytest.mark.skipif(not have_ref(), reason="reference library not built")


@pytest.mark.parametrize("iflux", [1, 0], ids=["roe", "vanleer"])
def test_config1_200_iterations_bit_exact(iflux):
    prm, a, (wp, pa) = config1(iflux=iflux, seed=0)
    b = a.copy()
    pb = [p.copy() for p in pa]
    oa, rb = OracleBlock(prm, a), RefBlock(prm, b)
    oa.set_walls(wp, pa)
    rb.set_walls(wp, pb)
    for it in range(200):
        r1 = oa.cycle()
        r2 = rb.cycle()
        assert np.array_equal(r1, r2), (it, r1, r2)
    assert np.all(np.isfinite(a.F))
    np.testing.assert_array_equal(a.F, b.F)
    assert_patches_equal(pa, pb, what="config 1 wall")
    np.testing.assert_array_equal(oa.forces(0), rb.forces(0))
    # the iterations did work: the density residual fell from its first value
    assert r1[0] > 0
