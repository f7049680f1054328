"""The INPUTS/BLOCKS namelists of finflo_amd.namelist (NAMELIST /INPUTS/, /BLOCKS/
main.f:5172-5204, defaults main.f:5240-5360):
* a Fortran program compiled with amdflang declares a subset of the reference's INPUTS and
  BLOCKS names, reads the file this module writes (the Fortran runtime's namelist reader)
  and writes its namelists back (the runtime's namelist writer), and this module reads that
  output to the same values;
* the reader against the forms the standard allows (repeat counts, nulls, array elements,
  quotes, logicals, D exponents, comments, $ and &END delimiters);
* the derived parameters against what the mathematics fixes: BCINI1's free stream
  (state.f:2001-2083: p = rho R (T - T0REF), the speed of sound, V = M a, Re = rho V L /
  (mu_Sutherland * RMULTV) whichever two of rho, Re, L are given; main.f does not build
  here) and main.f:3772-3826's turbulence limits (rho k = 1.5 rho (Tu U)^2, mu_t = r mu,
  rho omega = rho (rho k) / mu_t for the SST, the FRSMUD = 1e-3 branch when Tu = 0)."""
import os
import subprocess

import numpy as np
import pytest

from finflo_amd.namelist import bcini1, mgrid_levels, params_from_inputs, read_namelists, write_namelist

FC = "/opt/rocm/bin/amdflang"

INPUTS = {"ITURB": 6, "ISTRES": 0, "LUSGS": 1, "CFL": 5.0, "CFLL": 3.5, "RMACH": 0.2,
          "RE": 5.0e6, "CHLREF": 0.0, "FRSTEM": 288.15, "FRSPRE": 101325.0, "FRSTUR": 0.01,
          "FRSMUT": 1.0, "TURLIM": 5000.0, "CSIMPS": 5.0, "FULLNC": "NO", "TURCOC": "YES",
          "FLUXTY": "ROE", "PR": 0.72, "PRT": 0.9, "RMULTV": 1.0}
BLOCKS = {"MGRID": [3, 3], "INTERI": [-3, 2], "INTERJ": [-3, -3], "INTERK": [-3, 55],
          "IDI1": [1, 0]}


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")
def test_fortran_runtime_reads_and_writes_the_namelists(tmp_path):
    (tmp_path / "INPUT").write_text(write_namelist("INPUTS", INPUTS) + write_namelist("BLOCKS", BLOCKS))
    prog = """      PROGRAM RDINP
      IMPLICIT NONE
      INTEGER :: ITURB, ISTRES, LUSGS
      REAL(8) :: CFL, CFLL, RMACH, RE, CHLREF, FRSTEM, FRSPRE, FRSTUR,
     +  FRSMUT, TURLIM, CSIMPS, PR, PRT, RMULTV
      CHARACTER(LEN=10) :: FULLNC, TURCOC, FLUXTY
      INTEGER :: MGRID(2), INTERI(2), INTERJ(2), INTERK(2), IDI1(2)
      NAMELIST /INPUTS/ ITURB, ISTRES, LUSGS, CFL, CFLL, RMACH, RE,
     +  CHLREF, FRSTEM, FRSPRE, FRSTUR, FRSMUT, TURLIM, CSIMPS, FULLNC,
     +  TURCOC, FLUXTY, PR, PRT, RMULTV
      NAMELIST /BLOCKS/ MGRID, INTERI, INTERJ, INTERK, IDI1
      OPEN(10,FILE='INPUT',STATUS='OLD')
      READ(10,NML=INPUTS)
      READ(10,NML=BLOCKS)
      CLOSE(10)
      OPEN(11,FILE='OUT')
      WRITE(11,NML=INPUTS)
      WRITE(11,NML=BLOCKS)
      CLOSE(11)
      END PROGRAM RDINP
"""
    (tmp_path / "rdinp.f").write_text(prog)
    subprocess.run([FC, "-fdefault-real-8", "-o", str(tmp_path / "rdinp"), str(tmp_path / "rdinp.f")],
                   check=True, capture_output=True)
    subprocess.run([str(tmp_path / "rdinp")], check=True, cwd=tmp_path, capture_output=True)
    back = read_namelists((tmp_path / "OUT").read_text())
    for k, v in INPUTS.items():
        got = back["INPUTS"][k]
        if isinstance(v, str):
            assert got.strip() == v, k
        else:
            assert got == v, k
    for k, v in BLOCKS.items():
        assert back["BLOCKS"][k] == v, k


def test_reader_forms():
    text = """ $INPUTS  ITURB=6 , CFL = 2.5D0, CFLL=, RE=5.E+6 ! a comment
      NAME = 'it''s', FULLNC="YES" , CONVC = .true., STARTC=F
   $END
   &BLOCKS MGRID = 2*3, 1, INTERI(3) = -3, IDI1 = 1, , 4 /
   &OTHER X = 1 /"""
    g = read_namelists(text)
    i = g["INPUTS"]
    assert i["ITURB"] == 6 and i["CFL"] == 2.5 and i["RE"] == 5.0e6
    assert i["NAME"] == "it's" and i["FULLNC"] == "YES"
    assert i["CONVC"] is True and i["STARTC"] is False
    assert "CFLL" not in i                 # a null value: the variable keeps its value
    b = g["BLOCKS"]
    assert b["MGRID"] == [3, 3, 1]
    assert b["INTERI"] == [None, None, -3]
    assert b["IDI1"] == [1, None, 4]
    assert g["OTHER"]["X"] == 1
    assert mgrid_levels(g, 3) == 1 and mgrid_levels(g, 1) == 3


def test_bcini1_closed_forms():
    r, gam, t = 287.05287, 1.4, 288.15
    mu = 1.458e-6 * t ** 1.5 / (t + 110.4)
    # temperature, pressure, Mach and Re given, CHLREF = 0: the length follows
    fs = bcini1({"FRSTEM": t, "FRSPRE": 101325.0, "RMACH": 0.3, "RE": 2e6, "CHLREF": 0.0,
                 "RMULTV": 2.0})
    a = np.sqrt(gam * r * t)
    rho = 101325.0 / (r * t)
    np.testing.assert_allclose(fs["FRSSSP"], a, rtol=1e-15)
    np.testing.assert_allclose(fs["FRSVEL"], 0.3 * a, rtol=1e-15)
    np.testing.assert_allclose(fs["FRSDEN"], rho, rtol=1e-15)
    np.testing.assert_allclose(fs["FRSVIS"], 2.0 * mu, rtol=1e-15)
    np.testing.assert_allclose(fs["CHLREF"], 2e6 * 2.0 * mu / (rho * 0.3 * a), rtol=1e-14)
    np.testing.assert_allclose(fs["T0"], t * (1 + 0.2 * 0.09), rtol=1e-15)
    # density and length given: Re follows; velocity given: Mach follows
    fs = bcini1({"FRSTEM": t, "FRSDEN": 1.2, "FRSVEL": 50.0, "CHLREF": 2.0})
    np.testing.assert_allclose(fs["RE"], 1.2 * 50.0 * 2.0 / mu, rtol=1e-14)
    np.testing.assert_allclose(fs["RMACH"], 50.0 / a, rtol=1e-15)
    np.testing.assert_allclose(fs["FRSPRE"], 1.2 * r * t, rtol=1e-15)
    # Re and length given: the density follows
    fs = bcini1({"FRSTEM": t, "RE": 1e6, "CHLREF": 1.0, "RMACH": 0.5})
    np.testing.assert_allclose(fs["FRSDEN"], 1e6 * mu / (0.5 * a), rtol=1e-14)
    with pytest.raises(ValueError):
        bcini1({"FRSTEM": t, "FRSPRE": 1e5, "RMACH": 0.2, "RE": 1e6, "CHLREF": 1.0})
    with pytest.raises(ValueError):
        bcini1({"FRSTEM": t, "FRSPRE": 1e5, "RMACH": 0.2, "FRSVEL": 30.0, "CHLREF": 1.0})


def test_params_from_inputs():
    g = read_namelists(write_namelist("INPUTS", INPUTS) + write_namelist("BLOCKS", BLOCKS))
    p = params_from_inputs(g, block=2)
    fs = bcini1(INPUTS)
    assert p.iturb == 6 and p.ddadi == 0 and p.fullns == 0 and p.turcor == 1 and p.iflux == 1
    assert p.cfl == 5.0 and p.cfll == 3.5 and p.cdiff == 0.5
    assert list(p.inter) == [2, -3, 55]
    assert p.frsvel == fs["FRSVEL"] and p.frsden == fs["FRSDEN"] and p.frsvis == fs["FRSVIS"]
    assert p.tlolim == 0.05 * 288.15 and p.tuplim == 6 * 288.15
    rk = 1.5 * p.frsden * (0.01 * p.frsvel) ** 2
    np.testing.assert_allclose(p.rklim, rk, rtol=1e-15)
    np.testing.assert_allclose(p.frsmut, 1.0 * p.frsvis, rtol=1e-15)
    np.testing.assert_allclose(p.epslim, p.frsden * rk / p.frsmut, rtol=1e-15)
    # Tu = 0: mu_t = 1e-3 mu, rho k = mu_t U / L
    q = params_from_inputs({"INPUTS": dict(INPUTS, FRSTUR=0.0, LUSGS=0)})
    np.testing.assert_allclose(q.frsmut, 1e-3 * q.frsvis, rtol=1e-15)
    np.testing.assert_allclose(q.rklim, q.frsmut * q.frsvel / fs["CHLREF"], rtol=1e-14)
    assert q.ddadi == 1
