"""Closed-form and invariant checks of the oracle, independent of the reference build.

Each check is chosen so that a plausible slip in the restatement (a dropped term, a wrong
sign or index, a transposed operand) fails it:
  * MUSCL reproduces linear data exactly and is first order when asked;
  * SURFVECTS returns an orthonormal right-handed basis;
  * the Roe + viscous face flux of a uniform state is the physical Euler flux;
  * the residual of a uniform state on a curvilinear grid vanishes (freestream
    preservation: face vectors close every cell);
  * the residual telescopes: sum of VOL*D over the block = minus the net boundary flux;
  * COPR inverts the linearisation of (rho, rho u, E) in (h, u, v, w, p);
  * TIME3 on a Cartesian block equals its closed form;
  * the LU-SGS sweeps are linear in the right-hand side, reduce to diagonal scaling on a
    1x1x1 block and reproduce the residual-free identity (zero in, zero out).
"""
import ctypes as C

import numpy as np
import pytest

from finflo_amd.cases import make_block, fill_state
from finflo_amd.layout import default_params
from oracle.oracle_py import OracleBlock, oracle

D5 = ["DRO", "DM", "DN", "DW", "DE"]


def lib():
    L = oracle()
    if not hasattr(L, "_math_bound"):
        L.ffo_face_flux.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_int, C.c_void_p]
        L.ffo_xxtrap.argtypes = [C.c_void_p, C.c_long, C.c_int, C.c_void_p, C.c_void_p]
        L.ffo_surfvects.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ffo_time3.argtypes = [C.c_void_p, C.c_void_p, C.c_double]
        L.ffo_copr.argtypes = [C.c_void_p, C.c_void_p]
        L.ffo_lusgs.argtypes = [C.c_void_p, C.c_void_p]
        for f in ("ffo_face_flux", "ffo_xxtrap", "ffo_surfvects", "ffo_time3", "ffo_copr", "ffo_lusgs"):
            getattr(L, f).restype = None
        L._math_bound = True
    return L


def face_flux(prm, blk, L, d):
    out = np.zeros(5)
    v = blk.view_struct()
    lib().ffo_face_flux(C.byref(prm), C.byref(v), L, d, out.ctypes.data)
    return out


@pytest.mark.parametrize("intem", [-3, -1, -2, -4, -5, 0])
def test_muscl_unlimited_reproduces_linear_data(intem):
    q = 3.0 + 0.7 * np.arange(10.0)
    qr, ql = C.c_double(), C.c_double()
    lib().ffo_xxtrap(q.ctypes.data, 5, intem, C.byref(qr), C.byref(ql))
    face = 0.5 * (q[4] + q[5])
    assert abs(qr.value - face) < 1e-14 and abs(ql.value - face) < 1e-14


@pytest.mark.parametrize("intem", [1, 2, 3, 4])
def test_muscl_limited_linear_and_extremum(intem):
    q = 1.0 + 0.25 * np.arange(10.0)
    qr, ql = C.c_double(), C.c_double()
    lib().ffo_xxtrap(q.ctypes.data, 5, intem, C.byref(qr), C.byref(ql))
    face = 0.5 * (q[4] + q[5])
    assert abs(qr.value - face) < 1e-9 and abs(ql.value - face) < 1e-9  # limiter ~ 1 (eps)
    q2 = np.array([0., 0., 0., 1., 2., 1., 0., 0.])       # extremum at index 4
    lib().ffo_xxtrap(q2.ctypes.data, 4, intem, C.byref(qr), C.byref(ql))
    # left state at the face 3|4 is limited towards first order at the extremum side
    assert min(q2[3], q2[4]) - 1e-12 <= ql.value <= max(q2[3], q2[4]) + 1e-12


def test_muscl_first_order():
    q = np.random.default_rng(0).uniform(size=8)
    qr, ql = C.c_double(), C.c_double()
    lib().ffo_xxtrap(q.ctypes.data, 4, 55, C.byref(qr), C.byref(ql))
    assert qr.value == q[4] and ql.value == q[3]


def test_surfvects_orthonormal_right_handed():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        a2, a3 = np.zeros(3), np.zeros(3)
        lib().ffo_surfvects(n.ctypes.data, a2.ctypes.data, a3.ctypes.data)
        M = np.stack([n, a2, a3])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-14)
        assert np.isclose(np.linalg.det(M), 1.0, atol=1e-14)


def uniform_block(dims=(8, 6, 5), distort=0.2, prm=None, vel=(1.0, 0.3, -0.2)):
    prm = prm or default_params()
    blk = make_block(*dims, prm=prm, distort=distort, seed=4, amp=0.0, kind="density")
    d = np.array(vel) / np.linalg.norm(vel)
    blk["U"], blk["V"], blk["W"] = prm.frsvel * d[0], prm.frsvel * d[1], prm.frsvel * d[2]
    o = OracleBlock(prm, blk)
    o.state()
    o.turb()
    return prm, blk, o


# (IFLUX, free-stream speed, direction): Roe and van Leer, subsonic faces, supersonic faces
# in both directions (van Leer's split parts then vanish on one side)
UNIFORM_FLUX_CASES = [
    (1, None, (1.0, 0.3, -0.2)), (0, None, (1.0, 0.3, -0.2)), (1, 520.0, (1.0, 0.3, -0.2)),
    (0, 520.0, (1.0, 0.3, -0.2)), (0, 520.0, (-1.0, -0.2, 0.1)), (0, 250.0, (0.2, -1.0, 0.7)),
]


@pytest.mark.parametrize("iflux,speed,vel", UNIFORM_FLUX_CASES)
def test_uniform_face_flux_is_euler_flux(iflux, speed, vel):
    """Consistency: F(U, U) is the physical flux (Roe); F+(U) + F-(U) = F(U) (van Leer)."""
    prm = default_params(iflux=iflux)
    if speed:
        prm.frsvel = speed
    prm, blk, o = uniform_block(prm=prm, vel=vel)
    L = blk.cell(3, 2, 2)
    ro, u, v, w = blk["RO"][L], blk["U"][L], blk["V"][L], blk["W"][L]
    p, E = blk["P"][L], blk["E"][L]
    for d, an in enumerate(("A1", "A2", "A3")):
        A = blk[an][L]
        n = np.array([blk[an + "X"][L], blk[an + "Y"][L], blk[an + "Z"][L]])
        un = u * n[0] + v * n[1] + w * n[2]
        exact = A * np.array([ro * un, ro * u * un + p * n[0], ro * v * un + p * n[1],
                              ro * w * un + p * n[2], (E + p) * un])
        f = face_flux(prm, blk, L, d)
        scale = A * (abs(ro * un) * np.array([1, abs(u), abs(v), abs(w), E / ro]) + p)
        assert np.all(np.abs(f - exact) <= 1e-12 * scale + 1e-300), (d, f, exact)


@pytest.mark.parametrize("iflux", [1, 0])
def test_freestream_preservation_curvilinear(iflux):
    prm, blk, o = uniform_block(distort=0.25, prm=default_params(iflux=iflux))
    o.residual()
    scale = blk["RO"][0] * prm.frsvel * np.mean(blk.interior("A1")) / np.mean(blk.interior("VOL"))
    for n, s in zip(D5, (1, prm.frsvel, prm.frsvel, prm.frsvel, blk["E"][0] / blk["RO"][0])):
        assert np.abs(blk.interior(n)).max() <= 1e-10 * scale * s, n


def test_residual_telescopes_to_boundary_flux():
    prm = default_params()
    blk = make_block(6, 5, 4, prm, distort=0.15, seed=8, amp=0.05, kind="all")
    o = OracleBlock(prm, blk)
    o.state()
    o.turb()
    o.residual()
    total = np.array([np.sum(blk.interior("VOL") * blk.interior(n)) for n in D5])
    net = np.zeros(5)
    for j in range(1, blk.jmax + 1):
        for k in range(1, blk.kmax + 1):
            net += face_flux(prm, blk, blk.cell(blk.imax + 1, j, k), 0) - face_flux(prm, blk, blk.cell(1, j, k), 0)
    for i in range(1, blk.imax + 1):
        for k in range(1, blk.kmax + 1):
            net += face_flux(prm, blk, blk.cell(i, blk.jmax + 1, k), 1) - face_flux(prm, blk, blk.cell(i, 1, k), 1)
    for i in range(1, blk.imax + 1):
        for j in range(1, blk.jmax + 1):
            net += face_flux(prm, blk, blk.cell(i, j, blk.kmax + 1), 2) - face_flux(prm, blk, blk.cell(i, j, 1), 2)
    scale = np.array([np.sum(np.abs(blk.interior("VOL") * blk.interior(n))) for n in D5])
    assert np.all(np.abs(total + net) <= 1e-11 * scale), (total, -net)


def test_copr_inverts_the_linearisation():
    prm = default_params()
    blk = make_block(5, 4, 3, prm, seed=2, amp=0.05, kind="all")
    o = OracleBlock(prm, blk)
    o.state()
    rng = np.random.default_rng(3)
    sh = blk.interior("RO").shape
    dh, du, dv, dw, dp = (rng.normal(size=sh) * s for s in (1e3, 1.0, 1.0, 1.0, 1e2))
    ro, u, v, w = (blk.interior(n) for n in ("RO", "U", "V", "W"))
    drdp, drdh, P, E = (blk.interior(n) for n in ("DRDP", "DRDH", "P", "E"))
    h = (E + P) / ro - 0.5 * (u * u + v * v + w * w)      # static enthalpy
    q2 = u * u + v * v + w * w
    drho = drdp * dp + drdh * dh
    dm, dn, dww = ro * du + u * drho, ro * dv + v * drho, ro * dw + w * drho
    dE = ro * dh + h * drho - dp + 0.5 * q2 * drho + ro * (u * du + v * dv + w * dw)
    for n, x in zip(D5, (drho, dm, dn, dww, dE)):
        blk.interior(n)[...] = x
    lib().ffo_copr(C.byref(prm), C.byref(blk.view_struct()))
    for n, x in zip(D5, (dh, du, dv, dw, dp)):
        assert np.allclose(blk.interior(n), x, rtol=1e-9, atol=1e-9 * np.abs(x).max()), n


def test_time3_cartesian_closed_form():
    prm = default_params(rmult=50.0)
    blk = make_block(6, 5, 4, prm, seed=5, amp=0.02, kind="all")
    o = OracleBlock(prm, blk)
    o.state()
    o.turb()
    cfl = 3.0
    lib().ffo_time3(C.byref(prm), C.byref(blk.view_struct()), cfl)
    V = blk.interior("VOL")
    A1, A2, A3 = blk.interior("A1"), blk.interior("A2"), blk.interior("A3")   # uniform grid
    u, v, w, c, ro = (blk.interior(n) for n in ("U", "V", "W", "C", "RO"))
    b1 = (np.abs(u) * A1 + np.abs(v) * A2 + np.abs(w) * A3 + c * np.sqrt(A1**2 + A2**2 + A3**2)) / V / cfl
    z = np.maximum(np.maximum(A1, A2), A3)
    b2 = 2 * 1.8 * z**2 / V**2 / (ro + 1e-10) * blk.interior("VIS") / prm.cdiff
    P = blk.interior("P")
    den = np.maximum(1 + np.maximum(0, (prm.frspre - 2 * P) / P),
                     1 + np.maximum(0, (ro - 2 * prm.frsden) / (ro + 1e-10)) * 5)
    assert np.allclose(blk.interior("DTL"), 1 / (np.maximum(b1, b2) * den), rtol=1e-13)


def lusgs_block(dims=(6, 5, 4), seed=6):
    prm = default_params()
    blk = make_block(*dims, prm=prm, distort=0.1, seed=seed, amp=0.05, kind="all")
    o = OracleBlock(prm, blk)
    o.state()
    o.turb()
    lib().ffo_time3(C.byref(prm), C.byref(blk.view_struct()), prm.cfl)
    return prm, blk


def run_lusgs(prm, blk, rhs):
    b = blk.copy()
    for n, x in zip(D5, rhs):
        b.interior(n)[...] = x
    lib().ffo_lusgs(C.byref(prm), C.byref(b.view_struct()))
    return [b.interior(n).copy() for n in D5]


def test_lusgs_is_linear_in_the_rhs():
    prm, blk = lusgs_block()
    rng = np.random.default_rng(7)
    sh = blk.interior("RO").shape
    r1 = [rng.normal(size=sh) for _ in D5]
    r2 = [rng.normal(size=sh) for _ in D5]
    a, b = 1.7, -0.6
    x1, x2 = run_lusgs(prm, blk, r1), run_lusgs(prm, blk, r2)
    x12 = run_lusgs(prm, blk, [a * p + b * q for p, q in zip(r1, r2)])
    for p, q, s in zip(x1, x2, x12):
        assert np.allclose(s, a * p + b * q, rtol=1e-10, atol=1e-12 * np.abs(s).max())
    z = run_lusgs(prm, blk, [np.zeros(sh) for _ in D5])
    assert all(np.all(x == 0) for x in z)


def test_lusgs_single_cell_is_diagonal_scaling():
    prm, blk = lusgs_block(dims=(1, 1, 1), seed=9)
    rhs = [np.full((1, 1, 1), v) for v in (0.3, -1.2, 0.7, 2.0, 5.0)]
    out = run_lusgs(prm, blk, rhs)
    L = blk.cell(1, 1, 1)
    g = {n: blk[n] for n in ("A1", "A2", "A3", "A1X", "A1Y", "A1Z", "A2X", "A2Y", "A2Z",
                             "A3X", "A3Y", "A3Z", "VOL", "U", "V", "W", "C", "DTL", "RO",
                             "VIS", "VIST")}
    IS, IL = blk.istrid, blk.istrid * blk.jstrid
    vel = np.array([g["U"][L], g["V"][L], g["W"][L]])
    c, vol, dt1 = g["C"][L], g["VOL"][L], prm.theta * g["DTL"][L]

    def lam(ax, an, idx):
        return abs(np.dot([g[ax + "X"][idx], g[ax + "Y"][idx], g[ax + "Z"][idx]], vel)) + c

    def visc(ax, idx_a, nb, ff):
        rnum = (vol + g["VOL"][nb]) * (g["RO"][L] + g["RO"][nb])
        return ff * 2.0 * g[ax][idx_a] / rnum * (g["VIST"][L] + g["VIS"][L] + g["VIST"][nb] + g["VIS"][nb])

    # PREDGS: lower faces, viscous factor 4 at I=J=K=1
    xa = [lam(a, a, L) + visc(a, L, L - s, 4.0) for a, s in (("A1", 1), ("A2", IS), ("A3", IL))]
    rm = [lam(a, a, L) for a in ("A1", "A2", "A3")]
    sv1 = vol / (vol + dt1 * (xa[0] * g["A1"][L] + xa[1] * g["A2"][L] + xa[2] * g["A3"][L]))
    sp1 = vol / (vol + dt1 * (rm[0] * g["A1"][L] + rm[1] * g["A2"][L] + rm[2] * g["A3"][L]))
    # DIAGGS over all six faces
    sc = 0.5 * sum(g[a][L] * lam(a, a, L) + g[a][L + s] * lam(a, a, L + s)
                   for a, s in (("A1", 1), ("A2", IS), ("A3", IL)))
    fac = 1 + prm.theta * g["DTL"][L] / vol * sc
    # CORRGS: upper faces, viscous factor 2, viscous area of the lower face
    xa2 = [lam(a, a, L + s) + visc(a, L, L + s, 2.0) for a, s in (("A1", 1), ("A2", IS), ("A3", IL))]
    rm2 = [lam(a, a, L + s) for a, s in (("A1", 1), ("A2", IS), ("A3", IL))]
    sv2 = vol / (vol + dt1 * (xa2[0] * g["A1"][L + 1] + xa2[1] * g["A2"][L + IS] + xa2[2] * g["A3"][L + IL]))
    sp2 = vol / (vol + dt1 * (rm2[0] * g["A1"][L] + rm2[1] * g["A2"][L] + rm2[2] * g["A3"][L]))
    for q in range(5):
        expect = rhs[q].item() * fac * ((sv1 * sv2) if q < 4 else (sp1 * sp2))
        assert np.isclose(out[q].item(), expect, rtol=1e-12), (q, out[q].item(), expect)


def test_oracle_walls_do_not_outlive_their_block():
    """Wall patches set on a level belong to that block: a block made afterwards on the same
    level runs without them (the earlier patch arrays may be gone)."""
    import gc
    from oracle.oracle_py import OracleBlock
    from tests.parity import WALL_CASES, make_wall_case, RES
    name, dims, pover, ckw, specs, grav = WALL_CASES[0]
    prm, blk, wp, patches = make_wall_case(dims, pover, ckw, specs, grav)
    ow = OracleBlock(prm, blk.copy())
    ow.set_walls(wp, [p.copy() for p in patches])
    del ow, patches
    gc.collect()
    a, b = blk.copy(), blk.copy()
    oa = OracleBlock(prm, a)
    oa.save_old(); oa.state(); oa.turb(); oa.residual()
    from oracle.oracle_py import oracle
    import ctypes as C
    oracle().ffo_wall_set(1, None, None, 0)
    ob = OracleBlock(prm, b)
    ob.save_old(); ob.state(); ob.turb(); ob.residual()
    for n in RES:
        assert np.array_equal(a[n], b[n])
