"""Shared helpers of the parity tests: seeded cases and field comparisons."""
import numpy as np

from finflo_amd.cases import make_block
from finflo_amd.layout import WALL_ARRAYS, WALL_OUT, WallParams, WallPatch, default_params

# (name, dims, params overrides, case kwargs)
CASES = [
    ("cart_density", (8, 6, 5), {}, dict(kind="density", amp=0.01)),
    ("curv_all", (10, 7, 6), {}, dict(kind="all", amp=0.05, distort=0.15, seed=3)),
    ("curv_viscous", (7, 9, 6), dict(rmult=3000.0), dict(kind="all", amp=0.03, distort=0.1, seed=5)),
    ("limited_inter3", (9, 6, 7), dict(inter=(3, 2, 1)), dict(kind="all", amp=0.08, distort=0.1, seed=7)),
    ("first_order", (6, 8, 5), dict(inter=(55, 55, 55), cfl=2.0), dict(kind="all", amp=0.05, seed=9)),
    ("stretched_j", (12, 10, 4), dict(cfl=10.0), dict(kind="velocity", amp=0.02, jratio=1.3, seed=11)),
    # van Leer flux-vector splitting (IFLUX = 0): subsonic faces, faces on both sides of
    # Mach 1 (oblique supersonic stream with large perturbations) and reversed flow
    ("vanleer", (9, 7, 6), dict(iflux=0), dict(kind="all", amp=0.05, distort=0.12, seed=13)),
    ("vanleer_transonic", (8, 6, 7), dict(iflux=0, frsvel=380.0, cfl=2.0),
     dict(kind="all", amp=0.12, distort=0.1, seed=17, mach_dir=(1.0, 0.6, -0.3))),
    ("vanleer_reverse", (7, 8, 6), dict(iflux=0, frsvel=420.0, inter=(2, -3, 55), cfl=1.5),
     dict(kind="all", amp=0.1, distort=0.1, seed=19, mach_dir=(-1.0, -0.4, 0.5))),
]


# k-omega SST (ITURB = 6): curvilinear with F1 across (0,1), a stretched wall case whose
# large RE puts cells under FUN1KO's wall-distance limit, limited/first-order MUSCL mixes,
# and the full production form (KATOL = NO)
SST_CASES = [
    ("sst_curv", (10, 7, 6), dict(iturb=6),
     dict(kind="all", amp=0.05, distort=0.15, seed=23, turb=(1.0, 50.0))),
    ("sst_wall", (9, 8, 5), dict(iturb=6, re=1e14, cfl=2.0),
     dict(kind="velocity", amp=0.03, jratio=1.3, seed=29, turb=(0.5, 400.0))),
    ("sst_limited", (8, 7, 6), dict(iturb=6, inter=(3, 2, 55)),
     dict(kind="all", amp=0.04, distort=0.1, seed=31, turb=(2.0, 1000.0))),
    ("sst_fullprod", (7, 6, 8), dict(iturb=6, katol=0, rklim=0.05, epslim=30.0),
     dict(kind="all", amp=0.08, distort=0.12, seed=37, turb=(0.3, 80.0))),
]

# Wallin-Johansson EARSM on top of the SST (ISTRES = 1): EARSM1's cubic takes both of its
# branches on these cases (P2 >= 0 where rotation dominates, P2 < 0 under strain), and the
# ghost cells carry seeded b_ij so that ANISFLUX's boundary faces see non-zero stresses
EARSM_CASES = [
    ("earsm_curv", (10, 7, 6), dict(iturb=6, istres=1),
     dict(kind="all", amp=0.05, distort=0.15, seed=61, turb=(1.0, 50.0))),
    ("earsm_wall", (9, 8, 5), dict(iturb=6, istres=1, re=1e14, cfl=2.0),
     dict(kind="velocity", amp=0.03, jratio=1.3, seed=67, turb=(0.5, 400.0))),
    ("earsm_fullprod", (7, 6, 8), dict(iturb=6, istres=1, inter=(3, 2, 55)),
     dict(kind="all", amp=0.08, distort=0.12, seed=71, turb=(0.3, 80.0))),
]
BIJ = ["B11", "B12", "B13", "B22", "B23", "B33"]


def seed_bij(blk, seed=0, amp=0.05):
    """Seeded b_ij on every cell (trace-free: B33 = -B11 - B22)."""
    rng = np.random.default_rng(seed)
    for name in BIJ[:5]:
        blk[name] = amp * rng.uniform(-1, 1, blk.ntot)
    blk["B33"] = -blk["B11"] - blk["B22"]
    return blk


def make_earsm_case(dims, pover, ckw):
    prm, blk = make_case(dims, pover, ckw)
    seed_bij(blk, seed=ckw.get("seed", 0) + 500)
    return prm, blk


# Full viscous fluxes (FULLNS): FLUXBL with the TEMDER gradients, laminar and SST (FLKEBL),
# with and without the lambda term (IDIFF = 2 drops it), with the van Leer flux, and under
# the EARSM; curvilinear cells so that the cell-centre connection is not along the face
# normal and the Blazek correction contributes
FULLNS_CASES = [
    ("ns_curv", (10, 7, 6), dict(fullns=1, rmult=3000.0),
     dict(kind="all", amp=0.05, distort=0.15, seed=81)),
    ("ns_idiff2", (7, 9, 6), dict(fullns=1, idiff=2, rmult=3000.0),
     dict(kind="all", amp=0.03, distort=0.1, seed=83)),
    ("ns_vanleer", (9, 7, 6), dict(fullns=1, iflux=0, rmult=1000.0),
     dict(kind="all", amp=0.05, distort=0.12, seed=85)),
    ("ns_sst", (10, 7, 6), dict(iturb=6, fullns=1),
     dict(kind="all", amp=0.05, distort=0.15, seed=87, turb=(1.0, 50.0))),
    ("ns_sst_wall", (9, 8, 5), dict(iturb=6, fullns=1, re=1e14, cfl=2.0),
     dict(kind="velocity", amp=0.03, jratio=1.3, seed=89, turb=(0.5, 400.0))),
    ("ns_earsm", (8, 7, 6), dict(iturb=6, istres=1, fullns=1, inter=(3, 2, 55)),
     dict(kind="all", amp=0.06, distort=0.12, seed=91, turb=(0.3, 80.0))),
    # ghost-cell centres moved past the first interior cell on the i- and k-low sides: the
    # boundary faces there see ALFA < 0 (no correction in FLUXBL, ALFA = 0 in FLKEBL)
    ("ns_sst_backghost", (8, 6, 6), dict(iturb=6, fullns=1),
     dict(kind="all", amp=0.05, distort=0.1, seed=93, turb=(1.0, 50.0), back_ghost=True)),
    # one K-plane: TEMDER leaves out the K ghost layer when KMAX = 1, DERFUN does not
    ("ns_kplane", (10, 8, 1), dict(fullns=1, rmult=2000.0),
     dict(kind="all", amp=0.05, distort=0.12, seed=95)),
]
GRADT = ["DTDX", "DTDY", "DTDZ"]
GRADKW = ["DKDX", "DKDY", "DKDZ", "DEDX", "DEDY", "DEDZ"]


def make_fullns_case(dims, pover, ckw):
    ckw = dict(ckw)
    back = ckw.pop("back_ghost", False)
    prm, blk = make_case(dims, pover, ckw)
    if prm.istres:
        seed_bij(blk, seed=ckw.get("seed", 0) + 500)
    if back:
        x = {c: blk.view3(c) for c in ("XC", "YC", "ZC")}   # (k, j, i), ghosts included
        for c in x:
            x[c][:, :, 1] = x[c][:, :, 2] + 0.6 * (x[c][:, :, 3] - x[c][:, :, 2])
            x[c][1, :, :] = x[c][2, :, :] + 0.6 * (x[c][3, :, :] - x[c][2, :, :])
    return prm, blk


# fields each stage writes, for the stage-by-stage comparisons
STATE = ["P", "RO", "E", "C", "VIS", "CH", "CP", "DRDP", "DRDH", "RM", "RN", "RW"]
RES = ["DRO", "DM", "DN", "DW", "DE"]
PRIM = ["TEMP", "U", "V", "W", "PDIFF", "P"]
SST_TURB = ["OHMI", "W12", "W13", "W23", "S11", "S12", "S13", "S22", "S23", "S33", "STRAIN",
            "EPS2", "VIST"]
SST_RES = RES + ["DRK", "DEPS", "FUN1", "PTUR", "SRK", "SEPS"]
SST_PRIM = PRIM + ["RK", "REPS"]


def make_case(dims, pover, ckw):
    prm = default_params(**pover)
    blk = make_block(*dims, prm=prm, **ckw)
    return prm, blk


def assert_fields_equal(a, b, names, interior=True, rtol=0.0, what=""):
    for n in names:
        x = a.interior(n) if interior else a[n]
        y = b.interior(n) if interior else b[n]
        if rtol == 0.0:
            bad = np.count_nonzero(x != y)
            assert bad == 0, f"{what} {n}: {bad} values differ, max |diff| = {np.abs(x - y).max():.3e}"
        else:
            scale = max(np.abs(y).max(), 1e-300)
            err = np.abs(x - y).max() / scale
            assert err <= rtol, f"{what} {n}: rel err {err:.3e} > {rtol}"


# Solid-wall patches (BOTSKE): (iface, kx1, kx2, ky1, ky2, iheat) as ICON(3..7,L), ICON(20,L).
# Faces low and high in every direction, partial and whole faces, the four heat options,
# laminar and SST (with and without TURCOR), with FULLNS and with gravity in the forces
WALL_CASES = [
    ("wall_lam_jlow", (10, 8, 6), dict(rmult=500.0),
     dict(kind="all", amp=0.05, distort=0.12, seed=101), [(2, 2, 9, 1, 6, 0)], False),
    ("wall_lam_multi", (9, 8, 7), dict(rmult=500.0),
     dict(kind="all", amp=0.05, distort=0.1, seed=103),
     [(4, 1, 8, 2, 7, 1), (3, 1, 9, 1, 8, 2), (5, 2, 8, 1, 7, 3), (1, 3, 6, 1, 4, 0)], True),
    ("wall_sst_jlow", (10, 8, 6), dict(iturb=6),
     dict(kind="all", amp=0.05, distort=0.12, seed=105, turb=(1.0, 50.0)),
     [(2, 1, 10, 1, 6, 0), (6, 2, 9, 1, 8, 1)], False),
    ("wall_sst_noturcor", (9, 8, 7), dict(iturb=6, turcor=0),
     dict(kind="all", amp=0.05, distort=0.1, seed=107, turb=(0.5, 200.0)),
     [(2, 1, 9, 1, 7, 3), (4, 1, 8, 1, 7, 0)], True),
    ("wall_sst_fullns", (8, 9, 6), dict(iturb=6, fullns=1, re=1e14),
     dict(kind="velocity", amp=0.03, jratio=1.3, seed=109, turb=(0.5, 400.0)),
     [(2, 1, 8, 1, 6, 2), (1, 1, 9, 1, 6, 1), (5, 1, 8, 1, 6, 0)], False),
]


def make_walls(blk, prm, specs, seed, gravity=False):
    """Patches of the given ICON entries with seeded wall data (moving walls, wall
    temperatures and heat fluxes, roughness, face centres) and the BOTSKE constants."""
    rng = np.random.default_rng(seed)
    wp = WallParams()
    wp.t0 = 1.02 * prm.frstem
    wp.difpre = 0.01 * prm.frspre
    wp.xmom, wp.ymom, wp.zmom = 0.25, 0.1, -0.05
    if gravity:
        wp.gx, wp.gy, wp.gz, wp.ground, wp.altitude = 0.3, -9.81, 0.1, -0.2, -10000.0
    else:
        wp.ground, wp.altitude = 0.0, -10000.0
    patches = []
    for iface, kx1, kx2, ky1, ky2, iheat in specs:
        p = WallPatch(iface, kx1, kx2, ky1, ky2, iheat)
        n = p.size
        p.a["uwall"][:] = rng.uniform(-1, 1, n)
        p.a["vwall"][:] = rng.uniform(-1, 1, n)
        p.a["wwall"][:] = rng.uniform(-1, 1, n)
        p.a["twall"][:] = prm.frstem * rng.uniform(0.95, 1.1, n)
        p.a["hflux"][:] = rng.uniform(-50.0, 50.0, n)
        p.a["rbk"][:] = rng.uniform(0.0, 80.0, n)
        for c in ("xcp", "ycp", "zcp"):
            p.a[c][:] = rng.uniform(-1, 1, n)
        patches.append(p)
    return wp, patches


def make_wall_case(dims, pover, ckw, specs, gravity):
    prm, blk = make_case(dims, pover, ckw)
    wp, patches = make_walls(blk, prm, specs, ckw.get("seed", 0) + 900, gravity)
    return prm, blk, wp, patches


def assert_patches_equal(pa, pb, rtol=0.0, what=""):
    for l, (x, y) in enumerate(zip(pa, pb)):
        for n in WALL_OUT + ["twall"]:
            u, v = x.a[n], y.a[n]
            if rtol == 0.0:
                bad = np.count_nonzero(u != v)
                assert bad == 0, f"{what} patch {l} {n}: {bad} values differ, max {np.abs(u - v).max():.3e}"
            else:
                scale = max(np.abs(v).max(), 1e-300)
                err = np.abs(u - v).max() / scale
                assert err <= rtol, f"{what} patch {l} {n}: rel err {err:.3e} > {rtol}"
