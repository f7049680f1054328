"""Host-side mirror of include/ffhip.h: field slots, parameter and block structs.

A block is held as one packed FP64 array ``F[nfield, ntot]``; row ``f`` is the reference's
flat per-field array (MAIN_ARRAYS, charac.f:391-574) for that (level, block) slab:
i-fastest with two ghost layers, cell (I,J,K) at offset
``(K+1)*ISTRID*JSTRID + (J+1)*ISTRID + (I+1)`` (res3c.f:127-136).
"""
import ctypes as C

import numpy as np

GHOST = 2

FIELDS = [
    "PDIFF", "TEMP", "U", "V", "W", "RK", "REPS",
    "P", "RO", "E", "C", "VIS", "CH", "CP", "DRDP", "DRDH",
    "RM", "RN", "RW", "EPS2", "VIST",
    "DRO", "DM", "DN", "DW", "DE", "DRK", "DEPS", "DTL",
    "TOLD", "UOLD", "VOLD", "WOLD", "POLD", "RKOLD", "EPSOLD",
    "VOL", "A1", "A1X", "A1Y", "A1Z", "A2", "A2X", "A2Y", "A2Z",
    "A3", "A3X", "A3Y", "A3Z", "D1", "D2", "D3", "XC", "YC", "ZC", "DISTW",
    "SRK", "SEPS", "PTUR", "FUN1", "OHMI",
    "W12", "W13", "W23", "S11", "S12", "S13", "S22", "S23", "S33", "STRAIN", "TTS",
    "B11", "B12", "B13", "B22", "B23", "B33",
    "ROP2H", "RMP2H", "RNP2H", "RWP2H", "EP2H", "RKP2H", "EPSP2H",
    "DTDX", "DTDY", "DTDZ", "DKDX", "DKDY", "DKDZ", "DEDX", "DEDY", "DEDZ",
]
NF = len(FIELDS)
FID = {name: i for i, name in enumerate(FIELDS)}

# groups used by sync field lists and tests
SOLVED = ["PDIFF", "TEMP", "U", "V", "W", "RK", "REPS"]
DERIVED = ["P", "RO", "E", "C", "VIS", "CH", "CP", "DRDP", "DRDH", "RM", "RN", "RW",
           "EPS2", "VIST"]
RESID = ["DRO", "DM", "DN", "DW", "DE", "DRK", "DEPS", "DTL"]
OLD = ["TOLD", "UOLD", "VOLD", "WOLD", "POLD", "RKOLD", "EPSOLD"]
GEOM = ["VOL", "A1", "A1X", "A1Y", "A1Z", "A2", "A2X", "A2Y", "A2Z", "A3", "A3X", "A3Y",
        "A3Z", "D1", "D2", "D3", "XC", "YC", "ZC", "DISTW"]



class Params(C.Structure):
    """ff_params of include/ffhip.h."""
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "rgas", "pr", "prt", "visu0", "expsu", "tsu0",
        "frspre", "frsden", "frstem", "frsvel", "frsvis",
        "cfl", "cfll", "theta", "cdiff", "tuplim", "tlolim", "rklim", "epslim",
        "turlim", "frsmut", "rmult", "re", "csimps", "cmgk", "cmgeps")] + [
        ("reserved", C.c_double * 3),
        ("iturb", C.c_int), ("iflux", C.c_int), ("idiff", C.c_int),
        ("inter", C.c_int * 3), ("kover", C.c_int), ("katol", C.c_int),
        ("istres", C.c_int), ("fullns", C.c_int), ("turcor", C.c_int),
        ("ddadi", C.c_int), ("reserved_i", C.c_int * 4)]

    DOUBLES = ("gamma", "rgas", "pr", "prt", "visu0", "expsu", "tsu0",
               "frspre", "frsden", "frstem", "frsvel", "frsvis",
               "cfl", "cfll", "theta", "cdiff", "tuplim", "tlolim", "rklim", "epslim",
               "turlim", "frsmut", "rmult", "re", "csimps", "cmgk", "cmgeps")

    def dvec(self):
        """Parameters as the PRM vector of the reference driver (same order)."""
        return np.array([getattr(self, n) for n in self.DOUBLES], dtype=np.float64)


class BlockView(C.Structure):
    """ff_block of include/ffhip.h: dims + one pointer per field slot."""
    _fields_ = [("imax", C.c_int), ("jmax", C.c_int), ("kmax", C.c_int),
                ("reserved", C.c_int), ("f", C.c_void_p * NF)]


def default_params(**kw):
    """FINFLO defaults (main.f:5232-5394) for an ideal-gas air case."""
    p = Params()
    p.gamma, p.rgas, p.pr, p.prt = 1.4, 287.05287, 0.72, 0.9
    p.visu0, p.expsu, p.tsu0 = 1.458e-6, 1.5, 110.4
    p.frstem, p.frspre = 288.15, 101325.0
    p.frsden = p.frspre / (p.rgas * p.frstem)
    p.frsvel = 0.3 * np.sqrt(p.gamma * p.rgas * p.frstem)
    p.frsvis = 1.789e-5
    p.cfl, p.cfll = 5.0, 5.0
    p.theta = 1.5          # main.f:3499
    p.cdiff = 1.0 / 2.0    # CDIFF = 2 inverted at main.f:4886
    p.tuplim, p.tlolim = 6.0 * p.frstem, 0.05 * p.frstem   # main.f:3827-3828
    p.rklim, p.epslim = 1e-10, 1e-12
    p.turlim, p.frsmut, p.rmult = 1e5, 1e-10, 1.0
    p.re, p.csimps = 1e5, 5.0
    p.cmgk, p.cmgeps = 0.1, 0.2                            # main.f:5344-5345
    p.iturb, p.iflux, p.idiff = 0, 1, 1
    p.inter[0] = p.inter[1] = p.inter[2] = -3
    p.kover, p.katol = 0, 1
    p.istres = 0
    p.fullns = 0
    p.turcor = 1           # TURCOR = YES (main.f:5257)
    for k, v in kw.items():
        if k == "inter":
            for i in range(3):
                p.inter[i] = v[i]
        else:
            setattr(p, k, v)
    return p


class Block:
    """One (level, block) slab as a packed host array F[NF, ntot]."""

    def __init__(self, imax, jmax, kmax, F=None):
        self.imax, self.jmax, self.kmax = int(imax), int(jmax), int(kmax)
        self.istrid = self.imax + 2 * GHOST
        self.jstrid = self.jmax + 2 * GHOST
        self.kstrid = self.kmax + 2 * GHOST
        self.ntot = self.istrid * self.jstrid * self.kstrid
        if F is None:
            F = np.zeros((NF, self.ntot), dtype=np.float64)
        assert F.shape == (NF, self.ntot) and F.dtype == np.float64 and F.flags.c_contiguous
        self.F = F

    def copy(self):
        return Block(self.imax, self.jmax, self.kmax, self.F.copy())

    def __getitem__(self, name):
        return self.F[FID[name]]

    def __setitem__(self, name, value):
        self.F[FID[name]][...] = value

    def view3(self, name):
        """(KSTRID, JSTRID, ISTRID) view of a field, ghosts included."""
        return self.F[FID[name]].reshape(self.kstrid, self.jstrid, self.istrid)

    def interior(self, name):
        """(KMAX, JMAX, IMAX) view of the interior cells."""
        g = GHOST
        return self.view3(name)[g:g + self.kmax, g:g + self.jmax, g:g + self.imax]

    def cell(self, i, j, k):
        """0-based offset of the reference's 1-based cell (I,J,K)."""
        return (k + 1) * self.istrid * self.jstrid + (j + 1) * self.istrid + (i + 1)

    def view_struct(self, base=None, itemsize=8):
        """ff_block pointing into F (or into a device buffer laid out like F)."""
        v = BlockView()
        v.imax, v.jmax, v.kmax = self.imax, self.jmax, self.kmax
        addr = self.F.ctypes.data if base is None else base
        for i in range(NF):
            v.f[i] = addr + i * self.ntot * itemsize
        return v


# ---- solid-wall patches (ff_wall_params, ff_wall_patch of include/ffhip.h) ---------------

class WallParams(C.Structure):
    """ff_wall_params: the constants BOTSKE reads besides ff_params."""
    _fields_ = [(n, C.c_double) for n in (
        "t0", "difpre", "xmom", "ymom", "zmom", "gx", "gy", "gz", "ground", "altitude")] + [
        ("reserved", C.c_double * 6)]

    def dvec(self):
        return np.array([getattr(self, n) for n, _ in self._fields_[:10]] + [0.0] * 6)


WALL_IN = ["uwall", "vwall", "wwall", "twall", "hflux", "rbk", "xcp", "ycp", "zcp"]
WALL_OUT = ["cpwall", "qwall", "qwfric", "tauw1", "tauw2", "tauwx", "tauwy", "tauwz",
            "surfx", "surfy", "surfz", "surfpx", "surfpy", "surfpz"]
WALL_ARRAYS = WALL_IN + WALL_OUT
FORCES = ["CX", "CY", "CZ", "CMX", "CMY", "CMZ", "TOM", "DX", "DY", "DZ", "QX", "QY", "QZ",
          "VMX", "VMY", "VMZ", "BUX", "BUY", "BUZ"]


class WallPatchView(C.Structure):
    _fields_ = [("iface", C.c_int), ("kx1", C.c_int), ("kx2", C.c_int), ("ky1", C.c_int),
                ("ky2", C.c_int), ("iheat", C.c_int), ("reserved", C.c_int * 2)] + [
        (n, C.c_void_p) for n in WALL_ARRAYS]


class WallPatch:
    """A solid-wall patch: ICON(3..7,L), ICON(20,L) and the patch arrays, each
    (NA+4)*(NB+4) doubles with two ghost rows (na along the face's fastest axis: J on an
    I-face, K on a J-face, I on a K-face)."""

    LN = 2

    def __init__(self, iface, kx1, kx2, ky1, ky2, iheat=0, arrays=None):
        self.iface, self.kx1, self.kx2, self.ky1, self.ky2 = iface, kx1, kx2, ky1, ky2
        self.iheat = iheat
        self.dir = (iface - 3 if iface > 3 else iface) - 1
        if self.dir == 1:      # J-face: K fastest (FLXCYC swaps KX and KY)
            self.na, self.nb = ky2 - ky1 + 1, kx2 - kx1 + 1
        else:
            self.na, self.nb = kx2 - kx1 + 1, ky2 - ky1 + 1
        self.size = (self.na + 2 * self.LN) * (self.nb + 2 * self.LN)
        self.a = arrays if arrays is not None else {
            n: np.zeros(self.size) for n in WALL_ARRAYS}

    def copy(self):
        return WallPatch(self.iface, self.kx1, self.kx2, self.ky1, self.ky2, self.iheat,
                         {n: v.copy() for n, v in self.a.items()})

    def interior(self, name):
        """(NB, NA) view of a patch array without its ghost rows."""
        L = self.LN
        return self.a[name].reshape(self.nb + 2 * L, self.na + 2 * L)[L:L + self.nb, L:L + self.na]

    def view(self):
        v = WallPatchView()
        v.iface, v.kx1, v.kx2, v.ky1, v.ky2, v.iheat = (self.iface, self.kx1, self.kx2,
                                                        self.ky1, self.ky2, self.iheat)
        for n in WALL_ARRAYS:
            setattr(v, n, self.a[n].ctypes.data)
        return v


def wall_views(patches):
    arr = (WallPatchView * max(len(patches), 1))()
    for i, wp in enumerate(patches):
        arr[i] = wp.view()
    return arr


# ---- boundary patches (ff_bc_patch of include/ffhip.h) -----------------------------------

# the INL/OUT patch arrays of ff_bc_patch, in its order
BOUN_ARRAYS = ("bounr", "bounu", "bounv", "bounw", "bount", "bounp", "bounpd", "bounrk",
               "bounep")


class BcPatchView(C.Structure):
    """ff_bc_patch: one REFLEC / GHOSTEXT / INLET / OUTLET patch."""
    _fields_ = [("ibc", C.c_int), ("iface", C.c_int), ("kx1", C.c_int), ("kx2", C.c_int),
                ("ky1", C.c_int), ("ky2", C.c_int), ("iheat", C.c_int), ("viscous", C.c_int),
                ("tsurf", C.c_double), ("reserved", C.c_double * 3),
                ("uwall", C.c_void_p), ("vwall", C.c_void_p), ("wwall", C.c_void_p),
                ("inlout", C.c_int), ("reserved_i", C.c_int * 3)] + \
        [(n, C.c_void_p) for n in BOUN_ARRAYS]


class BcPatch:
    """A boundary patch: ICON(1,L) = ibc (2 EXT, 3 INL, 5 OUT, 7 symmetry, 8-10 solid wall),
    ICON(3..7,L), ICON(20,L), REFLEC's viscous-branch flag, RCON(5,L), the wall velocity
    arrays ((NA+4)*(NB+4), as WallPatch; None = at rest), and for INL/OUT patches ICON(8,L)
    (inlout) and the patch arrays BOUNR ... BOUNEP (boun: a dict of (NA+4)*(NB+4) arrays by
    the names of BOUN_ARRAYS; missing ones are None)."""

    def __init__(self, ibc, iface, kx1, kx2, ky1, ky2, iheat=0, viscous=1, tsurf=0.0,
                 uwall=None, vwall=None, wwall=None, inlout=0, boun=None):
        self.ibc, self.iface = ibc, iface
        self.kx1, self.kx2, self.ky1, self.ky2 = kx1, kx2, ky1, ky2
        self.iheat, self.viscous, self.tsurf = iheat, viscous, tsurf
        self.uwall, self.vwall, self.wwall = uwall, vwall, wwall
        self.size = (kx2 - kx1 + 5) * (ky2 - ky1 + 5)
        self.inlout = inlout
        self.boun = {n: None for n in BOUN_ARRAYS}
        for n, a in (boun or {}).items():
            if n not in self.boun:
                raise ValueError(f"unknown patch array {n}")
            self.boun[n] = None if a is None else np.ascontiguousarray(a, dtype=np.float64)

    def copy(self):
        """The same patch with its own copies of the arrays."""
        c = BcPatch(self.ibc, self.iface, self.kx1, self.kx2, self.ky1, self.ky2, self.iheat,
                    self.viscous, self.tsurf, self.uwall, self.vwall, self.wwall, self.inlout,
                    {n: (None if a is None else a.copy()) for n, a in self.boun.items()})
        return c

    def view(self):
        v = BcPatchView()
        v.ibc, v.iface, v.kx1, v.kx2, v.ky1, v.ky2 = (self.ibc, self.iface, self.kx1, self.kx2,
                                                      self.ky1, self.ky2)
        v.iheat, v.viscous, v.tsurf = self.iheat, self.viscous, self.tsurf
        for n in ("uwall", "vwall", "wwall"):
            a = getattr(self, n)
            setattr(v, n, None if a is None else a.ctypes.data)
        v.inlout = self.inlout
        for n in BOUN_ARRAYS:
            a = self.boun[n]
            setattr(v, n, None if a is None else a.ctypes.data)
        return v


def bc_views(patches):
    arr = (BcPatchView * max(len(patches), 1))()
    for i, bp in enumerate(patches):
        arr[i] = bp.view()
    return arr


# ---- block connections (ff_con_patch of include/ffhip.h) --------------------------------

class ConPatchView(C.Structure):
    """ff_con_patch: one side of a one-to-one block connection."""
    _fields_ = [("iface", C.c_int), ("ka1", C.c_int), ("ka2", C.c_int), ("kb1", C.c_int),
                ("kb2", C.c_int), ("nblock", C.c_int), ("nface", C.c_int), ("na0", C.c_int),
                ("nb0", C.c_int), ("swap", C.c_int), ("da", C.c_int), ("db", C.c_int),
                ("reserved", C.c_int * 4)]


def con_views(patches):
    """ff_con_patch array from tuples (iface, ka1, ka2, kb1, kb2, nblock, nface, na0, nb0,
    swap, da, db)."""
    arr = (ConPatchView * max(len(patches), 1))()
    for i, t in enumerate(patches):
        v = arr[i]
        (v.iface, v.ka1, v.ka2, v.kb1, v.kb2, v.nblock, v.nface, v.na0, v.nb0, v.swap, v.da,
         v.db) = t
    return arr
