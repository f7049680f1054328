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
        ("istres", C.c_int), ("fullns", C.c_int), ("reserved_i", C.c_int * 6)]

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
