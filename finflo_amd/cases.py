"""Synthetic blocks: node grids, finite-volume metrics and seeded initial states.

The reference reads its grid from an unformatted file (main.f:1366-1374) and forms face
areas, unit normals and volumes in AREVC1/VOLINC (gluebm.f:1586, 1889).  That
preprocessing is outside the hot path; this module builds the same arrays for synthetic
cases: face-area vectors from the face diagonals, volumes from the divergence theorem,
two ghost layers of linearly extrapolated nodes.  Face data of direction d is stored at the
upper cell of the face, as the reference stores it.
"""
import numpy as np

from .layout import GHOST, BcPatch, Block, WallParams, WallPatch, default_params


def stretched(n, length, first=None, ratio=None):
    """n+1 node coordinates on [0, length]: uniform, or geometric with the given ratio
    (first cell width follows from the ratio)."""
    if ratio is None or ratio == 1.0:
        return np.linspace(0.0, length, n + 1)
    w = ratio ** np.arange(n)
    x = np.concatenate([[0.0], np.cumsum(w)])
    return x * (length / x[-1])


def geometric_from_wall(n, first, ratio):
    """n+1 nodes with first cell width `first` growing by `ratio`."""
    w = first * ratio ** np.arange(n)
    return np.concatenate([[0.0], np.cumsum(w)])


def cosine_clustered(n, length):
    """n+1 nodes clustered towards both ends (cosine spacing)."""
    t = np.linspace(0.0, np.pi, n + 1)
    return 0.5 * length * (1.0 - np.cos(t))


def _extend(x):
    """Add two linearly extrapolated ghost nodes at each end of a 1-D node array."""
    lo1 = 2 * x[0] - x[1]
    lo2 = 2 * lo1 - x[0]
    hi1 = 2 * x[-1] - x[-2]
    hi2 = 2 * hi1 - x[-1]
    return np.concatenate([[lo2, lo1], x, [hi1, hi2]])


def _extend_axis(X, axis):
    """Two linearly extrapolated ghost node layers at both ends of `axis` of a node array."""
    X = np.moveaxis(X, axis, 0)
    lo1 = 2 * X[0] - X[1]
    lo2 = 2 * lo1 - X[0]
    hi1 = 2 * X[-1] - X[-2]
    hi2 = 2 * hi1 - X[-1]
    X = np.concatenate([lo2[None], lo1[None], X, hi1[None], hi2[None]], axis=0)
    return np.moveaxis(X, 0, axis)


def coarsen_nodes(nodes, planar=False):
    """Node array of the next multigrid level: every other interior node of the finer level
    (the reference's coarse levels take every second grid line, main.f:4600-4660), with two
    ghost node layers extrapolated again.  planar: a block one cell thick in K stays one cell
    thick (KMAXG = MAX0(1, KMAX/NDIV), main.f:1756); its coarse cell spans K node planes 1
    and 3, every second plane as in I and J, the third being the first ghost plane, so that
    it is twice as thick and TRANSF's planar sum, which counts each of the four fine cells
    twice (multi.f:185-209), is the volume average."""
    g = GHOST
    if planar:
        inner = nodes[g:g + 3:2, g:-g, g:-g][:, ::2, ::2]
    else:
        inner = nodes[g:-g, g:-g, g:-g][::2, ::2, ::2]
    for ax in range(3):
        inner = _extend_axis(inner, ax)
    return inner


def node_grid(xi, yj, zk, distort=0.0, seed=0):
    """Nodes (KS+1, JS+1, IS+1, 3) of a box with ghost nodes; `distort` adds a smooth
    interior deformation (fraction of the local spacing) to exercise curvilinear metrics."""
    x, y, z = _extend(np.asarray(xi, float)), _extend(np.asarray(yj, float)), _extend(np.asarray(zk, float))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    if distort:
        rng = np.random.default_rng(seed)
        ph = rng.uniform(0, 2 * np.pi, 6)
        lx, ly, lz = x[-1] - x[0], y[-1] - y[0], z[-1] - z[0]
        dx = np.gradient(x).min()
        dy = np.gradient(y).min()
        dz = np.gradient(z).min()
        s = distort
        X = X + s * dx * np.sin(2 * np.pi * Y / ly + ph[0]) * np.sin(2 * np.pi * Z / lz + ph[1])
        Y = Y + s * dy * np.sin(2 * np.pi * X / lx + ph[2]) * np.sin(2 * np.pi * Z / lz + ph[3])
        Z = Z + s * dz * np.sin(2 * np.pi * X / lx + ph[4]) * np.sin(2 * np.pi * Y / ly + ph[5])
    return np.stack([X, Y, Z], axis=-1)


def _face_vec(p00, p10, p11, p01):
    return 0.5 * np.cross(p11 - p00, p01 - p10)


def fill_geometry(blk, nodes):
    """Face areas/normals, volumes, spacings and centres of every cell (ghosts included)."""
    Xn = nodes
    ks, js, is_ = blk.kstrid, blk.jstrid, blk.istrid
    assert Xn.shape == (ks + 1, js + 1, is_ + 1, 3)
    # i-faces at node planes 0..IS (cells' lower faces plus the last upper face)
    Si = _face_vec(Xn[:-1, :-1, :], Xn[:-1, 1:, :], Xn[1:, 1:, :], Xn[1:, :-1, :])      # (KS,JS,IS+1)
    Sj = _face_vec(Xn[:-1, :, :-1], Xn[1:, :, :-1], Xn[1:, :, 1:], Xn[:-1, :, 1:])      # (KS,JS+1,IS)
    Sk = _face_vec(Xn[:, :-1, :-1], Xn[:, :-1, 1:], Xn[:, 1:, 1:], Xn[:, 1:, :-1])      # (KS+1,JS,IS)
    Ci = 0.25 * (Xn[:-1, :-1, :] + Xn[:-1, 1:, :] + Xn[1:, 1:, :] + Xn[1:, :-1, :])
    Cj = 0.25 * (Xn[:-1, :, :-1] + Xn[1:, :, :-1] + Xn[1:, :, 1:] + Xn[:-1, :, 1:])
    Ck = 0.25 * (Xn[:, :-1, :-1] + Xn[:, :-1, 1:] + Xn[:, 1:, 1:] + Xn[:, 1:, :-1])
    vol = (np.sum(Si[:, :, 1:] * Ci[:, :, 1:], -1) - np.sum(Si[:, :, :-1] * Ci[:, :, :-1], -1)
           + np.sum(Sj[:, 1:, :] * Cj[:, 1:, :], -1) - np.sum(Sj[:, :-1, :] * Cj[:, :-1, :], -1)
           + np.sum(Sk[1:, :, :] * Ck[1:, :, :], -1) - np.sum(Sk[:-1, :, :] * Ck[:-1, :, :], -1)) / 3.0
    if np.any(vol <= 0):
        raise ValueError("non-positive cell volume in synthetic grid")
    for name, S in (("A1", Si[:, :, :-1]), ("A2", Sj[:, :-1, :]), ("A3", Sk[:-1, :, :])):
        a = np.linalg.norm(S, axis=-1)
        blk.view3(name)[...] = a
        blk.view3(name + "X")[...] = S[..., 0] / a
        blk.view3(name + "Y")[...] = S[..., 1] / a
        blk.view3(name + "Z")[...] = S[..., 2] / a
    blk.view3("VOL")[...] = vol
    cen = 0.125 * (Xn[:-1, :-1, :-1] + Xn[:-1, :-1, 1:] + Xn[:-1, 1:, :-1] + Xn[:-1, 1:, 1:]
                   + Xn[1:, :-1, :-1] + Xn[1:, :-1, 1:] + Xn[1:, 1:, :-1] + Xn[1:, 1:, 1:])
    blk.view3("XC")[...] = cen[..., 0]
    blk.view3("YC")[...] = cen[..., 1]
    blk.view3("ZC")[...] = cen[..., 2]
    blk.view3("D1")[...] = np.linalg.norm(Ci[:, :, 1:] - Ci[:, :, :-1], axis=-1)
    blk.view3("D2")[...] = np.linalg.norm(Cj[:, 1:, :] - Cj[:, :-1, :], axis=-1)
    blk.view3("D3")[...] = np.linalg.norm(Ck[1:, :, :] - Ck[:-1, :, :], axis=-1)
    # wall distance to the j = 0 node plane (flat-plate / box cases)
    blk.view3("DISTW")[...] = np.abs(cen[..., 1] - Xn[:, GHOST:GHOST + 1, :, 1].mean())
    return blk


def fill_state(blk, prm, seed=0, amp=0.01, mach_dir=(1.0, 0.0, 0.0), kind="density",
               turb=None, turb_amp=None):
    """Free stream plus seeded perturbations on every cell (ghosts included).

    kind "density": T = T_inf / (1 + d), d ~ U(-amp, amp) at fixed p (the configs' density
    perturbation expressed in FINFLO's primitive variables); "velocity": u, v, w perturbed
    by amp*U_inf; "all": every primitive perturbed (parity tests).  `turb` = (RK, REPS) free
    stream values, perturbed relatively by `turb_amp` (default `amp`)."""
    rng = np.random.default_rng(seed)
    n = blk.ntot
    d = np.asarray(mach_dir, float)
    d = d / np.linalg.norm(d)
    U = prm.frsvel
    blk["PDIFF"] = 0.0
    blk["TEMP"] = prm.frstem
    blk["U"], blk["V"], blk["W"] = U * d[0], U * d[1], U * d[2]
    if kind == "density":
        delta = rng.uniform(-amp, amp, n)
        blk["TEMP"] = prm.frstem / (1.0 + delta)
    elif kind == "velocity":
        for name in ("U", "V", "W"):
            blk[name] = blk[name] + amp * U * rng.uniform(-1, 1, n)
    elif kind == "all":
        blk["PDIFF"] = amp * prm.frspre * rng.uniform(-1, 1, n)
        blk["TEMP"] = prm.frstem * (1.0 + amp * rng.uniform(-1, 1, n))
        for name in ("U", "V", "W"):
            blk[name] = blk[name] + amp * U * rng.uniform(-1, 1, n) * 3
    else:
        raise ValueError(kind)
    if turb is not None:
        rk, reps = turb
        ta = amp if turb_amp is None else turb_amp
        blk["RK"] = rk * (1.0 + ta * rng.uniform(-1, 1, n))
        blk["REPS"] = reps * (1.0 + ta * rng.uniform(-1, 1, n))
    blk["P"] = blk["PDIFF"] + prm.frspre
    # laminar eddy-viscosity ratio and eddy viscosity on every cell (the boundary routines
    # own the ghost values; TURBDI writes the interior)
    blk["EPS2"] = 1.0
    blk["VIST"] = 0.0
    return blk


def make_block(imax, jmax, kmax, prm=None, lengths=(2.0, 1.0, 1.0), jratio=None,
               distort=0.0, seed=0, amp=0.01, kind="density", turb=None,
               mach_dir=(1.0, 0.0, 0.0)):
    """A synthetic block with geometry and state (parity-test cases and workloads)."""
    prm = prm or default_params()
    blk = Block(imax, jmax, kmax)
    xi = stretched(imax, lengths[0])
    yj = stretched(jmax, lengths[1], ratio=jratio)
    zk = stretched(kmax, lengths[2])
    fill_geometry(blk, node_grid(xi, yj, zk, distort=distort, seed=seed + 1000))
    fill_state(blk, prm, seed=seed, amp=amp, kind=kind, turb=turb, mach_dir=mach_dir)
    return blk


def make_hierarchy(imax, jmax, kmax, nlev, prm=None, lengths=(2.0, 1.0, 1.0), jratio=None,
                   distort=0.0, seed=0, amp=0.01, kind="density", turb=None,
                   mach_dir=(1.0, 0.0, 0.0), nodes=None, turb_amp=None, surface=None):
    """A block and its nlev-1 coarser multigrid levels (dimensions divisible by 2^(nlev-1)).
    Every level has its own geometry; the state of the coarse levels is free stream (the
    restriction overwrites what the V-cycle uses).  `surface(x, z)`, when given, is the
    height of the J-low wall; the wall distance is then measured vertically from it."""
    prm = prm or default_params()
    f = 2 ** (nlev - 1)
    planar = kmax == 1          # one cell thick in K: KMAX = 1 on every level
    if imax % f or jmax % f or (kmax % f and not planar):
        raise ValueError(f"dimensions must be divisible by {f} for {nlev} levels")
    if nodes is None:
        xi = stretched(imax, lengths[0])
        yj = stretched(jmax, lengths[1], ratio=jratio)
        zk = stretched(kmax, lengths[2])
        nodes = node_grid(xi, yj, zk, distort=distort, seed=seed + 1000)
    blocks = []
    for m in range(nlev):
        b = Block(imax >> m, jmax >> m, kmax if planar else kmax >> m)
        fill_geometry(b, nodes)
        if surface is not None:
            b["DISTW"] = np.abs(b["YC"] - surface(b["XC"], b["ZC"]))
        fill_state(b, prm, seed=seed + m, amp=amp if m == 0 else 0.0, kind=kind,
                   turb=turb, mach_dir=mach_dir, turb_amp=turb_amp if m == 0 else 0.0)
        blocks.append(b)
        if m < nlev - 1:
            nodes = coarsen_nodes(nodes, planar)
    return blocks


def flat_plate_nodes(imax, jmax, kmax, first=1e-6, ratio=1.15, length=1.0, width=None):
    xi = cosine_clustered(imax, length)
    yj = geometric_from_wall(jmax, first * length, ratio)
    zk = np.linspace(0.0, width if width else 0.5 * length, kmax + 1)
    return node_grid(xi, yj, zk)


def flat_plate_block(imax, jmax, kmax, prm, first=1e-6, ratio=1.15, length=1.0, width=None,
                     seed=0, amp=0.005):
    """Flat-plate block: cosine-clustered i over the plate length, geometric j from the
    wall with first cell `first` and growth `ratio`, uniform k."""
    blk = Block(imax, jmax, kmax)
    xi = cosine_clustered(imax, length)
    yj = geometric_from_wall(jmax, first * length, ratio)
    zk = np.linspace(0.0, width if width else 0.5 * length, kmax + 1)
    fill_geometry(blk, node_grid(xi, yj, zk))
    fill_state(blk, prm, seed=seed, amp=amp, kind="velocity")
    return blk


def flat_plate_hierarchy(imax, jmax, kmax, nlev, prm, first=1e-6, ratio=1.15, length=1.0,
                         width=None, seed=0, amp=0.005, turb=None):
    """The flat-plate block with its nlev-1 coarser multigrid levels: velocity noise of
    relative amplitude `amp` on the finest level, unperturbed (RK, REPS) = `turb`."""
    nodes = flat_plate_nodes(imax, jmax, kmax, first, ratio, length, width)
    return make_hierarchy(imax, jmax, kmax, nlev, prm, seed=seed, amp=amp, kind="velocity",
                          turb=turb, nodes=nodes, turb_amp=0.0)


def plate_wall(blk, prm, iheat=0):
    """The no-slip wall of the flat-plate and box cases: one solid-wall patch over the whole
    J-low face (ICON(3) = 2, KX = 1..IMAX, KY = 1..KMAX), at rest, smooth (RBK = 0),
    adiabatic by default, moment arms from the face centres (the cell centres of J = 1
    projected onto the y = 0 plane), no gravity; force sums relative to the free-stream
    pressure."""
    wp = WallParams()
    wp.t0 = prm.frstem
    wp.difpre = prm.frspre
    wp.ground, wp.altitude = 0.0, 1.0          # ALTITUDE >= GROUND: no hydrostatic part
    p = WallPatch(2, 1, blk.imax, 1, blk.kmax, iheat)
    L = p.LN
    # J-face patch arrays: K fastest (na = KMAX), I slowest (nb = IMAX)
    for name, src in (("xcp", "XC"), ("zcp", "ZC")):
        a = p.a[name].reshape(p.nb + 2 * L, p.na + 2 * L)
        a[L:L + p.nb, L:L + p.na] = blk.interior(src)[:, 0, :].T
    p.a["twall"][:] = prm.frstem
    return wp, [p]


# ---- the benchmark configurations ----------------------------------------------------------

def flow_params(iturb=0, mach=0.3, reynolds=1e5, cfl=5.0, **kw):
    """FINFLO defaults with a case's free stream: the speed from the Mach number at the
    default temperature, and Sutherland's viscosity scaled (RMULTV) so that rho U L / mu is
    the Reynolds number per unit length."""
    prm = default_params(iturb=iturb, **kw)
    c = np.sqrt(prm.gamma * prm.rgas * prm.frstem)
    prm.frsvel = mach * c
    t = prm.frstem
    mu_suth = prm.visu0 * t * np.sqrt(t) / (t + prm.tsu0)
    prm.rmult = prm.frsden * prm.frsvel / reynolds / mu_suth
    prm.frsvis = mu_suth * prm.rmult
    prm.re = reynolds
    prm.cfl = prm.cfll = cfl
    return prm


def freestream_turbulence(prm, tu=0.01, mut_ratio=1.0, limits=False):
    """(RK, REPS) = (rho k, rho omega) of a free stream with turbulence intensity `tu` and
    eddy-to-molecular viscosity ratio `mut_ratio`.  With limits, the parameters take them as
    the reference does for a two-equation model (main.f:3772-3826: the input RKLIM and EPSLIM
    are Tu and mu_t/mu; RKLIM = 1.5 rho (Tu U)^2, FRSMUT = (mu_t/mu) mu, EPSLIM = rho RKLIM /
    FRSMUT for IDIS = 2): the floors of rho k and rho omega, the free-stream mu_t, and the
    values GHOSTEXT puts into the free-stream ghost cells."""
    rkii = 1.5 * prm.frsden * (tu * prm.frsvel) ** 2
    frsmut = mut_ratio * prm.frsvis
    repsi = prm.frsden * rkii / frsmut
    if limits:
        prm.rklim, prm.epslim, prm.frsmut = rkii, repsi, frsmut
    return rkii, repsi


def config1(iflux=1, seed=0):
    """The plain method's CPU-runnable case: one 48x24x24 block on [0,2]x[0,1]x[0,1],
    J spacing growing by 1.1 from the no-slip wall on the J-low face, laminar, M = 0.3,
    Re = 1e5 per unit length, CFL 5, free stream with a seeded +/-1 % density perturbation
    (T = T_inf/(1+d) at the free-stream pressure); the other faces keep their ghost cells
    as given (free stream).  iflux 1: Roe, 0: van Leer."""
    prm = flow_params(iturb=0, mach=0.3, reynolds=1e5, cfl=5.0, iflux=iflux)
    blk = make_block(48, 24, 24, prm, lengths=(2.0, 1.0, 1.0), jratio=1.1, seed=seed,
                     amp=0.01, kind="density")
    return prm, blk, plate_wall(blk, prm)


def config2(seed=0, nlev=3):
    """The flat plate: 192x96x96 cells, cosine-clustered I over the plate length, first wall
    cell 1e-6 of it and J growth 1.15, k-omega SST, M = 0.2, Re = 5e6, velocity noise
    +/-0.5 %, CFL 5, nlev multigrid levels, the no-slip wall on every level's J-low face."""
    prm = flow_params(iturb=6, mach=0.2, reynolds=5e6, cfl=5.0)
    turb = freestream_turbulence(prm, limits=True)
    blocks = flat_plate_hierarchy(192, 96, 96, nlev, prm, first=1e-6, ratio=1.15, seed=seed,
                                  amp=0.005, turb=turb)
    return prm, blocks, [plate_wall(b, prm) for b in blocks]


def gaussian_bump(h=0.05, w=0.2, x0=0.0, z0=0.0, sweep=0.0):
    """Height of a smoothed Gaussian wall bump, h exp(-((x-x0)^2 + (z-z0)^2)/w^2); with a
    sweep angle (radians) its crest line runs at that angle across the span instead."""
    def surface(x, z):
        if sweep:
            xs = (x - x0) - np.tan(sweep) * (z - z0)
            return h * np.exp(-(xs * xs) / (w * w))
        return h * np.exp(-((x - x0) ** 2 + (z - z0) ** 2) / (w * w))
    return surface


def bump_nodes(imax, jmax, kmax, surface, xlen=2.0, height=1.0, zlen=1.0, first=1e-5,
               ratio=None):
    """Nodes of a box over a wall bump: I and K uniform over [-xlen/2, xlen/2] and
    [-zlen/2, zlen/2], J lines from the bump surface to y = height, geometric from the wall
    with the first cell `first` of the local height: growing by the rate that fills the
    height when ratio is None, else by `ratio` up to a spacing cap that fills it."""
    xi = np.linspace(-0.5 * xlen, 0.5 * xlen, imax + 1)
    zk = np.linspace(-0.5 * zlen, 0.5 * zlen, kmax + 1)
    if ratio is None:
        lo, hi = 1.0 + 1e-9, 2.0
        for _ in range(200):                     # growth with first*(r^n - 1)/(r - 1) = 1
            r = 0.5 * (lo + hi)
            if first * (r ** jmax - 1.0) / (r - 1.0) > 1.0:
                hi = r
            else:
                lo = r
        ratio = 0.5 * (lo + hi)
        eta = geometric_from_wall(jmax, 1.0, ratio)
        eta = eta / eta[-1]
    else:
        # the first cell and the growth as given, the spacing capped where the cells would
        # otherwise overshoot the height: min(first * ratio^n, cap) summing to 1
        lo, hi = first, 1.0
        for _ in range(200):
            cap = 0.5 * (lo + hi)
            w = np.minimum(first * ratio ** np.arange(jmax), cap)
            if w.sum() > 1.0:
                hi = cap
            else:
                lo = cap
        w = np.minimum(first * ratio ** np.arange(jmax), 0.5 * (lo + hi))
        eta = np.concatenate([[0.0], np.cumsum(w)])
        eta = eta / eta[-1]
    X = _extend(xi)
    Z = _extend(zk)
    E = _extend(eta)
    Zg, Eg, Xg = np.meshgrid(Z, E, X, indexing="ij")
    ys = surface(Xg, Zg)
    Y = ys + (height - ys) * Eg
    return np.stack([Xg, Y, Zg], axis=-1)


def config4(seed=0):
    """The learned-closure case: one 256x128x128 block over a 3-D Gaussian bump (height
    0.05, width 0.2, on a 2 x 1 x 1 box, first wall cell 1e-5), k-omega SST with the EARSM
    (ISTRES = 1), M = 0.3, Re = 2e6, free stream with +/-5 % velocity and 30 % turbulence
    noise so that the anisotropy is not trivial everywhere."""
    prm = flow_params(iturb=6, mach=0.3, reynolds=2e6, cfl=5.0, istres=1)
    turb = freestream_turbulence(prm, limits=True)
    surf = gaussian_bump(0.05, 0.2)
    nodes = bump_nodes(256, 128, 128, surf, xlen=2.0, height=1.0, zlen=1.0, first=1e-5)
    blk = make_hierarchy(256, 128, 128, 1, prm, seed=seed, amp=0.05, kind="velocity",
                         turb=turb, nodes=nodes, turb_amp=0.3, surface=surf)[0]
    return prm, blk, plate_wall(blk, prm)


def plate_bcs(blk, wall_faces=(2,), outer="ext", prm=None):
    """The boundary patches of the flat-plate and box cases: a no-slip wall (REFLEC's
    viscous branch, at rest) on each face of `wall_faces`, the free stream (EXT, GHOSTEXT)
    on every other face.  outer = "inout" makes the I-low face a subsonic inlet (INLOUT 4:
    the free stream's density, velocity along I and temperature given, the pressure
    extrapolated; the k-omega free-stream limits) and the I-high face a pressure outlet
    (INLOUT 3: the free-stream pressure given, the velocity extrapolated); prm gives the
    free stream then."""
    ext = {1: (blk.jmax, blk.kmax), 2: (blk.imax, blk.kmax), 3: (blk.imax, blk.jmax)}
    out = []
    for face in range(1, 7):
        na, nb = ext[(face - 1) % 3 + 1]
        if face in wall_faces:
            out.append(BcPatch(8, face, 1, na, 1, nb, iheat=0, viscous=1))
        elif outer == "inout" and face in (1, 4):
            bp = BcPatch(3 if face == 1 else 5, face, 1, na, 1, nb, inlout=4 if face == 1 else 3)
            n = bp.size
            full = lambda v: np.full(n, float(v))
            bp.boun.update(bounr=full(prm.frsden), bounu=full(prm.frsvel), bounv=full(0.0),
                           bounw=full(0.0), bount=full(prm.frstem), bounp=full(prm.frspre),
                           bounpd=full(0.0), bounrk=full(prm.rklim), bounep=full(prm.epslim))
            out.append(bp)
        else:
            out.append(BcPatch(2, face, 1, na, 1, nb, iheat=0, viscous=1))
    return out


# ---- multi-block cases ----------------------------------------------------------------------

def _proper(perm, flip):
    """True when the axis map (sub-block axis n runs along global axis perm[n], reversed when
    flip[n]) is a rotation, so that volumes stay positive."""
    inv = sum(1 for a in range(3) for b in range(a + 1, 3) if perm[a] > perm[b])
    return (inv + sum(bool(f) for f in flip)) % 2 == 0


def sub_block(glob, nodes, lo, hi, perm=(0, 1, 2), flip=(False, False, False)):
    """The sub-block of global cells lo..hi (1-based (i, j, k) inclusive) of the block `glob`
    on node array `nodes`, with its own index axes: its axis n runs along global axis perm[n]
    (0 = i, 1 = j, 2 = k), reversed when flip[n].  Geometry from the sub-range's nodes, whose
    ghost nodes are the global block's on its boundary and extrapolated on a cut side (the
    connection exchange brings the neighbour's geometry there); every other field, ghost
    cells included, copied from the global block."""
    if not _proper(perm, flip):
        raise ValueError("the axis map must be a rotation")
    g = GHOST
    (i0, j0, k0), (i1, j1, k1) = lo, hi
    # the sub-range with two ghost node layers: the global block's own ghost nodes on its
    # boundary, linear extrapolation on a cut side
    sub = nodes[k0 - 1:g + k1 + 1 + g, j0 - 1:g + j1 + 1 + g, i0 - 1:g + i1 + 1 + g].copy()
    gmax = (glob.imax, glob.jmax, glob.kmax)
    for ax, (a0, a1) in zip((2, 1, 0), ((i0, i1), (j0, j1), (k0, k1))):
        gm = gmax[2 - ax]
        x = np.moveaxis(sub, ax, 0)
        if a0 > 1:
            x[1] = 2 * x[2] - x[3]
            x[0] = 2 * x[1] - x[2]
        if a1 < gm:
            x[-2] = 2 * x[-3] - x[-4]
            x[-1] = 2 * x[-2] - x[-3]
    A = np.transpose(sub, (2, 1, 0, 3))                   # [i][j][k]
    B = np.transpose(A, tuple(perm) + (3,))
    for n in range(3):
        if flip[n]:
            B = np.flip(B, axis=n)
    B = np.ascontiguousarray(np.transpose(B, (2, 1, 0, 3)))   # [k'][j'][i']
    dims = [(hi[perm[n]] - lo[perm[n]] + 1) for n in range(3)]
    blk = Block(*dims)
    fill_geometry(blk, B)
    geom = set(("VOL", "A1", "A1X", "A1Y", "A1Z", "A2", "A2X", "A2Y", "A2Z", "A3", "A3X",
                "A3Y", "A3Z", "D1", "D2", "D3", "XC", "YC", "ZC"))
    from .layout import FIELDS
    for name in FIELDS:
        if name in geom:
            continue
        v = glob.view3(name)[k0 - 1:k1 + 4, j0 - 1:j1 + 4, i0 - 1:i1 + 4]
        C = np.transpose(v, (2, 1, 0))
        C = np.transpose(C, perm)
        for n in range(3):
            if flip[n]:
                C = np.flip(C, axis=n)
        blk.view3(name)[...] = np.transpose(C, (2, 1, 0))
    return blk


def block_from_nodes(nodes, lo, hi, prm, seed=0, amp=0.0, kind="velocity", turb=None,
                     turb_amp=None, surface=None):
    """The block of global cells lo..hi (1-based, inclusive) of node array `nodes`, in the
    global index orientation: geometry from its nodes with extrapolated ghost nodes, and a
    seeded state of its own."""
    g = GHOST
    (i0, j0, k0), (i1, j1, k1) = lo, hi
    sub = nodes[g + k0 - 1:g + k1 + 1, g + j0 - 1:g + j1 + 1, g + i0 - 1:g + i1 + 1]
    for ax in range(3):
        sub = _extend_axis(sub, ax)
    blk = Block(i1 - i0 + 1, j1 - j0 + 1, k1 - k0 + 1)
    fill_geometry(blk, sub)
    if surface is not None:
        blk["DISTW"] = np.abs(blk["YC"] - surface(blk["XC"], blk["ZC"]))
    fill_state(blk, prm, seed=seed, amp=amp, kind=kind, turb=turb, turb_amp=turb_amp)
    return blk, sub


def segment_layout(counts, seed, target, tol=0.07, lo=16, hi=96, mult=4):
    """Block dimensions along each axis: counts[a] segments per axis a, each a multiple of
    `mult` in [lo, hi], drawn with a seeded generator until the total cell count is within
    `tol` of `target`."""
    rng = np.random.default_rng(seed)
    while True:
        segs = [rng.integers(lo // mult, hi // mult + 1, n) * mult for n in counts]
        total = int(np.prod([sg.sum() for sg in segs]))
        if abs(total - target) <= tol * target:
            return segs


def config3_layout(seed=3, target=8.0e6, tol=0.07):
    """The segment lengths of the 24-block case: 4 x 3 x 2 blocks whose dimensions are drawn
    from multiples of 4 in [16, 96] until the total is within `tol` of `target` cells."""
    return segment_layout((4, 3, 2), seed, target, tol)


def bump_blocks(prm, segs, surf, nlev, seed, first, ratio=None, turb=None, amp=0.005):
    """Blocks tiling a 2 x 1 x 1 box over the wall bump `surf`, one per combination of the
    axis segments `segs`, each with its nlev levels, in the global index orientation; their
    one-to-one connections (full faces), walls (the J-low faces of the bottom row: BOTSKE
    patches and REFLEC's viscous branch) and free-stream (EXT) patches on the other outer
    faces.  Block ids run I fastest, then J, then K, from 1."""
    si, sj, sk = segs
    nodes = bump_nodes(int(si.sum()), int(sj.sum()), int(sk.sum()), surf, xlen=2.0,
                       height=1.0, zlen=1.0, first=first, ratio=ratio)
    ci = np.concatenate([[0], np.cumsum(si)])
    cj = np.concatenate([[0], np.cumsum(sj)])
    ck = np.concatenate([[0], np.cumsum(sk)])
    ids = {}
    for c in range(len(sk)):
        for b in range(len(sj)):
            for a in range(len(si)):
                ids[(a, b, c)] = len(ids) + 1
    blocks, cons, walls, bcs = [], [], [], []
    for (a, b, c), n in ids.items():
        lo = (ci[a] + 1, cj[b] + 1, ck[c] + 1)
        hi = (ci[a + 1], cj[b + 1], ck[c + 1])
        fine, sub = block_from_nodes(nodes, lo, hi, prm, seed=seed * 100 + n, amp=amp,
                                     kind="velocity", turb=turb, turb_amp=0.0, surface=surf)
        lv = [fine]
        cn = sub
        for m in range(1, nlev):
            cn = coarsen_nodes(cn)
            cb = Block(fine.imax >> m, fine.jmax >> m, fine.kmax >> m)
            fill_geometry(cb, cn)
            cb["DISTW"] = np.abs(cb["YC"] - surf(cb["XC"], cb["ZC"]))
            fill_state(cb, prm, seed=0, amp=0.0, kind="velocity", turb=turb, turb_amp=0.0)
            lv.append(cb)
        blocks.append(lv)
        pc = []
        nb = {1: (a - 1, b, c), 4: (a + 1, b, c), 2: (a, b - 1, c), 5: (a, b + 1, c),
              3: (a, b, c - 1), 6: (a, b, c + 1)}
        for face, key in nb.items():
            if key in ids:
                ea = fine.jmax if face in (1, 4) else fine.imax
                eb = fine.jmax if face in (3, 6) else fine.kmax
                opp = face + 3 if face <= 3 else face - 3
                pc.append((face, 1, ea, 1, eb, ids[key], opp, 1, 1, 0, 1, 1))
        cons.append(pc)
        wl, bl = [], []
        for blk in lv:
            wl.append(plate_wall(blk, prm) if b == 0 else None)
            faces = []
            for face, key in nb.items():
                if key in ids:
                    continue
                ea = blk.jmax if face in (1, 4) else blk.imax
                eb = blk.jmax if face in (3, 6) else blk.kmax
                faces.append(BcPatch(8 if face == 2 else 2, face, 1, ea, 1, eb, iheat=0, viscous=1))
            bl.append(faces)
        walls.append(wl)
        bcs.append(bl)
    return blocks, cons, walls, bcs


def config3(seed=3, nlev=3):
    """The batched form's case: 24 blocks (4 x 3 x 2, seeded dimensions in [16, 96], about
    8M cells) tiling a 2 x 1 x 1 box over a smoothed Gaussian wall bump (height 0.05, width
    0.2, first wall cell 1e-5), k-omega SST, M = 0.5, Re = 1e6, CFL 5, nlev multigrid levels.
    Returns the parameters, the blocks (each a list of its levels, block ids 1..24), their
    connection patches (tuples of ff_con_patch), their walls (the J-low faces of the bottom
    row) and their boundary patches (walls and free stream)."""
    prm = flow_params(iturb=6, mach=0.5, reynolds=1e6, cfl=5.0)
    turb = freestream_turbulence(prm, limits=True)
    segs = config3_layout(seed)
    blocks, cons, walls, bcs = bump_blocks(prm, segs, gaussian_bump(0.05, 0.2), nlev, seed,
                                           first=1e-5, turb=turb)
    return prm, blocks, cons, walls, bcs


def config5(seed=5, nlev=4, target=4.0e7):
    """The full chosen form's case: 64 blocks (4 x 4 x 4, seeded dimensions, multiples of 8
    in [16, 128] so that four multigrid levels halve them, about 40M cells) around a swept
    bump (height 0.05, width 0.2, crest swept 30 degrees), first wall cell 1e-6 growing by
    1.12, k-omega SST with the EARSM (ISTRES = 1; the learned closure is the caller's
    choice), M = 0.6, Re = 1e7, CFL 5."""
    prm = flow_params(iturb=6, mach=0.6, reynolds=1e7, cfl=5.0, istres=1)
    turb = freestream_turbulence(prm, limits=True)
    segs = segment_layout((4, 4, 4), seed, target, tol=0.05, lo=16, hi=128, mult=8)
    blocks, cons, walls, bcs = bump_blocks(prm, segs, gaussian_bump(0.05, 0.2, sweep=np.pi / 6),
                                           nlev, seed, first=1e-6, ratio=1.12, turb=turb)
    return prm, blocks, cons, walls, bcs
