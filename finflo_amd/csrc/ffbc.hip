// ffbc.hip — the boundary routines of the path on the device: REFLEC (turb3.f:5294-5878),
// which every STATIM of CYCLES ends with (ns3c.f:10497-10512) and CYCLES calls once more
// at the end of a cycle (ns3c.f:2168), GHOSTEXT (ns3c.f:10979-11226), the free-stream
// ghost cells of EXT patches set at start-up (ns3c.f:294), and INLET/OUTLET (base3c.f:
// 9223-10216), the update of the inlet and outlet ghost cells that the level-1 residual
// makes before its fluxes (FLXCYC res3c.f:6639-6735, FLUCYC res3c.f:7724-7830).
// Included at the end of ffhip.hip (it uses the context and launch helpers defined there).

// cell of REFLEC's loop indices (I over KX, J over KY) on the patch's face plane
// (turb3.f:5528-5530): I-faces (i, j, k) = (K, I, J), J-faces (J, K, I), K-faces (I, J, K)
__device__ __forceinline__ int bc_cell(const Blk &b, const BcPatchDev &q, int I, int J)
{
    if (q.dir == 0) return cid(b, q.kpl, I, J);
    if (q.dir == 1) return cid(b, J, q.kpl, I);
    return cid(b, I, J, q.kpl);
}

// REFLEC for one patch, thread per face: loop 1200 (the ghost temperature, VIS, CH, C,
// DRDP, DRDH), then loop 500 on viscous surfaces or loop 1150 on inviscid ones and
// singularities.  Each face reads only its first two interior cells and writes only its own
// ghost cells, so the loops fuse per face; patches run in order, one launch each.
__global__ void __launch_bounds__(256) k_reflec(Prm p, Blk b, BcPatchDev q)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q.n) return;
    const int na = q.kx2 - q.kx1 + 1;
    const int I = q.kx1 + t % na, J = q.ky1 + t / na;
    const int L = bc_cell(b, q, I, J);
    const int LB = L - q.idir * q.kstr, LT = L + q.idir * q.kstr, LC = L - 2 * q.idir * q.kstr;
    if (q.ibc == 3) {
        // an inlet: its pressures extrapolated into both ghost layers (turb3.f:5493-5509)
        double *P = FLD(b, FF_P), *PD = FLD(b, FF_PDIFF);
        const double pl = P[L], pdl = PD[L];
        P[LB] = pl; PD[LB] = pdl;
        P[LC] = pl; PD[LC] = pdl;
        return;
    }
    double *U = FLD(b, FF_U), *V = FLD(b, FF_V), *W = FLD(b, FF_W), *T = FLD(b, FF_TEMP),
           *RO = FLD(b, FF_RO), *P = FLD(b, FF_P), *PD = FLD(b, FF_PDIFF), *RM = FLD(b, FF_RM),
           *RN = FLD(b, FF_RN), *RW = FLD(b, FF_RW), *E = FLD(b, FF_E), *VIS = FLD(b, FF_VIS),
           *VIST = FLD(b, FF_VIST), *CH = FLD(b, FF_CH), *EPS2 = FLD(b, FF_EPS2), *C = FLD(b, FF_C),
           *DRDP = FLD(b, FF_DRDP), *DRDH = FLD(b, FF_DRDH), *RK = FLD(b, FF_RK), *REPS = FLD(b, FF_REPS);
    const bool kw = p.iturb >= 3 && p.iturb != 8;
    const int ihfl = q.iheat == 1 ? 1 : 0;
    const double tsurf = q.iheat == 1 ? q.tsurf : 0.;
    // loop 1200 (turb3.f:5511-5535)
    double tb = (double)(1 - ihfl) * (1. * T[L] - 0. * T[LT]) + (double)ihfl * (2. * tsurf - T[L]);
    tb = dmin(tb, p.tuplim);
    tb = dmax(tb, p.tlolim);
    T[LB] = tb;
    T[LC] = tb;
    VIS[LB] = VIS[L];
    CH[LB] = CH[L];
    C[LB] = C[L];
    DRDP[LB] = DRDP[L];
    DRDH[LB] = DRDH[L];
    if (q.ibc != 7 && q.viscous) {
        // loop 500 (turb3.f:5545-5655): REFLECL = .FALSE., ISTATE /= 10
        const int LN = 2;
        const int NP = (LN + J - q.ky1) * (na + 2 * LN) + (I - q.kx1 + LN);
        const double uw = q.uw ? q.uw[NP] : 0., vw = q.vw ? q.vw[NP] : 0., ww = q.ww ? q.ww[NP] : 0.;
        const double ub = -U[L] + 2. * uw, vb = -V[L] + 2. * vw, wb = -W[L] + 2. * ww;
        U[LB] = ub; V[LB] = vb; W[LB] = wb;
        U[LC] = ub; V[LC] = vb; W[LC] = wb;
        // MAX(a, b) as the reference writes it: the first argument unless the second is larger
        const double rob = dmax(0.001 * p.frsden, 1. * RO[L] - 0. * RO[LT]);
        const double pb = dmax(0.001 * p.frspre, 2. * P[L] - P[LT]);
        const double pdb = dmax(-0.999 * p.frspre, 2. * PD[L] - PD[LT]);
        RO[LB] = rob; P[LB] = pb; PD[LB] = pdb;
        EPS2[LB] = 1.;
        VIST[LB] = 0.;
        const double rmb = rob * ub, rnb = rob * vb, rwb = rob * wb, eb = E[L];
        RM[LB] = rmb; RN[LB] = rnb; RW[LB] = rwb; E[LB] = eb;
        RM[LC] = rmb; RN[LC] = rnb; RW[LC] = rwb; E[LC] = eb;
        EPS2[LC] = 1.;
        VIST[LC] = 0.;
        PD[LC] = pdb; P[LC] = pb; RO[LC] = rob;
        if (kw) {
            RK[LB] = -RK[L];
            if (p.iturb != 6) REPS[LB] = -REPS[L] + 2. * 0.;
            else REPS[LB] = dmax(0., 2. * REPS[L] - REPS[LT]);
        }
    } else {
        // loop 1150 (turb3.f:5659-5768)
        RO[LB] = dmax(0.001 * p.frsden, 1. * RO[L] - 0. * RO[LT]);
        P[LB] = dmax(0.001 * p.frspre, 2. * P[L] - P[LT]);
        PD[LB] = dmax(-0.999 * p.frspre, 2. * PD[L] - PD[LT]);
        // the face normal at L (low faces) or L + KSTR (high faces); as in the reference the
        // K-faces take the I-face normals (turb3.f:5687-5707)
        const int la = q.iface >= 4 ? L + q.kstr : L;
        const bool jf = q.iface == 2 || q.iface == 5;
        const double a11 = jf ? FLD(b, FF_A2X)[la] : FLD(b, FF_A1X)[la];
        const double a12 = jf ? FLD(b, FF_A2Y)[la] : FLD(b, FF_A1Y)[la];
        const double a13 = jf ? FLD(b, FF_A2Z)[la] : FLD(b, FF_A1Z)[la];
        double a21, a22, a23, a31, a32, a33;
        surfvects(a11, a12, a13, a21, a22, a23, a31, a32, a33);
        const double u = U[L], v = V[L], w = W[L];
        double un = a11 * u + a12 * v + a13 * w;
        const double vn = a21 * u + a22 * v + a23 * w;
        const double wn = a31 * u + a32 * v + a33 * w;
        un = -un;
        U[LB] = a11 * un + a21 * vn + a31 * wn;
        V[LB] = a12 * un + a22 * vn + a32 * wn;
        W[LB] = a13 * un + a23 * vn + a33 * wn;
        EPS2[LB] = 1. * EPS2[L] - 0. * EPS2[LT];
        VIST[LB] = 1. * VIST[L] - 0. * VIST[LT];
        if (kw) {
            RK[LB] = dmax(0., 1. * RK[L] - 0. * RK[LT]);
            REPS[LB] = dmax(0., 1. * REPS[L] - 0. * REPS[LT]);
        }
    }
}

// GHOSTEXT for one EXT patch, thread per ghost cell (both layers): the free stream without
// the atmosphere model, FRSSIE = -E0REF + RGAS/GM1*(FRSTEM-T0REF) with E0REF = T0REF = 0
__global__ void __launch_bounds__(256) k_ghostext(Prm p, Blk b, int i0, int ni, int j0, int nj,
                                                  int k0, int nk, double u10, double v10,
                                                  double w10, double frsmut)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni * nj * nk) return;
    const int i = i0 + t % ni, j = j0 + (t / ni) % nj, k = k0 + t / (ni * nj);
    const int kr = cid(b, i, j, k);
    const double gm1 = p.gamma - 1.0;
    const double frssie = -0. + p.rgas / gm1 * (p.frstem - 0.);
    FLD(b, FF_RO)[kr] = p.frsden;
    FLD(b, FF_P)[kr] = p.frspre;
    FLD(b, FF_PDIFF)[kr] = p.frspre - p.frspre;
    FLD(b, FF_RM)[kr] = p.frsden * u10;
    FLD(b, FF_RN)[kr] = p.frsden * v10;
    FLD(b, FF_RW)[kr] = p.frsden * w10;
    FLD(b, FF_U)[kr] = u10;
    FLD(b, FF_V)[kr] = v10;
    FLD(b, FF_W)[kr] = w10;
    double e = p.frsden * (frssie + .5 * (u10 * u10 + v10 * v10 + w10 * w10));
    FLD(b, FF_TEMP)[kr] = p.frstem;
    FLD(b, FF_VIS)[kr] = p.frsvis;
    if (p.iturb >= 3 && p.iturb != 8) {
        FLD(b, FF_RK)[kr] = p.rklim;
        FLD(b, FF_REPS)[kr] = p.epslim;
        FLD(b, FF_VIST)[kr] = frsmut;
        FLD(b, FF_EPS2)[kr] = 1. + frsmut / p.frsvis;
        e = e + p.rklim;
    }
    FLD(b, FF_E)[kr] = e;
}

// The patch arrays of an INL/OUT patch in BcPatchDev::bn, in the order of ff_bc_patch
enum { BN_R, BN_U, BN_V, BN_W, BN_T, BN_P, BN_PD, BN_RK, BN_EP, BN_N };

// INLET (ibc 3, types 4 and 5) or OUTLET (ibc 5, types 2-5) on one patch, thread per face
// cell in FLXCYC's frame: II the first interior cell (REFLEC's L), II2 the second, IM1 and
// IM2 the ghost cells, IF1 the boundary face, NP the patch-array entry with LN = 2.  One
// phase, ISTATE = 1 (EFPT: RGAS/GM1*T; EFPRO: (P + 0)/(GM1*RO) - E0REF with E0REF = 0),
// WROT = 0, no scalars, no transition, ISTRES = 0, IPRESC = 0, RELAX = RK1 = 1, RK2 = 0
// (base3c.f:9282-9284, 9650-9652).  Each face reads its own two interior cells and writes
// only its own ghost cells and patch entries.  The literals 1 and 0 of RK1, RK2 and WROT
// stay in the arithmetic: the reference's products by them fix the signs of zeros.
__global__ void __launch_bounds__(256) k_inlout(Prm p, Blk b, BcPatchDev q, const int *icyc,
                                                double turlim)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q.n) return;
    const int na = q.kx2 - q.kx1 + 1;
    const int I = q.kx1 + t % na, J = q.ky1 + t / na;
    const int ii = bc_cell(b, q, I, J);
    const int lstrid = q.kstr * q.idir;
    const int ii2 = ii + lstrid, im1 = ii - lstrid, im2 = im1 - lstrid;
    const int if1 = ii + (q.idir == 1 ? 0 : -lstrid);              // II + NFL
    const int np = (2 + J - q.ky1) * (na + 4) + (I - q.kx1 + 2);
    const int icycle = icyc ? *icyc : 0;
    const bool kw = p.iturb >= 3 && p.iturb != 8;
    const double *a3x = FLD(b, FF_A1X + 4 * q.dir), *a3y = FLD(b, FF_A1Y + 4 * q.dir),
                 *a3z = FLD(b, FF_A1Z + 4 * q.dir);
    double *P = FLD(b, FF_P), *PD = FLD(b, FF_PDIFF), *U = FLD(b, FF_U), *V = FLD(b, FF_V),
           *W = FLD(b, FF_W), *E = FLD(b, FF_E), *RO = FLD(b, FF_RO), *RM = FLD(b, FF_RM),
           *RN = FLD(b, FF_RN), *RW = FLD(b, FF_RW), *T = FLD(b, FF_TEMP), *EPS2 = FLD(b, FF_EPS2),
           *VIST = FLD(b, FF_VIST), *VIS = FLD(b, FF_VIS), *RK = FLD(b, FF_RK), *REPS = FLD(b, FF_REPS);
    double *const *bn = q.bn;
    const double gm1 = p.gamma - 1.;
    const double relax = 1.0, rk1 = 1., rk2 = 0., wrot = 0.;
    const double idir = (double)q.idir;
    if (q.ibc == 3) {
        const double br = bn[BN_R][np];
        EPS2[im1] = rk1 * EPS2[ii] + rk2 * EPS2[ii2];
        EPS2[im2] = EPS2[im1];
        VIST[im1] = rk1 * VIST[ii] + rk2 * VIST[ii2];
        VIST[im2] = VIST[im1];
        double pim1, pdim1;
        if (q.inlout == 4) {
            // subsonic inflow: the pressure extrapolated (base3c.f:9300-9425)
            pim1 = relax * (rk1 * P[ii] + rk2 * P[ii2]) + (1. - relax) * P[im1];
            pdim1 = relax * (rk1 * PD[ii] + rk2 * PD[ii2]) + (1. - relax) * PD[im1];
            bn[BN_P][np] = pim1;
            bn[BN_PD][np] = pdim1;
        } else {
            // supersonic inflow: every variable given (base3c.f:9427-9526)
            pim1 = bn[BN_P][np];
            pdim1 = bn[BN_PD][np];
        }
        double rm = bn[BN_U][np] * br, rn = bn[BN_V][np] * br, rw = bn[BN_W][np] * br;
        double ro1 = RO[im1];
        if (icycle == 1) ro1 = br;
        const double ro2 = RO[im2];
        const double tb = bn[BN_T][np];
        if (icycle >= 1) {
            const double roave = ro1;
            rm = rm + a3x[if1] * wrot * (roave - br);
            rn = rn + a3y[if1] * wrot * (roave - br);
            rw = rw + a3z[if1] * wrot * (roave - br);
        }
        const double ei = p.rgas / gm1 * tb;                          // EFPT
        // type 4 divides by the ghost densities, type 5 by BOUNR (base3c.f:9415-9420, 9515-9520)
        const double d1 = q.inlout == 4 ? ro1 : br, d2 = q.inlout == 4 ? ro2 : br;
        const double u1 = rm / d1, u2 = rm / d2, v1 = rn / d1, v2 = rn / d2, w1 = rw / d1, w2 = rw / d2;
        double e1 = ro1 * (ei + .5 * (u1 * u1 + v1 * v1 + w1 * w1));
        double e2 = ro2 * (ei + .5 * (u2 * u2 + v2 * v2 + w2 * w2));
        RO[im1] = ro1;
        RM[im1] = rm; RN[im1] = rn; RW[im1] = rw;
        RM[im2] = rm; RN[im2] = rn; RW[im2] = rw;
        T[im1] = tb; T[im2] = tb;
        P[im1] = pim1; P[im2] = pim1;
        PD[im1] = pdim1; PD[im2] = pdim1;
        U[im1] = u1; U[im2] = u2;
        V[im1] = v1; V[im2] = v2;
        W[im1] = w1; W[im2] = w2;
        if (kw) {
            // base3c.f:9535-9545
            const double rkb = bn[BN_RK][np];
            RK[im1] = rkb; RK[im2] = rkb;
            REPS[im1] = bn[BN_EP][np]; REPS[im2] = bn[BN_EP][np];
            e1 = e1 + rkb;
            e2 = e2 + rkb;
        }
        E[im1] = e1;
        E[im2] = e2;
        return;
    }
    // OUTLET
    const double a11 = a3x[if1], a12 = a3y[if1], a13 = a3z[if1];
    double a21, a22, a23, a31, a32, a33;
    surfvects(a11, a12, a13, a21, a22, a23, a31, a32, a33);
    double veloc;
    if (q.inlout == 2 || q.inlout == 3) {
        // pressure given, velocity extrapolated (base3c.f:9694-9892)
        const double rk3 = q.inlout == 2 ? 0.7 : 1., rk4 = q.inlout == 2 ? 0.3 : 0.;
        const double uim1 = bn[BN_U][np], vim1 = bn[BN_V][np], wim1 = bn[BN_W][np];
        veloc = .5 * (a11 * (U[ii] + uim1) + a12 * (V[ii] + vim1) + a13 * (W[ii] + wim1));
        const double ploss3 = 0., ploss4 = 0.;
        const double pb = bn[BN_P][np], pdb = bn[BN_PD][np];
        const double pg = rk1 * (pb - ploss3) + rk2 * (pb - ploss4);
        const double pdg = rk1 * (pdb - ploss3) + rk2 * (pdb - ploss4);
        P[im1] = pg; P[im2] = pg;
        PD[im1] = pdg; PD[im2] = pdg;
        double u1, v1, w1, ro1, ro2, tg;
        double e2v = rk1 * EPS2[ii] + rk2 * EPS2[ii2];
        double vtv = rk1 * VIST[ii] + rk2 * VIST[ii2];
        e2v = dmin(e2v, turlim);
        vtv = dmin(vtv, turlim * VIS[im2]);
        if (-veloc * idir > 0.) {
            ro1 = RO[im1];
            ro2 = RO[im2];
            u1 = (rk1 * RO[ii] * U[ii] + rk2 * RO[ii2] * U[ii2]) / ro1;
            v1 = rk1 * V[ii] + rk2 * V[ii2];
            w1 = rk1 * W[ii] + rk2 * W[ii2];
            tg = rk1 * T[ii] + rk2 * T[ii2];
        } else {
            const double ukm1 = rk1 * U[ii] + rk2 * U[ii2];
            const double vkm1 = rk1 * V[ii] + rk2 * V[ii2];
            const double wkm1 = rk1 * W[ii] + rk2 * W[ii2];
            const double velc1 = a11 * ukm1 + a12 * vkm1 + a13 * wkm1;
            const double ukp1 = bn[BN_U][np], vkp1 = bn[BN_V][np], wkp1 = bn[BN_W][np];
            const double velc2 = a21 * ukp1 + a22 * vkp1 + a23 * wkp1;
            const double velc3 = a31 * ukp1 + a32 * vkp1 + a33 * wkp1;
            const double uu = a11 * velc1 + a21 * velc2 + a31 * velc3;
            const double vv = a12 * velc1 + a22 * velc2 + a32 * velc3;
            const double ww = a13 * velc1 + a23 * velc2 + a33 * velc3;
            u1 = rk3 * U[im1] + rk4 * uu;
            v1 = rk3 * V[im1] + rk4 * vv;
            w1 = rk3 * W[im1] + rk4 * ww;
            tg = bn[BN_T][np];
            ro1 = bn[BN_R][np];
            ro2 = ro1;
            RO[im1] = ro1; RO[im2] = ro2;
        }
        U[im1] = u1; U[im2] = u1;
        V[im1] = v1; V[im2] = v1;
        W[im1] = w1; W[im2] = w1;
        EPS2[im1] = e2v; EPS2[im2] = e2v;
        VIST[im1] = vtv; VIST[im2] = vtv;
        T[im1] = tg; T[im2] = tg;
        RM[im1] = ro1 * u1; RM[im2] = ro2 * u1;
        RN[im1] = ro1 * v1; RN[im2] = ro2 * v1;
        RW[im1] = ro1 * w1; RW[im2] = ro2 * w1;
    } else if (q.inlout == 4) {
        // every variable extrapolated (base3c.f:9894-9968)
        veloc = a11 * U[ii] + a12 * V[ii] + a13 * W[ii];
        const double p1 = rk1 * P[ii] + rk2 * P[ii2], p2 = rk1 * p1 + rk2 * P[ii];
        const double pd1 = rk1 * PD[ii] + rk2 * PD[ii2], pd2 = rk1 * pd1 + rk2 * PD[ii];
        const double t1 = rk1 * T[ii] + rk2 * T[ii2], t2 = rk1 * t1 + rk2 * T[ii];
        double e2v = rk1 * EPS2[ii] + rk2 * EPS2[ii2];
        double vtv = rk1 * VIST[ii] + rk2 * VIST[ii2];
        e2v = dmin(e2v, turlim);
        vtv = dmin(vtv, turlim * VIS[im1]);
        const double ro1 = rk1 * RO[ii] + rk2 * RO[ii2], ro2 = rk1 * ro1 + rk2 * RO[ii];
        const double rm1 = rk1 * RM[ii] + rk2 * RM[ii2], rm2 = rk1 * rm1 + rk2 * RM[ii];
        const double rn1 = rk1 * RN[ii] + rk2 * RN[ii2], rn2 = rk1 * rn1 + rk2 * RN[ii];
        const double rw1 = rk1 * RW[ii] + rk2 * RW[ii2], rw2 = rk1 * rw1 + rk2 * RW[ii];
        const double u1 = rm1 / ro1, u2 = rm2 / ro2, v1 = rn1 / ro1, v2 = rn2 / ro2;
        const double w1 = rw1 / ro1, w2 = rw2 / ro2;
        const double ei1 = (p1 + 0.) / (gm1 * ro1) - 0.;                 // EFPRO
        const double ei2 = (p2 + 0.) / (gm1 * ro2) - 0.;
        P[im1] = p1; P[im2] = p2; PD[im1] = pd1; PD[im2] = pd2; T[im1] = t1; T[im2] = t2;
        EPS2[im1] = e2v; EPS2[im2] = e2v; VIST[im1] = vtv; VIST[im2] = vtv;
        RO[im1] = ro1; RO[im2] = ro2; RM[im1] = rm1; RM[im2] = rm2;
        RN[im1] = rn1; RN[im2] = rn2; RW[im1] = rw1; RW[im2] = rw2;
        U[im1] = u1; U[im2] = u2; V[im1] = v1; V[im2] = v2; W[im1] = w1; W[im2] = w2;
        E[im1] = ro1 * (ei1 + .5 * (u1 * u1 + v1 * v1 + w1 * w1));
        E[im2] = ro2 * (ei2 + .5 * (u2 * u2 + v2 * v2 + w2 * w2));
    } else {
        // type 5: the x-momentum given (base3c.f:9970-10017); the pressure set from BOUNP
        // first is overwritten by the extrapolation
        veloc = a11 * U[im2] + a12 * V[im2] + a13 * W[im2];
        double e2v = rk1 * EPS2[ii] + rk2 * EPS2[ii2];
        double vtv = rk1 * VIST[ii] + rk2 * VIST[ii2];
        e2v = dmin(e2v, turlim);
        vtv = dmin(vtv, turlim * VIS[ii]);
        const double p1 = rk1 * P[ii] + rk2 * P[ii2];
        const double pd1 = rk1 * PD[ii] + rk2 * PD[ii2];
        const double ro1 = rk1 * RO[ii] + rk2 * RO[ii2];
        const double rm1 = 2. * bn[BN_R][np] * bn[BN_U][np] - RM[ii];
        const double rm2 = 2. * rm1 - RM[ii];
        const double rn1 = rk1 * RN[ii] + rk2 * RN[ii2];
        const double rw1 = rk1 * RW[ii] + rk2 * RW[ii2];
        const double u1 = rm1 / ro1, u2 = rm2 / ro1, v1 = rn1 / ro1, w1 = rw1 / ro1;
        double e = (p1 + 0.) / (gm1 * ro1) - 0.;                          // EFPRO
        // CONSVE base3c.f:668-701 on the one cell
        const double rkin = (u1 * u1 + v1 * v1 + w1 * w1) * .5;
        const double epot = 0.;
        e = kw ? (e + rkin + epot) * ro1 + RK[im1] : (e + rkin + epot) * ro1;
        EPS2[im1] = e2v; EPS2[im2] = e2v; VIST[im1] = vtv; VIST[im2] = vtv;
        P[im1] = p1; P[im2] = p1; PD[im1] = pd1; PD[im2] = pd1;
        RO[im1] = ro1; RO[im2] = ro1;
        RM[im1] = rm1; RM[im2] = rm2; U[im1] = u1; U[im2] = u2;
        RN[im1] = rn1; RN[im2] = rn1; V[im1] = v1; V[im2] = v1;
        RW[im1] = rw1; RW[im2] = rw1; W[im1] = w1; W[im2] = w1;
        E[im1] = e; E[im2] = e;
    }
    if (kw && -veloc * idir > 0.) {
        // base3c.f:10135-10150
        const double rk = rk1 * RK[ii] + rk2 * RK[ii2];
        const double reps = rk1 * REPS[ii] + rk2 * REPS[ii2];
        RK[im1] = rk; RK[im2] = rk;
        REPS[im1] = reps; REPS[im2] = reps;
        E[im1] = E[im1] + rk;
        E[im2] = E[im2] + rk;
    }
}

// FLXCYC's update of the INL and OUT patches of level 1 (every direction; the patches'
// ghost cells are disjoint, and no flux reads another direction's, so one pass before the
// fluxes does what the reference's per-direction updates do)
static int inlout_impl(ff_ctx *c, BlockRec *r, int level)
{
    if (!r->bc_icyc) return FF_OK;
    c->lev = level;
    const double turlim = c->prm.iturb == 6 ? c->prm.turlim : 1.E-6;   // FLUCYC res3c.f:7750
    for (const BcPatchDev &q : r->bcs) {
        if (q.ibc != 3 && q.ibc != 5) continue;
        int rc = launch(c, "k_inlout", k_inlout, dim3((q.n + 255) / 256), dim3(256),
                        make_prm(&c->prm, level), r->dev, q, (const int *)r->bc_icyc, turlim);
        if (rc) return rc;
    }
    return FF_OK;
}

static int reflec_impl(ff_ctx *c, BlockRec *r, int level)
{
    c->lev = level;
    for (const BcPatchDev &q : r->bcs) {
        if ((q.ibc < 7 || q.ibc > 10) && q.ibc != 3) continue;
        int rc = launch(c, "k_reflec", k_reflec, dim3((q.n + 255) / 256), dim3(256),
                        make_prm(&c->prm, level), r->dev, q);
        if (rc) return rc;
    }
    return FF_OK;
}

// The patch arrays an INL/OUT patch of INLOUT type t must bring (bit i: BN_i; the k-omega
// inlets also RK and EP), from what INLET and OUTLET read (base3c.f:9300-9545, 9694-10017)
static unsigned inlout_needs(int ibc, int t, bool kw)
{
    const unsigned R = 1u << BN_R, U = 1u << BN_U, V = 1u << BN_V, W = 1u << BN_W, T = 1u << BN_T,
                   P = 1u << BN_P, PD = 1u << BN_PD, KE = (1u << BN_RK) | (1u << BN_EP);
    if (ibc == 3) return (t == 4 ? R | U | V | W | T : R | U | V | W | T | P | PD) | (kw ? KE : 0u);
    if (t == 2 || t == 3) return R | U | V | W | T | P | PD;
    if (t == 4) return 0u;
    return R | U | P | PD;
}

extern "C" int ff_bc_register(ff_ctx *c, int level, int block, const ff_bc_patch *p, int n)
{
    if (c) c->gen++;
    if (!c || n < 0 || (n > 0 && !p)) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    const Blk &b = r->dev;
    const bool kw = c->prm.iturb >= 3 && c->prm.iturb != 8;
    std::vector<BcPatchDev> bcs;
    size_t nw = 0;
    bool inout = false;
    for (int l = 0; l < n; l++) {
        const ff_bc_patch &h = p[l];
        const bool io = h.ibc == 3 || h.ibc == 5;
        if (h.ibc != 2 && !io && (h.ibc < 7 || h.ibc > 10))
            return fail(c, FF_E_UNSUP, "ff_bc_register: ICON(1) = %s (2, 3, 5, 7-10 supported)", std::to_string(h.ibc).c_str());
        if (h.iface < 1 || h.iface > 6) return fail(c, FF_E_ARG, "ff_bc_register: IFACE %s", std::to_string(h.iface).c_str());
        BcPatchDev q{};
        q.ibc = h.ibc;
        q.iface = h.iface;
        q.dir = (h.iface - 1) % 3;
        q.idir = h.iface >= 4 ? -1 : 1;
        const int nd = q.dir == 0 ? b.imax : (q.dir == 1 ? b.jmax : b.kmax);
        q.kpl = h.iface <= 3 ? 1 : nd;
        q.kx1 = h.kx1; q.kx2 = h.kx2; q.ky1 = h.ky1; q.ky2 = h.ky2;
        if (q.dir == 1) { q.ky1 = h.kx1; q.ky2 = h.kx2; q.kx1 = h.ky1; q.kx2 = h.ky2; }
        q.kstr = q.dir == 0 ? 1 : (q.dir == 1 ? b.is : b.il);
        // the in-plane extent of each index (the first two are those of ICON(4..7))
        const int ea = q.dir == 0 ? b.jmax : b.imax, eb = q.dir == 2 ? b.jmax : b.kmax;
        if (h.kx1 < 1 || h.kx2 > ea || h.kx1 > h.kx2 || h.ky1 < 1 || h.ky2 > eb || h.ky1 > h.ky2)
            return fail(c, FF_E_ARG, "ff_bc_register: patch %s outside its face", std::to_string(l).c_str());
        if (h.ibc != 2 && nd < 2)
            return fail(c, FF_E_UNSUP, "ff_bc_register: REFLEC, INLET and OUTLET need two cells normal to the face (patch %s)",
                        std::to_string(l).c_str());
        q.iheat = h.iheat;
        q.viscous = h.viscous;
        q.tsurf = h.tsurf;
        q.n = (q.kx2 - q.kx1 + 1) * (q.ky2 - q.ky1 + 1);
        const size_t sz = (size_t)(q.kx2 - q.kx1 + 5) * (q.ky2 - q.ky1 + 5);
        if (io) {
            if (c->prm.istres >= 1)
                return fail(c, FF_E_UNSUP, "ff_bc_register: INLET/OUTLET with the EARSM (ISTRES = 1)%s");
            if (c->prm.iturb != 0 && c->prm.iturb != 6)
                return fail(c, FF_E_UNSUP, "ff_bc_register: INLET/OUTLET with ITURB %s", std::to_string(c->prm.iturb).c_str());
            int t = h.inlout;
            if (t > 10 && t <= 20) t = t - 10;           // FLXCYC res3c.f:6647-6648
            if (t > 20 && t <= 30) t = t - 20;
            if ((h.ibc == 3 && t != 4 && t != 5) || (h.ibc == 5 && (t < 2 || t > 5)))
                return fail(c, FF_E_UNSUP, "ff_bc_register: INLOUT type %s (inlets 4, 5; outlets 2-5)",
                            std::to_string(h.inlout).c_str());
            q.inlout = t;
            const double *src[BN_N] = {h.bounr, h.bounu, h.bounv, h.bounw, h.bount, h.bounp,
                                       h.bounpd, h.bounrk, h.bounep};
            const unsigned need = inlout_needs(h.ibc, t, kw);
            for (int a = 0; a < BN_N; a++)
                if ((need >> a & 1u) && !src[a])
                    return fail(c, FF_E_ARG, "ff_bc_register: patch %s lacks a patch array its INLOUT type reads",
                                std::to_string(l).c_str());
            nw += BN_N * sz;
            inout = true;
        }
        if (h.uwall || h.vwall || h.wwall) nw += 3 * sz;
        bcs.push_back(q);
    }
    double *mem = nullptr;
    if (nw && hipMalloc(&mem, sizeof(double) * nw) != hipSuccess)
        return fail(c, FF_E_NOMEM, "boundary patch arrays%s");
    int *icyc = r->bc_icyc;
    if (inout && !icyc && hipMalloc(&icyc, 256) != hipSuccess) {
        if (mem) (void)hipFree(mem);
        return fail(c, FF_E_NOMEM, "boundary cycle count%s");
    }
    auto put = [&](double *dst, const double *src, size_t sz) -> bool {
        return (src ? hipMemcpyAsync(dst, src, sizeof(double) * sz, hipMemcpyHostToDevice, c->stream)
                    : hipMemsetAsync(dst, 0, sizeof(double) * sz, c->stream)) == hipSuccess;
    };
    size_t o = 0;
    bool ok = true;
    for (int l = 0; l < n && ok; l++) {
        const ff_bc_patch &h = p[l];
        BcPatchDev &q = bcs[l];
        const size_t sz = (size_t)(q.kx2 - q.kx1 + 5) * (q.ky2 - q.ky1 + 5);
        if (h.uwall || h.vwall || h.wwall) {
            const double *src[3] = {h.uwall, h.vwall, h.wwall};
            const double **dst[3] = {&q.uw, &q.vw, &q.ww};
            for (int a = 0; a < 3; a++) {
                if (src[a]) {
                    ok = ok && put(mem + o, src[a], sz);
                    *dst[a] = mem + o;
                }
                o += sz;
            }
        }
        if (h.ibc == 3 || h.ibc == 5) {
            const double *src[BN_N] = {h.bounr, h.bounu, h.bounv, h.bounw, h.bount, h.bounp,
                                       h.bounpd, h.bounrk, h.bounep};
            for (int a = 0; a < BN_N; a++) {
                ok = ok && put(mem + o, src[a], sz);
                q.bn[a] = mem + o;
                o += sz;
            }
        }
    }
    // ICYCLE = 0: the first cycle after registration runs as ICYCLE = 1
    if (ok && inout) ok = hipMemsetAsync(icyc, 0, sizeof(int), c->stream) == hipSuccess;
    if (!ok || hipStreamSynchronize(c->stream) != hipSuccess) {
        if (mem) (void)hipFree(mem);
        if (icyc && icyc != r->bc_icyc) (void)hipFree(icyc);
        return fail(c, FF_E_HIP, "boundary patch arrays: copy failed%s");
    }
    if (r->bc_mem) (void)hipFree(r->bc_mem);
    r->bc_mem = mem;
    if (!inout && icyc) {
        (void)hipFree(icyc);
        icyc = nullptr;
    }
    r->bc_icyc = icyc;
    r->bcs = bcs;
    return FF_OK;
}

extern "C" int ff_reflec(ff_ctx *c, int level, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    return reflec_impl(c, r, level);
}

extern "C" int ff_inlout(ff_ctx *c, int level, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    return inlout_impl(c, r, level);
}

extern "C" int ff_bc_cycle(ff_ctx *c, int level, int block, int icycle)
{
    if (!c || icycle < 0) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    if (!r->bc_icyc) return fail(c, FF_E_STATE, "ff_bc_cycle: no INL/OUT patches registered%s");
    HIPCHK(c, hipMemcpyAsync(r->bc_icyc, &icycle, sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

extern "C" int ff_bc_read(ff_ctx *c, int level, int block, int patch, double *bounp, double *bounpd)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    if (patch < 0 || patch >= (int)r->bcs.size()) return FF_E_ARG;
    const BcPatchDev &q = r->bcs[(size_t)patch];
    if (q.ibc != 3 && q.ibc != 5) return fail(c, FF_E_ARG, "ff_bc_read: not an INL/OUT patch%s");
    const size_t sz = (size_t)(q.kx2 - q.kx1 + 5) * (q.ky2 - q.ky1 + 5);
    if (bounp) HIPCHK(c, hipMemcpyAsync(bounp, q.bn[BN_P], sizeof(double) * sz, hipMemcpyDeviceToHost, c->stream));
    if (bounpd) HIPCHK(c, hipMemcpyAsync(bounpd, q.bn[BN_PD], sizeof(double) * sz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

extern "C" int ff_ghostext(ff_ctx *c, int level, int block, double alpha, double beta)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    const Blk &b = r->dev;
    c->lev = level;
    const double fv = c->prm.frsvel;
    const double u10 = fv * cos(alpha) * cos(beta), v10 = fv * sin(alpha), w10 = fv * cos(alpha) * sin(beta);
    for (const BcPatchDev &q : r->bcs) {
        if (q.ibc != 2) continue;
        // the ghost range of GHOSTEXT (ns3c.f:11062-11104) from ICON(4..7) = IA, IM, JA, JM
        const int ia = q.dir == 1 ? q.ky1 : q.kx1, im = q.dir == 1 ? q.ky2 : q.kx2;
        const int ja = q.dir == 1 ? q.kx1 : q.ky1, jm = q.dir == 1 ? q.kx2 : q.ky2;
        int i0, i1, j0, j1, k0, k1;
        switch (q.iface) {
        case 1: i0 = -1; i1 = 0; j0 = ia; j1 = im; k0 = ja; k1 = jm; break;
        case 2: i0 = ia; i1 = im; j0 = -1; j1 = 0; k0 = ja; k1 = jm; break;
        case 3: i0 = ia; i1 = im; j0 = ja; j1 = jm; k0 = -1; k1 = 0; break;
        case 4: i0 = b.imax + 1; i1 = b.imax + 2; j0 = ia; j1 = im; k0 = ja; k1 = jm; break;
        case 5: i0 = ia; i1 = im; j0 = b.jmax + 1; j1 = b.jmax + 2; k0 = ja; k1 = jm; break;
        default: i0 = ia; i1 = im; j0 = ja; j1 = jm; k0 = b.kmax + 1; k1 = b.kmax + 2; break;
        }
        const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
        const int nt = ni * nj * nk;
        int rc = launch(c, "k_ghostext", k_ghostext, dim3((nt + 255) / 256), dim3(256),
                        make_prm(&c->prm, level), b, i0, ni, j0, nj, k0, nk, u10, v10, w10,
                        c->prm.frsmut);
        if (rc) return rc;
    }
    return FF_OK;
}
