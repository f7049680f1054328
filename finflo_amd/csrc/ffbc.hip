// ffbc.hip — the boundary routines of the path on the device: REFLEC (turb3.f:5294-5878),
// which every STATIM of CYCLES ends with (ns3c.f:10497-10512) and CYCLES calls once more
// at the end of a cycle (ns3c.f:2168), and GHOSTEXT (ns3c.f:10979-11226), the free-stream
// ghost cells of EXT patches set at start-up (ns3c.f:294).
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

static int reflec_impl(ff_ctx *c, BlockRec *r, int level)
{
    c->lev = level;
    for (const BcPatchDev &q : r->bcs) {
        if (q.ibc < 7 || q.ibc > 10) continue;
        int rc = launch(c, "k_reflec", k_reflec, dim3((q.n + 255) / 256), dim3(256),
                        make_prm(&c->prm, level), r->dev, q);
        if (rc) return rc;
    }
    return FF_OK;
}

extern "C" int ff_bc_register(ff_ctx *c, int level, int block, const ff_bc_patch *p, int n)
{
    if (c) c->gen++;
    if (!c || n < 0 || (n > 0 && !p)) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    const Blk &b = r->dev;
    std::vector<BcPatchDev> bcs;
    size_t nw = 0;
    for (int l = 0; l < n; l++) {
        const ff_bc_patch &h = p[l];
        if (h.ibc != 2 && (h.ibc < 7 || h.ibc > 10))
            return fail(c, FF_E_UNSUP, "ff_bc_register: ICON(1) = %s (2, 7-10 supported)", std::to_string(h.ibc).c_str());
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
            return fail(c, FF_E_UNSUP, "ff_bc_register: REFLEC needs two cells normal to the face (patch %s)",
                        std::to_string(l).c_str());
        q.iheat = h.iheat;
        q.viscous = h.viscous;
        q.tsurf = h.tsurf;
        q.n = (q.kx2 - q.kx1 + 1) * (q.ky2 - q.ky1 + 1);
        if (h.uwall || h.vwall || h.wwall) nw += 3 * (size_t)(q.kx2 - q.kx1 + 5) * (q.ky2 - q.ky1 + 5);
        bcs.push_back(q);
    }
    double *mem = nullptr;
    if (nw && hipMalloc(&mem, sizeof(double) * nw) != hipSuccess)
        return fail(c, FF_E_NOMEM, "boundary patch arrays%s");
    size_t o = 0;
    for (int l = 0; l < n; l++) {
        const ff_bc_patch &h = p[l];
        BcPatchDev &q = bcs[l];
        if (!(h.uwall || h.vwall || h.wwall)) continue;
        const size_t sz = (size_t)(q.kx2 - q.kx1 + 5) * (q.ky2 - q.ky1 + 5);
        const double *src[3] = {h.uwall, h.vwall, h.wwall};
        const double **dst[3] = {&q.uw, &q.vw, &q.ww};
        for (int a = 0; a < 3; a++) {
            if (src[a]) {
                if (hipMemcpyAsync(mem + o, src[a], sizeof(double) * sz, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
                    (void)hipFree(mem);
                    return fail(c, FF_E_HIP, "boundary patch arrays: copy failed%s");
                }
                *dst[a] = mem + o;
            }
            o += sz;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (r->bc_mem) (void)hipFree(r->bc_mem);
    r->bc_mem = mem;
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
