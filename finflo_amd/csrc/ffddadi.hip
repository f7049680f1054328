// ffddadi.hip — the DDADI implicit stage (BLKS%LUSGS = 0, IMPS imps3.f:321-373) on the
// device.  Included at the end of ffhip.hip (it uses the context and launch helpers
// defined there).
//
// DDADI replaces the LU-SGS sweeps by three direction-split sweeps, K, I and J, each a
// backward pass PRED (imps4.f:788-905; k-omega PREDKE imps4.f:2084-2229), a pointwise
// diagonal DIAG (imps4.f:911-971; DIAGKE imps4.f:2235-2309) and a forward pass CORR
// (imps4.f:671-787; CORRKE imps4.f:1939-2078).  Each pass is a recurrence along one grid
// line only, so the lines are independent: a thread owns one line and marches along it.
// The reference sorts its arrays so that the swept direction runs slowest (SORT38, SORTZZ
// of multi.f: IJK for K, JKI for I, KIJ for J) and loops over the entries of its sorted
// (a, b) plane from entry IN+1 of the first interior row to the end of the last (DO I =
// IN+1, JMAX*ISTRID: ghost entries of the fast index a included, every 2*IN+1-th when the
// fast dimension is 1).  The kernels address the unsorted IJK arrays through the strides of
// that order and take the same plane entries, so every line does the reference's
// operations on the reference's operands.  IDI = all cells viscous (CDIFF = 2), UROT = 0,
// DTURB = 1.

struct DdDir {
    int sa, sb, sm;               // strides of the sorted fast, middle and swept indices
    int na, nb, nm;               // their interior sizes
    const double *A, *AX, *AY, *AZ;   // the face arrays of the swept direction
};

static DdDir dd_dir(const Blk &b, int dir)
{
    DdDir r;
    if (dir == 0) {          // K sweep: IJK
        r.sa = 1; r.na = b.imax; r.sb = b.is; r.nb = b.jmax; r.sm = b.il; r.nm = b.kmax;
        r.A = b.f[FF_A3]; r.AX = b.f[FF_A3X]; r.AY = b.f[FF_A3Y]; r.AZ = b.f[FF_A3Z];
    } else if (dir == 1) {   // I sweep: JKI
        r.sa = b.is; r.na = b.jmax; r.sb = b.il; r.nb = b.kmax; r.sm = 1; r.nm = b.imax;
        r.A = b.f[FF_A1]; r.AX = b.f[FF_A1X]; r.AY = b.f[FF_A1Y]; r.AZ = b.f[FF_A1Z];
    } else {                 // J sweep: KIJ
        r.sa = b.il; r.na = b.kmax; r.sb = 1; r.nb = b.imax; r.sm = b.is; r.nm = b.jmax;
        r.A = b.f[FF_A2]; r.AX = b.f[FF_A2X]; r.AY = b.f[FF_A2Y]; r.AZ = b.f[FF_A2Z];
    }
    return r;
}

// the number of plane entries (lines) of a direction
static int dd_lines(const DdDir &r)
{
    const int ist = r.na + 2 * FF_GHOST, step = r.na == 1 ? 2 * FF_GHOST + 1 : 1;
    return (r.nb * ist - FF_GHOST + step - 1) / step;
}

// SYIKE imps4.f:982-1035 (its DRH and DRP are formed but unused there)
__device__ __forceinline__ void syike(double w[7], const double x[7], double c, double ro,
                                      double a2x, double a2y, double a2z, double &a21,
                                      double &a22, double &a23, double &a31, double &a32,
                                      double &a33)
{
    const double z1 = x[0], z5 = x[4], z6 = x[5], z7 = x[6];
    surfvects(a2x, a2y, a2z, a21, a22, a23, a31, a32, a33);
    const double z2 = a2x * x[1] + a2y * x[2] + a2z * x[3];
    const double z3 = a21 * x[1] + a22 * x[2] + a23 * x[3];
    const double z4 = a31 * x[1] + a32 * x[2] + a33 * x[3];
    const double roc = ro * c;
    const double pc2 = 1. / ro;
    const double dpres = z5;
    w[0] = z1 - pc2 * dpres;
    w[1] = roc * z2 + dpres;
    w[2] = z3;
    w[3] = z4;
    w[5] = z6;
    w[6] = z7;
    w[4] = -roc * z2 + dpres;
}

// SYM1KE imps4.f:1041-1072
__device__ __forceinline__ void sym1ke(double x[7], const double w[7], double rk, double c,
                                       double ro, double drdp, double drdh, double a2x,
                                       double a2y, double a2z, double a21, double a22,
                                       double a23, double a31, double a32, double a33)
{
    const double ypro = 1. / ro;
    const double pc2 = .5 * ypro;
    const double proc = pc2 / c;
    const double drh = .667 * rk * (ypro * ypro) * drdh;
    const double ydrp = 1. / (1. + .667 * rk * ypro * drdp);
    const double dh = w[0] + pc2 * (w[1] + w[4]);
    x[0] = dh;
    const double x2i = proc * (w[1] - w[4]);
    const double x3i = w[2], x4i = w[3];
    const double x5 = .5 * (w[1] + w[4]);
    x[4] = ydrp * (x5 - drh * ro * dh - .667 * ro * w[5]);
    x[5] = w[5];
    x[6] = w[6];
    x[1] = a2x * x2i + a21 * x3i + a31 * x4i;
    x[2] = a2y * x2i + a22 * x3i + a32 * x4i;
    x[3] = a2z * x2i + a23 * x3i + a33 * x4i;
}

// One DDADI pass along every line of one direction: PH 0 PRED (backward, the upper
// neighbour's flux entering through the upper face), 1 DIAG, 2 CORR (forward, the lower
// neighbour's flux through the lower face).  Thread t owns plane entry f = IN + t*step.
template <bool KE, int PH>
__global__ void __launch_bounds__(128) k_ddadi(Prm p, Blk b, DdDir r, int nlines)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlines) return;
    constexpr int NV = KE ? 7 : 5;
    const int ist = r.na + 2 * FF_GHOST, step = r.na == 1 ? 2 * FF_GHOST + 1 : 1;
    const int f = FF_GHOST + t * step;
    const int base = (f % ist) * r.sa + (f / ist + FF_GHOST) * r.sb;
    double *D[7] = {FLD(b, FF_DRO), FLD(b, FF_DM), FLD(b, FF_DN), FLD(b, FF_DW), FLD(b, FF_DE),
                    FLD(b, FF_DRK), FLD(b, FF_DEPS)};
    const double *dtl = FLD(b, FF_DTL), *vol = FLD(b, FF_VOL), *u = FLD(b, FF_U),
                 *v = FLD(b, FF_V), *w = FLD(b, FF_W), *c = FLD(b, FF_C), *ro = FLD(b, FF_RO),
                 *vis = FLD(b, FF_VIS), *vist = FLD(b, FF_VIST), *rk = FLD(b, FF_RK),
                 *drdp = FLD(b, FF_DRDP), *drdh = FLD(b, FF_DRDH);
    const double cdiff = 2., dturb = 1., vrel = 1.E-21;
    double bd[7] = {0., 0., 0., 0., 0., 0., 0.};
    for (int s = 0; s < r.nm; s++) {
        const int m = PH == 0 ? r.nm - s : s + 1;
        const int n = base + (m + FF_GHOST - 1) * r.sm, nu = n + r.sm, nl = n - r.sm;
        // the face whose normal the characteristic variables use: the lower face in PRED,
        // the upper face in DIAG and CORR
        const int nf = PH == 0 ? n : nu;
        const double fx = r.AX[nf], fy = r.AY[nf], fz = r.AZ[nf];
        double x[7], y[7], a21, a22, a23, a31, a32, a33;
        const double dt1 = p.theta * dtl[n];
        if (PH == 1) {
#pragma unroll
            for (int q = 0; q < NV; q++) x[q] = D[q][n];
        } else {
            // the flux carried along the line enters through the upper face in PRED, the
            // lower one in CORR
            const double af = PH == 0 ? r.A[nu] : r.A[n];
#pragma unroll
            for (int q = 0; q < 5; q++) x[q] = dt1 * af * bd[q] + vol[n] * D[q][n];
            if (KE) {
#pragma unroll
                for (int q = 5; q < 7; q++) x[q] = dturb * dt1 * af * bd[q] + vol[n] * D[q][n];
            }
        }
        if (KE)
            syike(y, x, c[n], ro[n], fx, fy, fz, a21, a22, a23, a31, a32, a33);
        else
            syi(y[0], y[1], y[2], y[3], y[4], x[0], x[1], x[2], x[3], x[4], c[n], ro[n], fx, fy,
                fz, a21, a22, a23, a31, a32, a33);
        double xs[7];
        if (PH == 1) {
            double ucon = r.AX[nu] * u[n] + r.AY[nu] * v[n] + r.AZ[nu] * w[n] - 0.;
            const double rl2n = dmax(-(ucon + c[n]), 0.), rl5n = dmax(-(ucon - c[n]), 0.);
            const double rl1n = dmax(-ucon, 0.);
            ucon = r.AX[n] * u[n] + r.AY[n] * v[n] + r.AZ[n] * w[n] - 0.;
            const double rl2p = dmax(ucon + c[n], 0.), rl5p = dmax(ucon - c[n], 0.);
            const double rl1p = dmax(ucon, 0.);
            const double dt = KE ? dt1 / (vol[n] + 1.E-21) : dt1 / vol[n];
            const double s1 = r.A[n] * rl1p + r.A[nu] * rl1n, s2 = r.A[n] * rl2p + r.A[nu] * rl2n;
            const double s5 = r.A[n] * rl5p + r.A[nu] * rl5n;
            y[0] = y[0] * (1. + dt * s1);
            y[1] = y[1] * (1. + dt * s2);
            y[2] = y[2] * (1. + dt * s1);
            y[3] = y[3] * (1. + dt * s1);
            y[4] = y[4] * (1. + dt * s5);
            if (KE) {
                y[5] = y[5] * (1. + dturb * dt * (s1 + 0.));
                y[6] = y[6] * (1. + dturb * dt * (s1 + 0.));
            }
        } else {
            const double rl1 = fx * u[n] + fy * v[n] + fz * w[n] - 0.;
            const double rl2 = rl1 + c[n], rl5 = rl1 - c[n];
            const int nn = PH == 0 ? nl : nu;           // the line neighbour of the face
            const double ff = PH == 0 ? (m == 1 ? .25 : .5) : .5;
            const double rnum = ff * (vol[n] + vol[nn]) * (ro[n] + ro[nn]);
            const double af = PH == 0 ? r.A[n] : r.A[nu];
            const double tt = cdiff * af / rnum * (vist[n] + vis[n] + vist[nn] + vis[nn]);
            double x1, x2, x5;
            if (PH == 0) {
                x1 = dmax(-rl1, 0.) + tt; x2 = dmax(-rl2, 0.) + tt; x5 = dmax(-rl5, 0.) + tt;
            } else {
                x1 = dmax(rl1, 0.) + tt; x2 = dmax(rl2, 0.) + tt; x5 = dmax(rl5, 0.) + tt;
            }
            const double x1a = fabs(rl1) + tt, x2a = fabs(rl2) + tt, x5a = fabs(rl5) + tt;
            const double xa[5] = {x1a, x2a, x1a, x1a, x5a};
            if (KE) {
#pragma unroll
                for (int q = 0; q < 5; q++) y[q] = y[q] / (vol[n] + vrel + dt1 * xa[q] * af);
#pragma unroll
                for (int q = 5; q < 7; q++) y[q] = y[q] / (vol[n] + vrel + dturb * dt1 * (x1a * af + 0.));
            } else {
#pragma unroll
                for (int q = 0; q < 5; q++) y[q] = y[q] / (vol[n] + dt1 * xa[q] * af);
            }
            xs[0] = x1; xs[1] = x2; xs[2] = x1; xs[3] = x1; xs[4] = x5; xs[5] = x1; xs[6] = x1;
        }
        // the new increments of the cell
        if (KE) {
            double o[7];
            sym1ke(o, y, rk[n], c[n], ro[n], drdp[n], drdh[n], fx, fy, fz, a21, a22, a23, a31,
                   a32, a33);
#pragma unroll
            for (int q = 0; q < 7; q++) D[q][n] = o[q];
        } else {
            double o0, o1, o2, o3, o4;
            sym1i(o0, o1, o2, o3, o4, y[0], y[1], y[2], y[3], y[4], c[n], ro[n], fx, fy, fz, a21,
                  a22, a23, a31, a32, a33);
            D[0][n] = o0; D[1][n] = o1; D[2][n] = o2; D[3][n] = o3; D[4][n] = o4;
        }
        if (PH == 1) continue;
        // the flux handed to the next cell of the line
#pragma unroll
        for (int q = 0; q < NV; q++) y[q] = xs[q] * y[q];
        if (KE)
            sym1ke(bd, y, rk[n], c[n], ro[n], drdp[n], drdh[n], fx, fy, fz, a21, a22, a23, a31, a32,
                   a33);
        else
            sym1i(bd[0], bd[1], bd[2], bd[3], bd[4], y[0], y[1], y[2], y[3], y[4], c[n], ro[n], fx,
                  fy, fz, a21, a22, a23, a31, a32, a33);
    }
}

// the end of IMPS on the interior cells: dh -> dT (DIVV12 imps3.f:483) and, with the
// k-omega model, D(rho k), D(rho omega) times RO (MULV12 imps3.f:497-500)
template <bool KE>
__global__ void __launch_bounds__(256) k_ddadi_end(Blk b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.imax) return;
    const int n = cid(b, i, j, k);
    FLD(b, FF_DRO)[n] = FLD(b, FF_DRO)[n] / FLD(b, FF_CP)[n];
    if (KE) {
        FLD(b, FF_DRK)[n] = FLD(b, FF_DRK)[n] * FLD(b, FF_RO)[n];
        FLD(b, FF_DEPS)[n] = FLD(b, FF_DEPS)[n] * FLD(b, FF_RO)[n];
    }
}

// the three DDADI directions in IMPS's order, K, I, J, each PRED, DIAG, CORR, then the end
// of IMPS
static int ddadi_impl(ff_ctx *c, const Prm &kp, const Blk &b)
{
    const bool ke = c->prm.iturb == 6;
    for (int dir = 0; dir < 3; dir++) {
        const DdDir r = dd_dir(b, dir);
        const int nl = dd_lines(r);
        const dim3 g((unsigned)((nl + 127) / 128)), t(128);
        int rc;
        if (ke) {
            if ((rc = launch(c, "k_ddadi", k_ddadi<true, 0>, g, t, kp, b, r, nl)) ||
                (rc = launch(c, "k_ddadi", k_ddadi<true, 1>, g, t, kp, b, r, nl)) ||
                (rc = launch(c, "k_ddadi", k_ddadi<true, 2>, g, t, kp, b, r, nl)))
                return rc;
        } else {
            if ((rc = launch(c, "k_ddadi", k_ddadi<false, 0>, g, t, kp, b, r, nl)) ||
                (rc = launch(c, "k_ddadi", k_ddadi<false, 1>, g, t, kp, b, r, nl)) ||
                (rc = launch(c, "k_ddadi", k_ddadi<false, 2>, g, t, kp, b, r, nl)))
                return rc;
        }
    }
    const dim3 ge((unsigned)((b.imax + 255) / 256), b.jmax, b.kmax);
    return ke ? launch(c, "k_ddadi_end", k_ddadi_end<true>, ge, dim3(256), b)
              : launch(c, "k_ddadi_end", k_ddadi_end<false>, ge, dim3(256), b);
}
