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

// The sweep's line of thread t, as the flat offset of its cell m = 1 minus one step: the
// reference's plane entry f = IN + t*step on the K and I sweeps (consecutive threads take
// neighbouring I columns on the K sweep; on the I sweep each thread walks its own
// contiguous row); on the J sweep the entries are renumbered with the middle index b (= I)
// fastest so that consecutive threads again take neighbouring I columns.  The set of
// entries is the reference's: f = a + b*(NA+4) >= IN, every 2*IN+1-th with NA = 1.
__device__ __forceinline__ bool dd_line(const DdDir &r, int dir, int t, int &base)
{
    const int ist = r.na + 2 * FF_GHOST;
    int a0, b0;
    if (dir == 2) {
        b0 = t % r.nb;
        a0 = t / r.nb;
        if (a0 >= ist || (b0 == 0 && a0 < FF_GHOST) || (r.na == 1 && a0 != FF_GHOST)) return false;
    } else {
        const int step = r.na == 1 ? 2 * FF_GHOST + 1 : 1;
        const int f = FF_GHOST + t * step;
        if (f >= r.nb * ist) return false;
        a0 = f % ist;
        b0 = f / ist;
    }
    base = a0 * r.sa + (b0 + FF_GHOST) * r.sb + (FF_GHOST - 1) * r.sm;
    return true;
}

// threads of a sweep launch
static int dd_threads(const DdDir &r, int dir)
{
    return dir == 2 ? (r.na + 2 * FF_GHOST) * r.nb : dd_lines(r);
}

// What a PRED (FWD = false) or CORR (FWD = true) step reads of its cell: the neighbour nn
// across the face whose viscous term T and diagonal it forms (the lower neighbour in PRED,
// the upper one in CORR), the area of the face the carried flux crosses (PRED: the upper
// face, CORR: the lower face) and of the diagonal's face, and the normal of the face whose
// characteristic variables it uses (PRED: the lower face, CORR: the upper face).
struct DdIn {
    double dt1, vol, voln, u, v, w, c, ro, ron, vis, vist, visn, vistn;
    double aflux, adiag, fx, fy, fz;
    double d[7], rk, drdp, drdh;
};

template <bool KE, bool FWD>
__device__ __forceinline__ void dd_load(const Prm &p, const Blk &b, const DdDir &r, int n, DdIn &x)
{
    const int nu = n + r.sm, nn = FWD ? nu : n - r.sm, nf = FWD ? nu : n;
    x.dt1 = p.theta * FLD(b, FF_DTL)[n];
    x.vol = FLD(b, FF_VOL)[n];
    x.voln = FLD(b, FF_VOL)[nn];
    x.u = FLD(b, FF_U)[n];
    x.v = FLD(b, FF_V)[n];
    x.w = FLD(b, FF_W)[n];
    x.c = FLD(b, FF_C)[n];
    x.ro = FLD(b, FF_RO)[n];
    x.ron = FLD(b, FF_RO)[nn];
    x.vis = FLD(b, FF_VIS)[n];
    x.vist = FLD(b, FF_VIST)[n];
    x.visn = FLD(b, FF_VIS)[nn];
    x.vistn = FLD(b, FF_VIST)[nn];
    x.aflux = FWD ? r.A[n] : r.A[nu];
    x.adiag = FWD ? r.A[nu] : r.A[n];
    x.fx = r.AX[nf];
    x.fy = r.AY[nf];
    x.fz = r.AZ[nf];
#pragma unroll
    for (int q = 0; q < (KE ? 7 : 5); q++) x.d[q] = FLD(b, FF_DRO + q)[n];
    if (KE) {
        x.rk = FLD(b, FF_RK)[n];
        x.drdp = FLD(b, FF_DRDP)[n];
        x.drdh = FLD(b, FF_DRDH)[n];
    }
}

// PRED (FWD = false: backward along the line) or CORR (FWD = true: forward) along every line
// of one direction.  The reads of the next cell are issued before the current cell's
// arithmetic (they do not depend on the recurrence), so their latency overlaps it.
template <bool KE, bool FWD>
__global__ void __launch_bounds__(128) k_ddadi_sweep(Prm p, Blk b, DdDir r, int dir, int nthreads)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int base;
    if (t >= nthreads || !dd_line(r, dir, t, base)) return;
    constexpr int NV = KE ? 7 : 5;
    double *D[7] = {FLD(b, FF_DRO), FLD(b, FF_DM), FLD(b, FF_DN), FLD(b, FF_DW), FLD(b, FF_DE),
                    FLD(b, FF_DRK), FLD(b, FF_DEPS)};
    const double cdiff = 2., dturb = 1., vrel = 1.E-21;
    double bd[7] = {0., 0., 0., 0., 0., 0., 0.};
    int m = FWD ? 1 : r.nm;
    DdIn cur, nxt;
    dd_load<KE, FWD>(p, b, r, base + m * r.sm, cur);
    for (int s = 0; s < r.nm; s++) {
        const int n = base + m * r.sm;
        const int mn = FWD ? m + 1 : m - 1;
        if (s + 1 < r.nm) dd_load<KE, FWD>(p, b, r, base + mn * r.sm, nxt);
        double x[7], y[7], a21, a22, a23, a31, a32, a33;
#pragma unroll
        for (int q = 0; q < 5; q++) x[q] = cur.dt1 * cur.aflux * bd[q] + cur.vol * cur.d[q];
        if (KE) {
#pragma unroll
            for (int q = 5; q < 7; q++) x[q] = dturb * cur.dt1 * cur.aflux * bd[q] + cur.vol * cur.d[q];
            syike(y, x, cur.c, cur.ro, cur.fx, cur.fy, cur.fz, a21, a22, a23, a31, a32, a33);
        } else {
            syi(y[0], y[1], y[2], y[3], y[4], x[0], x[1], x[2], x[3], x[4], cur.c, cur.ro, cur.fx,
                cur.fy, cur.fz, a21, a22, a23, a31, a32, a33);
        }
        const double rl1 = cur.fx * cur.u + cur.fy * cur.v + cur.fz * cur.w - 0.;
        const double rl2 = rl1 + cur.c, rl5 = rl1 - cur.c;
        const double ff = FWD ? .5 : (m == 1 ? .25 : .5);
        const double rnum = ff * (cur.vol + cur.voln) * (cur.ro + cur.ron);
        const double tt = cdiff * cur.adiag / rnum * (cur.vist + cur.vis + cur.vistn + cur.visn);
        double x1, x2, x5;
        if (FWD) {
            x1 = dmax(rl1, 0.) + tt; x2 = dmax(rl2, 0.) + tt; x5 = dmax(rl5, 0.) + tt;
        } else {
            x1 = dmax(-rl1, 0.) + tt; x2 = dmax(-rl2, 0.) + tt; x5 = dmax(-rl5, 0.) + tt;
        }
        const double x1a = fabs(rl1) + tt, x2a = fabs(rl2) + tt, x5a = fabs(rl5) + tt;
        const double xa[5] = {x1a, x2a, x1a, x1a, x5a};
        if (KE) {
#pragma unroll
            for (int q = 0; q < 5; q++) y[q] = y[q] / (cur.vol + vrel + cur.dt1 * xa[q] * cur.adiag);
#pragma unroll
            for (int q = 5; q < 7; q++) y[q] = y[q] / (cur.vol + vrel + dturb * cur.dt1 * (x1a * cur.adiag + 0.));
            double o[7];
            sym1ke(o, y, cur.rk, cur.c, cur.ro, cur.drdp, cur.drdh, cur.fx, cur.fy, cur.fz, a21, a22,
                   a23, a31, a32, a33);
#pragma unroll
            for (int q = 0; q < 7; q++) D[q][n] = o[q];
        } else {
#pragma unroll
            for (int q = 0; q < 5; q++) y[q] = y[q] / (cur.vol + cur.dt1 * xa[q] * cur.adiag);
            double o0, o1, o2, o3, o4;
            sym1i(o0, o1, o2, o3, o4, y[0], y[1], y[2], y[3], y[4], cur.c, cur.ro, cur.fx, cur.fy,
                  cur.fz, a21, a22, a23, a31, a32, a33);
            D[0][n] = o0; D[1][n] = o1; D[2][n] = o2; D[3][n] = o3; D[4][n] = o4;
        }
        // the flux handed to the next cell of the line
        const double xs[7] = {x1, x2, x1, x1, x5, x1, x1};
#pragma unroll
        for (int q = 0; q < NV; q++) y[q] = xs[q] * y[q];
        if (KE)
            sym1ke(bd, y, cur.rk, cur.c, cur.ro, cur.drdp, cur.drdh, cur.fx, cur.fy, cur.fz, a21, a22,
                   a23, a31, a32, a33);
        else
            sym1i(bd[0], bd[1], bd[2], bd[3], bd[4], y[0], y[1], y[2], y[3], y[4], cur.c, cur.ro,
                  cur.fx, cur.fy, cur.fz, a21, a22, a23, a31, a32, a33);
        cur = nxt;
        m = mn;
    }
}

// DIAG (DIAGKE) on every entry of the direction's plane and line, one thread each: x runs
// along I (stride 1), y and z over the other two sorted indices
template <bool KE>
__global__ void __launch_bounds__(256) k_ddadi_diag(Prm p, Blk b, DdDir r, int dir)
{
    const int gx = blockIdx.x * blockDim.x + threadIdx.x, gy = blockIdx.y, gz = blockIdx.z;
    int a0, b0, m;
    if (dir == 0) { a0 = gx; b0 = gy; m = gz + 1; }      // a = I, b = J, m = K
    else if (dir == 1) { m = gx + 1; a0 = gy; b0 = gz; } // a = J, b = K, m = I
    else { b0 = gx; a0 = gy; m = gz + 1; }                // a = K, b = I, m = J
    const int ist = r.na + 2 * FF_GHOST;
    if (a0 >= ist || b0 >= r.nb || m > r.nm) return;
    if ((b0 == 0 && a0 < FF_GHOST) || (r.na == 1 && a0 != FF_GHOST)) return;
    const int n = a0 * r.sa + (b0 + FF_GHOST) * r.sb + (m + FF_GHOST - 1) * r.sm, nu = n + r.sm;
    double *D[7] = {FLD(b, FF_DRO), FLD(b, FF_DM), FLD(b, FF_DN), FLD(b, FF_DW), FLD(b, FF_DE),
                    FLD(b, FF_DRK), FLD(b, FF_DEPS)};
    const double *vol = FLD(b, FF_VOL), *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W),
                 *c = FLD(b, FF_C), *ro = FLD(b, FF_RO);
    const double dturb = 1.;
    const double dt1 = p.theta * FLD(b, FF_DTL)[n];
    const double fx = r.AX[nu], fy = r.AY[nu], fz = r.AZ[nu];
    double xv[7], y[7], a21, a22, a23, a31, a32, a33;
#pragma unroll
    for (int q = 0; q < (KE ? 7 : 5); q++) xv[q] = D[q][n];
    if (KE)
        syike(y, xv, c[n], ro[n], fx, fy, fz, a21, a22, a23, a31, a32, a33);
    else
        syi(y[0], y[1], y[2], y[3], y[4], xv[0], xv[1], xv[2], xv[3], xv[4], c[n], ro[n], fx, fy, fz,
            a21, a22, a23, a31, a32, a33);
    double ucon = fx * u[n] + fy * v[n] + fz * w[n] - 0.;
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
        double o[7];
        sym1ke(o, y, FLD(b, FF_RK)[n], c[n], ro[n], FLD(b, FF_DRDP)[n], FLD(b, FF_DRDH)[n], fx, fy,
               fz, a21, a22, a23, a31, a32, a33);
#pragma unroll
        for (int q = 0; q < 7; q++) D[q][n] = o[q];
    } else {
        double o0, o1, o2, o3, o4;
        sym1i(o0, o1, o2, o3, o4, y[0], y[1], y[2], y[3], y[4], c[n], ro[n], fx, fy, fz, a21, a22,
              a23, a31, a32, a33);
        D[0][n] = o0; D[1][n] = o1; D[2][n] = o2; D[3][n] = o3; D[4][n] = o4;
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
        const int nt = dd_threads(r, dir);
        const dim3 g((unsigned)((nt + 127) / 128)), t(128);
        // DIAG: x along I, y and z over the other two indices of the sorted order
        const int ist = r.na + 2 * FF_GHOST;
        const int xn = dir == 0 ? ist : dir == 1 ? r.nm : r.nb;
        const dim3 gd((unsigned)((xn + 255) / 256), dir == 0 ? r.nb : ist, dir == 1 ? r.nb : r.nm);
        int rc;
        if (ke) {
            if ((rc = launch(c, "k_ddadi", k_ddadi_sweep<true, false>, g, t, kp, b, r, dir, nt)) ||
                (rc = launch(c, "k_ddadi", k_ddadi_diag<true>, gd, dim3(256), kp, b, r, dir)) ||
                (rc = launch(c, "k_ddadi", k_ddadi_sweep<true, true>, g, t, kp, b, r, dir, nt)))
                return rc;
        } else {
            if ((rc = launch(c, "k_ddadi", k_ddadi_sweep<false, false>, g, t, kp, b, r, dir, nt)) ||
                (rc = launch(c, "k_ddadi", k_ddadi_diag<false>, gd, dim3(256), kp, b, r, dir)) ||
                (rc = launch(c, "k_ddadi", k_ddadi_sweep<false, true>, g, t, kp, b, r, dir, nt)))
                return rc;
        }
    }
    const dim3 ge((unsigned)((b.imax + 255) / 256), b.jmax, b.kmax);
    return ke ? launch(c, "k_ddadi_end", k_ddadi_end<true>, ge, dim3(256), b)
              : launch(c, "k_ddadi_end", k_ddadi_end<false>, ge, dim3(256), b);
}
