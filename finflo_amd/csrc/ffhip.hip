// ffhip.hip — MI355X (gfx950) kernels and C-ABI of FINFLO's per-iteration hot path.
//
// Entry points are declared in include/ffhip.h.  All arithmetic is FP64 and follows the
// reference's operation order (see ffhip_device.h); built with -ffp-contract=off.
//
// Kernels (one block slab per launch, i-fastest SoA, two ghost layers):
//   k_state        STATIM: PDTOP + PERFGA + CONSVA on every cell (ns3c.f:10380-10447)
//   k_turb_lam     TURBDI ITURB=0 (ns3c.f:2987-2991)
//   k_residual     RESIDU: the k, i, j face fluxes (FLUXDO + FLUXKE) of a cell and their
//                  SWEEP differences fused, thread per interior cell (res3c.f:194-559)
//   k_resl2_*      RESL2 sums (res3c.f:8979-9006), deterministic two-stage reduction
//   k_time3_copr   TIME3 + LOCDT + COPR (base3c.f:79, imps3.f:784, imps3.f:891)
//   k_pred_plane   PREDGS on one hyperplane I+J+K = m (imps4.f:345-598)
//   k_corr_plane   DIAGGS + CORRGS on one hyperplane (imps4.f:604-664, 76-339)
//   k_divcp        DRO/CP (imps3.f:483)
//   k_fin          FIN (res3c.f:1122-1126)
//   k_limits       UPLIM/LOLIM of T, U, V, W, PDTOP, LOLIMP (ns3c.f:1553-1596)
//
// LU-SGS: the reference sweeps K-planes in order and, inside a plane, the diagonals
// M = I+J (imps4.f:488-570, 221-308).  Every cell's update reads only the final increments
// of its three upstream neighbours and its own data, so processing the hyperplanes
// I+J+K = m in order gives the same results bit for bit.  One quirk of PREDGS is kept:
// its line-buffer shift (imps4.f:556-562) starts at IALKU, so a cell in the row J = JMAX
// never receives the contribution of its I+1 neighbour.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <mutex>
#include <string>
#include <vector>

#include "ffhip_device.h"

using namespace ff;

// ------------------------------------------------------------------------------------
// kernels
// ------------------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_state(Prm p, Blk b)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= b.ntot) return;
    const double gm1 = p.gamma - 1.;
    const double cp2 = p.gamma / gm1 * p.rgas;
    const double cmin = 1.;
    double pd = FLD(b, FF_PDIFF)[n], t = FLD(b, FF_TEMP)[n];
    double u = FLD(b, FF_U)[n], v = FLD(b, FF_V)[n], w = FLD(b, FF_W)[n];
    double P = pd + p.frspre;                                   // PDTOP state.f:1707
    double trgs = t * p.rgas;                                   // PERFGA state.f:212-229
    double drdp = 1. / trgs;
    double ro = P * drdp;
    double vis = p.visu0 * t * p.rmult * sqrt(t) / (t + p.tsu0);
    double e = trgs / gm1;
    double drdh = -ro / e / p.gamma;
    double ch = vis * cp2 / p.pr;
    double rkin = (u * u + v * v + w * w) * .5;                 // CONSVA base3c.f:519-528
    double c;
    if (p.iturb < 3 || p.iturb == 8) {
        double x = 1. / (drdh / ro + drdp);
        c = sqrt(x > cmin ? x : cmin);
        e = (e + rkin + 0.) * ro;
    } else {                                                    // base3c.f:533-545
        const double r23 = 2. / 3.;
        double rk = FLD(b, FF_RK)[n];
        double ypro = 1. / ro;
        double xturb = 1. + r23 * rk * ypro * drdp;
        double x = xturb / (drdh * ypro + drdp);
        c = sqrt(x > cmin ? x : cmin);
        e = (e + rkin + 0.) * ro + rk;
    }
    FLD(b, FF_P)[n] = P;
    FLD(b, FF_RO)[n] = ro;
    FLD(b, FF_E)[n] = e;
    FLD(b, FF_C)[n] = c;
    FLD(b, FF_VIS)[n] = vis;
    FLD(b, FF_CH)[n] = ch;
    FLD(b, FF_CP)[n] = cp2;
    FLD(b, FF_DRDP)[n] = drdp;
    FLD(b, FF_DRDH)[n] = drdh;
    FLD(b, FF_RM)[n] = ro * u;
    FLD(b, FF_RN)[n] = ro * v;
    FLD(b, FF_RW)[n] = ro * w;
}

// ADJUIN (res3c.f:9447-9467): interior cells only; ghost values belong to the caller
__global__ void __launch_bounds__(256) k_turb_lam(Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > b.imax) return;
    int n = cid(b, i, j, k);
    FLD(b, FF_EPS2)[n] = 1.;
    FLD(b, FF_VIST)[n] = 0.;
}

// ------------------------------------------------------------------------------------
// k-omega SST (ITURB = 6): TURBEQ and the RESIKE terms outside the face fluxes
// ------------------------------------------------------------------------------------

__device__ __forceinline__ double pow4(double x) { return x * x * x * x; }   // ((x*x)*x)*x

// DERFUN (turb3.f:2539-2665, IDERI = 1) at cell L: velocity gradients by distance-weighted
// Green-Gauss, stored as vorticity, strain rate, |Omega| (OHMI) and strain magnitude
__device__ __forceinline__ double derfun_cell(const Blk &b, int L)
{
    const int IS = b.is, IL = b.il;
    const double *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W);
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3);
    const double *d1 = FLD(b, FF_D1), *d2 = FLD(b, FF_D2), *d3 = FLD(b, FF_D3);
    double w1, w2;
    w1 = 2.0 * (d1[L] / (d1[L] + d1[L - 1])); w2 = 2.0 - w1;
    double u1m = a1[L] * (w2 * u[L] + w1 * u[L - 1]);
    double v1m = a1[L] * (w2 * v[L] + w1 * v[L - 1]);
    double w1m = a1[L] * (w2 * w[L] + w1 * w[L - 1]);
    w1 = 2.0 * (d1[L] / (d1[L] + d1[L + 1])); w2 = 2.0 - w1;
    double u1p = a1[L + 1] * (w2 * u[L] + w1 * u[L + 1]);
    double v1p = a1[L + 1] * (w2 * v[L] + w1 * v[L + 1]);
    double w1p = a1[L + 1] * (w2 * w[L] + w1 * w[L + 1]);
    w1 = 2.0 * (d2[L] / (d2[L] + d2[L - IS])); w2 = 2.0 - w1;
    double u2m = a2[L] * (w2 * u[L] + w1 * u[L - IS]);
    double v2m = a2[L] * (w2 * v[L] + w1 * v[L - IS]);
    double w2m = a2[L] * (w2 * w[L] + w1 * w[L - IS]);
    w1 = 2.0 * (d2[L] / (d2[L] + d2[L + IS])); w2 = 2.0 - w1;
    double u2p = a2[L + IS] * (w2 * u[L] + w1 * u[L + IS]);
    double v2p = a2[L + IS] * (w2 * v[L] + w1 * v[L + IS]);
    double w2p = a2[L + IS] * (w2 * w[L] + w1 * w[L + IS]);
    w1 = 2.0 * (d3[L] / (d3[L] + d3[L - IL])); w2 = 2.0 - w1;
    double u3m = a3[L] * (w2 * u[L] + w1 * u[L - IL]);
    double v3m = a3[L] * (w2 * v[L] + w1 * v[L - IL]);
    double w3m = a3[L] * (w2 * w[L] + w1 * w[L - IL]);
    w1 = 2.0 * (d3[L] / (d3[L] + d3[L + IL])); w2 = 2.0 - w1;
    double u3p = a3[L + IL] * (w2 * u[L] + w1 * u[L + IL]);
    double v3p = a3[L + IL] * (w2 * v[L] + w1 * v[L + IL]);
    double w3p = a3[L + IL] * (w2 * w[L] + w1 * w[L + IL]);
    const double pvol = .5 / FLD(b, FF_VOL)[L];
    double g[3][3];   // g[c][q]: d(u,v,w)_q / d(x,y,z)_c
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
        const double *c1 = FLD(b, FF_A1X + cc), *c2 = FLD(b, FF_A2X + cc), *c3 = FLD(b, FF_A3X + cc);
        double e3p = c3[L + IL], e3m = c3[L], e2p = c2[L + IS], e2m = c2[L], e1p = c1[L + 1], e1m = c1[L];
        g[cc][0] = pvol * (e3p * u3p - e3m * u3m + e2p * u2p - e2m * u2m + e1p * u1p - e1m * u1m);
        g[cc][1] = pvol * (e3p * v3p - e3m * v3m + e2p * v2p - e2m * v2m + e1p * v1p - e1m * v1m);
        g[cc][2] = pvol * (e3p * w3p - e3m * w3m + e2p * w2p - e2m * w2m + e1p * w1p - e1m * w1m);
    }
    const double dudx = g[0][0], dudy = g[1][0], dudz = g[2][0];
    const double dvdx = g[0][1], dvdy = g[1][1], dvdz = g[2][1];
    const double dwdx = g[0][2], dwdy = g[1][2], dwdz = g[2][2];
    double w12 = .5 * (dudy - dvdx), w13 = .5 * (dudz - dwdx), w23 = .5 * (dvdz - dwdy);
    FLD(b, FF_W12)[L] = w12; FLD(b, FF_W13)[L] = w13; FLD(b, FF_W23)[L] = w23;
    FLD(b, FF_OHMI)[L] = 2. * sqrt(w12 * w12 + w13 * w13 + w23 * w23);
    double s11 = .5 * (dudx + dudx), s22 = .5 * (dvdy + dvdy), s33 = .5 * (dwdz + dwdz);
    double s12 = .5 * (dudy + dvdx), s13 = .5 * (dudz + dwdx), s23 = .5 * (dvdz + dwdy);
    double st = s11 * s11 + s12 * s12 + s13 * s13 + s12 * s12 + s22 * s22 + s23 * s23
              + s13 * s13 + s23 * s23 + s33 * s33;
    const double strain = sqrt(2. * st);
    FLD(b, FF_STRAIN)[L] = strain;
    FLD(b, FF_S11)[L] = s11; FLD(b, FF_S12)[L] = s12; FLD(b, FF_S13)[L] = s13;
    FLD(b, FF_S22)[L] = s22; FLD(b, FF_S23)[L] = s23; FLD(b, FF_S33)[L] = s33;
    return strain;
}

// EARSM1 earsm.f:60-219 at one interior cell: the Wallin-Johansson extra anisotropy b_ij
// from the time scale TTS, the strain-rate tensor SIJ and the vorticity tensor W12, W13, W23.
// The curvature correction is off (PA0 = 0, earsm.f:47; its WIR term is 0*0), and the
// routine's EPS2/VIST update is dead code (earsm.f:222), so only BIJ is written.  The
// association of every sum and product is the routine's; x**4 is a left-to-right product,
// x**P3 and x**P6 are pow.  Three parts: the invariants and tensor terms (earsm.f:72-167),
// the cubic solve, N correction and betas (earsm.f:170-205), and the b_ij sums
// (earsm.f:209-219); the learned closure replaces the middle part only.
struct EarsmT {
    double sii, wii, swwiv, sswwv, iiis;
    double t[4][5];               // terms of beta3, beta4, beta6, beta9: 11, 22, 12, 13, 23
};

__device__ __forceinline__ void earsm_terms(const Blk &b, int L, EarsmT &e)
{
    const double p3 = 1.0 / 3.0, tt = 2.0 * p3;
    const double pa0 = 0., wir = 0.;
    const double ts = FLD(b, FF_TTS)[L];
    double s11 = ts * FLD(b, FF_S11)[L], s22 = ts * FLD(b, FF_S22)[L], s33 = ts * FLD(b, FF_S33)[L];
    double s12 = ts * FLD(b, FF_S12)[L], s13 = ts * FLD(b, FF_S13)[L], s23 = ts * FLD(b, FF_S23)[L];
    double w12 = ts * FLD(b, FF_W12)[L] - pa0 * wir;
    double w13 = ts * FLD(b, FF_W13)[L] + pa0 * wir;
    double w23 = ts * FLD(b, FF_W23)[L] - pa0 * wir;
    double s11sq = s11 * s11, s22sq = s22 * s22, s33sq = s33 * s33;
    double s12sq = s12 * s12, s13sq = s13 * s13, s23sq = s23 * s23;
    double w12sq = w12 * w12, w13sq = w13 * w13, w23sq = w23 * w23;
    double sii = s11sq + s22sq + s33sq + 2.0 * (s12sq + s13sq + s23sq);
    double wii = -(2.0 * (w12sq + w13sq + w23sq));
    double wiip3 = p3 * wii;
    double swwiv = -(s11 * (w12sq + w13sq)) - s22 * (w12sq + w23sq) - s33 * (w13sq + w23sq)
                 + 2.0 * (-(s12 * w13 * w23) + s13 * w12 * w23 - s23 * w12 * w13);
    double swwivtt = tt * swwiv;
    double sswwv = 2.0 * (-((s12 * s13 + s22 * s23 + s23 * s33) * w12 * w13)
                          + (s11 * s13 + s12 * s23 + s13 * s33) * w12 * w23
                          - (s11 * s12 + s12 * s22 + s13 * s23) * w13 * w23)
                 - (s11sq + s12sq + s13sq) * (w12sq + w13sq)
                 - (s12sq + s22sq + s23sq) * (w12sq + w23sq)
                 - (s13sq + s23sq + s33sq) * (w13sq + w23sq);
    e.sii = sii; e.wii = wii; e.swwiv = swwiv; e.sswwv = sswwv;
    // tr(S^3), an input of the learned closure only
    e.iiis = s11 * s11sq + s22 * s22sq + s33 * s33sq
           + 3.0 * (s11 * (s12sq + s13sq) + s22 * (s12sq + s23sq) + s33 * (s13sq + s23sq))
           + 6.0 * s12 * s13 * s23;
    e.t[0][0] = -w12sq - w13sq - wiip3;
    e.t[0][1] = -w12sq - w23sq - wiip3;
    e.t[0][2] = -(w13 * w23);
    e.t[0][3] = w12 * w23;
    e.t[0][4] = -(w12 * w13);
    e.t[1][0] = -(2.0 * (s12 * w12 + s13 * w13));
    e.t[1][1] = 2.0 * (s12 * w12 - s23 * w23);
    e.t[1][2] = (s11 - s22) * w12 - s23 * w13 - s13 * w23;
    e.t[1][3] = -(s23 * w12) + (s11 - s33) * w13 + s12 * w23;
    e.t[1][4] = s13 * w12 + s12 * w13 + (s22 - s33) * w23;
    e.t[2][0] = -(2.0 * ((s12 * w13 - s13 * w12) * w23 + s11 * (w12sq + w13sq)))
              - swwivtt - wii * s11;
    e.t[2][1] = -(2.0 * ((s23 * w12 + s12 * w23) * w13 + s22 * (w12sq + w23sq)))
              - swwivtt - wii * s22;
    e.t[2][2] = -(s12 * (2.0 * w12sq + w13sq + w23sq))
              - (s13 * w13 - s23 * w23) * w12 - (s11 + s22) * w13 * w23 - wii * s12;
    e.t[2][3] = -(s13 * (w12sq + 2.0 * w13sq + w23sq))
              - (s12 * w12 + s23 * w23) * w13 + (s11 + s33) * w12 * w23 - wii * s13;
    e.t[2][4] = -(s23 * (w12sq + w13sq + 2.0 * w23sq))
              + (s12 * w12 - s13 * w13) * w23 - (s22 + s33) * w12 * w13 - wii * s23;
    e.t[3][0] = -(2.0 * ((s12 * w12 + s13 * w13 - s23 * w23) * w12sq
                         + (s12 * w12 + s13 * w13 + s23 * w23) * w13sq
                         + (s22 - s33) * w12 * w13 * w23));
    e.t[3][1] = -(2.0 * ((-(s12 * w12) - s13 * w13 + s23 * w23) * w12sq
                         + (-(s12 * w12) + s13 * w13 + s23 * w23) * w23sq
                         + (-s11 + s33) * w12 * w13 * w23));
    e.t[3][2] = ((s11 - s22) * w12 - 2.0 * (s13 * w23 + s23 * w13)) * w12sq
              + ((s11 - s33) * w12 - 2.0 * s13 * w23) * w13sq
              + ((s33 - s22) * w12 - 2.0 * s23 * w13) * w23sq;
    e.t[3][3] = ((s11 - s22) * w13 + 2.0 * s12 * w23) * w12sq
              + ((s11 - s33) * w13 + 2.0 * (s12 * w23 - s23 * w12)) * w13sq
              + ((s22 - s33) * w13 - 2.0 * s23 * w12) * w23sq;
    e.t[3][4] = ((s22 - s11) * w23 + 2.0 * s12 * w13) * w12sq
              + ((s11 - s33) * w23 + 2.0 * s13 * w12) * w13sq
              + ((s22 - s33) * w23 + 2.0 * (s12 * w13 + s13 * w12)) * w23sq;
}

// beta3, beta4, beta6, beta9 from N (earsm.f:196-205)
__device__ __forceinline__ void earsm_betas_n(const EarsmT &e, double fn, double beta[4])
{
    const double p6 = 0.5 * (1.0 / 3.0);
    const double wii = e.wii, swwiv = e.swwiv;
    double fnsq = fn * fn;
    double q = 5.0 * p6 * (fnsq - 2.0 * wii) * (2.0 * fnsq - wii);
    double pq = 1.0 / q;
    beta[0] = -(pq * 12.0 * swwiv / fn);
    beta[1] = -(pq * 2.0 * (fnsq - 2.0 * wii));
    beta[2] = -(pq * 6.0 * fn);
    beta[3] = pq * 6.0;
}

// the exact coefficients beta3, beta4, beta6, beta9 (earsm.f:170-205)
__device__ __forceinline__ void earsm_betas(const EarsmT &e, double beta[4])
{
    const double p3 = 1.0 / 3.0, p6 = 0.5 * p3, tt = 2.0 * p3;
    const double c1 = 1.8, c1p = 9.0 * (c1 - 1.0) / 4.0, c1p3 = c1p * p3, c1psq = c1p * c1p;
    const double sii = e.sii, wii = e.wii, swwiv = e.swwiv, sswwv = e.sswwv;
    // N_c from the cubic (earsm.f:170-184)
    double p1 = (c1psq / 27.0 + 0.45 * sii - tt * wii) * c1p;
    double cb = c1psq * (p3 * p3) + 0.9 * sii + tt * wii;
    double p2 = p1 * p1 - cb * cb * cb;
    double fnc;
    if (p2 >= 0.0) {
        double sqp2 = sqrt(p2);
        double pm = p1 - sqp2;
        double pmp = pow(fabs(pm), p3);
        fnc = c1p3 + pow(p1 + sqp2, p3) + copysign(fabs(pmp), pm);
    } else {
        double pm = p1 * p1 - p2;
        double sqpm = sqrt(pm);
        double alpo = p1 / sqpm;
        alpo = dmin(alpo, 1.);
        alpo = dmax(alpo, -1.);
        double facos = p3 * acos(alpo);
        fnc = c1p3 + 2.0 * pow(pm, p6) * cos(facos);
    }
    // the N correction and the betas (earsm.f:187-205)
    double dd = 20.0 * pow4(fnc) * (fnc - 0.5 * c1p)
              - wii * (10.0 * fnc + 15.0 * c1p) * (fnc * fnc)
              + 10.0 * c1p * (wii * wii);
    double fii1 = swwiv * swwiv;
    double fii2 = sswwv - 0.5 * sii * wii;
    double fn = fnc + 162.0 * (fii1 + fii2 * (fnc * fnc)) / dd;
    earsm_betas_n(e, fn, beta);
}

// b_ij from the betas and the tensor terms (earsm.f:209-219)
__device__ __forceinline__ void earsm_store(const Blk &b, int L, const EarsmT &e, const double beta[4])
{
    const double beta3 = beta[0], beta4 = beta[1], beta6 = beta[2], beta9 = beta[3];
    double b1 = 0.5 * (beta3 * e.t[0][0] + beta4 * e.t[1][0] + beta6 * e.t[2][0] + beta9 * e.t[3][0]);
    double b4 = 0.5 * (beta3 * e.t[0][1] + beta4 * e.t[1][1] + beta6 * e.t[2][1] + beta9 * e.t[3][1]);
    FLD(b, FF_B11)[L] = b1;
    FLD(b, FF_B12)[L] = 0.5 * (beta3 * e.t[0][2] + beta4 * e.t[1][2] + beta6 * e.t[2][2] + beta9 * e.t[3][2]);
    FLD(b, FF_B13)[L] = 0.5 * (beta3 * e.t[0][3] + beta4 * e.t[1][3] + beta6 * e.t[2][3] + beta9 * e.t[3][3]);
    FLD(b, FF_B22)[L] = b4;
    FLD(b, FF_B23)[L] = 0.5 * (beta3 * e.t[0][4] + beta4 * e.t[1][4] + beta6 * e.t[2][4] + beta9 * e.t[3][4]);
    FLD(b, FF_B33)[L] = -b1 - b4;
}

__device__ __forceinline__ void earsm_cell(const Blk &b, int L)
{
    EarsmT e;
    earsm_terms(b, L, e);
    double beta[4];
    earsm_betas(e, beta);
    earsm_store(b, L, e, beta);
}

// TURBEQ (res3c.f:5951-6371) on the cells of level 1 plus one ghost layer: DERFUN
// (turb3.f:2476-2665, IDERI = 1: distance-weighted Green-Gauss) and TTIMES (earsm.f:671-684)
// everywhere there, EPS2KO (turb3.f:7454-7537, limiter argument STRAIN, FRSMUT =
// 1E-3*FRSVIS) on the interior.  Thread per cell, I fastest.
__global__ void __launch_bounds__(256) k_turb_sst(Prm p, Blk b)
{
    const int ni = b.imax + 2;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y, k = blockIdx.z;
    if (i >= ni) return;
    const int L = cid(b, i, j, k);
    const double strain = derfun_cell(b, L);
    // TTIMES
    const double ro = FLD(b, FF_RO)[L], vis = FLD(b, FF_VIS)[L];
    const double rk = FLD(b, FF_RK)[L], rom = FLD(b, FF_REPS)[L];
    {
        const double ct = 6.0, bstar = 0.09, eps = 1.E-20;
        double bnum = 1.0 / fabs(bstar * rom + eps);
        double tstur = ro * bnum;
        double tsvis = ct * sqrt(ro * vis * bnum / fabs(rk + eps));
        FLD(b, FF_TTS)[L] = dmax(tstur, tsvis);
    }
    if (i < 1 || i > b.imax || j < 1 || j > b.jmax || k < 1 || k > b.kmax) return;
    // EARSM1 after TTIMES (res3c.f:5987-5990); with the learned closure a later pass forms BIJ
    if (p.istres == 1 && !p.earsm_mlp) earsm_cell(b, L);
    // EPS2KO
    const SstC s = coefko(p.kover);
    const double eps = 1.E-20;
    const double sst = p.kover < 5 ? 1.0 : 0.0, bsl = p.kover < 5 ? 0.0 : 1.0;
    const double frsmut = 1.E-3 * p.frsvis;
    const double distw = FLD(b, FF_DISTW)[L];
    double prol = 1. / ro;
    double rkl = dmax(rk * prol, 0.0);
    double oml = rom * prol;
    double dis2 = distw * distw;
    double arg21 = 2. * sqrt(rkl) / dmax(eps, s.bstar * oml * distw);
    double auxvi = vis / dmax(eps, rom * dis2);
    double arg22 = 500. * auxvi;
    double arg2 = dmax(arg21, arg22);
    double arg4 = 150. * auxvi;
    double fun2 = tanh(arg2 * arg2);
    double fun3 = 1. - tanh(pow4(arg4));
    double rmubas = rk / oml;
    double rmumen = s.a1kleb * rk / dmax(eps, strain * fun2 * fun3);
    double rmutrb = dmin(rmubas, rmumen);
    rmutrb = rmutrb * sst + rmubas * bsl;
    rmutrb = dmax(rmutrb, frsmut);
    rmutrb = dmin(rmutrb, p.turlim * vis);
    FLD(b, FF_VIST)[L] = rmutrb;
    double rmutpv = rmutrb / vis;
    FLD(b, FF_EPS2)[L] = 1.0 + rmutpv;
}

// RESIKE with M2 >= 2 (res3c.f:3749-3769): OHMI, W12 and SIJ zeroed on every cell, then
// DERFUN on level 1 with the first ghost layer (g0 = 1) or on a coarse level's interior
// (g0 = 0).  Thread per cell of the range.
constexpr int NGRAD = 10;
__global__ void __launch_bounds__(256) k_grad_zero(Blk b)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= b.ntot) return;
    FLD(b, FF_OHMI)[n] = 0.;
#pragma unroll
    for (int q = 0; q < 9; q++) FLD(b, FF_W12 + q)[n] = 0.;
}

__global__ void __launch_bounds__(256) k_derfun(Blk b, int g0)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1 - g0;
    int j = blockIdx.y + 1 - g0, k = blockIdx.z + 1 - g0;
    if (i > b.imax + g0) return;
    derfun_cell(b, cid(b, i, j, k));
}

// TEMDER turb3.f:3456-3559, IDERI = 1, of the cell scalar q (gq picks it) at cell L: the
// distance-weighted Green-Gauss gradient
template <typename Q>
__device__ __forceinline__ void temder_cell(const Blk &b, int L, Q gq, double &gx, double &gy,
                                            double &gz)
{
    const int IS = b.is, IL = b.il;
    const double *d1 = FLD(b, FF_D1), *d2 = FLD(b, FF_D2), *d3 = FLD(b, FF_D3);
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3);
    const double qc = gq(L);
    double w1, w2;
    w1 = 2.0 * (d1[L] / (d1[L] + d1[L - 1])); w2 = 2.0 - w1;
    const double t1m = w2 * qc + w1 * gq(L - 1);
    w1 = 2.0 * (d1[L] / (d1[L] + d1[L + 1])); w2 = 2.0 - w1;
    const double t1p = w2 * qc + w1 * gq(L + 1);
    w1 = 2.0 * (d2[L] / (d2[L] + d2[L - IS])); w2 = 2.0 - w1;
    const double t2m = w2 * qc + w1 * gq(L - IS);
    w1 = 2.0 * (d2[L] / (d2[L] + d2[L + IS])); w2 = 2.0 - w1;
    const double t2p = w2 * qc + w1 * gq(L + IS);
    w1 = 2.0 * (d3[L] / (d3[L] + d3[L - IL])); w2 = 2.0 - w1;
    const double t3m = w2 * qc + w1 * gq(L - IL);
    w1 = 2.0 * (d3[L] / (d3[L] + d3[L + IL])); w2 = 2.0 - w1;
    const double t3p = w2 * qc + w1 * gq(L + IL);
    const double pvol = .5 / FLD(b, FF_VOL)[L];
    double g[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double *n1 = FLD(b, FF_A1X + c), *n2 = FLD(b, FF_A2X + c), *n3 = FLD(b, FF_A3X + c);
        g[c] = pvol * (a3[L + IL] * n3[L + IL] * t3p - a3[L] * n3[L] * t3m
                       + a2[L + IS] * n2[L + IS] * t2p - a2[L] * n2[L] * t2m
                       + a1[L + 1] * n1[L + 1] * t1p - a1[L] * n1[L] * t1m);
    }
    gx = g[0]; gy = g[1]; gz = g[2];
}

// The FULLNS gradients on level 1, first ghost layer included (the K ghost layer only when
// KMAX > 1, turb3.f:3498-3507): TEMDER of T (RESIDU res3c.f:257-259, RESIKE
// res3c.f:3991-3993) and, KW, of k/rho and omega/rho (DIVV12 + TEMDER res3c.f:3995-4005;
// the quotients are formed per neighbour, the same division as DIVV12's).  Thread per cell.
template <bool KW>
__global__ void __launch_bounds__(256) k_temder(Blk b, int gk)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, k = blockIdx.z + 1 - gk;
    if (i > b.imax + 1) return;
    const int L = cid(b, i, j, k);
    const double *t = FLD(b, FF_TEMP);
    double gx, gy, gz;
    temder_cell(b, L, [=](int n) { return t[n]; }, gx, gy, gz);
    FLD(b, FF_DTDX)[L] = gx; FLD(b, FF_DTDY)[L] = gy; FLD(b, FF_DTDZ)[L] = gz;
    if (KW) {
        const double *ro = FLD(b, FF_RO), *rk = FLD(b, FF_RK), *reps = FLD(b, FF_REPS);
        temder_cell(b, L, [=](int n) { return rk[n] / ro[n]; }, gx, gy, gz);
        FLD(b, FF_DKDX)[L] = gx; FLD(b, FF_DKDY)[L] = gy; FLD(b, FF_DKDZ)[L] = gz;
        temder_cell(b, L, [=](int n) { return reps[n] / ro[n]; }, gx, gy, gz);
        FLD(b, FF_DEDX)[L] = gx; FLD(b, FF_DEDY)[L] = gy; FLD(b, FF_DEDZ)[L] = gz;
    }
}

// EXTEND (turb3.f:3342-3450) of OHMI, W12 and SIJ on a coarse level: the interior value
// next to each face copied into the first ghost layer (no edges or corners)
__global__ void __launch_bounds__(256) k_extend(Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.imax) return;
    const int n = cid(b, i, j, k);
    const int IS = b.is, IL = b.il;
    int t[6], nt = 0;
    if (k == 1) t[nt++] = n - IL;
    if (k == b.kmax) t[nt++] = n + IL;
    if (j == 1) t[nt++] = n - IS;
    if (j == b.jmax) t[nt++] = n + IS;
    if (i == 1) t[nt++] = n - 1;
    if (i == b.imax) t[nt++] = n + 1;
    if (nt == 0) return;
    for (int q = 0; q < NGRAD; q++) {
        double *x = FLD(b, q == 0 ? FF_OHMI : FF_W12 + q - 1);
        double v = x[n];
        for (int r = 0; r < nt; r++) x[t[r]] = v;
    }
}

// ------------------------------------------------------------------------------------
// multigrid transfers (multi.f)
// ------------------------------------------------------------------------------------

// Volume-weighted sum of the 8 fine cells of a coarse cell in TRANSF's order
// (multi.f:214-218): IPJPK, IJK, IPJK, IJPK, IPJPKP, IJKP, IPJKP, IJPKP.  kp is the K
// stride, or 0 on a planar fine level (TRANSF's IMJM2 = 0, multi.f:185-191), where the
// four "KP" cells are the four K cells again.
__device__ __forceinline__ double restrict8(const Blk &f, const double *vol, const double *x,
                                            int ijk, int kp, double pvol)
{
    const int ipjk = ijk + 1, ijpk = ijk + f.is, ipjpk = ijk + f.is + 1;
    const int ijkp = ijk + kp, ipjpkp = ipjpk + kp, ijpkp = ijpk + kp, ipjkp = ipjk + kp;
    return (vol[ipjpk] * x[ipjpk] + vol[ijk] * x[ijk] + vol[ipjk] * x[ipjk] +
            vol[ijpk] * x[ijpk] + vol[ipjpkp] * x[ipjpkp] + vol[ijkp] * x[ijkp] +
            vol[ipjkp] * x[ipjkp] + vol[ijpkp] * x[ijpkp]) * pvol;
}

// TRANSF (+ TRANKE, KW) loop 1000 and the OLD copies (multi.f:194-278, 453-510): coarse cell
// per thread over the interior and the first ghost layer.  A planar fine level (KMAX = 1)
// maps coarse K to fine K (KIND = 1, multi.f:185-196).
template <bool KW>
__global__ void __launch_bounds__(256) k_restrict(Blk f, Blk c)
{
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    int J = blockIdx.y, K = blockIdx.z;
    if (I > c.imax + 1) return;
    const int n = cid(c, I, J, K);
    const bool planar = f.kmax == 1;
    const int ijk = cid(f, 2 * I - 1, 2 * J - 1, planar ? K : 2 * K - 1), kp = planar ? 0 : f.il;
    const double *vol = FLD(f, FF_VOL);
    const double pvol = 1. / FLD(c, FF_VOL)[n];
    const int st[7] = {FF_TEMP, FF_U, FF_V, FF_W, FF_PDIFF, FF_EPS2, FF_VIST};
    double x[7];
#pragma unroll
    for (int q = 0; q < 7; q++) {
        x[q] = restrict8(f, vol, FLD(f, st[q]), ijk, kp, pvol);
        FLD(c, st[q])[n] = x[q];
    }
#pragma unroll
    for (int q = 0; q < 5; q++) FLD(c, FF_ROP2H + q)[n] = restrict8(f, vol, FLD(f, FF_DRO + q), ijk, kp, pvol);
    double rk = 0., re = 0.;
    if (KW) {
        rk = restrict8(f, vol, FLD(f, FF_RK), ijk, kp, pvol);
        re = restrict8(f, vol, FLD(f, FF_REPS), ijk, kp, pvol);
        FLD(c, FF_RK)[n] = rk;
        FLD(c, FF_REPS)[n] = re;
        FLD(c, FF_RKP2H)[n] = restrict8(f, vol, FLD(f, FF_DRK), ijk, kp, pvol);
        FLD(c, FF_EPSP2H)[n] = restrict8(f, vol, FLD(f, FF_DEPS), ijk, kp, pvol);
    }
    if (I >= 1 && I <= c.imax && J >= 1 && J <= c.jmax && K >= 1 && K <= c.kmax) {
        FLD(c, FF_TOLD)[n] = x[0];
        FLD(c, FF_UOLD)[n] = x[1];
        FLD(c, FF_VOLD)[n] = x[2];
        FLD(c, FF_WOLD)[n] = x[3];
        FLD(c, FF_POLD)[n] = x[4];
        if (KW) {
            FLD(c, FF_RKOLD)[n] = rk;
            FLD(c, FF_EPSOLD)[n] = re;
        }
    }
}

// The ghost treatment of TRANSF/TRANKE (loops 2100-2600, multi.f:283-425) for one pair of
// opposite faces: face 0 = I, 1 = J, 2 = K.  Residual transfers are zeroed on both ghost
// layers and the variables copied from the first to the second; the launches for I, J and
// K run in that order, as the reference's loops do.
template <bool KW>
__global__ void __launch_bounds__(256) k_restrict_ghosts(Blk c, int face)
{
    const int na = (face == 0 ? c.jmax : c.imax) + 4;          // index a runs over -1..N+2
    int a = blockIdx.x * blockDim.x + threadIdx.x - 1;
    int bb = blockIdx.y - 1;                                   // the other index
    if (a > na - 2) return;
    int lo, hi, st;
    if (face == 0) {        // (J = a, K = bb)
        lo = cid(c, 0, a, bb); hi = cid(c, c.imax + 1, a, bb); st = 1;
    } else if (face == 1) { // (I = a, K = bb)
        lo = cid(c, a, 0, bb); hi = cid(c, a, c.jmax + 1, bb); st = c.is;
    } else {                // (I = a, J = bb)
        lo = cid(c, a, bb, 0); hi = cid(c, a, bb, c.kmax + 1); st = c.il;
    }
    const int nz = KW ? 7 : 5;
    for (int q = 0; q < nz; q++) {
        double *x = FLD(c, FF_ROP2H + q);
        x[lo] = 0.; x[lo - st] = 0.; x[hi] = 0.; x[hi + st] = 0.;
    }
    const int cp[9] = {FF_TEMP, FF_U, FF_V, FF_W, FF_PDIFF, FF_EPS2, FF_VIST, FF_RK, FF_REPS};
    const int ncp = KW ? 9 : 7;
    for (int q = 0; q < ncp; q++) {
        double *x = FLD(c, cp[q]);
        x[lo - st] = x[lo];
        x[hi + st] = x[hi];
    }
}

// FORCFU (phase 0, P = P - D, multi.f:1630-1648) or ADDRES (phase 1, D = P + D,
// multi.f:1707-1725) on every cell
__global__ void __launch_bounds__(256) k_forcing(Blk b, int nq, int phase)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= b.ntot) return;
    for (int q = 0; q < nq; q++) {
        double *P = FLD(b, FF_ROP2H + q), *D = FLD(b, FF_DRO + q);
        if (phase == 0) P[n] = P[n] - D[n];
        else D[n] = P[n] + D[n];
    }
}

// INTERP (multi.f:1790-1819) and, KW, INTERT/INTERO (multi.f:2007-2077): fine interior cell
// per thread, the correction of coarse cell ((I+1)/2, (J+1)/2, (K+1)/2)
template <bool KW>
__global__ void __launch_bounds__(256) k_prolong(Prm p, Blk c, Blk f)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > f.imax) return;
    const int ijk = cid(f, i, j, k);
    const int n = cid(c, (i + 1) / 2, (j + 1) / 2, (k + 1) / 2);
    double dtok = FLD(c, FF_TEMP)[n] - FLD(c, FF_TOLD)[n];
    double duk = FLD(c, FF_U)[n] - FLD(c, FF_UOLD)[n];
    double dvk = FLD(c, FF_V)[n] - FLD(c, FF_VOLD)[n];
    double dwk = FLD(c, FF_W)[n] - FLD(c, FF_WOLD)[n];
    double dpk = FLD(c, FF_PDIFF)[n] - FLD(c, FF_POLD)[n];
    FLD(f, FF_TEMP)[ijk] = FLD(f, FF_TEMP)[ijk] + dtok;
    FLD(f, FF_U)[ijk] = FLD(f, FF_U)[ijk] + duk;
    FLD(f, FF_V)[ijk] = FLD(f, FF_V)[ijk] + dvk;
    FLD(f, FF_W)[ijk] = FLD(f, FF_W)[ijk] + dwk;
    FLD(f, FF_P)[ijk] = FLD(f, FF_P)[ijk] + dpk;
    FLD(f, FF_PDIFF)[ijk] = FLD(f, FF_PDIFF)[ijk] + dpk;
    if (KW) {
        const double eps = 1.E-7;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const double cc1 = q == 0 ? p.cmgk : p.cmgeps, cc2 = cc1 + eps;
            double *ro = FLD(f, q == 0 ? FF_RK : FF_REPS);
            double drok = FLD(c, q == 0 ? FF_RK : FF_REPS)[n] - FLD(c, q == 0 ? FF_RKOLD : FF_EPSOLD)[n];
            double alpo = fabs(drok);
            double r = ro[ijk];
            ro[ijk] = r + cc1 * r * drok / (cc2 * r + alpo);
        }
    }
}

// The limiters CYCLES applies on a level other than the finest after FIN (ns3c.f:1936-1985:
// UPLIM/LOLIM of TEMP, LOLIMP with the level's last pressure) or on a level after the
// prolongation into it (ns3c.f:2073-2118: for two-equation models the TEMP limits and LOLIM of
// RK and REPS, then LOLIMP).  Every cell.
__global__ void __launch_bounds__(256) k_limits_mg(Prm p, Blk b, int temp_limits, int kw_limits)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= b.ntot) return;
    if (temp_limits) {
        int kk = n / b.il;
        int jj = (n / b.is) % b.js;
        bool up = kk >= 2 && kk < 2 + b.kmax && jj >= 2 && jj < 2 + b.jmax;
        double t = FLD(b, FF_TEMP)[n];
        if (up) t = t < p.tuplim ? t : p.tuplim;
        t = t > p.tlolim ? t : p.tlolim;
        FLD(b, FF_TEMP)[n] = t;
    }
    if (kw_limits) {
        double rk = FLD(b, FF_RK)[n], re = FLD(b, FF_REPS)[n];
        FLD(b, FF_RK)[n] = rk > p.rklim ? rk : p.rklim;
        FLD(b, FF_REPS)[n] = re > p.epslim ? re : p.epslim;
    }
    double P = FLD(b, FF_P)[n];
    double y = P - 0.01 * p.frspre;
    double prelim = 0. < y ? 0. : y;
    FLD(b, FF_P)[n] = P - prelim;
    FLD(b, FF_PDIFF)[n] = FLD(b, FF_PDIFF)[n] - prelim;
}

// k = RK/RO and omega = REPS/RO as CROSKO forms them in place (turb3.f:7817-7821)
__device__ __forceinline__ double per_ro(const double *x, const double *ro, int n)
{
    double prol = 1. / ro[n];
    return x[n] * prol;
}

// CROSKO turb3.f:7824-7892 on the interior: SEPS = 2*SIGOM2*RO*grad(k).grad(omega)/omega,
// with k and omega formed from the unrounded RK, REPS of each stencil cell
__global__ void __launch_bounds__(256) k_crosko(Prm p, Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.imax) return;
    const int IS = b.is, IL = b.il;
    const int L = cid(b, i, j, k);
    const SstC s = coefko(p.kover);
    const double sigo22 = 2.0 * s.sigom2;
    const double *RK = FLD(b, FF_RK), *OM = FLD(b, FF_REPS), *ro = FLD(b, FF_RO);
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3);
    const double k0 = per_ro(RK, ro, L), o0 = per_ro(OM, ro, L);
    double rk1m = a1[L] * (k0 + per_ro(RK, ro, L - 1)), rk1p = a1[L + 1] * (k0 + per_ro(RK, ro, L + 1));
    double om1m = a1[L] * (o0 + per_ro(OM, ro, L - 1)), om1p = a1[L + 1] * (o0 + per_ro(OM, ro, L + 1));
    double rk2m = a2[L] * (k0 + per_ro(RK, ro, L - IS)), rk2p = a2[L + IS] * (k0 + per_ro(RK, ro, L + IS));
    double om2m = a2[L] * (o0 + per_ro(OM, ro, L - IS)), om2p = a2[L + IS] * (o0 + per_ro(OM, ro, L + IS));
    double rk3m = a3[L] * (k0 + per_ro(RK, ro, L - IL)), rk3p = a3[L + IL] * (k0 + per_ro(RK, ro, L + IL));
    double om3m = a3[L] * (o0 + per_ro(OM, ro, L - IL)), om3p = a3[L + IL] * (o0 + per_ro(OM, ro, L + IL));
    const double pvol = 0.5 / FLD(b, FF_VOL)[L];
    double dk[3], dom[3];
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
        const double *c1 = FLD(b, FF_A1X + cc), *c2 = FLD(b, FF_A2X + cc), *c3 = FLD(b, FF_A3X + cc);
        double e3p = c3[L + IL], e3m = c3[L], e2p = c2[L + IS], e2m = c2[L], e1p = c1[L + 1], e1m = c1[L];
        dk[cc] = pvol * (e3p * rk3p - e3m * rk3m + e2p * rk2p - e2m * rk2m + e1p * rk1p - e1m * rk1m);
        dom[cc] = pvol * (e3p * om3p - e3m * om3m + e2p * om2p - e2m * om2m + e1p * om1p - e1m * om1m);
    }
    double divdot = dk[0] * dom[0] + dk[1] * dom[1] + dk[2] * dom[2];
    FLD(b, FF_SEPS)[L] = sigo22 * ro[L] * divdot / o0;
}

// The rounding CROSKO leaves on RK and REPS of every cell: (X*(1/RO))*RO (turb3.f:7895-7898).
// Also zeroes FUN1 and SRK (RESIKE res3c.f:3673, 3730) and forms per-workgroup maxima of
// SEPS over the linear index range FUN1KO scans (turb3.f:7423-7428).
__global__ void __launch_bounds__(256) k_kw_round(Blk b, double *part)
{
    __shared__ double sh[256];
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (n < b.ntot) {
        const double ro = FLD(b, FF_RO)[n];
        FLD(b, FF_RK)[n] = per_ro(FLD(b, FF_RK), FLD(b, FF_RO), n) * ro;
        FLD(b, FF_REPS)[n] = per_ro(FLD(b, FF_REPS), FLD(b, FF_RO), n) * ro;
        FLD(b, FF_FUN1)[n] = 0.;
        FLD(b, FF_SRK)[n] = 0.;
        if (n >= cid(b, 1, 1, 1) && n <= cid(b, b.imax, b.jmax, b.kmax)) m = FLD(b, FF_SEPS)[n];
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = dmax(sh[threadIdx.x + h], sh[threadIdx.x]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// Maximum of the partial maxima (max is exact, so the order does not matter)
__global__ void __launch_bounds__(256) k_max_final(const double *part, int nparts, double *out)
{
    __shared__ double sh[256];
    double m = 0.0;
    for (int t = threadIdx.x; t < nparts; t += blockDim.x) m = dmax(part[t], m);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) sh[threadIdx.x] = dmax(sh[threadIdx.x + h], sh[threadIdx.x]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// FUN1KO (turb3.f:7430-7446) and PRODKE (turb3.f:4720-4787, with MAKE_EPS turb3.f:7913)
// on the interior
__global__ void __launch_bounds__(256) k_fun1_prod(Prm p, Blk b, const double *cdmax)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1, k = blockIdx.z + 1;
    if (i > b.imax) return;
    const int L = cid(b, i, j, k);
    const SstC s = coefko(p.kover);
    const double eps = 1.E-20, sigo24 = 4. * s.sigom2;
    const double ro = FLD(b, FF_RO)[L], rk = FLD(b, FF_RK)[L], om = FLD(b, FF_REPS)[L];
    const double vis = FLD(b, FF_VIS)[L], distw = FLD(b, FF_DISTW)[L];
    const double cross = FLD(b, FF_SEPS)[L];
    double cdmin = 1E-8 * cdmax[0];
    double prol = 1. / ro;
    double rkl = dmax(rk * prol, 0.0);
    double oml = om * prol;
    double dis2 = distw * distw;
    double cdkom = dmax(cross, cdmin);
    double arg11 = sqrt(rkl) / dmax(eps, s.bstar * oml * distw);
    double arg12 = 500. * vis / dmax(eps, om * dis2);
    double arg13 = dmax(arg11, arg12);
    double arg14 = sigo24 * rk / dmax(eps, cdkom * dis2);
    double arg1 = dmin(arg13, arg14);
    double f1 = tanh(pow4(arg1));
    double dislim = .40559E-7 * sqrt(p.re);
    if (distw < dislim) f1 = 1.;
    FLD(b, FF_FUN1)[L] = f1;
    if (p.istres >= 1) {
        // PEEKOO earsm.f:592-622 (res3c.f:3976-3978): production with the EARSM stresses
        const double p3 = 1.0 / 3.0, epsc = 1.E-20;
        double s11 = FLD(b, FF_S11)[L], s22 = FLD(b, FF_S22)[L], s33 = FLD(b, FF_S33)[L];
        double s12 = FLD(b, FF_S12)[L], s13 = FLD(b, FF_S13)[L], s23 = FLD(b, FF_S23)[L];
        double b1 = FLD(b, FF_B11)[L], b2 = FLD(b, FF_B12)[L], b3 = FLD(b, FF_B13)[L];
        double b4 = FLD(b, FF_B22)[L], b5 = FLD(b, FF_B23)[L];
        double rmut = dmax(FLD(b, FF_EPS2)[L] - 1.0, epsc) * vis;
        double divvp3 = p3 * (s11 + s22 + s33);
        double huu = dmax((rk * (p3 + b1) - rmut * (s11 - divvp3)), 0.0);
        double hvv = dmax((rk * (p3 + b4) - rmut * (s22 - divvp3)), 0.0);
        double hww = dmax((rk * (p3 - b1 - b4) - rmut * (s33 - divvp3)), 0.0);
        double huv = rk * b2 - rmut * s12;
        double huw = rk * b3 - rmut * s13;
        double hvw = rk * b5 - rmut * s23;
        double pt = -(2.0 * (huu * s11 + hvv * s22 + hww * s33
                             + 2.0 * (huv * s12 + huw * s13 + hvw * s23)));
        FLD(b, FF_PTUR)[L] = pt * 1.0;
        return;
    }
    // PRODKE
    const double r2o3 = 2. / 3.;
    double s11 = FLD(b, FF_S11)[L], s12 = FLD(b, FF_S12)[L], s13 = FLD(b, FF_S13)[L];
    double s22 = FLD(b, FF_S22)[L], s23 = FLD(b, FF_S23)[L], s33 = FLD(b, FF_S33)[L];
    double w12 = FLD(b, FF_W12)[L], w13 = FLD(b, FF_W13)[L], w23 = FLD(b, FF_W23)[L];
    double dudx = s11, dvdy = s22, dwdz = s33;
    double dudy = s12 + w12, dudz = s13 + w13;
    double dvdx = s12 - w12, dvdz = s23 + w23;
    double dwdx = s13 - w13, dwdy = s23 - w23;
    double strain = s11 * s11 + s12 * s12 + s13 * s13 + s12 * s12 + s22 * s22 + s23 * s23
                  + s13 * s13 + s23 * s23 + s33 * s33;
    strain = sqrt(2. * strain);
    double trans = 0.;
    double div = dudx + dvdy + dwdz;
    double visco = (FLD(b, FF_EPS2)[L] - 1. + trans) * vis;
    double pt;
    if (p.katol)
        pt = visco * strain * FLD(b, FF_OHMI)[L];
    else
        pt = visco * (2. * (dudx * dudx + dvdy * dvdy + dwdz * dwdz)
                      + (dudy + dvdx) * (dudy + dvdx) + (dudz + dwdx) * (dudz + dwdx)
                      + (dvdz + dwdy) * (dvdz + dwdy) - r2o3 * (div * div))
             - r2o3 * rk * div;
    double trueps = 0.09 * om * rk / ro;
    pt = dmin(pt, 20. * trueps);
    FLD(b, FF_PTUR)[L] = pt * 1.;
}

// SOURKO (turb3.f:7693-7778, RANS, KOVER 0/5) for one interior cell: returns SRK and the
// new SEPS (SEPS holding the cross-diffusion on entry)
__device__ __forceinline__ void sourko_cell(const Prm &p, const Blk &b, int L, double &srk,
                                            double &som)
{
    const SstC s = coefko(p.kover);
    const double eps = 1.0E-10;
    const double pbstar = 1.0 / s.bstar, persbs = 1.0 / sqrt(s.bstar);
    const double rotco = 0.0, rich = 0.0, bero = 3.6;
    const double f1 = FLD(b, FF_FUN1)[L], ro = FLD(b, FF_RO)[L], rk = FLD(b, FF_RK)[L];
    const double rom = FLD(b, FF_REPS)[L], ptur = FLD(b, FF_PTUR)[L];
    double onemf1 = dmax(0.0, (1.0 - f1));
    double beta = f1 * s.beta1 + onemf1 * s.beta2;
    double sigom = f1 * s.sigom1 + onemf1 * s.sigom2;
    double gamkom = beta * pbstar - sigom * (s.ckappa * s.ckappa) * persbs;
    double rmutrb = dmax(eps, FLD(b, FF_EPS2)[L] - 1.0) * FLD(b, FF_VIS)[L];
    double fco = dmax(0.0, (1.0 - rotco)) + dmax(0.0, rotco / (1.0 + bero * rich));
    double trueomega = rom;
    double prol = 1.0 / ro;
    som = onemf1 * FLD(b, FF_SEPS)[L];
    srk = ptur - s.bstar * rk * trueomega * prol;
    double omepro = gamkom * ro * ptur / (rmutrb * rom);
    double omedis = fco * beta * prol * trueomega;
    som = som + rom * (omepro - omedis);
}

// ------------------------------------------------------------------------------------
// XCD-aware workgroup numbering.  The dispatcher deals workgroups to the eight XCDs round-
// robin (XCD = linear id mod 8), each XCD with its own L2.  A grid padded to a multiple of 8
// is renumbered so that every XCD runs one contiguous range of the logical tiles: tiles that
// share halo cells (neighbours along I, J and the next K slab) then meet in the same L2.
// ------------------------------------------------------------------------------------
constexpr int NXCD = 8;
struct TileGrid {
    int nx, ny, nz;   // logical tiles along the three axes
    int per;          // logical tiles per XCD (the launch has NXCD * per workgroups)
};

static inline TileGrid tile_grid(int nx, int ny, int nz)
{
    const int n = nx * ny * nz;
    return TileGrid{nx, ny, nz, (n + NXCD - 1) / NXCD};
}

// logical tile of this workgroup, or -1 for the padding
__device__ __forceinline__ int xcd_tile(const TileGrid &g)
{
    const int w = blockIdx.x;
    const int l = (w % NXCD) * g.per + w / NXCD;
    return l < g.nx * g.ny * g.nz ? l : -1;
}

// ------------------------------------------------------------------------------------
// Solid-wall patches: BOTSKE base3c.f:6147-7095.  Before the residual kernel, k_wall forms
// per patch face the wall flux that replaces the face flux (loop 1000), the patch arrays,
// PTUR next to the wall and, with TURCOR and the SST, the corrections of the second and
// third faces (base3c.f:6763-6897).  The residual kernels then take these at the patch
// faces (wall_fix).  The force sums (loop 2000) follow in the reference's order on a side
// stream (k_wall_forces).
// ------------------------------------------------------------------------------------
constexpr int NWARR = 23;
enum { WA_UWALL, WA_VWALL, WA_WWALL, WA_TWALL, WA_HFLUX, WA_RBK, WA_XCP, WA_YCP, WA_ZCP,
       WA_CPWALL, WA_QWALL, WA_QWFRIC, WA_TAUW1, WA_TAUW2, WA_TAUWX, WA_TAUWY, WA_TAUWZ,
       WA_SURFX, WA_SURFY, WA_SURFZ, WA_SURFPX, WA_SURFPY, WA_SURFPZ };

// Per block level: the six block faces' cells stacked (face f = dir, dir + 3 for the high
// face; in-plane index wall_idx), a flag per cell that lies on a wall patch, the wall flux
// (7 components), the corrections of the faces two and three cells in (2 x 4 components)
// and the loop-1000 terms of the QY, QZ sums (2).
struct WallDev {
    int on, corr_on, nt;
    int off[6];
    const unsigned char *flag;
    double *flux;       // [7][nt]
    double *corr;       // [2][4][nt]
    double *qterm;      // [3][nt]: the loop-1000 QY, QZ terms and the wall pressure PB
};

// One patch in BOTSKE's frame: face (0..5), direction, ILMO, KBOTOP, the KX/KY ranges
// after the J-face swap, the rotated strides ISTR/JSTR/KSTR, IHEAT and the patch arrays
struct WallPatchDev {
    int face, dir, idir, kbot, a1, a2, b1, b2, na, sa, sb, sw, iheat, n;
    double *arr[NWARR];
};

struct WallEnv {
    double t0, difpre, xmom, ymom, zmom, gx, gy, gz, ground, altitude;
};

// in-plane index of cell (i, j, k) on the block faces normal to dir
__device__ __forceinline__ int wall_idx(const Blk &b, int dir, int i, int j, int k)
{
    return dir == 0 ? (k - 1) * b.jmax + (j - 1)
                    : (dir == 1 ? (k - 1) * b.imax + (i - 1) : (j - 1) * b.imax + (i - 1));
}

// The face fi (1..nd+1) along dir with in-plane index idx: a wall patch face takes the wall
// flux; with TURCOR the second and third faces from a wall add its corrections (F3RM(IF2) =
// F3RM(IF2) + F2VIG ..., base3c.f:6801-6804, 6849-6852).  Registration keeps nd >= 5 where
// both faces of a direction carry corrections, so they never meet on one face.
template <int NV>
__device__ __forceinline__ void wall_fix(const WallDev &w, int dir, int fi, int nd, int idx,
                                         FluxN<NV> &f)
{
    if (!w.on) return;
    if (fi == 1 || fi == nd + 1) {
        const int g = w.off[fi == 1 ? dir : dir + 3] + idx;
        if (w.flag[g]) {
#pragma unroll
            for (int q = 0; q < NV; q++) f.f[q] = w.flux[q * w.nt + g];
        }
        return;
    }
    if (!w.corr_on) return;
    int g = -1, s = 0;
    if (fi == 2 || fi == 3) {
        g = w.off[dir] + idx;
        s = fi - 2;
    } else if (fi == nd || fi == nd - 1) {
        g = w.off[dir + 3] + idx;
        s = nd - fi;
    }
    if (g >= 0 && w.flag[g]) {
#pragma unroll
        for (int c = 0; c < 4; c++) f.f[1 + c] = f.f[1 + c] + w.corr[(s * 4 + c) * w.nt + g];
    }
}

// BOTSKE loop 1000 (base3c.f:6322-6758) for one face of patch wp, IFLUX = 1, the viscous
// branch, POROS = 0, no wall functions, UROT = 0, and with TURCOR and ITURB < 9 the
// corrections of loop 6766-6897.  KW: the FLXCYC call (IDIS = 2, PSIGK1/PSIGE1 = SIGRK1/
// SIGOM1, CMU = 0.09); otherwise FLUCYC's (IDIS = 0, PSIG = CMU = 0, TURCOR = .FALSE.).
template <bool KW>
__global__ void __launch_bounds__(256) k_wall(Prm p, Blk b, WallDev w, WallPatchDev wp,
                                              WallEnv ev, int level)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= wp.n) return;
    const int nai = wp.a2 - wp.a1 + 1;
    const int I = wp.a1 + t % nai, J = wp.b1 + t / nai;
    const int LN = 2;
    const int II = (1 + I) * wp.sa + (1 + J) * wp.sb + (1 + wp.kbot) * wp.sw;
    const int lstrid = wp.sw * wp.idir;
    const int II2 = II + lstrid, IM1 = II - lstrid, IM2 = IM1 - lstrid;
    const int IF1 = wp.idir == 1 ? II : II + wp.sw;
    const int NP = (LN + J - wp.b1) * wp.na + (I - wp.a1 + LN);
    // natural (i, j, k) of the wall cell and its in-plane index
    int ci, cj, ck;
    if (wp.dir == 0) { ci = wp.kbot; cj = I; ck = J; }
    else if (wp.dir == 1) { ci = J; cj = wp.kbot; ck = I; }
    else { ci = I; cj = J; ck = wp.kbot; }
    const int g = w.off[wp.face] + wall_idx(b, wp.dir, ci, cj, ck);
    const double *A3 = FLD(b, FF_A1 + 4 * wp.dir), *A3X = FLD(b, FF_A1X + 4 * wp.dir),
                 *A3Y = FLD(b, FF_A1Y + 4 * wp.dir), *A3Z = FLD(b, FF_A1Z + 4 * wp.dir);
    const double *P = FLD(b, FF_P), *U = FLD(b, FF_U), *V = FLD(b, FF_V), *W = FLD(b, FF_W),
                 *VOL = FLD(b, FF_VOL), *VIS = FLD(b, FF_VIS), *EPS2 = FLD(b, FF_EPS2),
                 *VIST = FLD(b, FF_VIST), *PDI = FLD(b, FF_PDIFF), *RO = FLD(b, FF_RO),
                 *XC = FLD(b, FF_XC), *YC = FLD(b, FF_YC), *ZC = FLD(b, FF_ZC),
                 *TEMP = FLD(b, FF_TEMP), *CH = FLD(b, FF_CH), *STRAIN = FLD(b, FF_STRAIN),
                 *RK = FLD(b, FF_RK), *REPS = FLD(b, FF_REPS);
    double *const *pa = wp.arr;
    const int turcor = KW ? p.turcor : 0;
    const double eps3 = 1.E-20, xkappa = 0.4;
    double rc1 = 1.500, rc2 = rc1 - 1.;
    if (p.inter[wp.dir] == 6) { rc1 = 1.; rc2 = rc1 - 1.; }
    double psigk = 0., psige = 0., romfac = 0.0;
    if (KW) {
        const SstC sc = coefko(p.kover);
        psigk = sc.sigrk1; psige = sc.sigom1; romfac = 1.0;
    }
    const double repfac = 1.0 - romfac;
    const double cmu = KW ? 0.09 : 0.;
    const double prs = p.pr / p.prt;
    const double ymr = .166666667, yp1 = 1., yp4 = 0.;
    const double rsum = level == 1 ? 1. : 0.;
    const double eps = 1.E-3;
    const double t0 = ev.t0;
    int ihfl = 0, iht = 1;
    if (wp.iheat == 2 || wp.iheat >= 12) ihfl = 1;
    if (wp.iheat == 0) iht = 0;
    const double vislim = 0.01 * p.frsvis, rolim = .001 * p.frsden;
    const double xaux1 = KW ? 1. : 0.0, xaux2 = 0.0, xaux3 = 0.0;
    const double stra2 = STRAIN[II2], strai = STRAIN[II];
    const double dist = (A3[IF1] + A3[IF1 + lstrid]) / VOL[II];
    const double surg = -dist * ymr;
    const double eps2b = 1., vistb = 0.;
    const double visb = dmax(vislim, rc1 * VIS[II] - rc2 * VIS[II2]);
    const double chsb = dmax(eps, rc1 * CH[II] - rc2 * CH[II2]);
    const double straib = dmax(eps, rc1 * strai - rc2 * stra2);
    const double pb1 = dmax(0.01 * p.frspre, rc1 * P[II] - rc2 * P[II2]);
    const double rob = dmax(rolim, rc1 * RO[II] - rc2 * RO[II2]);
    const double epsur2 = dmax(eps, rc1 * REPS[II] - rc2 * REPS[II2]);
    const double tempb = dmax(0.01 * t0, rc1 * TEMP[II] - rc2 * TEMP[II2]);
    const double surf = surg * eps2b * visb;
    const double surfe = surg * (prs * vistb / visb + 1.) * chsb;
    const double rapub = dmax(1.0E-20, eps2b - 1.0);
    const double surfrk = surg * (1.0 + psigk * rapub) * visb;
    const double surfep = surg * (1.0 + psige * rapub) * visb;
    const double ypro1 = 1. / RO[II], ypro2 = 1. / RO[II2];
    double rksurf = 0.0;
    const double sqrk = (sqrt(RK[II] * ypro1) - sqrt(rksurf / rob)) * dist;
    const double epsur3 = 2. * visb * (sqrk * sqrk);
    double epsurf = xaux2 * epsur2 / rob + xaux3 * epsur3 / rob;
    const double rnub = visb / rob;
    const double utau2 = rnub * straib;
    const double plusfc = sqrt(utau2) / rnub;
    const double yplus1 = 2.0 * plusfc / dist;
    double skpmin = dmin(2.4 * pow(yplus1, 0.85), 8.0);
    skpmin = dmax(skpmin, eps);
    const double skplus = dmax(pa[WA_RBK][NP] * plusfc, skpmin);
    const double pskp = 1.0 / skplus;
    const double sr1 = 2500.0 * (pskp * pskp);
    const double sr2 = 100.0 * pskp;
    const double srcoef = dmax(sr1, sr2);
    const double omsurf = straib * srcoef;
    rksurf = 0.0;
    epsurf = romfac * omsurf + repfac * epsurf;
    const double usurf = pa[WA_UWALL][NP], vsurf = pa[WA_VWALL][NP], wsurf = pa[WA_WWALL][NP];
    const double temp1 = TEMP[II], temp2 = TEMP[II + lstrid];
    double tsurf = t0;
    if (wp.iheat == 1) {
        tsurf = pa[WA_TWALL][NP];
    } else if (wp.iheat != 3 && wp.iheat != 4) {
        tsurf = t0;
        pa[WA_TWALL][NP] = tempb;
    } else if (wp.iheat == 3) {
        tsurf = tempb;
        pa[WA_TWALL][NP] = tempb;
    }
    const double ihtd = (double)(iht * wp.idir), idd = (double)wp.idir;
    const double cvi2 = ihtd * (9.000 * temp1 - 1.000 * temp2 - 8. * tsurf);
    const double uvi2 = idd * (9.000 * U[II] - 1.000 * U[II2] - 8. * usurf);
    const double vvi2 = idd * (9.000 * V[II] - 1.000 * V[II2] - 8. * vsurf);
    const double wvi2 = idd * (9.000 * W[II] - 1.000 * W[II2] - 8. * wsurf);
    const double cvi1 = ihtd * (temp1 - tsurf);
    const double uvi1 = idd * (U[II] - usurf) * 6.;
    const double vvi1 = idd * (V[II] - vsurf) * 6.;
    const double wvi1 = idd * (W[II] - wsurf) * 6.;
    const double pb4 = dmax(-.99 * p.frspre, rc1 * PDI[II] - rc2 * PDI[II2]);
    const double pb = yp1 * pb1 + yp4 * pb4;
    pa[WA_CPWALL][NP] = pb4;
    // the force pass runs on a side stream while later stages rewrite P and PDIFF: it reads
    // the wall pressure of this residual from here
    w.qterm[2 * w.nt + g] = pb;
    double alpha = 1.;
    const double yp = 1. / dist;
    if (turcor) {
        const double du = U[II] - usurf, dv = V[II] - vsurf, dw = W[II] - wsurf;
        const double uvel = sqrt(du * du + dv * dv + dw * dw);
        const double para = yp * uvel * RO[II] / VIS[II];
        alpha = dmin(para / 140., 1.);
    }
    const double cvii = (1. - alpha) * cvi1 + alpha * cvi2;
    const double uvii = (1. - alpha) * uvi1 + alpha * uvi2;
    const double vvii = (1. - alpha) * vvi1 + alpha * vvi2;
    const double wvii = (1. - alpha) * wvi1 + alpha * wvi2;
    const double nx = A3X[IF1], ny = A3Y[IF1], nz = A3Z[IF1], an = A3[IF1];
    const double duc = .33333 * (ny * vvii + nz * wvii + nx * uvii);
    const double f2v = surf * (uvii + nx * duc);
    const double f3v = surf * (vvii + ny * duc);
    const double f4v = surf * (wvii + nz * duc);
    const double f5vf = usurf * f2v + vsurf * f3v + wsurf * f4v;
    const double f5vq = (double)(1 - ihfl) * surfe * cvii + (double)ihfl * pa[WA_HFLUX][NP];
    const double f5v = f5vf + f5vq;
    w.qterm[g] = an * rsum * f5v;
    w.qterm[w.nt + g] = an * rsum * f5vq;
    double evii = idd * (9.0 * REPS[II] * ypro1 - 1.0 * REPS[II2] * ypro2 - 8. * epsurf);
    double rkii = idd * (9.000 * RK[II] * ypro1 - 1.000 * RK[II2] * ypro2 - 8. * rksurf);
    rkii = idd * dmax(idd * rkii, 0.0);
    const double utau = sqrt(sqrt(f2v * f2v + f3v * f3v + f4v * f4v) / rob);
    double f6v, f7v;
    if (!turcor) {
        f6v = surfrk * rkii * dmax(romfac, xaux1);
        f7v = surfep * evii * dmax(romfac, xaux1);
    } else {
        const double t1 = rob * (utau * utau * utau) / xkappa / (yp * yp);
        const double t2 = 4. * visb * (utau * utau) / sqrt(cmu) / (yp * yp * yp);
        const double evik = idd * 6. * yp * dmin(0., -t1 + t2) * (1. + VIST[II] / VIS[II] * psige);
        f6v = surf * rkii * .5 * (1. + EPS2[II]);
        f7v = surf * (romfac * evii + repfac * evik);
        const double rou2 = RO[II] * (utau * utau);
        const double vv = VIS[II] + VIST[II];
        FLD(b, FF_PTUR)[II] = VIST[II] * (rou2 * rou2) / (vv * vv);
    }
    const double anxi = XC[II + wp.sa] - XC[II - wp.sa], anyi = YC[II + wp.sa] - YC[II - wp.sa];
    const double anzi = ZC[II + wp.sa] - ZC[II - wp.sa];
    const double xnori = 1. / sqrt(anxi * anxi + anyi * anyi + anzi * anzi + eps3);
    const double anxj = XC[II + wp.sb] - XC[II - wp.sb], anyj = YC[II + wp.sb] - YC[II - wp.sb];
    const double anzj = ZC[II + wp.sb] - ZC[II - wp.sb];
    const double xnorj = 1. / sqrt(anxj * anxj + anyj * anyj + anzj * anzj + eps3);
    pa[WA_TAUW1][NP] = -(anxi * f2v + anyi * f3v + anzi * f4v) * xnori;
    pa[WA_TAUW2][NP] = -(anxj * f2v + anyj * f3v + anzj * f4v) * xnorj;
    pa[WA_TAUWX][NP] = -f2v;
    pa[WA_TAUWY][NP] = -f3v;
    pa[WA_TAUWZ][NP] = -f4v;
    double fx[7];
    fx[0] = 0.;
    fx[1] = an * (nx * pb + f2v);
    fx[2] = an * (ny * pb + f3v);
    fx[3] = an * (nz * pb + f4v);
    fx[4] = an * (0. * pb1 + f5v + f6v);          // UROT = 0
    fx[5] = an * f6v;
    fx[6] = an * f7v;
#pragma unroll
    for (int q = 0; q < 7; q++) w.flux[q * w.nt + g] = fx[q];
    pa[WA_QWFRIC][NP] = idd * (f5vf + 0. * pb);
    pa[WA_QWALL][NP] = idd * f5vq;
    if (KW && turcor && p.iturb < 9) {
        // the second and third faces from the wall (base3c.f:6766-6897)
        const int II3 = II + 2 * lstrid, IF2 = IF1 + lstrid, IF3 = IF1 + 2 * lstrid;
        double ypd2 = A3[IF2] / (VOL[II2] + VOL[II]);
        double sg = -A3[IF2] * ypd2;
        double sf = sg * (-.33333 * VIST[II2] + VIST[II]);
        double sfe = sg * ((-.33333 * prs * VIST[II2] / VIS[II2]) * CH[II2]
                           + (prs * VIST[II] / VIS[II]) * CH[II]);
        double dc = .333333 * (A3X[IF2] * (U[II2] - U[II]) + A3Y[IF2] * (V[II2] - V[II])
                               + A3Z[IF2] * (W[II2] - W[II]));
        double c2 = sf * (U[II2] - U[II] + A3X[IF2] * dc);
        double c3 = sf * (V[II2] - V[II] + A3Y[IF2] * dc);
        double c4 = sf * (W[II2] - W[II] + A3Z[IF2] * dc);
        double c5 = .5 * ((U[II2] + U[II]) * c2 + (V[II2] + V[II]) * c3 + (W[II2] + W[II]) * c4)
                  + sfe * (TEMP[II2] - TEMP[II]);
        w.corr[0 * w.nt + g] = c2; w.corr[1 * w.nt + g] = c3;
        w.corr[2 * w.nt + g] = c4; w.corr[3 * w.nt + g] = c5;
        ypd2 = A3[IF3] / (VOL[II3] + VOL[II2]);
        sg = -A3[IF3] * ypd2;
        sf = sg * (-.2 * VIST[II3] + .3333 * VIST[II2]);
        sfe = sg * ((-.2 * prs * VIST[II3] / VIS[II3]) * CH[II3]
                    + (+.3333 * prs * VIST[II2] / VIS[II2]) * CH[II2]);
        dc = .2 * (A3X[IF3] * (U[II3] - U[II2]) + A3Y[IF3] * (V[II3] - V[II2])
                   + A3Z[IF3] * (W[II3] - W[II2]));
        c2 = sf * (U[II3] - U[II2] + A3X[IF3] * dc);
        c3 = sf * (V[II3] - V[II2] + A3Y[IF3] * dc);
        c4 = sf * (W[II3] - W[II2] + A3Z[IF3] * dc);
        c5 = .5 * ((U[II3] + U[II2]) * c2 + (V[II3] + V[II2]) * c3 + (W[II3] + W[II2]) * c4)
           + sfe * (TEMP[II3] - TEMP[II2]);
        w.corr[4 * w.nt + g] = c2; w.corr[5 * w.nt + g] = c3;
        w.corr[6 * w.nt + g] = c4; w.corr[7 * w.nt + g] = c5;
    }
}

// BOTSKE loop 2000 (base3c.f:6988-7093) on level 1: the force sums of every patch (one
// workgroup per patch) in the reference's face order.  A chunk of 256 faces is prepared by
// the threads in parallel (moment-arm products, pressure parts) into LDS; then lane q < 19
// of the first wave carries sum q through the chunk's faces in order, with the reference's
// association (e.g. CMX = (CMX + BZ*YD) - BY*ZD, taken as two additions of the signed
// terms, which is the same IEEE operation), so the 19 sequential chains run side by side;
// every lane leaves its running sum per face in LDS, and every thread forms SURFPX.. of its
// face from the running DX, DY, DZ.
constexpr int WF_T = 29;
constexpr int WF_N = 256;
// per force sum (enum ff_force order): first term and its sign, second term and its sign.
// A sum of one term takes slot 28 (zero) as its second: adding +0. leaves every value the
// sum can hold unchanged (it starts at +0. and so never holds -0.), and QY's "+ 0." of the
// rotating-frame term (UROT = 0) is exactly that.
__constant__ signed char kWfTerms[FF_NFORCE][4] = {
    {6, 1, 28, 1}, {7, 1, 28, 1}, {8, 1, 28, 1},                     // CX CY CZ
    {0, 1, 1, -1}, {2, 1, 3, -1}, {4, 1, 5, -1},                     // CMX CMY CMZ
    {9, -1, 28, 1},                                                  // TOM
    {10, -1, 28, 1}, {11, -1, 28, 1}, {12, -1, 28, 1},               // DX DY DZ
    {22, -1, 28, 1}, {23, 1, 28, 1}, {24, 1, 28, 1},                 // QX QY QZ
    {13, 1, 14, -1}, {15, 1, 16, -1}, {17, 1, 18, -1},               // VMX VMY VMZ
    {19, 1, 28, 1}, {20, 1, 28, 1}, {21, 1, 28, 1}};                 // BUX BUY BUZ
__global__ void __launch_bounds__(WF_N) k_wall_forces(Prm p, Blk b, WallDev w,
                                                      const WallPatchDev *wps, WallEnv ev,
                                                      double *sums)
{
    const WallPatchDev wp = wps[blockIdx.x];
    const int tid = threadIdx.x;
    // rows padded by one so that the 19 lanes, each reading its own row at the same face,
    // hit different banks
    __shared__ double tm[WF_T][WF_N + 1];
    __shared__ double run[FF_NFORCE][WF_N + 1];
    __shared__ double fin[FF_NFORCE];
    const double *A3 = FLD(b, FF_A1 + 4 * wp.dir), *A3X = FLD(b, FF_A1X + 4 * wp.dir),
                 *A3Y = FLD(b, FF_A1Y + 4 * wp.dir), *A3Z = FLD(b, FF_A1Z + 4 * wp.dir);
    double *const *pa = wp.arr;
    const int LN = 2, nai = wp.a2 - wp.a1 + 1;
    const double ompor = 1. - 0.0, ypg = ev.difpre;
    const double idd = (double)wp.idir;
    // the sum this lane carries
    const int q = tid < FF_NFORCE ? tid : 0;
    const int ia = kWfTerms[q][0], ib = kWfTerms[q][2];
    const double sa = (double)kWfTerms[q][1], sb = (double)kWfTerms[q][3];
    double acc = 0.;
    for (int c0 = 0; c0 < wp.n; c0 += WF_N) {
        const int t = c0 + tid;
        int NP = 0;
        double an = 0.;
        if (t < wp.n) {
            const int I = wp.a1 + t % nai, J = wp.b1 + t / nai;
            const int II = (1 + I) * wp.sa + (1 + J) * wp.sb + (1 + wp.kbot) * wp.sw;
            const int IF1 = wp.idir == 1 ? II : II + wp.sw;
            NP = (LN + J - wp.b1) * wp.na + (I - wp.a1 + LN);
            int ci, cj, ck;
            if (wp.dir == 0) { ci = wp.kbot; cj = I; ck = J; }
            else if (wp.dir == 1) { ci = J; cj = wp.kbot; ck = I; }
            else { ci = I; cj = J; ck = wp.kbot; }
            const int g = w.off[wp.face] + wall_idx(b, wp.dir, ci, cj, ck);
            an = A3[IF1];
            const double nx = A3X[IF1], ny = A3Y[IF1], nz = A3Z[IF1];
            const double xd = pa[WA_XCP][NP] - ev.xmom, yd = pa[WA_YCP][NP] - ev.ymom;
            const double zd = pa[WA_ZCP][NP] - ev.zmom;
            double deltap = p.frsden * (ev.gx * (pa[WA_XCP][NP] - ev.ground)
                                        + ev.gy * (pa[WA_YCP][NP] - ev.ground)
                                        + ev.gz * (pa[WA_ZCP][NP] - ev.ground));
            if (ev.altitude >= ev.ground) deltap = 0.0;
            const double frm = w.flux[1 * w.nt + g], frn = w.flux[2 * w.nt + g];
            const double frw = w.flux[3 * w.nt + g], fe = w.flux[4 * w.nt + g];
            const double frk = w.flux[5 * w.nt + g];
            double bx = -idd * (frm - an * ompor * nx * ypg);
            double by = -idd * (frn - an * ompor * ny * ypg);
            double bz = -idd * (frw - an * ompor * nz * ypg);
            const double bxdp = -idd * (-an * ompor * nx * (-deltap));
            const double bydp = -idd * (-an * ompor * ny * (-deltap));
            const double bzdp = -idd * (-an * ompor * nz * (-deltap));
            bx = bx + bxdp;
            by = by + bydp;
            bz = bz + bzdp;
            pa[WA_SURFX][NP] = bx / (an + 1.E-20);
            pa[WA_SURFY][NP] = by / (an + 1.E-20);
            pa[WA_SURFZ][NP] = bz / (an + 1.E-20);
            // PB = YP1*PBB1 + YP4*PBB4 as k_wall formed it from P and PDIFF at II, II2
            const double pb = w.qterm[2 * w.nt + g];
            const double qx_ = an * ompor * nx * (pb - ypg), qy_ = an * ompor * ny * (pb - ypg);
            const double qz_ = an * ompor * nz * (pb - ypg);
            const double bx2 = bx + idd * qx_, by2 = by + idd * qy_, bz2 = bz + idd * qz_;
            const double v[WF_T] = {bz * yd, by * zd, bx * zd, bz * xd, by * xd, bx * yd,
                                    bx, by, bz, idd * (fe - frk), idd * qx_, idd * qy_, idd * qz_,
                                    bz2 * yd, by2 * zd, bx2 * zd, bz2 * xd, by2 * xd, bx2 * yd,
                                    bxdp, bydp, bzdp, idd * fe, w.qterm[g], w.qterm[w.nt + g],
                                    -idd * qx_, -idd * qy_, -idd * qz_, 0.};
#pragma unroll
            for (int k = 0; k < WF_T; k++) tm[k][tid] = v[k];
        }
        __syncthreads();
        const int m = min(WF_N, wp.n - c0);
        if (tid < FF_NFORCE) {
            // every lane runs the same two additions per face: no divergent paths
#pragma unroll 4
            for (int r = 0; r < m; r++) {
                acc = acc + sa * tm[ia][r];
                acc = acc + sb * tm[ib][r];
                run[q][r] = acc;
            }
        }
        __syncthreads();
        if (t < wp.n) {
            pa[WA_SURFPX][NP] = (run[FF_DX][tid] + tm[25][tid]) / (an + 1.E-20);
            pa[WA_SURFPY][NP] = (run[FF_DY][tid] + tm[26][tid]) / (an + 1.E-20);
            pa[WA_SURFPZ][NP] = (run[FF_DZ][tid] + tm[27][tid]) / (an + 1.E-20);
        }
        __syncthreads();
    }
    if (tid < FF_NFORCE) fin[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        double s[FF_NFORCE];
#pragma unroll
        for (int k = 0; k < FF_NFORCE; k++) s[k] = fin[k];
        s[FF_TOM] = s[FF_TOM] + idd * s[FF_QZ];
        s[FF_QZ] = idd * s[FF_QZ];
        s[FF_QY] = idd * s[FF_QY];
        s[FF_DX] = s[FF_DX] + s[FF_BUX];
        s[FF_DY] = s[FF_DY] + s[FF_BUY];
        s[FF_DZ] = s[FF_DZ] + s[FF_BUZ];
#pragma unroll
        for (int k = 0; k < FF_NFORCE; k++) sums[blockIdx.x * FF_NFORCE + k] = s[k];
    }
}

// RESIKE face fluxes (FLXCYC res3c.f:6424-7566 for k, i, j with SWEEPK) and SOURKO/ADDSKE
// (res3c.f:4957-5068), with every face flux formed once: the layout of k_residual_tile with
// seven components.  FNS: the FULLNS face fluxes FLKEBL + FLUXBL (level 1).
template <int RIW, int RJW, int RKW, bool FNS = false>
__global__ void __launch_bounds__(RIW * RJW) k_residual_kw(Prm p, Blk b, WallDev wd, TileGrid tg)
{
    __shared__ double fi[7][RJW][RIW + 1];
    __shared__ double fj[7][RJW + 1][RIW];
    __shared__ double fk[7][RIW * RJW];
    const SstC sc = coefko(p.kover);
    const int ti = threadIdx.x % RIW, tj = threadIdx.x / RIW;
    const int tl = xcd_tile(tg);
    if (tl < 0) return;
    const int bx = tl % tg.nx, by = (tl / tg.nx) % tg.ny, bz = tl / (tg.nx * tg.ny);
    const int i0 = bx * RIW + 1, j0 = by * RJW + 1, k0 = bz * RKW + 1;
    const int i = i0 + ti, j = j0 + tj;
    const bool cell = i <= b.imax && j <= b.jmax;
    const int il = min(RIW, b.imax - i0 + 1), jl = min(RJW, b.jmax - j0 + 1);
    const int k1 = min(k0 + RKW - 1, b.kmax);
    const double *a3 = FLD(b, FF_A3), *a3x = FLD(b, FF_A3X), *a3y = FLD(b, FF_A3Y), *a3z = FLD(b, FF_A3Z);
    const double *a1 = FLD(b, FF_A1), *a1x = FLD(b, FF_A1X), *a1y = FLD(b, FF_A1Y), *a1z = FLD(b, FF_A1Z);
    const double *a2 = FLD(b, FF_A2), *a2x = FLD(b, FF_A2X), *a2y = FLD(b, FF_A2Y), *a2z = FLD(b, FF_A2Z);
    const double *d1 = FLD(b, FF_D1), *d2 = FLD(b, FF_D2), *d3 = FLD(b, FF_D3);
    // One call site per face kind keeps the kernel's code within the instruction cache: the
    // plane loop starts one plane early to form the tile's lowest K-faces, and a plane's I-
    // and J-faces go through one rolled loop.
    for (int k = k0 - 1; k <= k1; k++) {
        if (k < k0) {
            if (cell) {
                Flux7 kf = face_flux_kw<true, FNS>(p, sc, b, cid(b, i, j, k + 1), b.il, p.inter[2], a3, a3x, a3y, a3z, d3);
                wall_fix<7>(wd, 2, k + 1, b.kmax, wall_idx(b, 2, i, j, 0), kf);
#pragma unroll
                for (int q = 0; q < 7; q++) fk[q][threadIdx.x] = kf.f[q];
            }
            continue;
        }
        const int n = cid(b, i, j, k);
        // t = 0: the cell's lower I-face, 1: its lower J-face, 2: the tile's upper J-face
        // (row 0, one wave) or upper I-face (row 1)
#pragma unroll 1
        for (int t = 0; t < 3; t++) {
            bool act = cell;
            int L = n, dir = t, fx = i, fy = j;
            double *dst = t == 0 ? &fi[0][tj][ti] : &fj[0][tj][ti];
            if (t == 2) {
                act = (tj == 0 && ti < il) || (tj == 1 && ti < jl);
                dir = tj == 0 ? 1 : 0;
                fx = tj == 0 ? i : i0 + il;
                fy = tj == 0 ? j0 + jl : j0 + ti;
                L = cid(b, fx, fy, k);
                dst = tj == 0 ? &fj[0][jl][ti] : &fi[0][ti][il];
            }
            if (!act) continue;
            const bool di = dir == 0;
            const int qs = di ? RJW * (RIW + 1) : (RJW + 1) * RIW;
            Flux7 f = face_flux_kw<false, FNS>(p, sc, b, L, di ? 1 : b.is, di ? p.inter[0] : p.inter[1],
                                               di ? a1 : a2, di ? a1x : a2x, di ? a1y : a2y,
                                               di ? a1z : a2z, di ? d1 : d2);
            wall_fix<7>(wd, dir, di ? fx : fy, di ? b.imax : b.jmax,
                        di ? wall_idx(b, 0, 0, fy, k) : wall_idx(b, 1, fx, 0, k), f);
#pragma unroll
            for (int q = 0; q < 7; q++) dst[q * qs] = f.f[q];
        }
        __syncthreads();
        if (cell) {
            // the cell's upper K-face after its I- and J-faces: the K-face of the plane below
            // waits in LDS, so no flux is held in registers while the others are formed
            Flux7 khi = face_flux_kw<true, FNS>(p, sc, b, cid(b, i, j, k + 1), b.il, p.inter[2], a3, a3x, a3y, a3z, d3);
            wall_fix<7>(wd, 2, k + 1, b.kmax, wall_idx(b, 2, i, j, 0), khi);
            double dt1 = 1. / FLD(b, FF_VOL)[n];
            double d[7];
#pragma unroll
            for (int q = 0; q < 7; q++) {
                double x = 0.;
                x = x - (khi.f[q] - fk[q][threadIdx.x]);                // SWEEPK, K
                x = x - (fi[q][tj][ti + 1] - fi[q][tj][ti]);             // SWEEPK, I
                d[q] = (x - (fj[q][tj + 1][ti] - fj[q][tj][ti])) * dt1;  // SWEEPK, J, /VOL
                fk[q][threadIdx.x] = khi.f[q];
            }
            double srk, som;
            sourko_cell(p, b, n, srk, som);
            FLD(b, FF_SRK)[n] = srk;
            FLD(b, FF_SEPS)[n] = som;
            d[5] = d[5] + srk;                                          // ADDSKE
            d[6] = d[6] + som;
#pragma unroll
            for (int q = 0; q < 7; q++) FLD(b, FF_DRO + q)[n] = d[q];
        }
        __syncthreads();
    }
}

// Thread per interior cell.  The residual of a cell is
//   D = ((0 - (Fk(k+1) - Fk(k))) - (Fi(i+1) - Fi(i)) - (Fj(j+1) - Fj(j))) * (1/VOL)
// in the reference's order (SWEEP after the k, i and j directions).  Both faces of the
// cell are evaluated by the same code as the neighbour evaluates them, so the shared face
// values agree bit for bit.
template <int IFLUX, bool FNS = false>
__global__ void __launch_bounds__(128) k_residual(Prm p, Blk b, WallDev wd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > b.imax) return;
    int n = cid(b, i, j, k);
    double d[5] = {0., 0., 0., 0., 0.};
    {   // k-direction (res3c.f:270, SWEEP res3c.f:330)
        const int S = b.il;
        Flux5 lo = face_flux<true, IFLUX, FNS>(p, b, n, S, p.inter[2], FLD(b, FF_A3), FLD(b, FF_A3X),
                                   FLD(b, FF_A3Y), FLD(b, FF_A3Z), FLD(b, FF_D3));
        Flux5 hi = face_flux<true, IFLUX, FNS>(p, b, n + S, S, p.inter[2], FLD(b, FF_A3), FLD(b, FF_A3X),
                                   FLD(b, FF_A3Y), FLD(b, FF_A3Z), FLD(b, FF_D3));
        wall_fix<5>(wd, 2, k, b.kmax, wall_idx(b, 2, i, j, 0), lo);
        wall_fix<5>(wd, 2, k + 1, b.kmax, wall_idx(b, 2, i, j, 0), hi);
#pragma unroll
        for (int q = 0; q < 5; q++) d[q] = d[q] - (hi.f[q] - lo.f[q]);
    }
    {   // i-direction (res3c.f:402, SWEEP res3c.f:448)
        const int S = 1;
        Flux5 lo = face_flux<false, IFLUX, FNS>(p, b, n, S, p.inter[0], FLD(b, FF_A1), FLD(b, FF_A1X),
                                    FLD(b, FF_A1Y), FLD(b, FF_A1Z), FLD(b, FF_D1));
        Flux5 hi = face_flux<false, IFLUX, FNS>(p, b, n + S, S, p.inter[0], FLD(b, FF_A1), FLD(b, FF_A1X),
                                    FLD(b, FF_A1Y), FLD(b, FF_A1Z), FLD(b, FF_D1));
        wall_fix<5>(wd, 0, i, b.imax, wall_idx(b, 0, 0, j, k), lo);
        wall_fix<5>(wd, 0, i + 1, b.imax, wall_idx(b, 0, 0, j, k), hi);
#pragma unroll
        for (int q = 0; q < 5; q++) d[q] = d[q] - (hi.f[q] - lo.f[q]);
    }
    {   // j-direction, last: divide by VOL (res3c.f:515, SWEEP res3c.f:559)
        const int S = b.is;
        Flux5 lo = face_flux<false, IFLUX, FNS>(p, b, n, S, p.inter[1], FLD(b, FF_A2), FLD(b, FF_A2X),
                                    FLD(b, FF_A2Y), FLD(b, FF_A2Z), FLD(b, FF_D2));
        Flux5 hi = face_flux<false, IFLUX, FNS>(p, b, n + S, S, p.inter[1], FLD(b, FF_A2), FLD(b, FF_A2X),
                                    FLD(b, FF_A2Y), FLD(b, FF_A2Z), FLD(b, FF_D2));
        wall_fix<5>(wd, 1, j, b.jmax, wall_idx(b, 1, i, 0, k), lo);
        wall_fix<5>(wd, 1, j + 1, b.jmax, wall_idx(b, 1, i, 0, k), hi);
        double dt1 = 1. / FLD(b, FF_VOL)[n];
#pragma unroll
        for (int q = 0; q < 5; q++) d[q] = (d[q] - (hi.f[q] - lo.f[q])) * dt1;
    }
    FLD(b, FF_DRO)[n] = d[0];
    FLD(b, FF_DM)[n] = d[1];
    FLD(b, FF_DN)[n] = d[2];
    FLD(b, FF_DW)[n] = d[3];
    FLD(b, FF_DE)[n] = d[4];
}

// Residual with every face flux formed once.  A workgroup is a tile of RI x RJ cells in
// (I,J) that marches over RK consecutive K-planes.  At each plane a lane forms the K-face
// above its cell (kept in registers: it is the lower K-face of the next plane), its lower
// I-face and its lower J-face; the tile's upper I- and J-faces are formed by two waves in an
// extra phase.  I- and J-faces meet their two cells through LDS.  A face flux is the same
// function of the same operands whichever cell forms it, so the differences are those of
// k_residual bit for bit.
constexpr int RI = 64, RJ = 4, RKC = 8;
static_assert(RJ >= 2 && RJ <= RI, "the tile's upper I-faces are formed by the lanes of row 1");

template <int IFLUX, bool FNS = false, int RI = ::RI, int RJ = ::RJ, int RKC = ::RKC>
__global__ void __launch_bounds__(RI * RJ) k_residual_tile(Prm p, Blk b, WallDev wd, TileGrid tg)
{
    static_assert(RJ >= 2 && RI >= RJ, "the tile's upper faces need two rows and RI >= RJ");
    __shared__ double fi[5][RJ][RI + 1];
    __shared__ double fj[5][RJ + 1][RI];
    __shared__ double fk[5][RI * RJ];
    const int ti = threadIdx.x % RI, tj = threadIdx.x / RI;
    const int tl = xcd_tile(tg);
    if (tl < 0) return;
    const int bx = tl % tg.nx, by = (tl / tg.nx) % tg.ny, bz = tl / (tg.nx * tg.ny);
    const int i0 = bx * RI + 1, j0 = by * RJ + 1, k0 = bz * RKC + 1;
    const int i = i0 + ti, j = j0 + tj;
    const bool cell = i <= b.imax && j <= b.jmax;
    const int il = min(RI, b.imax - i0 + 1), jl = min(RJ, b.jmax - j0 + 1);   // cells in the tile
    const int k1 = min(k0 + RKC - 1, b.kmax);
    const double *a3 = FLD(b, FF_A3), *a3x = FLD(b, FF_A3X), *a3y = FLD(b, FF_A3Y), *a3z = FLD(b, FF_A3Z);
    const double *a1 = FLD(b, FF_A1), *a1x = FLD(b, FF_A1X), *a1y = FLD(b, FF_A1Y), *a1z = FLD(b, FF_A1Z);
    const double *a2 = FLD(b, FF_A2), *a2x = FLD(b, FF_A2X), *a2y = FLD(b, FF_A2Y), *a2z = FLD(b, FF_A2Z);
    const double *d1 = FLD(b, FF_D1), *d2 = FLD(b, FF_D2), *d3 = FLD(b, FF_D3);
    // one call site per face kind (the instruction cache): the plane loop starts one plane
    // early to form the tile's lowest K-faces; a plane's I- and J-faces share a rolled loop.
    // A cell's upper K-face is formed after its I- and J-faces and kept in LDS (fk) for the
    // plane above, so no flux is held in registers while another is formed.
    for (int k = k0 - 1; k <= k1; k++) {
        if (k < k0) {
            if (cell) {
                Flux5 kf = face_flux<true, IFLUX, FNS>(p, b, cid(b, i, j, k + 1), b.il, p.inter[2], a3, a3x, a3y, a3z, d3);
                wall_fix<5>(wd, 2, k + 1, b.kmax, wall_idx(b, 2, i, j, 0), kf);
#pragma unroll
                for (int q = 0; q < 5; q++) fk[q][threadIdx.x] = kf.f[q];
            }
            continue;
        }
        const int n = cid(b, i, j, k);
        // t = 0: the cell's lower I-face, 1: its lower J-face, 2: the tile's upper faces,
        // the J-face above row jl-1 (wave 0) and the I-face after column il-1 of every row
        // (wave 1)
#pragma unroll 1
        for (int t = 0; t < 3; t++) {
            bool act = cell;
            int L = n, dir = t, fx = i, fy = j;
            double *dst = t == 0 ? &fi[0][tj][ti] : &fj[0][tj][ti];
            if (t == 2) {
                act = (tj == 0 && ti < il) || (tj == 1 && ti < jl);
                dir = tj == 0 ? 1 : 0;
                fx = tj == 0 ? i : i0 + il;
                fy = tj == 0 ? j0 + jl : j0 + ti;
                L = cid(b, fx, fy, k);
                dst = tj == 0 ? &fj[0][jl][ti] : &fi[0][ti][il];
            }
            if (!act) continue;
            const bool di = dir == 0;
            const int qs = di ? RJ * (RI + 1) : (RJ + 1) * RI;
            Flux5 f = face_flux<false, IFLUX, FNS>(p, b, L, di ? 1 : b.is, di ? p.inter[0] : p.inter[1],
                                                   di ? a1 : a2, di ? a1x : a2x, di ? a1y : a2y,
                                                   di ? a1z : a2z, di ? d1 : d2);
            wall_fix<5>(wd, dir, di ? fx : fy, di ? b.imax : b.jmax,
                        di ? wall_idx(b, 0, 0, fy, k) : wall_idx(b, 1, fx, 0, k), f);
#pragma unroll
            for (int q = 0; q < 5; q++) dst[q * qs] = f.f[q];
        }
        __syncthreads();
        if (cell) {
            Flux5 khi = face_flux<true, IFLUX, FNS>(p, b, cid(b, i, j, k + 1), b.il, p.inter[2], a3, a3x, a3y, a3z, d3);
            wall_fix<5>(wd, 2, k + 1, b.kmax, wall_idx(b, 2, i, j, 0), khi);
            double dt1 = 1. / FLD(b, FF_VOL)[n];
#pragma unroll
            for (int q = 0; q < 5; q++) {
                double d = 0.;
                d = d - (khi.f[q] - fk[q][threadIdx.x]);                // SWEEP, K (res3c.f:330)
                d = d - (fi[q][tj][ti + 1] - fi[q][tj][ti]);             // SWEEP, I (res3c.f:448)
                d = (d - (fj[q][tj + 1][ti] - fj[q][tj][ti])) * dt1;     // SWEEP, J, /VOL (res3c.f:559)
                FLD(b, FF_DRO + q)[n] = d;
                fk[q][threadIdx.x] = khi.f[q];
            }
        }
        __syncthreads();
    }
}

// ZEROZZ (res3c.f:9330) of the nq residual fields DRO.. on every cell, one launch for all
__global__ void __launch_bounds__(256) k_zero(Blk b, int nq)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b.ntot) return;
    for (int q = 0; q < nq; q++) FLD(b, FF_DRO + q)[t] = 0.;
}

// RESL2 stage 1: per-workgroup partial sums of D^2 over interior cells (res3c.f:8979-9006);
// NQ = 7 adds D(rho k)^2 and MIN(|D(rho omega)|,1E15)^2 of the two-equation models
constexpr int RED_THREADS = 256;
template <int NQ>
__global__ void __launch_bounds__(RED_THREADS) k_resl2_part(Blk b, double *part)
{
    __shared__ double sh[NQ][RED_THREADS];
    int ncell = b.imax * b.jmax * b.kmax;
    double s[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) s[q] = 0.;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += gridDim.x * blockDim.x) {
        int i = t % b.imax + 1;
        int r = t / b.imax;
        int j = r % b.jmax + 1;
        int k = r / b.jmax + 1;
        int n = cid(b, i, j, k);
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double x = FLD(b, FF_DRO + q)[n];
            if (q == 6) x = dmin(fabs(x), 1.E15);
            s[q] += x * x;
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int h = RED_THREADS / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
#pragma unroll
            for (int q = 0; q < NQ; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) part[7 * blockIdx.x + q] = sh[q][0];
    }
}

// RESL2 stage 2: fixed-order sum of the partials (run-to-run deterministic): thread t sums
// partials t, t + RED_THREADS, ... in order, then the same halving tree as stage 1
__global__ void __launch_bounds__(RED_THREADS) k_resl2_final(const double *part, int nparts, int nq,
                                                            double *out)
{
    __shared__ double sh[7][RED_THREADS];
    const int t = threadIdx.x;
    double s[7];
#pragma unroll
    for (int q = 0; q < 7; q++) s[q] = 0.;
    for (int i = t; i < nparts; i += RED_THREADS) {
#pragma unroll
        for (int q = 0; q < 7; q++)
            if (q < nq) s[q] += part[7 * i + q];
    }
#pragma unroll
    for (int q = 0; q < 7; q++) sh[q][t] = s[q];
    __syncthreads();
    for (int h = RED_THREADS / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int q = 0; q < 7; q++) sh[q][t] += sh[q][t + h];
        }
        __syncthreads();
    }
    if (t < nq) out[t] = sh[t][0];
}

// TIME3 (base3c.f:79-220) + LOCDT (imps3.f:792-798) + COPR (imps3.f:902-927), thread per
// interior cell
__global__ void __launch_bounds__(256) k_time3_copr(Prm p, Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > b.imax) return;
    const int IS = b.is, IL = b.il;
    int kk = cid(b, i, j, k);
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3),
                 *a1x = FLD(b, FF_A1X), *a1y = FLD(b, FF_A1Y), *a1z = FLD(b, FF_A1Z),
                 *a2x = FLD(b, FF_A2X), *a2y = FLD(b, FF_A2Y), *a2z = FLD(b, FF_A2Z),
                 *a3x = FLD(b, FF_A3X), *a3y = FLD(b, FF_A3Y), *a3z = FLD(b, FF_A3Z);
    const double eps = 1.E-10, pvisc = 1.8;
    const double pcfl = 1. / p.cfl;
    const double prs = p.pr / p.prt;
    double z = .5 * (a1[kk] + a1[kk + 1]);
    double r2 = .5 * (a2[kk] + a2[kk + IS]);
    z = r2 > z ? r2 : z;
    double r3 = .5 * (a3[kk] + a3[kk + IL]);
    z = r3 > z ? r3 : z;
    double vol = FLD(b, FF_VOL)[kk], ro = FLD(b, FF_RO)[kk];
    double u = FLD(b, FF_U)[kk], v = FLD(b, FF_V)[kk], w = FLD(b, FF_W)[kk];
    double c = FLD(b, FF_C)[kk], P = FLD(b, FF_P)[kk];
    double pvol = 1. / vol;
    double ypro = 1. / (ro + eps);
    double a1k = .5 * (a1[kk] + a1[kk + 1]);
    double a2k = .5 * (a2[kk] + a2[kk + IS]);
    double a3k = .5 * (a3[kk] + a3[kk + IL]);
    double a1xk = .5 * (a1x[kk] * a1[kk] + a1x[kk + 1] * a1[kk + 1]);
    double a1yk = .5 * (a1y[kk] * a1[kk] + a1y[kk + 1] * a1[kk + 1]);
    double a1zk = .5 * (a1z[kk] * a1[kk] + a1z[kk + 1] * a1[kk + 1]);
    double a2xk = .5 * (a2x[kk] * a2[kk] + a2x[kk + IS] * a2[kk + IS]);
    double a2yk = .5 * (a2y[kk] * a2[kk] + a2y[kk + IS] * a2[kk + IS]);
    double a2zk = .5 * (a2z[kk] * a2[kk] + a2z[kk + IS] * a2[kk + IS]);
    double a3xk = .5 * (a3x[kk] * a3[kk] + a3x[kk + IL] * a3[kk + IL]);
    double a3yk = .5 * (a3y[kk] * a3[kk] + a3y[kk + IL] * a3[kk + IL]);
    double a3zk = .5 * (a3z[kk] * a3[kk] + a3z[kk + IL] * a3[kk + IL]);
    double ucii = a1xk * u + a1yk * v + a1zk * w - 0.;
    double vcii = a2xk * u + a2yk * v + a2zk * w - 0.;
    double wcii = a3xk * u + a3yk * v + a3zk * w - 0.;
    double b1 = pvol * (fabs(ucii) + fabs(vcii) + fabs(wcii)
                        + c * sqrt(a1k * a1k + a2k * a2k + a3k * a3k)) * pcfl;
    double rnum = (z * z) * (pvol * pvol) * ypro;
    double t = rnum * (FLD(b, FF_VIS)[kk] + prs * FLD(b, FF_VIST)[kk]) / p.cdiff;
    double b2 = 2. * t * pvisc;
    double b11 = b1 > b2 ? b1 : b2;
    double x = (p.frspre - 2. * P) / P;
    double denlow = 1. + (0. > x ? 0. : x);
    double y = (ro - 2. * p.frsden) * ypro;
    double denmax = 1. + (0. > y ? 0. : y) * 5.;
    denlow = denlow > denmax ? denlow : denmax;
    double dtl = 1. / (b11 * denlow);
    FLD(b, FF_DTL)[kk] = dtl;
    if (p.iturb == 6) {
        // LOCDTK (imps3.f:820-826, DTURB = 1)
        const double dturb = 1.;
        double dri = dtl * FLD(b, FF_DRO)[kk];
        double dmi = dtl * FLD(b, FF_DM)[kk];
        double dni = dtl * FLD(b, FF_DN)[kk];
        double dwi = dtl * FLD(b, FF_DW)[kk];
        double dre = dtl * FLD(b, FF_DE)[kk];
        double dki = dturb * dtl * FLD(b, FF_DRK)[kk];
        double dep = dturb * dtl * FLD(b, FF_DEPS)[kk];
        // IMSOKO (imps3.f:972-1020)
        const double rk = FLD(b, FF_RK)[kk], rom = FLD(b, FF_REPS)[kk];
        {
            const SstC s = coefko(p.kover);
            const double epsi = 1.0E-10;
            const double gamkom = s.beta1 / s.bstar - s.sigom1 * (s.ckappa * s.ckappa) / (sqrt(s.bstar));
            const double ckmax = 1.0 / (2.0 * p.csimps), comax = 1.0 / (2.0 * p.csimps);
            double emeanf = FLD(b, FF_E)[kk] - rk;
            double drkmax = dmax(ckmax * dmin(rk, fabs(0.1 * emeanf - 0.9 * rk)), epsi);
            double dommax = comax * rom;
            double prol = 1. / ro;
            double atprod = fabs(FLD(b, FF_PTUR)[kk]);
            double rnutrb = dmax((FLD(b, FF_EPS2)[kk] - 1.0), epsi) * FLD(b, FF_VIS)[kk] * prol;
            double drkden = 1.0 + dtl * (atprod / drkmax + prol * s.bstar * rom);
            double domden = 1.0 + dtl * (gamkom * atprod / (rnutrb * dommax) + 2. * s.beta1 * rom);
            dki = dki / drkden;
            dep = dep / domden;
        }
        // COPRKE (imps3.f:1145-1205, SOLUTION_TYPE = FLUID)
        double vel2 = (u * u + v * v + w * w);
        double drodh = FLD(b, FF_DRDH)[kk], drodp = FLD(b, FF_DRDP)[kk];
        double yp = 1. / ro;
        double rki = rk * yp;
        double dkin = u * dmi + v * dni + w * dwi + dki;
        FLD(b, FF_DM)[kk] = yp * (-u * dri + dmi);
        FLD(b, FF_DN)[kk] = yp * (-v * dri + dni);
        FLD(b, FF_DW)[kk] = yp * (-w * dri + dwi);
        FLD(b, FF_DRK)[kk] = yp * (-rki * dri + dki);
        FLD(b, FF_DEPS)[kk] = yp * (-rom * yp * dri + dep);
        double hi = yp * (FLD(b, FF_E)[kk] + P);
        double dei = dre - (hi - vel2 - rki) * dri - dkin;
        double dpi = (c * c) * (dri - drodh * yp * dei);
        double dhi = (dri - drodp * dpi) / drodh;
        FLD(b, FF_DRO)[kk] = dhi;
        FLD(b, FF_DE)[kk] = dpi + .667 * dki;
        return;
    }
    // LOCDT
    double dri = dtl * FLD(b, FF_DRO)[kk];
    double dmi = dtl * FLD(b, FF_DM)[kk];
    double dni = dtl * FLD(b, FF_DN)[kk];
    double dwi = dtl * FLD(b, FF_DW)[kk];
    double dre = dtl * FLD(b, FF_DE)[kk];
    // COPR
    double vel2 = (u * u + v * v + w * w);
    double beta = .5 * vel2;
    double drodh = FLD(b, FF_DRDH)[kk], drodp = FLD(b, FF_DRDP)[kk];
    double yp = 1. / ro;
    double dkin = u * dmi + v * dni + w * dwi;
    FLD(b, FF_DM)[kk] = yp * (-u * dri + dmi);
    FLD(b, FF_DN)[kk] = yp * (-v * dri + dni);
    FLD(b, FF_DW)[kk] = yp * (-w * dri + dwi);
    double hi = yp * (FLD(b, FF_E)[kk] + P);
    double dei = dre - (hi - 2 * beta) * dri - dkin;
    double dpi = (c * c) * (dri - drodh * yp * dei);
    double dhi = (dri - drodp * dpi) / drodh;
    FLD(b, FF_DRO)[kk] = dhi;
    FLD(b, FF_DE)[kk] = dpi;
}

__device__ __forceinline__ void load_dw(const Blk &b, int n, double dw[5])
{
#pragma unroll
    for (int q = 0; q < 5; q++) dw[q] = FLD(b, FF_DRO + q)[n];
}

// Hyperplane cell enumeration: thread t -> (j, k) with i = m - j - k inside the block.
__device__ __forceinline__ bool plane_cell(const Blk &b, int m, int jlo, int klo, int jn,
                                           int &i, int &j, int &k)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    j = jlo + t % jn;
    k = klo + t / jn;
    i = m - j - k;
    return (k <= b.kmax && j <= b.jmax && i >= 1 && i <= b.imax);
}

// PREDGS (imps4.f:345-598) for the cells of hyperplane I+J+K = m
__global__ void __launch_bounds__(128) k_pred_plane(Prm p, Blk b, int m, int jlo, int klo, int jn)
{
    int i, j, k;
    if (!plane_cell(b, m, jlo, klo, jn, i, j, k)) return;
    const int IS = b.is, IL = b.il;
    const double cdiff = 2.;
    int n = cid(b, i, j, k);
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3);
    const double *c = FLD(b, FF_C), *ro = FLD(b, FF_RO), *vol = FLD(b, FF_VOL);
    // own coefficients on the lower faces (loop 1100)
    double ffi = i == 1 ? 4. : 2., ffj = j == 1 ? 4. : 2., ffk = k == 1 ? 4. : 2.;
    Coef ci = lusgs_coef<true>(b, n, 1, ffi * cdiff, a1, FLD(b, FF_A1X), FLD(b, FF_A1Y), FLD(b, FF_A1Z));
    Coef cj = lusgs_coef<true>(b, n, IS, cdiff * ffj, a2, FLD(b, FF_A2X), FLD(b, FF_A2Y), FLD(b, FF_A2Z));
    Coef ck = lusgs_coef<true>(b, n, IL, ffk * cdiff, a3, FLD(b, FF_A3X), FLD(b, FF_A3Y), FLD(b, FF_A3Z));
    double dt1 = p.theta * FLD(b, FF_DTL)[n];
    double scinv = 1. / (vol[n] + dt1 * (ci.xa * a1[n] + cj.xa * a2[n] + ck.xa * a3[n]) + vol[n] * 0.);
    double scinp = 1. / (vol[n] + dt1 * (ci.rmax * a1[n] + cj.rmax * a2[n] + ck.rmax * a3[n]) + vol[n] * 0.);
    // upstream contributions: k+1 (loop 1600 of the previous slab), i+1 and j+1 (line buffers)
    double ckc[5] = {0., 0., 0., 0., 0.}, yi[5] = {0., 0., 0., 0., 0.}, bj[5] = {0., 0., 0., 0., 0.};
    if (k < b.kmax) {
        int nb = n + IL;
        Coef kb = lusgs_coef<true>(b, nb, IL, 2. * cdiff, a3, FLD(b, FF_A3X), FLD(b, FF_A3Y), FLD(b, FF_A3Z));
        double dw[5];
        load_dw(b, nb, dw);
        face_contrib(dw, c[nb], ro[nb], FLD(b, FF_A3X)[nb], FLD(b, FF_A3Y)[nb], FLD(b, FF_A3Z)[nb],
                     kb.x1, kb.x2, kb.x5, ckc);
    }
    if (i < b.imax && j < b.jmax) {   // J = JMAX: the shift of imps4.f:1460 leaves 0
        int nb = n + 1;
        Coef ib = lusgs_coef<true>(b, nb, 1, 2. * cdiff, a1, FLD(b, FF_A1X), FLD(b, FF_A1Y), FLD(b, FF_A1Z));
        double dw[5];
        load_dw(b, nb, dw);
        face_contrib(dw, c[nb], ro[nb], FLD(b, FF_A1X)[nb], FLD(b, FF_A1Y)[nb], FLD(b, FF_A1Z)[nb],
                     ib.x1, ib.x2, ib.x5, yi);
    }
    if (j < b.jmax) {
        int nb = n + IS;
        Coef jb = lusgs_coef<true>(b, nb, IS, cdiff * 2., a2, FLD(b, FF_A2X), FLD(b, FF_A2Y), FLD(b, FF_A2Z));
        double dw[5];
        load_dw(b, nb, dw);
        face_contrib(dw, c[nb], ro[nb], FLD(b, FF_A2X)[nb], FLD(b, FF_A2Y)[nb], FLD(b, FF_A2Z)[nb],
                     jb.x1, jb.x2, jb.x5, bj);
    }
    double a3u = a3[n + IL], a1u = a1[n + 1], a2u = a2[n + IS];
    double cd[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        cd[q] = -dt1 * a3u * ckc[q] + vol[n] * FLD(b, FF_DRO + q)[n];
        cd[q] = cd[q] - dt1 * (a1u * yi[q] + a2u * bj[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) FLD(b, FF_DRO + q)[n] = cd[q] * scinv;
    FLD(b, FF_DE)[n] = cd[4] * scinp;
}

// DIAGGS factor (imps4.f:621-659) of an interior cell
__device__ __forceinline__ double diag_factor(const Prm &p, const Blk &b, int n)
{
    const int IS = b.is, IL = b.il;
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3),
                 *a1x = FLD(b, FF_A1X), *a1y = FLD(b, FF_A1Y), *a1z = FLD(b, FF_A1Z),
                 *a2x = FLD(b, FF_A2X), *a2y = FLD(b, FF_A2Y), *a2z = FLD(b, FF_A2Z),
                 *a3x = FLD(b, FF_A3X), *a3y = FLD(b, FF_A3Y), *a3z = FLD(b, FF_A3Z);
    double u = FLD(b, FF_U)[n], v = FLD(b, FF_V)[n], w = FLD(b, FF_W)[n], c = FLD(b, FF_C)[n];
    double ucon;
    ucon = a1x[n + 1] * u + a1y[n + 1] * v + a1z[n + 1] * w - 0.;
    double r1n = 1.00 * (fabs(ucon) + c);
    ucon = a1x[n] * u + a1y[n] * v + a1z[n] * w - 0.;
    double r1p = 1.00 * (fabs(ucon) + c);
    ucon = a2x[n + IS] * u + a2y[n + IS] * v + a2z[n + IS] * w - 0.;
    double r2n = 1.00 * (fabs(ucon) + c);
    ucon = a2x[n] * u + a2y[n] * v + a2z[n] * w - 0.;
    double r2p = 1.00 * (fabs(ucon) + c);
    ucon = a3x[n + IL] * u + a3y[n + IL] * v + a3z[n + IL] * w - 0.;
    double r3n = 1.00 * (fabs(ucon) + c);
    ucon = a3x[n] * u + a3y[n] * v + a3z[n] * w - 0.;
    double r3p = 1.00 * (fabs(ucon) + c);
    double dt1 = p.theta * FLD(b, FF_DTL)[n] / FLD(b, FF_VOL)[n];
    double sc = .5 * (a1[n] * r1p + a1[n + 1] * r1n + a2[n] * r2p + a2[n + IS] * r2n +
                      a3[n] * r3p + a3[n + IL] * r3n);
    return 1. + dt1 * sc + 0.;
}

// DIAGGS + CORRGS (imps4.f:76-339) for the cells of hyperplane I+J+K = m
__global__ void __launch_bounds__(128) k_corr_plane(Prm p, Blk b, int m, int jlo, int klo, int jn)
{
    int i, j, k;
    if (!plane_cell(b, m, jlo, klo, jn, i, j, k)) return;
    const int IS = b.is, IL = b.il;
    const double cdiff = 2.;
    int n = cid(b, i, j, k);
    const double *a1 = FLD(b, FF_A1), *a2 = FLD(b, FF_A2), *a3 = FLD(b, FF_A3);
    const double *c = FLD(b, FF_C), *ro = FLD(b, FF_RO), *vol = FLD(b, FF_VOL);
    double fac = diag_factor(p, b, n);
    Coef ci = lusgs_coef<false>(b, n, 1, 2. * cdiff, a1, FLD(b, FF_A1X), FLD(b, FF_A1Y), FLD(b, FF_A1Z));
    Coef cj = lusgs_coef<false>(b, n, IS, cdiff * 2., a2, FLD(b, FF_A2X), FLD(b, FF_A2Y), FLD(b, FF_A2Z));
    Coef ck = lusgs_coef<false>(b, n, IL, 2. * cdiff, a3, FLD(b, FF_A3X), FLD(b, FF_A3Y), FLD(b, FF_A3Z));
    double dt1 = p.theta * FLD(b, FF_DTL)[n];
    double scinv = 1. / (vol[n] + dt1 * (ci.xa * a1[n + 1] + cj.xa * a2[n + IS] + ck.xa * a3[n + IL]) + vol[n] * 0.);
    double scinp = 1. / (vol[n] + dt1 * (ci.rmax * a1[n] + cj.rmax * a2[n] + ck.rmax * a3[n]) + vol[n] * 0.);
    double ckc[5] = {0., 0., 0., 0., 0.}, yi[5] = {0., 0., 0., 0., 0.}, bj[5] = {0., 0., 0., 0., 0.};
    if (k > 1) {
        int nb = n - IL;
        Coef kb = lusgs_coef<false>(b, nb, IL, 2. * cdiff, a3, FLD(b, FF_A3X), FLD(b, FF_A3Y), FLD(b, FF_A3Z));
        double dw[5];
        load_dw(b, nb, dw);
        face_contrib(dw, c[nb], ro[nb], FLD(b, FF_A3X)[n], FLD(b, FF_A3Y)[n], FLD(b, FF_A3Z)[n],
                     kb.x1, kb.x2, kb.x5, ckc);
    }
    if (i > 1) {
        int nb = n - 1;
        Coef ib = lusgs_coef<false>(b, nb, 1, 2. * cdiff, a1, FLD(b, FF_A1X), FLD(b, FF_A1Y), FLD(b, FF_A1Z));
        double dw[5];
        load_dw(b, nb, dw);
        face_contrib(dw, c[nb], ro[nb], FLD(b, FF_A1X)[n], FLD(b, FF_A1Y)[n], FLD(b, FF_A1Z)[n],
                     ib.x1, ib.x2, ib.x5, yi);
    }
    if (j > 1) {
        int nb = n - IS;
        Coef jb = lusgs_coef<false>(b, nb, IS, cdiff * 2., a2, FLD(b, FF_A2X), FLD(b, FF_A2Y), FLD(b, FF_A2Z));
        double dw[5];
        load_dw(b, nb, dw);
        face_contrib(dw, c[nb], ro[nb], FLD(b, FF_A2X)[n], FLD(b, FF_A2Y)[n], FLD(b, FF_A2Z)[n],
                     jb.x1, jb.x2, jb.x5, bj);
    }
    double a3l = a3[n], a1l = a1[n], a2l = a2[n];
    double cd[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        double dq = FLD(b, FF_DRO + q)[n] * fac;          // DIAGGS
        cd[q] = dt1 * a3l * ckc[q] + vol[n] * dq;
        cd[q] = cd[q] + dt1 * (a1l * yi[q] + a2l * bj[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) FLD(b, FF_DRO + q)[n] = cd[q] * scinv;
    FLD(b, FF_DE)[n] = cd[4] * scinp;
}

// ------------------------------------------------------------------------------------
// LU-SGS as a tile wavefront (the default path).
//
// A workgroup is one wave that owns an 8x8 tile of (J,K) columns and marches along I.
// Lane (jj,kk) runs one step behind its upstream lanes, so at every step the wave's cells
// lie on one hyperplane of the tile and the three contributions a cell needs were produced
// one step earlier: the I-contribution by the lane itself (a register), the J- and
// K-contributions by the neighbouring lanes (LDS) or, across a tile face, by the
// neighbouring tile (halo slots in HBM).  Each cell forms its own split-eigenvalue
// coefficients once and sends SYM1I(XI*SYI(dW)) downstream, as PREDGS/CORRGS do with their
// line buffers; the arithmetic per cell is the reference's, so results are bit for bit
// those of the K-plane/diagonal order.
//
// Halo slots hold a signalling-NaN pattern when empty.  The producer stores a value with a
// relaxed device-scope atomic store; the consumer polls it until it is filled, then empties
// it again for the next sweep.  Arithmetic never produces a signalling NaN, so the slots
// need no fences.  Workgroups take tiles through an atomic ticket in a topological order,
// so a tile only ever waits on tiles that have already started.  Waits are bounded: after
// SPIN_LIMIT polls (seconds) a wait gives up and raises the error word the host checks.
// ------------------------------------------------------------------------------------
constexpr int TJ = 8, TK = 8, TW = TJ * TK;
constexpr unsigned long long HALO_EMPTY = 0x7FF4DEADBEEF0001ULL;
constexpr int SPIN_LIMIT = 1 << 22;
constexpr int HS = 7;              // halo words per cell face (7 unknowns of the k-omega form)

struct Sweep {
    int ntj, ntk;                   // tiles along J and K
    unsigned long long *ticket;     // monotonic; ticket % (ntj*ntk) orders the tiles
    unsigned int *err;              // non-zero after a halo wait gave up
    unsigned long long *hj, *hk;    // halo slots of the J-faces and K-faces between tiles
};

// Halo slots laid out by anti-diagonal.  The edge lanes that exchange across a tile face on
// one step hold cells whose I plus offset along the face is the same (the step's hyperplane),
// so with the layout [face f][tile along the face][I + offset][word q][offset] the eight
// lanes' values of one word are one contiguous, 64-byte aligned run: a poll or a put of a
// word is a single memory request instead of eight.
constexpr int HQS = TK;             // stride between the words of one slot
static_assert(TJ == TK, "one slot stride for both faces");

__device__ __forceinline__ unsigned long long *hj_slot(const Sweep &s, const Blk &b, int f, int i,
                                                       int k)
{
    const int kk = (k - 1) % TK, t = (k - 1) / TK;
    return s.hj + (((size_t)f * s.ntk + t) * (b.imax + TK) + (i + kk - 1)) * (HS * TK) + kk;
}

__device__ __forceinline__ unsigned long long *hk_slot(const Sweep &s, const Blk &b, int f, int i,
                                                       int j)
{
    const int jj = (j - 1) % TJ, t = (j - 1) / TJ;
    return s.hk + (((size_t)f * s.ntj + t) * (b.imax + TJ) + (i + jj - 1)) * (HS * TJ) + jj;
}

// slot words are QS apart
template <int NV, int QS = 1>
__device__ __forceinline__ void halo_put(unsigned long long *slot, const double v[NV])
{
#pragma unroll
    for (int q = 0; q < NV; q++)
        __hip_atomic_store(slot + q * QS, (unsigned long long)__double_as_longlong(v[q]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NV, int QS = 1>
__device__ __forceinline__ void halo_issue(unsigned long long *slot, unsigned long long h[NV])
{
#pragma unroll
    for (int q = 0; q < NV; q++)
        h[q] = __hip_atomic_load(slot + q * QS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Complete a halo read begun by halo_issue: poll empty slots, then empty them.  The words
// are tested together first, so a slot that is already filled costs no polling branches.
template <int NV, int QS = 1>
__device__ __forceinline__ void halo_finish(unsigned long long *slot, unsigned long long h[NV],
                                            unsigned int *err, double v[NV])
{
    bool empty = false;
#pragma unroll
    for (int q = 0; q < NV; q++) empty |= h[q] == HALO_EMPTY;
    if (empty) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            unsigned long long x = h[q];
            int spins = 0;
            while (x == HALO_EMPTY) {
                if ((++spins & 255) == 0) {
                    if (spins > SPIN_LIMIT) {
                        __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        x = 0;
                        break;
                    }
                    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                        x = 0;
                        break;
                    }
                }
                __builtin_amdgcn_s_sleep(1);
                x = __hip_atomic_load(slot + q * QS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            h[q] = x;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
        v[q] = __longlong_as_double((long long)h[q]);
        __hip_atomic_store(slot + q * QS, HALO_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The wave's LDS exchange point between two steps (one wave per workgroup: no s_barrier
// is needed, only LDS ordering).
__device__ __forceinline__ void wave_step_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ int tile_ticket(const Sweep &s)
{
    unsigned long long t = 0;
    if (threadIdx.x == 0) t = atomicAdd(s.ticket, 1ULL);
    t = __shfl(t, 0);
    return (int)(t % (unsigned long long)(s.ntj * s.ntk));
}

// Values cell n of a sweep reads: its increments and state, the neighbours' VOL, RO, VIS,
// VIST across the three faces its coefficients use, and the face geometry.  `own` holds the
// faces the cell's coefficients and outgoing contributions use (lower faces in PREDGS, upper
// in CORRGS), `oth` the other three.
struct SweepRaw {
    double d[7], u, v, w, c, ro, vol, vis, vist, dtl, cp, rk, drdp, drdh;
    double vol1, ro1, vis1, vist1, vol2, ro2, vis2, vist2, vol3, ro3, vis3, vist3;
    double a1, a2, a3, n1x, n1y, n1z, n2x, n2y, n2z, n3x, n3y, n3z;   // lower faces
    double a1u, a2u, a3u, u1x, u1y, u1z, u2x, u2y, u2z, u3x, u3y, u3z; // upper faces
};

// FWD = false: PREDGS (neighbours n-1, n-IS, n-IL); true: CORRGS (n+1, n+IS, n+IL)
template <bool FWD, bool KE>
__device__ __forceinline__ void sweep_load(const Blk &b, int n, SweepRaw &r)
{
    const int IS = b.is, IL = b.il;
    constexpr int NV = KE ? 7 : 5;
    const int s1 = FWD ? 1 : -1, s2 = FWD ? IS : -IS, s3 = FWD ? IL : -IL;
#pragma unroll
    for (int q = 0; q < NV; q++) r.d[q] = FLD(b, FF_DRO + q)[n];
    r.u = FLD(b, FF_U)[n]; r.v = FLD(b, FF_V)[n]; r.w = FLD(b, FF_W)[n];
    r.c = FLD(b, FF_C)[n]; r.ro = FLD(b, FF_RO)[n]; r.vol = FLD(b, FF_VOL)[n];
    r.vis = FLD(b, FF_VIS)[n]; r.vist = FLD(b, FF_VIST)[n]; r.dtl = FLD(b, FF_DTL)[n];
    if (FWD) r.cp = FLD(b, FF_CP)[n];
    if (KE) {
        r.rk = FLD(b, FF_RK)[n]; r.drdp = FLD(b, FF_DRDP)[n]; r.drdh = FLD(b, FF_DRDH)[n];
    }
    r.vol1 = FLD(b, FF_VOL)[n + s1]; r.ro1 = FLD(b, FF_RO)[n + s1];
    r.vis1 = FLD(b, FF_VIS)[n + s1]; r.vist1 = FLD(b, FF_VIST)[n + s1];
    r.vol2 = FLD(b, FF_VOL)[n + s2]; r.ro2 = FLD(b, FF_RO)[n + s2];
    r.vis2 = FLD(b, FF_VIS)[n + s2]; r.vist2 = FLD(b, FF_VIST)[n + s2];
    r.vol3 = FLD(b, FF_VOL)[n + s3]; r.ro3 = FLD(b, FF_RO)[n + s3];
    r.vis3 = FLD(b, FF_VIS)[n + s3]; r.vist3 = FLD(b, FF_VIST)[n + s3];
    r.a1 = FLD(b, FF_A1)[n]; r.a2 = FLD(b, FF_A2)[n]; r.a3 = FLD(b, FF_A3)[n];
    r.n1x = FLD(b, FF_A1X)[n]; r.n1y = FLD(b, FF_A1Y)[n]; r.n1z = FLD(b, FF_A1Z)[n];
    r.n2x = FLD(b, FF_A2X)[n]; r.n2y = FLD(b, FF_A2Y)[n]; r.n2z = FLD(b, FF_A2Z)[n];
    r.n3x = FLD(b, FF_A3X)[n]; r.n3y = FLD(b, FF_A3Y)[n]; r.n3z = FLD(b, FF_A3Z)[n];
    r.a1u = FLD(b, FF_A1)[n + 1]; r.a2u = FLD(b, FF_A2)[n + IS]; r.a3u = FLD(b, FF_A3)[n + IL];
    if (FWD) {
        r.u1x = FLD(b, FF_A1X)[n + 1]; r.u1y = FLD(b, FF_A1Y)[n + 1]; r.u1z = FLD(b, FF_A1Z)[n + 1];
        r.u2x = FLD(b, FF_A2X)[n + IS]; r.u2y = FLD(b, FF_A2Y)[n + IS]; r.u2z = FLD(b, FF_A2Z)[n + IS];
        r.u3x = FLD(b, FF_A3X)[n + IL]; r.u3y = FLD(b, FF_A3Y)[n + IL]; r.u3z = FLD(b, FF_A3Z)[n + IL];
    }
}

// Recurrence-independent part of a cell's LU-SGS update
struct SweepCell {
    double d[7], dt1, vol, ro, roc, pc2, proc, scinv, scinp;
    double yp, ydrp, drhro, ro667;  // SYIKE/SYM1KE scalars (k-omega form)
    double wi, wj, wk;              // areas weighting the incoming I-, J-, K-contributions
    FaceBasis fi, fj, fk;           // faces the outgoing contributions cross
    Coef ci, cj, ck;
    double cp, rho;                 // CORRGS only: DRO/CP and D(rho k), D(rho omega) * RO
    double fac;                     // CORRGS only: the DIAGGS factor of the increments
};

template <bool KE>
__device__ __forceinline__ void cell_scalars(SweepCell &s, const SweepRaw &r)
{
    const double ro = r.ro, c = r.c;
    s.ro = ro;
    s.roc = ro * c;                // SYI imps4.f:31, SYIKE imps4.f:1015
    double ypro = 1. / ro;         // SYM1I imps4.f:57-59, SYM1KE imps4.f:1050-1052
    s.pc2 = .5 * ypro;
    s.proc = s.pc2 / c;
    if (KE) {
        s.yp = ypro;               // SYIKE's PC2 = 1./RO
        double drh = .667 * r.rk * (ypro * ypro) * r.drdh;   // SYM1KE imps4.f:1053-1054
        s.ydrp = 1. / (1. + .667 * r.rk * ypro * r.drdp);
        s.drhro = drh * ro;
        s.ro667 = .667 * ro;
    }
}

// SYIKE / SYM1KE (imps4.f:982-1072) around the split-eigenvalue scaling, with the basis and
// the cell scalars formed beforehand
__device__ __forceinline__ void face_contrib_ke_b(const double dw[7], const SweepCell &s,
                                                  const FaceBasis &f, double x1, double x2,
                                                  double x5, double out[7])
{
    double z1 = dw[0], z5 = dw[4], z6 = dw[5], z7 = dw[6];
    double z2 = f.nx * dw[1] + f.ny * dw[2] + f.nz * dw[3];
    double z3 = f.a21 * dw[1] + f.a22 * dw[2] + f.a23 * dw[3];
    double z4 = f.a31 * dw[1] + f.a32 * dw[2] + f.a33 * dw[3];
    double w1 = z1 - s.yp * z5;
    double w2 = s.roc * z2 + z5;
    double w3 = z3, w4 = z4, w6 = z6, w7 = z7;
    double w5 = -s.roc * z2 + z5;
    w1 = x1 * w1;
    w2 = x2 * w2;
    w3 = x1 * w3;
    w4 = x1 * w4;
    w5 = x5 * w5;
    w6 = x1 * w6;
    w7 = x1 * w7;
    double dh = w1 + s.pc2 * (w2 + w5);
    double x2i = s.proc * (w2 - w5);
    double x5v = .5 * (w2 + w5);
    x5v = s.ydrp * (x5v - s.drhro * dh - s.ro667 * w6);
    out[0] = dh;
    out[1] = f.nx * x2i + f.a21 * w3 + f.a31 * w4;
    out[2] = f.ny * x2i + f.a22 * w3 + f.a32 * w4;
    out[3] = f.nz * x2i + f.a23 * w3 + f.a33 * w4;
    out[4] = x5v;
    out[5] = w6;
    out[6] = w7;
}

template <bool KE>
__device__ __forceinline__ void face_xfer(const double *dn, const SweepCell &s, const FaceBasis &f,
                                          const Coef &cf, double *out)
{
    if (KE)
        face_contrib_ke_b(dn, s, f, cf.x1, cf.x2, cf.x5, out);
    else
        face_contrib_b(dn, s.ro, s.roc, s.pc2, s.proc, f, cf.x1, cf.x2, cf.x5, out);
}

// PREDGS's recurrence-independent coefficients of a cell (imps4.f:396-466): the
// split-eigenvalue coefficients of its three lower faces, SCINV/SCINP, THETA*DT
template <bool KE>
__device__ __forceinline__ void pred_coef(const Prm &p, const SweepRaw &r, int i, int j, int k,
                                          SweepCell &s)
{
    const double cdiff = 2.;
    double ffi = i == 1 ? 4. : 2., ffj = j == 1 ? 4. : 2., ffk = k == 1 ? 4. : 2.;
    s.ci = coef_vals<true>(r.n1x, r.n1y, r.n1z, r.u, r.v, r.w, r.c, r.a1, r.vol, r.vol1, r.ro,
                           r.ro1, r.vis, r.vist, r.vis1, r.vist1, ffi * cdiff);
    s.cj = coef_vals<true>(r.n2x, r.n2y, r.n2z, r.u, r.v, r.w, r.c, r.a2, r.vol, r.vol2, r.ro,
                           r.ro2, r.vis, r.vist, r.vis2, r.vist2, cdiff * ffj);
    s.ck = coef_vals<true>(r.n3x, r.n3y, r.n3z, r.u, r.v, r.w, r.c, r.a3, r.vol, r.vol3, r.ro,
                           r.ro3, r.vis, r.vist, r.vis3, r.vist3, ffk * cdiff);
    s.dt1 = p.theta * r.dtl;
    s.vol = r.vol;
    s.scinv = 1. / (r.vol + s.dt1 * (s.ci.xa * r.a1 + s.cj.xa * r.a2 + s.ck.xa * r.a3) + r.vol * 0.);
    s.scinp = 1. / (r.vol + s.dt1 * (s.ci.rmax * r.a1 + s.cj.rmax * r.a2 + s.ck.rmax * r.a3) + r.vol * 0.);
#pragma unroll
    for (int q = 0; q < (KE ? 7 : 5); q++) s.d[q] = r.d[q];
    s.wi = r.a1u; s.wj = r.a2u; s.wk = r.a3u;
}

// PREDGS's face bases (the lower faces) and SYI/SYM1I cell scalars
template <bool KE>
__device__ __forceinline__ void pred_basis(const SweepRaw &r, SweepCell &s)
{
    cell_scalars<KE>(s, r);
    s.fi = face_basis(r.n1x, r.n1y, r.n1z);
    s.fj = face_basis(r.n2x, r.n2y, r.n2z);
    s.fk = face_basis(r.n3x, r.n3y, r.n3z);
}

template <bool KE>
__device__ __forceinline__ void pred_prep(const Prm &p, const SweepRaw &r, int i, int j, int k,
                                          SweepCell &s)
{
    pred_coef<KE>(p, r, i, j, k, s);
    pred_basis<KE>(r, s);
}

// DIAGGS + CORRGS's recurrence-independent coefficients of a cell: the diagonal factor
// applied to the increments (imps4.f:621-659 / DIKEGS imps4.f:1877-1912), the coefficients
// of its three upper faces, SCINV/SCINP, THETA*DT
template <bool KE>
__device__ __forceinline__ void corr_coef(const Prm &p, const SweepRaw &r, SweepCell &s)
{
    constexpr int NV = KE ? 7 : 5;
    const double cdiff = 2.;
    double ucon;
    ucon = r.u1x * r.u + r.u1y * r.v + r.u1z * r.w - 0.;
    double r1n = 1.00 * (fabs(ucon) + r.c);
    ucon = r.n1x * r.u + r.n1y * r.v + r.n1z * r.w - 0.;
    double r1p = 1.00 * (fabs(ucon) + r.c);
    ucon = r.u2x * r.u + r.u2y * r.v + r.u2z * r.w - 0.;
    double r2n = 1.00 * (fabs(ucon) + r.c);
    ucon = r.n2x * r.u + r.n2y * r.v + r.n2z * r.w - 0.;
    double r2p = 1.00 * (fabs(ucon) + r.c);
    ucon = r.u3x * r.u + r.u3y * r.v + r.u3z * r.w - 0.;
    double r3n = 1.00 * (fabs(ucon) + r.c);
    ucon = r.n3x * r.u + r.n3y * r.v + r.n3z * r.w - 0.;
    double r3p = 1.00 * (fabs(ucon) + r.c);
    double dtv = p.theta * r.dtl / r.vol;
    double sc = .5 * (r.a1 * r1p + r.a1u * r1n + r.a2 * r2p + r.a2u * r2n + r.a3 * r3p + r.a3u * r3n);
    double fac = 1. + dtv * sc + 0.;
    s.fac = fac;
#pragma unroll
    for (int q = 0; q < NV; q++) s.d[q] = r.d[q] * fac;
    // CORRGS / COKEGS own coefficients (upper faces, viscous area A(n))
    s.ci = coef_vals<false>(r.u1x, r.u1y, r.u1z, r.u, r.v, r.w, r.c, r.a1, r.vol, r.vol1, r.ro,
                            r.ro1, r.vis, r.vist, r.vis1, r.vist1, 2. * cdiff);
    s.cj = coef_vals<false>(r.u2x, r.u2y, r.u2z, r.u, r.v, r.w, r.c, r.a2, r.vol, r.vol2, r.ro,
                            r.ro2, r.vis, r.vist, r.vis2, r.vist2, cdiff * 2.);
    s.ck = coef_vals<false>(r.u3x, r.u3y, r.u3z, r.u, r.v, r.w, r.c, r.a3, r.vol, r.vol3, r.ro,
                            r.ro3, r.vis, r.vist, r.vis3, r.vist3, 2. * cdiff);
    s.dt1 = p.theta * r.dtl;
    s.vol = r.vol;
    s.scinv = 1. / (r.vol + s.dt1 * (s.ci.xa * r.a1u + s.cj.xa * r.a2u + s.ck.xa * r.a3u) + r.vol * 0.);
    s.scinp = 1. / (r.vol + s.dt1 * (s.ci.rmax * r.a1 + s.cj.rmax * r.a2 + s.ck.rmax * r.a3) + r.vol * 0.);
    s.wi = r.a1; s.wj = r.a2; s.wk = r.a3;
}

// CORRGS's face bases (the upper faces), cell scalars, CP and RO for the end of IMPS
template <bool KE>
__device__ __forceinline__ void corr_basis(const SweepRaw &r, SweepCell &s)
{
    cell_scalars<KE>(s, r);
    s.fi = face_basis(r.u1x, r.u1y, r.u1z);
    s.fj = face_basis(r.u2x, r.u2y, r.u2z);
    s.fk = face_basis(r.u3x, r.u3y, r.u3z);
    s.cp = r.cp;
    s.rho = r.ro;
}

template <bool KE>
__device__ __forceinline__ void corr_prep(const Prm &p, const SweepRaw &r, SweepCell &s)
{
    corr_coef<KE>(p, r, s);
    corr_basis<KE>(r, s);
}

// Backward sweep PREDGS (imps4.f:345-598) or PRKEGS (imps4.f:1462-1841, KE) as a tile
// wavefront; I, J, K descending.
template <bool KE>
__global__ void __launch_bounds__(TW) k_pred_tiles(Prm p, Blk b, Sweep sw)
{
    constexpr int NV = KE ? 7 : 5;
    __shared__ double xj[2][NV][TW], xk[2][NV][TW];
    const int lane = threadIdx.x, jj = lane % TJ, kk = lane / TJ;
    const int t = tile_ticket(sw);
    const int tj = sw.ntj - 1 - t % sw.ntj, tk = sw.ntk - 1 - t / sw.ntj;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1;
    const bool col = j <= b.jmax && k <= b.kmax;
    const int lag = (TJ - 1 - jj) + (TK - 1 - kk);
    const int nsteps = b.imax + TJ + TK - 2;
    double yi[NV];
#pragma unroll
    for (int r = 0; r < NV; r++) yi[r] = 0.;
    SweepRaw cur;
    if (col) sweep_load<false, KE>(b, cid(b, b.imax, j, k), cur);
    for (int s = 0; s < nsteps; s++) {
        const int q = s - lag;
        if (col && q >= 0 && q < b.imax) {
            const int i = b.imax - q;
            const int n = cid(b, i, j, k);
            const bool hjin = jj == TJ - 1 && j < b.jmax, hkin = kk == TK - 1 && k < b.kmax;
            unsigned long long hv[NV], hw[NV];
            if (hjin) halo_issue<NV, HQS>(hj_slot(sw, b, tj + 1, i, k), hv);
            if (hkin) halo_issue<NV, HQS>(hk_slot(sw, b, tk + 1, i, j), hw);
            SweepRaw nxt;
            if (q + 1 < b.imax) sweep_load<false, KE>(b, n - 1, nxt);
            SweepCell sc;
            pred_prep<KE>(p, cur, i, j, k, sc);
            const int rb = (s + 1) & 1;
            double bj[NV], ckc[NV];
            if (j == b.jmax) {
#pragma unroll
                for (int r = 0; r < NV; r++) bj[r] = 0.;
            } else if (!hjin) {
#pragma unroll
                for (int r = 0; r < NV; r++) bj[r] = xj[rb][r][lane + 1];
            } else {
                halo_finish<NV, HQS>(hj_slot(sw, b, tj + 1, i, k), hv, sw.err, bj);
            }
            if (k == b.kmax) {
#pragma unroll
                for (int r = 0; r < NV; r++) ckc[r] = 0.;
            } else if (!hkin) {
#pragma unroll
                for (int r = 0; r < NV; r++) ckc[r] = xk[rb][r][lane + TJ];
            } else {
                halo_finish<NV, HQS>(hk_slot(sw, b, tk + 1, i, j), hw, sw.err, ckc);
            }
            // the line-buffer shift of PREDGS (imps4.f:556-562) never hands the I+1
            // contribution to a cell of the row J = JMAX; PRKEGS's (imps4.f:1762) does
            if (!KE && j == b.jmax) {
#pragma unroll
                for (int r = 0; r < NV; r++) yi[r] = 0.;
            }
            double dn[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) {
                double cd = -sc.dt1 * sc.wk * ckc[r] + sc.vol * sc.d[r];
                cd = cd - sc.dt1 * (sc.wi * yi[r] + sc.wj * bj[r]);
                dn[r] = r != 4 ? cd * sc.scinv : cd * sc.scinp;
            }
#pragma unroll
            for (int r = 0; r < NV; r++) FLD(b, FF_DRO + r)[n] = dn[r];
            double oj[NV], ok[NV];
            face_xfer<KE>(dn, sc, sc.fi, sc.ci, yi);
            face_xfer<KE>(dn, sc, sc.fj, sc.cj, oj);
            face_xfer<KE>(dn, sc, sc.fk, sc.ck, ok);
            const int wb = s & 1;
            if (jj > 0) {
#pragma unroll
                for (int r = 0; r < NV; r++) xj[wb][r][lane] = oj[r];
            } else if (j > 1) {
                halo_put<NV, HQS>(hj_slot(sw, b, tj, i, k), oj);
            }
            if (kk > 0) {
#pragma unroll
                for (int r = 0; r < NV; r++) xk[wb][r][lane] = ok[r];
            } else if (k > 1) {
                halo_put<NV, HQS>(hk_slot(sw, b, tk, i, j), ok);
            }
            if (q + 1 < b.imax) cur = nxt;
        }
        wave_step_sync();
    }
}

// DIAGGS + forward sweep CORRGS (imps4.f:604-664, 76-339), or DIKEGS + COKEGS
// (imps4.f:1841-1933, 1078-1456, KE), as a tile wavefront, I, J, K ascending, with the end
// of IMPS applied to the final increment: DRO/CP (imps3.f:483) and, KE, D(rho k) and
// D(rho omega) times RO (imps3.f:497-500).
template <bool KE>
__global__ void __launch_bounds__(TW) k_corr_tiles(Prm p, Blk b, Sweep sw)
{
    constexpr int NV = KE ? 7 : 5;
    __shared__ double xj[2][NV][TW], xk[2][NV][TW];
    const int lane = threadIdx.x, jj = lane % TJ, kk = lane / TJ;
    const int t = tile_ticket(sw);
    const int tj = t % sw.ntj, tk = t / sw.ntj;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1;
    const bool col = j <= b.jmax && k <= b.kmax;
    const int lag = jj + kk;
    const int nsteps = b.imax + TJ + TK - 2;
    double yi[NV];
#pragma unroll
    for (int r = 0; r < NV; r++) yi[r] = 0.;
    SweepRaw cur;
    if (col) sweep_load<true, KE>(b, cid(b, 1, j, k), cur);
    for (int s = 0; s < nsteps; s++) {
        const int q = s - lag;
        if (col && q >= 0 && q < b.imax) {
            const int i = q + 1;
            const int n = cid(b, i, j, k);
            const bool hjin = jj == 0 && j > 1, hkin = kk == 0 && k > 1;
            unsigned long long hv[NV], hw[NV];
            if (hjin) halo_issue<NV, HQS>(hj_slot(sw, b, tj, i, k), hv);
            if (hkin) halo_issue<NV, HQS>(hk_slot(sw, b, tk, i, j), hw);
            SweepRaw nxt;
            if (q + 1 < b.imax) sweep_load<true, KE>(b, n + 1, nxt);
            SweepCell sc;
            corr_prep<KE>(p, cur, sc);
            const int rb = (s + 1) & 1;
            double bj[NV], ckc[NV];
            if (j == 1) {
#pragma unroll
                for (int r = 0; r < NV; r++) bj[r] = 0.;
            } else if (!hjin) {
#pragma unroll
                for (int r = 0; r < NV; r++) bj[r] = xj[rb][r][lane - 1];
            } else {
                halo_finish<NV, HQS>(hj_slot(sw, b, tj, i, k), hv, sw.err, bj);
            }
            if (k == 1) {
#pragma unroll
                for (int r = 0; r < NV; r++) ckc[r] = 0.;
            } else if (!hkin) {
#pragma unroll
                for (int r = 0; r < NV; r++) ckc[r] = xk[rb][r][lane - TJ];
            } else {
                halo_finish<NV, HQS>(hk_slot(sw, b, tk, i, j), hw, sw.err, ckc);
            }
            double dn[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) {
                double cd = sc.dt1 * sc.wk * ckc[r] + sc.vol * sc.d[r];
                cd = cd + sc.dt1 * (sc.wi * yi[r] + sc.wj * bj[r]);
                dn[r] = r != 4 ? cd * sc.scinv : cd * sc.scinp;
            }
            FLD(b, FF_DRO)[n] = dn[0] / sc.cp;
#pragma unroll
            for (int r = 1; r < 5; r++) FLD(b, FF_DRO + r)[n] = dn[r];
            if (KE) {
                FLD(b, FF_DRK)[n] = dn[NV > 5 ? 5 : 0] * sc.rho;
                FLD(b, FF_DEPS)[n] = dn[NV > 6 ? 6 : 0] * sc.rho;
            }
            double oj[NV], ok[NV];
            face_xfer<KE>(dn, sc, sc.fi, sc.ci, yi);
            face_xfer<KE>(dn, sc, sc.fj, sc.cj, oj);
            face_xfer<KE>(dn, sc, sc.fk, sc.ck, ok);
            const int wb = s & 1;
            if (jj < TJ - 1) {
#pragma unroll
                for (int r = 0; r < NV; r++) xj[wb][r][lane] = oj[r];
            } else if (j < b.jmax) {
                halo_put<NV, HQS>(hj_slot(sw, b, tj + 1, i, k), oj);
            }
            if (kk < TK - 1) {
#pragma unroll
                for (int r = 0; r < NV; r++) xk[wb][r][lane] = ok[r];
            } else if (k < b.kmax) {
                halo_put<NV, HQS>(hk_slot(sw, b, tk + 1, i, j), ok);
            }
            if (q + 1 < b.imax) cur = nxt;
        }
        wave_step_sync();
    }
}

// Wait until at most n vector-memory operations of this wave are outstanding (they
// complete in issue order).  The LDS rings below are filled by LDS-DMA loads, which the
// compiler does not order against the LDS reads of the rings; the memory clobber keeps
// those reads after the wait.
template <int N>
__device__ __forceinline__ void vm_wait()
{
    static_assert(N >= 0 && N <= 63, "vmcnt range");
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// One LDS-DMA load: every lane fetches the 4 bytes at its `src` into LDS byte address
// lds + 4*lane (M0 holds the wave's base).  Issued from inline assembly so that the
// compiler, which cannot tell which LDS bytes such a load writes, does not follow it with
// a full vmcnt wait before every LDS read; the kernels wait for these loads themselves
// (vm_wait).  SC1 selects the device-coherent form used for the halo slots.  The s_nop is
// the wait state an LDS-DMA load needs after an SALU write of M0.
template <bool SC1>
__device__ __forceinline__ void lds_dma4(const void *src, unsigned lds)
{
    if (SC1)
        asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off sc1"
                     ::"v"(src), "s"(lds) : "memory", "m0");
    else
        asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, off"
                     ::"v"(src), "s"(lds) : "memory", "m0");
}

__device__ __forceinline__ unsigned lds_addr(const void *p)
{
    return (unsigned)(size_t)(const __attribute__((address_space(3))) void *)p;
}

// ------------------------------------------------------------------------------------
// LU-SGS with staged cell records.  The (J,K)-tile wavefront above spends most of a step
// forming the cell's coefficients from some sixty scattered 8-byte loads per lane (each lane
// its own cache line).  Here a parallel pass (k_sweep_prep) forms, for every cell, the
// predictor's and the corrector's recurrence-independent values (the same operations as
// pred_coef/pred_basis, corr_coef/corr_basis) and stores them in the order the sweeps
// consume them: for tile t and corrector step s the 64 lanes' values of one record field are
// one contiguous 512-byte row.  A lane (jj,kk) reaches the cell of corrector step s when
// i = s - (jj + kk) + 1; the predictor reaches the same cell at step NSTEP - 1 - s, so both
// sweeps read whole rows.  The sweeps then carry only the recurrence and the three outgoing
// contributions per step.  The predictor's increments go to a row buffer the corrector reads
// the same way.  Arithmetic is that of the tile kernels, operation for operation.  A face
// basis is stored as its unit normal and first tangent: the second tangent is their cross
// product, which the sweeps form again with SURFVECTS's operations (basis_get), so the
// record is 48 fields instead of 56 and the results are the same bit for bit.  (Storing the
// normal alone and forming both tangents in the sweeps saved more bytes but put SURFVECTS's
// square root and division on the sweeps' step latency, and the V-cycle was no faster.)
// ------------------------------------------------------------------------------------
constexpr int NRF = 48;            // record fields per cell and sweep (a multiple of 4)
// predictor record: vol*d (7), -dt1*wk, dt1, wi, wj, SCINV, SCINP, 7 scalars, 3 x (normal,
// first tangent) (18), 3 x (X1, X2, X5), 1 unused
enum { RP_VD = 0, RP_MK = 7, RP_DT1, RP_WI, RP_WJ, RP_SCINV, RP_SCINP, RP_SC = 13,
       RP_N = 20, RP_X = 38 };
// corrector record: DIAGGS factor, vol, dt1*wk, dt1, wi, wj, SCINV, SCINP, 7 scalars,
// 3 x (normal, first tangent), 3 x (X1, X2, X5), CP, RO, 4 unused
enum { RC_FAC = 0, RC_VOL, RC_MK, RC_DT1, RC_WI, RC_WJ, RC_SCINV, RC_SCINP, RC_SC = 8,
       RC_N = 15, RC_X = 33, RC_CP = 42, RC_RHO = 43 };
constexpr int NB = 6;              // record fields of one face basis
static_assert(NRF % 4 == 0, "a record is a whole number of 1 KiB LDS-DMA loads in FP32 too");

constexpr int NDN = 8;             // rows of predictor increments per step (7 used)
struct SweepSt {
    int nstep;                      // IMAX + TJ + TK - 2
    void *recp, *recc;              // [tile][step][NRF][TW] of the record type (FP64 or FP32)
    double *dn;                     // predictor increments [tile][step][NDN][TW]
};

__device__ __forceinline__ size_t st_row(const SweepSt &st, int tile, int s, int f)
{
    return (((size_t)tile * st.nstep + s) * NRF + f) * TW;
}

// the seven SYI/SYM1I (SYIKE/SYM1KE) scalars of a SweepCell in record order
template <bool KE>
__device__ __forceinline__ void sc_put(const SweepCell &c, double *v)
{
    if (KE) {
        v[0] = c.yp; v[1] = c.roc; v[2] = c.pc2; v[3] = c.proc; v[4] = c.ydrp; v[5] = c.drhro;
        v[6] = c.ro667;
    } else {
        v[0] = c.ro; v[1] = c.roc; v[2] = c.pc2; v[3] = c.proc; v[4] = v[5] = v[6] = 0.;
    }
}

template <bool KE>
__device__ __forceinline__ void sc_get(SweepCell &c, const double *v)
{
    if (KE) {
        c.yp = v[0]; c.roc = v[1]; c.pc2 = v[2]; c.proc = v[3]; c.ydrp = v[4]; c.drhro = v[5];
        c.ro667 = v[6];
    } else {
        c.ro = v[0]; c.roc = v[1]; c.pc2 = v[2]; c.proc = v[3];
    }
}

// a face basis in a record: the unit normal and the first tangent; the second tangent is
// their cross product, with the operations of SURFVECTS (a31 = a12*b3 - a13*b2, ...)
__device__ __forceinline__ void basis_put(const FaceBasis &f, double *v)
{
    v[0] = f.nx; v[1] = f.ny; v[2] = f.nz; v[3] = f.a21; v[4] = f.a22; v[5] = f.a23;
}

__device__ __forceinline__ FaceBasis basis_get(const double *v)
{
    FaceBasis f;
    f.nx = v[0]; f.ny = v[1]; f.nz = v[2]; f.a21 = v[3]; f.a22 = v[4]; f.a23 = v[5];
    f.a31 = f.ny * f.a23 - f.nz * f.a22;
    f.a32 = f.nz * f.a21 - f.nx * f.a23;
    f.a33 = f.nx * f.a22 - f.ny * f.a21;
    return f;
}

// Both records of every cell: thread (lane, step) of tile blockIdx.x, PSW steps per
// workgroup, so that a workgroup's loads of a column run along I and its record stores are
// whole rows.  The records are read back only by the sweeps, long after: non-temporal
// stores keep them from displacing the caches' working set on the way out.
constexpr int PSW = 4;             // record-pass steps per workgroup, and waves per workgroup

template <bool KE, typename RT = double>
__global__ void __launch_bounds__(PSW * TW) k_sweep_prep(Prm p, Blk b, Sweep sw, SweepSt st,
                                                        TileGrid tg)
{
    RT *const recp = (RT *)st.recp, *const recc = (RT *)st.recc;
    constexpr int NV = KE ? 7 : 5;
    // logical tiles: (step chunk, tile), chunks of one tile consecutive, so that a tile's
    // columns stay in one XCD's L2 while their cells are read
    const int tl = xcd_tile(tg);
    if (tl < 0) return;
    // wave w takes TW/PSW lanes over the workgroup's PSW steps, the steps fastest: the PSW
    // consecutive steps of a lane are PSW consecutive cells along I, so a load instruction
    // touches TW/PSW short runs instead of TW separate cache lines
    static_assert(TW % PSW == 0, "the lanes of a wave split over the record-pass steps");
    const int li = threadIdx.x % TW;
    const int lane = (TW / PSW) * (int)(threadIdx.x / TW) + li / PSW;
    const int s = (tl % tg.nx) * PSW + li % PSW;
    if (s >= st.nstep) return;
    const int tile = tl / tg.nx, tj = tile % sw.ntj, tk = tile / sw.ntj;
    const int jj = lane % TJ, kk = lane / TJ;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1, i = s - (jj + kk) + 1;
    if (j > b.jmax || k > b.kmax || i < 1 || i > b.imax) return;
    const int n = cid(b, i, j, k);
    double v[NRF];
    {   // predictor (PREDGS / PRKEGS): lower faces
        SweepRaw r;
        sweep_load<false, KE>(b, n, r);
        SweepCell c;
        pred_prep<KE>(p, r, i, j, k, c);
#pragma unroll
        for (int q = 0; q < 7; q++) v[RP_VD + q] = q < NV ? c.vol * c.d[q] : 0.;
        v[RP_MK] = -c.dt1 * c.wk;
        v[RP_DT1] = c.dt1; v[RP_WI] = c.wi; v[RP_WJ] = c.wj;
        v[RP_SCINV] = c.scinv; v[RP_SCINP] = c.scinp;
        sc_put<KE>(c, v + RP_SC);
        basis_put(c.fi, v + RP_N); basis_put(c.fj, v + RP_N + NB); basis_put(c.fk, v + RP_N + 2 * NB);
        const Coef *cf[3] = {&c.ci, &c.cj, &c.ck};
#pragma unroll
        for (int f = 0; f < 3; f++) {
            v[RP_X + 3 * f] = cf[f]->x1; v[RP_X + 3 * f + 1] = cf[f]->x2; v[RP_X + 3 * f + 2] = cf[f]->x5;
        }
        static_assert(RP_X + 9 <= NRF, "the predictor record fits NRF fields");
        for (int f = RP_X + 9; f < NRF; f++) v[f] = 0.;
#pragma unroll
        for (int f = 0; f < NRF; f++) __builtin_nontemporal_store((RT)v[f], &recp[st_row(st, tile, s, f) + lane]);
    }
    {   // corrector (DIAGGS + CORRGS / DIKEGS + COKEGS): upper faces
        SweepRaw r;
        sweep_load<true, KE>(b, n, r);
        SweepCell c;
        corr_prep<KE>(p, r, c);
        v[RC_FAC] = c.fac; v[RC_VOL] = c.vol; v[RC_MK] = c.dt1 * c.wk; v[RC_DT1] = c.dt1;
        v[RC_WI] = c.wi; v[RC_WJ] = c.wj; v[RC_SCINV] = c.scinv; v[RC_SCINP] = c.scinp;
        sc_put<KE>(c, v + RC_SC);
        basis_put(c.fi, v + RC_N); basis_put(c.fj, v + RC_N + NB); basis_put(c.fk, v + RC_N + 2 * NB);
        const Coef *cf[3] = {&c.ci, &c.cj, &c.ck};
#pragma unroll
        for (int f = 0; f < 3; f++) {
            v[RC_X + 3 * f] = cf[f]->x1; v[RC_X + 3 * f + 1] = cf[f]->x2; v[RC_X + 3 * f + 2] = cf[f]->x5;
        }
        v[RC_CP] = c.cp; v[RC_RHO] = c.rho;
        static_assert(RC_RHO < NRF, "the corrector record fits NRF fields");
        for (int f = RC_RHO + 1; f < NRF; f++) v[f] = 0.;
#pragma unroll
        for (int f = 0; f < NRF; f++) __builtin_nontemporal_store((RT)v[f], &recc[st_row(st, tile, s, f) + lane]);
    }
}

// The sweeps stream their records through an LDS ring: at step s the wave starts the copy of
// the record of step s + RS - 1 (and, in the corrector, the predictor's increments of that
// step) with 16-byte LDS-DMA loads, one instruction per 1 KiB of the contiguous record, and
// reads the record of step s from its slot.  The copies are issued from inline assembly, so
// the compiler does not wait for all of them before each LDS read; the sweep waits itself
// until the copy of the slot it reads is complete (vm_wait: loads complete in issue order,
// and the copies of the RS - 2 later steps may stay in flight).  For the same reason a step
// loads its halo words before it issues its copy: the wait for them then covers only the
// copies of earlier steps, which have had a step or more to arrive.
template <typename RT>
constexpr int rbytes() { return NRF * TW * (int)sizeof(RT); }   // one record: 24 or 12 KiB
constexpr int DBYTES = NDN * TW * 8;            // one step of increments: 4 KiB

// four 1 KiB loads from one address pair: the instruction offset advances the global and the
// LDS address alike.  The records are read once: non-temporal (nt), as they were written.
__device__ __forceinline__ void lds_dma16x4(const char *src, unsigned lds)
{
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\t"
                 "global_load_lds_dwordx4 %0, off nt\n\t"
                 "global_load_lds_dwordx4 %0, off offset:1024 nt\n\t"
                 "global_load_lds_dwordx4 %0, off offset:2048 nt\n\t"
                 "global_load_lds_dwordx4 %0, off offset:3072 nt"
                 ::"v"(src), "s"(lds) : "memory", "m0");
}

// one 1 KiB load (the remainder of a record that is not a whole number of 4 KiB groups)
__device__ __forceinline__ void lds_dma16x1(const char *src, unsigned lds)
{
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\t"
                 "global_load_lds_dwordx4 %0, off nt"
                 ::"v"(src), "s"(lds) : "memory", "m0");
}

// copy `bytes` (a multiple of 1 KiB) from global src to LDS dst, 16 bytes per lane per load:
// groups of four loads, then single loads
template <int BYTES>
__device__ __forceinline__ void ring_copy(const void *src, void *dst, int lane)
{
    static_assert(BYTES % 1024 == 0, "whole 1 KiB loads");
    const char *g = (const char *)src + lane * 16;
    const unsigned l = lds_addr(dst);
#pragma unroll
    for (int m = 0; m < BYTES / 4096; m++) lds_dma16x4(g + m * 4096, l + m * 4096);
#pragma unroll
    for (int m = (BYTES / 4096) * 4; m < BYTES / 1024; m++) lds_dma16x1(g + m * 1024, l + m * 1024);
}

// Halo words through LDS as well.  At a step, the edge lanes that read across one tile face
// hold cells on one anti-diagonal d = I + offset - 1, so the step's words of that face are
// one contiguous block of the slot layout (7 words x 8 lanes, 448 bytes).  The sweeps copy
// the block of step s + 2 to an LDS ring at step s (lanes 0-27, 16 bytes each, device-coherent
// loads), so the wait for it covers loads issued two steps earlier.  A word that was still
// empty when copied (either 32-bit half equal to the marker's, so a copy racing the
// producer's store is caught too) is polled again with 64-bit loads.
constexpr int HBLK = HS * TK;                   // words per block (448 bytes)
constexpr int HR = 3;                           // LDS halo slots: steps s, s + 1, s + 2

__device__ __forceinline__ const unsigned long long *hj_blk(const Sweep &s, const Blk &b, int f,
                                                             int t, int d)
{
    return s.hj + (((size_t)f * s.ntk + t) * (b.imax + TK) + d) * HBLK;
}

__device__ __forceinline__ const unsigned long long *hk_blk(const Sweep &s, const Blk &b, int f,
                                                             int t, int d)
{
    return s.hk + (((size_t)f * s.ntj + t) * (b.imax + TJ) + d) * HBLK;
}

__device__ __forceinline__ void halo_dma(const unsigned long long *blk, double *dst, int lane)
{
    static_assert(HBLK * 8 % 16 == 0 && HBLK * 8 / 16 <= TW, "one 16-byte load per lane");
    if (lane < HBLK * 8 / 16)
        asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off sc1"
                     ::"v"((const char *)blk + lane * 16), "s"(lds_addr(dst)) : "memory", "m0");
}

__device__ __forceinline__ bool half_empty(unsigned long long x)
{
    return (unsigned)x == (unsigned)HALO_EMPTY || (unsigned)(x >> 32) == (unsigned)(HALO_EMPTY >> 32);
}

// The NV words a lane takes across one tile face at a step, all from LDS: the in-tile
// neighbour's output of the previous step (words TW apart), the upstream tile's halo block
// (words TK apart) or a zero block at the block edge (words TK apart).  One address and
// stride per lane replace a branch per source.  marker_hit reads the words and tests them;
// halo_poll polls the halo words copied while still empty (either 32-bit half equal to the
// marker's) from their slot, in a branch that is rarely taken: every round reloads all the
// words still missing at once, so a lane that has caught up with its upstream tile pays one
// memory round trip per step, not one per word (the producer stores the words together).
// Marker halves in words from the other two sources are ordinary values and are not
// polled.  The staged sweeps leave their slots filled: a fill kernel empties all of them
// after the sweep, which costs less than a device-coherent store per word on the
// wavefront's critical path.
template <int NV>
__device__ __forceinline__ void halo_poll(unsigned long long *slot, unsigned int *err,
                                          unsigned long long h[NV])
{
    unsigned need = 0;                   // bit q: word q still to be read from its slot
#pragma unroll
    for (int q = 0; q < NV; q++)
        if (half_empty(h[q])) need |= 1u << q;
    int spins = 0;
    while (need) {
        if ((++spins & 255) == 0 &&
            (spins > SPIN_LIMIT || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            if (spins > SPIN_LIMIT) __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int q = 0; q < NV; q++)
                if (need >> q & 1u) h[q] = 0;
            break;
        }
        unsigned long long x[NV];
#pragma unroll
        for (int q = 0; q < NV; q++)
            x[q] = (need >> q & 1u) ? __hip_atomic_load(slot + q * HQS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : 0ull;
        unsigned left = 0;
#pragma unroll
        for (int q = 0; q < NV; q++) {
            if (!(need >> q & 1u)) continue;
            if (x[q] == HALO_EMPTY) left |= 1u << q;
            else h[q] = x[q];
        }
        need = left;
        if (need) __builtin_amdgcn_s_sleep(1);
    }
}

template <int NV>
__device__ __forceinline__ unsigned marker_hit(const double *src, int stride, unsigned long long h[NV])
{
    unsigned hit = 0xffffffffu;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        h[q] = (unsigned long long)__double_as_longlong(src[q * stride]);
        const unsigned lo = (unsigned)h[q] ^ (unsigned)HALO_EMPTY;
        const unsigned hi = (unsigned)(h[q] >> 32) ^ (unsigned)(HALO_EMPTY >> 32);
        hit = min(hit, min(lo, hi));
    }
    return hit;
}

// Both faces' inputs of a step (J into vj, K into vk): all 2 x NV LDS reads are issued
// together, and one wave-uniform vote guards both polling regions.
template <int NV, typename SlotJ, typename SlotK>
__device__ __forceinline__ void faces_in(const double *sj, int strj, bool haloj, SlotJ slotj,
                                         const double *sk, int strk, bool halok, SlotK slotk,
                                         unsigned int *err, double vj[NV], double vk[NV])
{
    unsigned long long hj[NV], hk[NV];
    const unsigned hitj = marker_hit<NV>(sj, strj, hj);
    const unsigned hitk = marker_hit<NV>(sk, strk, hk);
    const bool pj = haloj && hitj == 0, pk = halok && hitk == 0;
    // a wave-uniform test first, so that the common case skips both regions in one branch
    if (__builtin_amdgcn_ballot_w64(pj || pk)) {
        if (pj) halo_poll<NV>(slotj(), err, hj);
        if (pk) halo_poll<NV>(slotk(), err, hk);
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
        vj[q] = __longlong_as_double((long long)hj[q]);
        vk[q] = __longlong_as_double((long long)hk[q]);
    }
}

// Where a lane takes one face's inputs from: a lane-constant LDS base (its neighbour's
// column of the exchange buffers, its column of the halo ring, or the zero block) to which
// each step adds the uniform offset of the exchange buffer or of the halo ring slot under a
// lane mask, so that the choice costs no branch.
struct FaceSrc {
    const char *base;
    unsigned xmask, hmask;       // all ones on in-tile lanes / on halo lanes
    int stride;                  // doubles between words
    __device__ FaceSrc(bool edge, bool in, const double *x, const double *h, const double *z)
    {
        base = (const char *)(edge ? z : in ? x : h);
        xmask = !edge && in ? ~0u : 0u;
        hmask = !edge && !in ? ~0u : 0u;
        stride = !edge && in ? TW : TK;
    }
    __device__ const double *at(unsigned xoff, unsigned hoff) const
    {
        return (const double *)(base + ((xoff & xmask) | (hoff & hmask)));
    }
};

// PREDGS / PRKEGS on staged records, the (J,K)-tile wavefront of k_pred_tiles.  Issue order
// within step s: the step's stores, the halo blocks of step s + 2, the record of step s + 3.
// At the top of step s at least the two later record copies (2 NLD loads) follow the halo
// blocks of step s, which follow the record of step s, so vmcnt(2 NLD) covers both.
template <bool KE, typename RT = double>
__global__ void __launch_bounds__(TW) k_pred_st(Prm p, Blk b, Sweep sw, SweepSt st)
{
    constexpr int NV = KE ? 7 : 5;
    constexpr int RS = 4;                               // ring slots: 3 steps of lead
    constexpr int RBYTES = rbytes<RT>();
    constexpr int NLD = RBYTES / 1024;                  // record loads per step
    static_assert(NLD * (RS - 2) <= 63, "vmcnt range");
    __shared__ __attribute__((aligned(16))) RT ring[RS][NRF * TW];
    const RT *const recp = (const RT *)st.recp;
    __shared__ __attribute__((aligned(16))) double hring[HR][2][HBLK];
    __shared__ double xj[2][NV][TW], xk[2][NV][TW], zx[NV * TK];
    const int lane = threadIdx.x, jj = lane % TJ, kk = lane / TJ;
    if (lane < NV * TK) zx[lane] = 0.;
    wave_step_sync();
    const int t = tile_ticket(sw);
    const int tj = sw.ntj - 1 - t % sw.ntj, tk = sw.ntk - 1 - t / sw.ntj;
    const int tile = tk * sw.ntj + tj;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1;
    const bool col = j <= b.jmax && k <= b.kmax;
    const bool hasj = tj + 1 < sw.ntj, hask = tk + 1 < sw.ntk;   // upstream tiles
    const int lag = (TJ - 1 - jj) + (TK - 1 - kk);
    const int nsteps = st.nstep;
    // the edge lanes of step s read anti-diagonal d = IMAX + 6 - s of the upstream faces
    const unsigned long long *hj0 = hj_blk(sw, b, tj + 1, tk, 0), *hk0 = hk_blk(sw, b, tk + 1, tj, 0);
    auto halo_copy = [&](int s) {
        const int d = b.imax + TK - 2 - s;
        if (d < 0 || d > b.imax + TK - 2) return;
        if (hasj) halo_dma(hj0 + (size_t)d * HBLK, hring[s % HR][0], lane);
        if (hask) halo_dma(hk0 + (size_t)d * HBLK, hring[s % HR][1], lane);
    };
    auto rec_copy = [&](int s) {
        ring_copy<RBYTES>(recp + st_row(st, tile, nsteps - 1 - s, 0), ring[s % RS], lane);
    };
    const FaceSrc fj(j == b.jmax, jj < TJ - 1, &xj[0][0][min(lane + 1, TW - 1)], hring[0][0] + kk, zx + kk);
    const FaceSrc fk(k == b.kmax, kk < TK - 1, &xk[0][0][min(lane + TJ, TW - 1)], hring[0][1] + jj, zx + jj);
    double yi[NV];
#pragma unroll
    for (int r = 0; r < NV; r++) yi[r] = 0.;
    rec_copy(0);
    halo_copy(0);
    if (1 < nsteps) rec_copy(1);
    halo_copy(1);
    if (2 < nsteps) rec_copy(2);
    for (int s = 0; s < nsteps; s++) {
        const int row = nsteps - 1 - s;
        if (s + RS - 2 < nsteps) vm_wait<NLD * (RS - 2)>(); else vm_wait<0>();
        const RT *cur = ring[s % RS] + lane;
        const int q = s - lag;
        if (col && q >= 0 && q < b.imax) {
            const int i = b.imax - q;
            const int rb = (s + 1) & 1;
            double bj[NV], ckc[NV];
            const unsigned xoff = rb * (NV * TW * 8), hoff = (s % HR) * (2 * HBLK * 8);
            faces_in<NV>(fj.at(xoff, hoff), fj.stride, fj.hmask != 0,
                         [&] { return hj_slot(sw, b, tj + 1, i, k); },
                         fk.at(xoff, hoff), fk.stride, fk.hmask != 0,
                         [&] { return hk_slot(sw, b, tk + 1, i, j); }, sw.err, bj, ckc);
            // the line-buffer shift of PREDGS (imps4.f:556-562) never hands the I+1
            // contribution to a cell of the row J = JMAX; PRKEGS's (imps4.f:1762) does
            if (!KE && j == b.jmax) {
#pragma unroll
                for (int r = 0; r < NV; r++) yi[r] = 0.;
            }
            double dn[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) {
                double cd = cur[RP_MK * TW] * ckc[r] + cur[(RP_VD + r) * TW];
                cd = cd - cur[RP_DT1 * TW] * (cur[RP_WI * TW] * yi[r] + cur[RP_WJ * TW] * bj[r]);
                dn[r] = r != 4 ? cd * cur[RP_SCINV * TW] : cd * cur[RP_SCINP * TW];
            }
            double *dno = st.dn + (((size_t)tile * nsteps + row) * NDN) * TW + lane;
#pragma unroll
            for (int r = 0; r < NV; r++) dno[(size_t)r * TW] = dn[r];
            SweepCell sc;
            double scv[7];
#pragma unroll
            for (int m = 0; m < 7; m++) scv[m] = cur[(RP_SC + m) * TW];
            sc_get<KE>(sc, scv);
            double bv[3 * NB], xv[9];
#pragma unroll
            for (int m = 0; m < 3 * NB; m++) bv[m] = cur[(RP_N + m) * TW];
#pragma unroll
            for (int m = 0; m < 9; m++) xv[m] = cur[(RP_X + m) * TW];
            Coef ci, cj, ck;
            ci.x1 = xv[0]; ci.x2 = xv[1]; ci.x5 = xv[2];
            cj.x1 = xv[3]; cj.x2 = xv[4]; cj.x5 = xv[5];
            ck.x1 = xv[6]; ck.x2 = xv[7]; ck.x5 = xv[8];
            double oj[NV], ok[NV];
            face_xfer<KE>(dn, sc, basis_get(bv), ci, yi);
            face_xfer<KE>(dn, sc, basis_get(bv + NB), cj, oj);
            face_xfer<KE>(dn, sc, basis_get(bv + 2 * NB), ck, ok);
            // every lane leaves its outputs in LDS (an edge lane's are never read there)
            const int wb = s & 1;
#pragma unroll
            for (int r = 0; r < NV; r++) {
                xj[wb][r][lane] = oj[r];
                xk[wb][r][lane] = ok[r];
            }
            if (jj == 0 && j > 1) halo_put<NV, HQS>(hj_slot(sw, b, tj, i, k), oj);
            if (kk == 0 && k > 1) halo_put<NV, HQS>(hk_slot(sw, b, tk, i, j), ok);
        }
        // the LDS reads of this step are complete before its slots are refilled
        wave_step_sync();
        halo_copy(s + 2);
        if (s + RS - 1 < nsteps) rec_copy(s + RS - 1);
    }
    vm_wait<0>();
}

// DIAGGS + CORRGS / DIKEGS + COKEGS on staged records, the wavefront of k_corr_tiles, with
// the end of IMPS (DRO/CP, D(rho k) and D(rho omega) times RO).  Issue order within step s:
// the step's stores, the halo blocks of step s + 2, the record and increments of step s + 2;
// at the top of step s the copy of step s + 1 (32 loads) follows that of step s, which
// follows the halo blocks of step s, so vmcnt(32) covers both.
template <bool KE, typename RT = double>
__global__ void __launch_bounds__(TW) k_corr_st(Prm p, Blk b, Sweep sw, SweepSt st)
{
    constexpr int NV = KE ? 7 : 5;
    constexpr int RS = 3;                               // ring slots: 2 steps of lead
    constexpr int RBYTES = rbytes<RT>();
    constexpr int NLD = (RBYTES + DBYTES) / 1024;       // record loads per step
    static_assert(NLD * (RS - 2) <= 63, "vmcnt range");
    __shared__ __attribute__((aligned(16))) RT ring[RS][NRF * TW];
    const RT *const recc = (const RT *)st.recc;
    __shared__ __attribute__((aligned(16))) double dring[RS][NDN * TW];
    __shared__ __attribute__((aligned(16))) double hring[HR][2][HBLK];
    __shared__ double xj[2][NV][TW], xk[2][NV][TW], zx[NV * TK];
    const int lane = threadIdx.x, jj = lane % TJ, kk = lane / TJ;
    if (lane < NV * TK) zx[lane] = 0.;
    wave_step_sync();
    const int t = tile_ticket(sw);
    const int tj = t % sw.ntj, tk = t / sw.ntj;
    const int tile = tk * sw.ntj + tj;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1;
    const bool col = j <= b.jmax && k <= b.kmax;
    const bool hasj = tj > 0, hask = tk > 0;                  // upstream tiles
    const int lag = jj + kk;
    const int nsteps = st.nstep;
    // the edge lanes of step s read anti-diagonal d = s of the upstream faces
    const unsigned long long *hj0 = hj_blk(sw, b, tj, tk, 0), *hk0 = hk_blk(sw, b, tk, tj, 0);
    auto halo_copy = [&](int s) {
        if (s > b.imax + TK - 2) return;
        if (hasj) halo_dma(hj0 + (size_t)s * HBLK, hring[s % HR][0], lane);
        if (hask) halo_dma(hk0 + (size_t)s * HBLK, hring[s % HR][1], lane);
    };
    const FaceSrc fj(j == 1, jj > 0, &xj[0][0][max(lane - 1, 0)], hring[0][0] + kk, zx + kk);
    const FaceSrc fk(k == 1, kk > 0, &xk[0][0][max(lane - TJ, 0)], hring[0][1] + jj, zx + jj);
    auto rec_copy = [&](int s) {
        ring_copy<RBYTES>(recc + st_row(st, tile, s, 0), ring[s % RS], lane);
        ring_copy<DBYTES>(st.dn + (((size_t)tile * nsteps + s) * NDN) * TW, dring[s % RS], lane);
    };
    double yi[NV];
#pragma unroll
    for (int r = 0; r < NV; r++) yi[r] = 0.;
    halo_copy(0);
    rec_copy(0);
    halo_copy(1);
    if (1 < nsteps) rec_copy(1);
    for (int s = 0; s < nsteps; s++) {
        if (s + RS - 2 < nsteps) vm_wait<NLD * (RS - 2)>(); else vm_wait<0>();
        const RT *cur = ring[s % RS] + lane;
        const double *dni = dring[s % RS] + lane;
        const int q = s - lag;
        if (col && q >= 0 && q < b.imax) {
            const int i = q + 1;
            const int n = cid(b, i, j, k);
            const int rb = (s + 1) & 1;
            double bj[NV], ckc[NV];
            const unsigned xoff = rb * (NV * TW * 8), hoff = (s % HR) * (2 * HBLK * 8);
            faces_in<NV>(fj.at(xoff, hoff), fj.stride, fj.hmask != 0,
                         [&] { return hj_slot(sw, b, tj, i, k); },
                         fk.at(xoff, hoff), fk.stride, fk.hmask != 0,
                         [&] { return hk_slot(sw, b, tk, i, j); }, sw.err, bj, ckc);
            double d[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) d[r] = dni[r * TW] * cur[RC_FAC * TW];
            double dn[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) {
                double cd = cur[RC_MK * TW] * ckc[r] + cur[RC_VOL * TW] * d[r];
                cd = cd + cur[RC_DT1 * TW] * (cur[RC_WI * TW] * yi[r] + cur[RC_WJ * TW] * bj[r]);
                dn[r] = r != 4 ? cd * cur[RC_SCINV * TW] : cd * cur[RC_SCINP * TW];
            }
            FLD(b, FF_DRO)[n] = dn[0] / cur[RC_CP * TW];
#pragma unroll
            for (int r = 1; r < 5; r++) FLD(b, FF_DRO + r)[n] = dn[r];
            if (KE) {
                FLD(b, FF_DRK)[n] = dn[NV > 5 ? 5 : 0] * cur[RC_RHO * TW];
                FLD(b, FF_DEPS)[n] = dn[NV > 6 ? 6 : 0] * cur[RC_RHO * TW];
            }
            SweepCell sc;
            double scv[7];
#pragma unroll
            for (int m = 0; m < 7; m++) scv[m] = cur[(RC_SC + m) * TW];
            sc_get<KE>(sc, scv);
            double bv[3 * NB], xv[9];
#pragma unroll
            for (int m = 0; m < 3 * NB; m++) bv[m] = cur[(RC_N + m) * TW];
#pragma unroll
            for (int m = 0; m < 9; m++) xv[m] = cur[(RC_X + m) * TW];
            Coef ci, cj, ck;
            ci.x1 = xv[0]; ci.x2 = xv[1]; ci.x5 = xv[2];
            cj.x1 = xv[3]; cj.x2 = xv[4]; cj.x5 = xv[5];
            ck.x1 = xv[6]; ck.x2 = xv[7]; ck.x5 = xv[8];
            double oj[NV], ok[NV];
            face_xfer<KE>(dn, sc, basis_get(bv), ci, yi);
            face_xfer<KE>(dn, sc, basis_get(bv + NB), cj, oj);
            face_xfer<KE>(dn, sc, basis_get(bv + 2 * NB), ck, ok);
            // every lane leaves its outputs in LDS (an edge lane's are never read there)
            const int wb = s & 1;
#pragma unroll
            for (int r = 0; r < NV; r++) {
                xj[wb][r][lane] = oj[r];
                xk[wb][r][lane] = ok[r];
            }
            if (jj == TJ - 1 && j < b.jmax) halo_put<NV, HQS>(hj_slot(sw, b, tj + 1, i, k), oj);
            if (kk == TK - 1 && k < b.kmax) halo_put<NV, HQS>(hk_slot(sw, b, tk + 1, i, j), ok);
        }
        // the LDS reads of this step are complete before its slots are refilled
        wave_step_sync();
        halo_copy(s + 2);
        if (s + RS - 1 < nsteps) rec_copy(s + RS - 1);
    }
    vm_wait<0>();
}

// ------------------------------------------------------------------------------------
// The staged sweeps with four waves per tile.  Timestamps taken inside the one-wave kernels
// above (s_memtime at the phase boundaries of every step, 192x96x96 level) split a step of
// about 4,000 cycles into the face inputs (about 900: LDS reads of the exchange buffers and
// the halo ring, the marker tests), the recurrence and the three outgoing face
// contributions (about 2,000) and the issue of the next record's LDS-DMA copies (about
// 600: 28 loads of 1 KiB, which the memory pipeline accepts at about 64 bytes a clock).
// All three are latencies on the step's critical path.  Here, per tile:
// * waves 0-2 all form dn of the step from the same inputs, and wave f then forms only the
//   contribution across face f (0: I, 1: J, 2: K) and leaves it in LDS; wave 1 publishes
//   the J-face halo words and wave 2 the K-face ones;
// * wave 3 issues every LDS-DMA copy (records, increments, halo blocks) and waits for them,
//   so the three compute waves never wait on memory;
// * one workgroup barrier per step separates a step's LDS writes from the next step's
//   reads (the exchange buffers are double-buffered by step parity), and the copy wave
//   reaches it only once the copies of the next step are complete.
// Arithmetic is that of k_pred_st / k_corr_st, operation for operation, so the results
// are the same bit for bit.
// ------------------------------------------------------------------------------------
constexpr int SW = 3;              // compute waves per workgroup: one per face
constexpr int WGW = SW + 1;        // and the copy wave

// The workgroup's exchange point between two steps: LDS writes before it are visible to
// every wave after it.  Global memory is not ordered (the halo slots carry their own
// protocol, and the copy wave waits for its LDS-DMA copies with vm_wait before it).
__device__ __forceinline__ void wg_step_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// the tile ticket of a multi-wave workgroup: taken by its first lane, shared through LDS
__device__ __forceinline__ int tile_ticket_wg(const Sweep &s, int *slot)
{
    if (threadIdx.x == 0) *slot = (int)(atomicAdd(s.ticket, 1ULL) % (unsigned long long)(s.ntj * s.ntk));
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// PREDGS / PRKEGS on staged records, four waves per (J,K) tile.  The copy wave's issue order
// within step s: the halo blocks of step s + 2, the record of step s + 3; at the top of step
// s the two later records (2 NLD loads) follow the halo blocks of step s, which follow the
// record of step s, so waiting until 2 NLD loads remain covers both.
template <bool KE, typename RT = double>
__global__ void __launch_bounds__(WGW * TW) k_pred_wg(Prm p, Blk b, Sweep sw, SweepSt st)
{
    constexpr int NV = KE ? 7 : 5;
    constexpr int RS = 4;                               // ring slots: 3 steps of lead
    constexpr int RBYTES = rbytes<RT>();
    constexpr int NLD = RBYTES / 1024;                  // record loads per step
    static_assert(NLD * (RS - 2) <= 63, "vmcnt range");
    __shared__ __attribute__((aligned(16))) RT ring[RS][NRF * TW];
    __shared__ __attribute__((aligned(16))) double hring[HR][2][HBLK];
    __shared__ double xo[SW][2][NV][TW], zx[NV * TK];
    __shared__ int tkt;
    const RT *const recp = (const RT *)st.recp;
    const int lane = threadIdx.x % TW, w = __builtin_amdgcn_readfirstlane(threadIdx.x / TW);
    const int jj = lane % TJ, kk = lane / TJ;
    if (w == 0) {
        if (lane < NV * TK) zx[lane] = 0.;
        // a lane's I-face input before its first cell is zero
#pragma unroll
        for (int r = 0; r < NV; r++) xo[0][0][r][lane] = xo[0][1][r][lane] = 0.;
    }
    const int t = tile_ticket_wg(sw, &tkt);
    const int tj = sw.ntj - 1 - t % sw.ntj, tk = sw.ntk - 1 - t / sw.ntj;
    const int tile = tk * sw.ntj + tj;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1;
    const bool col = j <= b.jmax && k <= b.kmax;
    const bool hasj = tj + 1 < sw.ntj, hask = tk + 1 < sw.ntk;   // upstream tiles
    const int lag = (TJ - 1 - jj) + (TK - 1 - kk);
    const int nsteps = st.nstep;
    const unsigned long long *hj0 = hj_blk(sw, b, tj + 1, tk, 0), *hk0 = hk_blk(sw, b, tk + 1, tj, 0);
    auto halo_copy = [&](int s) {
        const int d = b.imax + TK - 2 - s;
        if (d < 0 || d > b.imax + TK - 2) return;
        if (hasj) halo_dma(hj0 + (size_t)d * HBLK, hring[s % HR][0], lane);
        if (hask) halo_dma(hk0 + (size_t)d * HBLK, hring[s % HR][1], lane);
    };
    auto rec_copy = [&](int s) {
        ring_copy<RBYTES>(recp + st_row(st, tile, nsteps - 1 - s, 0), ring[s % RS], lane);
    };
    const FaceSrc fj(j == b.jmax, jj < TJ - 1, &xo[1][0][0][min(lane + 1, TW - 1)], hring[0][0] + kk, zx + kk);
    const FaceSrc fk(k == b.kmax, kk < TK - 1, &xo[2][0][0][min(lane + TJ, TW - 1)], hring[0][1] + jj, zx + jj);
    const bool copier = w == SW;
    if (copier) {
        rec_copy(0);
        halo_copy(0);
        if (1 < nsteps) rec_copy(1);
        halo_copy(1);
        if (2 < nsteps) rec_copy(2);
    }
    for (int s = 0; s < nsteps; s++) {
        const int row = nsteps - 1 - s;
        if (copier) {
            if (s + RS - 2 < nsteps) vm_wait<NLD * (RS - 2)>(); else vm_wait<0>();
        }
        wg_step_sync();
        if (copier) {
            // the slots of step s - 1, which every wave has finished reading
            halo_copy(s + 2);
            if (s + RS - 1 < nsteps) rec_copy(s + RS - 1);
            continue;
        }
        const RT *cur = ring[s % RS] + lane;
        const int q = s - lag;
        if (col && q >= 0 && q < b.imax) {
            const int i = b.imax - q;
            const int rb = (s + 1) & 1;
            double yi[NV], bj[NV], ckc[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) yi[r] = xo[0][rb][r][lane];
            const unsigned xoff = rb * (NV * TW * 8), hoff = (s % HR) * (2 * HBLK * 8);
            faces_in<NV>(fj.at(xoff, hoff), fj.stride, fj.hmask != 0,
                         [&] { return hj_slot(sw, b, tj + 1, i, k); },
                         fk.at(xoff, hoff), fk.stride, fk.hmask != 0,
                         [&] { return hk_slot(sw, b, tk + 1, i, j); }, sw.err, bj, ckc);
            // the line-buffer shift of PREDGS (imps4.f:556-562) never hands the I+1
            // contribution to a cell of the row J = JMAX; PRKEGS's (imps4.f:1762) does
            if (!KE && j == b.jmax) {
#pragma unroll
                for (int r = 0; r < NV; r++) yi[r] = 0.;
            }
            double dn[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) {
                double cd = cur[RP_MK * TW] * ckc[r] + cur[(RP_VD + r) * TW];
                cd = cd - cur[RP_DT1 * TW] * (cur[RP_WI * TW] * yi[r] + cur[RP_WJ * TW] * bj[r]);
                dn[r] = r != 4 ? cd * cur[RP_SCINV * TW] : cd * cur[RP_SCINP * TW];
            }
            // the increments for the corrector, their stores dealt between the waves
            double *dno = st.dn + (((size_t)tile * nsteps + row) * NDN) * TW + lane;
#pragma unroll
            for (int r = 0; r < NV; r++)
                if (r % SW == w) dno[(size_t)r * TW] = dn[r];
            SweepCell sc;
            double scv[7];
#pragma unroll
            for (int m = 0; m < 7; m++) scv[m] = cur[(RP_SC + m) * TW];
            sc_get<KE>(sc, scv);
            double bv[NB];
#pragma unroll
            for (int m = 0; m < NB; m++) bv[m] = cur[(RP_N + NB * w + m) * TW];
            Coef cf;
            cf.x1 = cur[(RP_X + 3 * w) * TW];
            cf.x2 = cur[(RP_X + 3 * w + 1) * TW];
            cf.x5 = cur[(RP_X + 3 * w + 2) * TW];
            double o[NV];
            face_xfer<KE>(dn, sc, basis_get(bv), cf, o);
            const int wb = s & 1;
#pragma unroll
            for (int r = 0; r < NV; r++) xo[w][wb][r][lane] = o[r];
            if (w == 1 && jj == 0 && j > 1) halo_put<NV, HQS>(hj_slot(sw, b, tj, i, k), o);
            if (w == 2 && kk == 0 && k > 1) halo_put<NV, HQS>(hk_slot(sw, b, tk, i, j), o);
        }
    }
    vm_wait<0>();
}

// DIAGGS + CORRGS / DIKEGS + COKEGS on staged records, four waves per tile, with the end of
// IMPS (DRO/CP, D(rho k) and D(rho omega) times RO).  The copy wave's issue order within
// step s: the halo blocks of step s + 2, the record and increments of step s + 2 (NLD
// loads); at the top of step s the copy of step s + 1 follows that of step s, which follows
// the halo blocks of step s, so waiting until NLD loads remain covers both.
template <bool KE, typename RT = double>
__global__ void __launch_bounds__(WGW * TW) k_corr_wg(Prm p, Blk b, Sweep sw, SweepSt st)
{
    constexpr int NV = KE ? 7 : 5;
    constexpr int RS = 3;                               // ring slots: 2 steps of lead
    constexpr int RBYTES = rbytes<RT>();
    constexpr int NLD = (RBYTES + DBYTES) / 1024;       // record loads per step
    static_assert(NLD * (RS - 2) <= 63, "vmcnt range");
    __shared__ __attribute__((aligned(16))) RT ring[RS][NRF * TW];
    __shared__ __attribute__((aligned(16))) double dring[RS][NDN * TW];
    __shared__ __attribute__((aligned(16))) double hring[HR][2][HBLK];
    __shared__ double xo[SW][2][NV][TW], zx[NV * TK];
    __shared__ int tkt;
    const RT *const recc = (const RT *)st.recc;
    const int lane = threadIdx.x % TW, w = __builtin_amdgcn_readfirstlane(threadIdx.x / TW);
    const int jj = lane % TJ, kk = lane / TJ;
    if (w == 0) {
        if (lane < NV * TK) zx[lane] = 0.;
#pragma unroll
        for (int r = 0; r < NV; r++) xo[0][0][r][lane] = xo[0][1][r][lane] = 0.;
    }
    const int t = tile_ticket_wg(sw, &tkt);
    const int tj = t % sw.ntj, tk = t / sw.ntj;
    const int tile = tk * sw.ntj + tj;
    const int j = tj * TJ + jj + 1, k = tk * TK + kk + 1;
    const bool col = j <= b.jmax && k <= b.kmax;
    const bool hasj = tj > 0, hask = tk > 0;                  // upstream tiles
    const int lag = jj + kk;
    const int nsteps = st.nstep;
    const unsigned long long *hj0 = hj_blk(sw, b, tj, tk, 0), *hk0 = hk_blk(sw, b, tk, tj, 0);
    auto halo_copy = [&](int s) {
        if (s > b.imax + TK - 2) return;
        if (hasj) halo_dma(hj0 + (size_t)s * HBLK, hring[s % HR][0], lane);
        if (hask) halo_dma(hk0 + (size_t)s * HBLK, hring[s % HR][1], lane);
    };
    const FaceSrc fj(j == 1, jj > 0, &xo[1][0][0][max(lane - 1, 0)], hring[0][0] + kk, zx + kk);
    const FaceSrc fk(k == 1, kk > 0, &xo[2][0][0][max(lane - TJ, 0)], hring[0][1] + jj, zx + jj);
    auto rec_copy = [&](int s) {
        ring_copy<RBYTES>(recc + st_row(st, tile, s, 0), ring[s % RS], lane);
        ring_copy<DBYTES>(st.dn + (((size_t)tile * nsteps + s) * NDN) * TW, dring[s % RS], lane);
    };
    const bool copier = w == SW;
    if (copier) {
        halo_copy(0);
        rec_copy(0);
        halo_copy(1);
        if (1 < nsteps) rec_copy(1);
    }
    for (int s = 0; s < nsteps; s++) {
        if (copier) {
            if (s + RS - 2 < nsteps) vm_wait<NLD * (RS - 2)>(); else vm_wait<0>();
        }
        wg_step_sync();
        if (copier) {
            halo_copy(s + 2);
            if (s + RS - 1 < nsteps) rec_copy(s + RS - 1);
            continue;
        }
        const RT *cur = ring[s % RS] + lane;
        const double *dni = dring[s % RS] + lane;
        const int q = s - lag;
        if (col && q >= 0 && q < b.imax) {
            const int i = q + 1;
            const int n = cid(b, i, j, k);
            const int rb = (s + 1) & 1;
            double yi[NV], bj[NV], ckc[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) yi[r] = xo[0][rb][r][lane];
            const unsigned xoff = rb * (NV * TW * 8), hoff = (s % HR) * (2 * HBLK * 8);
            faces_in<NV>(fj.at(xoff, hoff), fj.stride, fj.hmask != 0,
                         [&] { return hj_slot(sw, b, tj, i, k); },
                         fk.at(xoff, hoff), fk.stride, fk.hmask != 0,
                         [&] { return hk_slot(sw, b, tk, i, j); }, sw.err, bj, ckc);
            double d[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) d[r] = dni[r * TW] * cur[RC_FAC * TW];
            double dn[NV];
#pragma unroll
            for (int r = 0; r < NV; r++) {
                double cd = cur[RC_MK * TW] * ckc[r] + cur[RC_VOL * TW] * d[r];
                cd = cd + cur[RC_DT1 * TW] * (cur[RC_WI * TW] * yi[r] + cur[RC_WJ * TW] * bj[r]);
                dn[r] = r != 4 ? cd * cur[RC_SCINV * TW] : cd * cur[RC_SCINP * TW];
            }
            // the end of IMPS, its stores dealt between the waves
            if (w == 0) FLD(b, FF_DRO)[n] = dn[0] / cur[RC_CP * TW];
#pragma unroll
            for (int r = 1; r < 5; r++)
                if (r % SW == w) FLD(b, FF_DRO + r)[n] = dn[r];
            if (KE) {
                if (w == 2) FLD(b, FF_DRK)[n] = dn[NV > 5 ? 5 : 0] * cur[RC_RHO * TW];
                if (w == 0) FLD(b, FF_DEPS)[n] = dn[NV > 6 ? 6 : 0] * cur[RC_RHO * TW];
            }
            SweepCell sc;
            double scv[7];
#pragma unroll
            for (int m = 0; m < 7; m++) scv[m] = cur[(RC_SC + m) * TW];
            sc_get<KE>(sc, scv);
            double bv[NB];
#pragma unroll
            for (int m = 0; m < NB; m++) bv[m] = cur[(RC_N + NB * w + m) * TW];
            Coef cf;
            cf.x1 = cur[(RC_X + 3 * w) * TW];
            cf.x2 = cur[(RC_X + 3 * w + 1) * TW];
            cf.x5 = cur[(RC_X + 3 * w + 2) * TW];
            double o[NV];
            face_xfer<KE>(dn, sc, basis_get(bv), cf, o);
            const int wb = s & 1;
#pragma unroll
            for (int r = 0; r < NV; r++) xo[w][wb][r][lane] = o[r];
            if (w == 1 && jj == TJ - 1 && j < b.jmax) halo_put<NV, HQS>(hj_slot(sw, b, tj + 1, i, k), o);
            if (w == 2 && kk == TK - 1 && k < b.kmax) halo_put<NV, HQS>(hk_slot(sw, b, tk + 1, i, j), o);
        }
    }
    vm_wait<0>();
}

// ------------------------------------------------------------------------------------
// LU-SGS as a tile wavefront marching along K.  A wave owns a KI x KJ tile of (I, J)
// columns (16 x 4: each field load of a step is four coalesced 128-byte rows) and walks
// K; lane (ii, jj) runs ii + jj steps behind the tile's corner (fewer in the backward
// sweep's mirrored order), so at every step its I- and J-upstream neighbours inside the
// tile finished one step earlier and left their contributions in LDS.  The K-upstream
// contribution stays in registers.  Contributions that cross a tile edge go through HBM
// halo slots with the same empty-marker protocol as the J/K tiles above.
// ------------------------------------------------------------------------------------
constexpr int KI = 16, KJ = 4;
static_assert(KI * KJ == TW, "one wave per tile");

struct SweepK {
    int nti, ntj;                   // tiles along I and J
    unsigned long long *ticket;
    unsigned int *err;
    unsigned long long *hi, *hj;    // halo slots: I-faces (nti+1, KMAX, JMAX, HS), J-faces
};

__device__ __forceinline__ unsigned long long *hi_slot(const SweepK &s, const Blk &b, int f, int j,
                                                       int k)
{
    return s.hi + HS * (((size_t)f * b.kmax + (k - 1)) * b.jmax + (j - 1));
}

__device__ __forceinline__ unsigned long long *hjk_slot(const SweepK &s, const Blk &b, int f, int i,
                                                        int k)
{
    return s.hj + HS * (((size_t)f * b.kmax + (k - 1)) * b.imax + (i - 1));
}

__device__ __forceinline__ int tile_ticket_k(const SweepK &s)
{
    unsigned long long t = 0;
    if (threadIdx.x == 0) t = atomicAdd(s.ticket, 1ULL);
    t = __shfl(t, 0);
    return (int)(t % (unsigned long long)(s.nti * s.ntj));
}

// Halo words a lane has loaded for its next cell (I- and J-face contributions)
struct HaloIn {
    unsigned long long i[HS], j[HS];
};

// Halo words of one face taken from the LDS copy an LDS-DMA load made of the slot (two
// 32-bit halves per word, hb[2q][lane] and hb[2q+1][lane]).  The halves are separate
// loads, so a copy taken while the producer stored the word can mix the empty marker with
// the value: a copy with either half equal to the marker's is not used, and the word is
// then polled in HBM with 64-bit loads, as is a word still empty.  Every word is emptied
// again after use.
template <int NV>
__device__ __forceinline__ void halo_take(unsigned long long *slot, const unsigned int *hb,
                                          int lane, unsigned int *err, double v[NV])
{
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const unsigned int lo = hb[(2 * q) * TW + lane], hi = hb[(2 * q + 1) * TW + lane];
        unsigned long long x = (unsigned long long)lo | ((unsigned long long)hi << 32);
        if (lo == (unsigned int)HALO_EMPTY || hi == (unsigned int)(HALO_EMPTY >> 32)) x = HALO_EMPTY;
        int spins = 0;
        while (x == HALO_EMPTY) {
            if ((++spins & 255) == 0) {
                if (spins > SPIN_LIMIT) {
                    __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    x = 0;
                    break;
                }
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    x = 0;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(1);
            x = __hip_atomic_load(slot + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        v[q] = __longlong_as_double((long long)x);
        __hip_atomic_store(slot + q, HALO_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Raw values of a cell for the coefficient half of a sweep's preparation (pred_coef /
// corr_coef) and for the basis half (pred_basis / corr_basis)
template <bool FWD, bool KE>
__device__ __forceinline__ void sweep_load_coef(const Blk &b, int n, SweepRaw &r)
{
    const int IS = b.is, IL = b.il;
    constexpr int NV = KE ? 7 : 5;
    const int s1 = FWD ? 1 : -1, s2 = FWD ? IS : -IS, s3 = FWD ? IL : -IL;
#pragma unroll
    for (int q = 0; q < NV; q++) r.d[q] = FLD(b, FF_DRO + q)[n];
    r.u = FLD(b, FF_U)[n]; r.v = FLD(b, FF_V)[n]; r.w = FLD(b, FF_W)[n];
    r.c = FLD(b, FF_C)[n]; r.ro = FLD(b, FF_RO)[n]; r.vol = FLD(b, FF_VOL)[n];
    r.vis = FLD(b, FF_VIS)[n]; r.vist = FLD(b, FF_VIST)[n]; r.dtl = FLD(b, FF_DTL)[n];
    r.vol1 = FLD(b, FF_VOL)[n + s1]; r.ro1 = FLD(b, FF_RO)[n + s1];
    r.vis1 = FLD(b, FF_VIS)[n + s1]; r.vist1 = FLD(b, FF_VIST)[n + s1];
    r.vol2 = FLD(b, FF_VOL)[n + s2]; r.ro2 = FLD(b, FF_RO)[n + s2];
    r.vis2 = FLD(b, FF_VIS)[n + s2]; r.vist2 = FLD(b, FF_VIST)[n + s2];
    r.vol3 = FLD(b, FF_VOL)[n + s3]; r.ro3 = FLD(b, FF_RO)[n + s3];
    r.vis3 = FLD(b, FF_VIS)[n + s3]; r.vist3 = FLD(b, FF_VIST)[n + s3];
    r.a1 = FLD(b, FF_A1)[n]; r.a2 = FLD(b, FF_A2)[n]; r.a3 = FLD(b, FF_A3)[n];
    r.n1x = FLD(b, FF_A1X)[n]; r.n1y = FLD(b, FF_A1Y)[n]; r.n1z = FLD(b, FF_A1Z)[n];
    r.n2x = FLD(b, FF_A2X)[n]; r.n2y = FLD(b, FF_A2Y)[n]; r.n2z = FLD(b, FF_A2Z)[n];
    r.n3x = FLD(b, FF_A3X)[n]; r.n3y = FLD(b, FF_A3Y)[n]; r.n3z = FLD(b, FF_A3Z)[n];
    r.a1u = FLD(b, FF_A1)[n + 1]; r.a2u = FLD(b, FF_A2)[n + IS]; r.a3u = FLD(b, FF_A3)[n + IL];
    if (FWD) {
        r.u1x = FLD(b, FF_A1X)[n + 1]; r.u1y = FLD(b, FF_A1Y)[n + 1]; r.u1z = FLD(b, FF_A1Z)[n + 1];
        r.u2x = FLD(b, FF_A2X)[n + IS]; r.u2y = FLD(b, FF_A2Y)[n + IS]; r.u2z = FLD(b, FF_A2Z)[n + IS];
        r.u3x = FLD(b, FF_A3X)[n + IL]; r.u3y = FLD(b, FF_A3Y)[n + IL]; r.u3z = FLD(b, FF_A3Z)[n + IL];
    }
}

template <bool FWD, bool KE>
__device__ __forceinline__ void sweep_load_basis(const Blk &b, int n, SweepRaw &r)
{
    const int IS = b.is, IL = b.il;
    r.c = FLD(b, FF_C)[n]; r.ro = FLD(b, FF_RO)[n];
    if (KE) {
        r.rk = FLD(b, FF_RK)[n]; r.drdp = FLD(b, FF_DRDP)[n]; r.drdh = FLD(b, FF_DRDH)[n];
    }
    if (FWD) {
        r.cp = FLD(b, FF_CP)[n];
        r.u1x = FLD(b, FF_A1X)[n + 1]; r.u1y = FLD(b, FF_A1Y)[n + 1]; r.u1z = FLD(b, FF_A1Z)[n + 1];
        r.u2x = FLD(b, FF_A2X)[n + IS]; r.u2y = FLD(b, FF_A2Y)[n + IS]; r.u2z = FLD(b, FF_A2Z)[n + IS];
        r.u3x = FLD(b, FF_A3X)[n + IL]; r.u3y = FLD(b, FF_A3Y)[n + IL]; r.u3z = FLD(b, FF_A3Z)[n + IL];
    } else {
        r.n1x = FLD(b, FF_A1X)[n]; r.n1y = FLD(b, FF_A1Y)[n]; r.n1z = FLD(b, FF_A1Z)[n];
        r.n2x = FLD(b, FF_A2X)[n]; r.n2y = FLD(b, FF_A2Y)[n]; r.n2z = FLD(b, FF_A2Z)[n];
        r.n3x = FLD(b, FF_A3X)[n]; r.n3y = FLD(b, FF_A3Y)[n]; r.n3z = FLD(b, FF_A3Z)[n];
    }
}

// Rows of the LDS cell buffer through which the preparing waves hand a SweepCell to the
// recurrence wave (64 cells per row)
enum {
    CB_D = 0, CB_DT1 = 7, CB_VOL, CB_WI, CB_WJ, CB_WK, CB_SCINV, CB_SCINP,
    CB_X = 14,      // x1, x2, x5 of the I, J, K faces
    CB_B = 23,      // nx ny nz a21 a22 a23 a31 a32 a33 of the I, J, K faces
    CB_S = 50,      // ro roc pc2 proc yp ydrp drhro ro667
    CB_CP = 58, CB_RHO, CB_N
};

__device__ __forceinline__ void cb_put_coef(double *cb, int lane, const SweepCell &s, int nv)
{
    for (int q = 0; q < nv; q++) cb[(CB_D + q) * TW + lane] = s.d[q];
    cb[CB_DT1 * TW + lane] = s.dt1;
    cb[CB_VOL * TW + lane] = s.vol;
    cb[CB_WI * TW + lane] = s.wi;
    cb[CB_WJ * TW + lane] = s.wj;
    cb[CB_WK * TW + lane] = s.wk;
    cb[CB_SCINV * TW + lane] = s.scinv;
    cb[CB_SCINP * TW + lane] = s.scinp;
    const Coef *c3[3] = {&s.ci, &s.cj, &s.ck};
#pragma unroll
    for (int f = 0; f < 3; f++) {
        cb[(CB_X + 3 * f + 0) * TW + lane] = c3[f]->x1;
        cb[(CB_X + 3 * f + 1) * TW + lane] = c3[f]->x2;
        cb[(CB_X + 3 * f + 2) * TW + lane] = c3[f]->x5;
    }
}

__device__ __forceinline__ void cb_put_basis(double *cb, int lane, const SweepCell &s)
{
    const FaceBasis *b3[3] = {&s.fi, &s.fj, &s.fk};
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const FaceBasis &B = *b3[f];
        const double v[9] = {B.nx, B.ny, B.nz, B.a21, B.a22, B.a23, B.a31, B.a32, B.a33};
#pragma unroll
        for (int x = 0; x < 9; x++) cb[(CB_B + 9 * f + x) * TW + lane] = v[x];
    }
    const double sv[8] = {s.ro, s.roc, s.pc2, s.proc, s.yp, s.ydrp, s.drhro, s.ro667};
#pragma unroll
    for (int x = 0; x < 8; x++) cb[(CB_S + x) * TW + lane] = sv[x];
    cb[CB_CP * TW + lane] = s.cp;
    cb[CB_RHO * TW + lane] = s.rho;
}

__device__ __forceinline__ void cb_get(const double *cb, int lane, SweepCell &s, int nv)
{
    for (int q = 0; q < nv; q++) s.d[q] = cb[(CB_D + q) * TW + lane];
    s.dt1 = cb[CB_DT1 * TW + lane];
    s.vol = cb[CB_VOL * TW + lane];
    s.wi = cb[CB_WI * TW + lane];
    s.wj = cb[CB_WJ * TW + lane];
    s.wk = cb[CB_WK * TW + lane];
    s.scinv = cb[CB_SCINV * TW + lane];
    s.scinp = cb[CB_SCINP * TW + lane];
    Coef *c3[3] = {&s.ci, &s.cj, &s.ck};
#pragma unroll
    for (int f = 0; f < 3; f++) {
        c3[f]->x1 = cb[(CB_X + 3 * f + 0) * TW + lane];
        c3[f]->x2 = cb[(CB_X + 3 * f + 1) * TW + lane];
        c3[f]->x5 = cb[(CB_X + 3 * f + 2) * TW + lane];
    }
    FaceBasis *b3[3] = {&s.fi, &s.fj, &s.fk};
#pragma unroll
    for (int f = 0; f < 3; f++) {
        FaceBasis &B = *b3[f];
        const double *r = cb + (CB_B + 9 * f) * TW + lane;
        B.nx = r[0]; B.ny = r[TW]; B.nz = r[2 * TW];
        B.a21 = r[3 * TW]; B.a22 = r[4 * TW]; B.a23 = r[5 * TW];
        B.a31 = r[6 * TW]; B.a32 = r[7 * TW]; B.a33 = r[8 * TW];
    }
    const double *sv = cb + CB_S * TW + lane;
    s.ro = sv[0]; s.roc = sv[TW]; s.pc2 = sv[2 * TW]; s.proc = sv[3 * TW];
    s.yp = sv[4 * TW]; s.ydrp = sv[5 * TW]; s.drhro = sv[6 * TW]; s.ro667 = sv[7 * TW];
    s.cp = cb[CB_CP * TW + lane];
    s.rho = cb[CB_RHO * TW + lane];
}

// s_waitcnt immediate of gfx9 that waits for vmcnt = 0 only (expcnt 7, lgkmcnt 15)
constexpr int VMCNT0_ONLY = 0x0F70;

// End of a step of a warp-specialised workgroup: this wave's LDS writes are complete,
// then the hardware barrier.  (A workgroup-scope fence would also wait for the waves'
// outstanding global loads and stores, which must stay in flight across steps.)
__device__ __forceinline__ void lds_step_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// One LU-SGS sweep marching K: FWD = false is PREDGS / PRKEGS (imps4.f:345-598, 1462-1841),
// K descending; FWD = true is DIAGGS + CORRGS / DIKEGS + COKEGS (imps4.f:604-664, 76-339,
// 1841-1933, 1078-1456), K ascending, with the end of IMPS on the final increment
// (imps3.f:483, 497-500).
//
// A workgroup of four waves owns one 16 x 4 tile of (I, J) columns; its waves specialise
// and meet at one LDS barrier per step:
//   wave 0  the recurrence: the cell's increments from its prepared coefficients and its
//           three upstream contributions, and the contributions it hands downstream
//           (LDS inside the tile, halo slots across tile faces);
//   wave 1  the coefficient half of the next cell's preparation (pred_coef / corr_coef);
//   wave 2  the basis half (face bases, cell scalars);
//   wave 3  the halo words of the step after next, copied into LDS by device-coherent
//           LDS-DMA loads.
// Waves 1 and 2 load the raw values of the step after next while they prepare the next
// one (two register sets, exchanged by writing the loop body twice).  Wave 0 reads only
// LDS, so no load latency sits on the recurrence.
template <bool FWD, bool KE>
__global__ void __launch_bounds__(4 * TW) k_sweep_km(Prm p, Blk b, SweepK sw)
{
    constexpr int NV = KE ? 7 : 5;
    __shared__ double cbuf[2][CB_N * TW];
    __shared__ unsigned int hbuf[3][2 * 2 * NV * TW];
    __shared__ double xi[NV][TW], xj[NV][TW];
    __shared__ int tshare;
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / TW);
    const int lane = threadIdx.x % TW, ii = lane % KI, jj = lane / KI;
    if (threadIdx.x == 0)
        tshare = (int)(atomicAdd(sw.ticket, 1ULL) % (unsigned long long)(sw.nti * sw.ntj));
    __syncthreads();
    const int t = tshare;
    const int ti = FWD ? t % sw.nti : sw.nti - 1 - t % sw.nti;
    const int tj = FWD ? t / sw.nti : sw.ntj - 1 - t / sw.nti;
    const int i = ti * KI + ii + 1, j = tj * KJ + jj + 1;
    const bool col = i <= b.imax && j <= b.jmax;
    const int lag = FWD ? ii + jj : (KI - 1 - ii) + (KJ - 1 - jj);
    const int nsteps = b.kmax + KI + KJ - 2;
    const int nit = (nsteps + 3) & ~1;       // steps -2 .. nsteps-1, rounded up to even
    const bool hin_i = FWD ? (ii == 0 && i > 1) : (ii == KI - 1 && i < b.imax);
    const bool hin_j = FWD ? (jj == 0 && j > 1) : (jj == KJ - 1 && j < b.jmax);
    const int fi_in = FWD ? ti : ti + 1, fj_in = FWD ? tj : tj + 1;     // slot faces read
    const int fi_out = FWD ? ti + 1 : ti, fj_out = FWD ? tj + 1 : tj;   // slot faces written
    auto kof = [&](int q) { return FWD ? q + 1 : b.kmax - q; };
    auto live = [&](int q) { return col && q >= 0 && q < b.kmax; };

    if (wid == 0) {
        // ---- the recurrence -----------------------------------------------------------
        double yk[NV];
#pragma unroll
        for (int r = 0; r < NV; r++) yk[r] = 0.;
        // The data registers of a global store stay locked until the store completes; the
        // values this step stores are kept live through the next step, so that the next
        // step's LDS reads land in other registers and do not wait for the stores.
        double kept[3 * NV], held[3 * NV];
#pragma unroll
        for (int r = 0; r < 3 * NV; r++) kept[r] = held[r] = 0.;
        for (int it = 0; it < nit; it++) {
            const int s = it - 2;
            const int q = s - lag;
            if (s >= 0 && live(q)) {
                const int k = kof(q);
                const int n = cid(b, i, j, k);
                SweepCell sc;
                cb_get(cbuf[s & 1], lane, sc, NV);
                const unsigned int *hb = hbuf[s % 3];
                double bi[NV], bj[NV];
                if (FWD ? i == 1 : i == b.imax) {
#pragma unroll
                    for (int r = 0; r < NV; r++) bi[r] = 0.;
                } else if (!hin_i) {
#pragma unroll
                    for (int r = 0; r < NV; r++) bi[r] = xi[r][FWD ? lane - 1 : lane + 1];
                } else {
                    halo_take<NV>(hi_slot(sw, b, fi_in, j, k), hb, lane, sw.err, bi);
                }
                if (FWD ? j == 1 : j == b.jmax) {
#pragma unroll
                    for (int r = 0; r < NV; r++) bj[r] = 0.;
                } else if (!hin_j) {
#pragma unroll
                    for (int r = 0; r < NV; r++) bj[r] = xj[r][FWD ? lane - KI : lane + KI];
                } else {
                    halo_take<NV>(hjk_slot(sw, b, fj_in, i, k), hb + 2 * NV * TW, lane, sw.err, bj);
                }
                double dn[NV];
                if (FWD) {
#pragma unroll
                    for (int r = 0; r < NV; r++) {
                        double cd = sc.dt1 * sc.wk * yk[r] + sc.vol * sc.d[r];
                        cd = cd + sc.dt1 * (sc.wi * bi[r] + sc.wj * bj[r]);
                        dn[r] = r != 4 ? cd * sc.scinv : cd * sc.scinp;
                    }
                    kept[0] = dn[0] / sc.cp;
#pragma unroll
                    for (int r = 1; r < 5; r++) kept[r] = dn[r];
                    if (KE) {
                        kept[NV > 5 ? 5 : 0] = dn[NV > 5 ? 5 : 0] * sc.rho;
                        kept[NV > 6 ? 6 : 0] = dn[NV > 6 ? 6 : 0] * sc.rho;
                    }
                } else {
                    // the line-buffer shift of PREDGS (imps4.f:556-562) never hands the I+1
                    // contribution to a cell of the row J = JMAX; PRKEGS's (imps4.f:1762) does
                    if (!KE && j == b.jmax) {
#pragma unroll
                        for (int r = 0; r < NV; r++) bi[r] = 0.;
                    }
#pragma unroll
                    for (int r = 0; r < NV; r++) {
                        double cd = -sc.dt1 * sc.wk * yk[r] + sc.vol * sc.d[r];
                        cd = cd - sc.dt1 * (sc.wi * bi[r] + sc.wj * bj[r]);
                        dn[r] = r != 4 ? cd * sc.scinv : cd * sc.scinp;
                    }
#pragma unroll
                    for (int r = 0; r < NV; r++) kept[r] = dn[r];
                }
#pragma unroll
                for (int r = 0; r < NV; r++) FLD(b, FF_DRO + r)[n] = kept[r];
                double *oi = kept + NV, *oj = kept + 2 * NV;
                face_xfer<KE>(dn, sc, sc.fi, sc.ci, oi);
                face_xfer<KE>(dn, sc, sc.fj, sc.cj, oj);
                face_xfer<KE>(dn, sc, sc.fk, sc.ck, yk);
                // the neighbours' values of this step were read above, before these writes
                if (FWD ? ii < KI - 1 : ii > 0) {
#pragma unroll
                    for (int r = 0; r < NV; r++) xi[r][lane] = oi[r];
                } else if (FWD ? i < b.imax : i > 1) {
                    halo_put<NV>(hi_slot(sw, b, fi_out, j, k), oi);
                }
                if (FWD ? jj < KJ - 1 : jj > 0) {
#pragma unroll
                    for (int r = 0; r < NV; r++) xj[r][lane] = oj[r];
                } else if (FWD ? j < b.jmax : j > 1) {
                    halo_put<NV>(hjk_slot(sw, b, fj_out, i, k), oj);
                }
            }
            // the previous step's stored values end their life here
#pragma unroll
            for (int r = 0; r < 3 * NV; r++) asm volatile("" ::"v"(held[r]));
#pragma unroll
            for (int r = 0; r < 3 * NV; r++) held[r] = kept[r];
            lds_step_barrier();
        }
    } else if (wid == 1) {
        // ---- coefficients of the next cell --------------------------------------------
        auto iter = [&](int s, const SweepRaw &use, SweepRaw &fill) __attribute__((always_inline)) {
            const int q1 = s + 1 - lag, q2 = s + 2 - lag;
            // the loads of the previous step (the values used now) had a whole step to land;
            // an explicit wait tells the compiler so, and it adds none inside the step
            __builtin_amdgcn_s_waitcnt(VMCNT0_ONLY);
            if (live(q2)) sweep_load_coef<FWD, KE>(b, cid(b, i, j, kof(q2)), fill);
            if (live(q1)) {
                SweepCell sc;
                if (FWD)
                    corr_coef<KE>(p, use, sc);
                else
                    pred_coef<KE>(p, use, i, j, kof(q1), sc);
                cb_put_coef(cbuf[(s + 1) & 1], lane, sc, NV);
            }
            lds_step_barrier();
        };
        SweepRaw ra, rb;
        for (int it = 0; it < nit; it += 2) {
            iter(it - 2, rb, ra);
            iter(it - 1, ra, rb);
        }
    } else if (wid == 2) {
        // ---- face bases and cell scalars of the next cell ------------------------------
        auto iter = [&](int s, const SweepRaw &use, SweepRaw &fill) __attribute__((always_inline)) {
            const int q1 = s + 1 - lag, q2 = s + 2 - lag;
            __builtin_amdgcn_s_waitcnt(VMCNT0_ONLY);
            if (live(q2)) sweep_load_basis<FWD, KE>(b, cid(b, i, j, kof(q2)), fill);
            if (live(q1)) {
                SweepCell sc;
                if (FWD)
                    corr_basis<KE>(use, sc);
                else
                    pred_basis<KE>(use, sc);
                cb_put_basis(cbuf[(s + 1) & 1], lane, sc);
            }
            lds_step_barrier();
        };
        SweepRaw ra, rb;
        for (int it = 0; it < nit; it += 2) {
            iter(it - 2, rb, ra);
            iter(it - 1, ra, rb);
        }
    } else {
        // ---- halo words of the step after next -----------------------------------------
        for (int it = 0; it < nit; it++) {
            const int s2 = it;                     // step whose halo words are copied
            const int q = s2 - lag;
            const bool ok = live(q);
            const int k = ok ? kof(q) : 1;
            const unsigned long long *si = hin_i && ok ? hi_slot(sw, b, fi_in, j, k) : sw.hi;
            const unsigned long long *sj = hin_j && ok ? hjk_slot(sw, b, fj_in, i, k) : sw.hi;
            unsigned int *hb = hbuf[s2 % 3];
#pragma unroll
            for (int w = 0; w < 2 * NV; w++) {
                lds_dma4<true>((const unsigned int *)si + w, lds_addr(hb + w * TW));
                lds_dma4<true>((const unsigned int *)sj + w, lds_addr(hb + (2 * NV + w) * TW));
            }
            // the copies of step s2 - 1 are complete: 4*NV loads were issued after them
            vm_wait<4 * NV>();
            lds_step_barrier();
        }
        vm_wait<0>();
    }
}

// Fill 64-bit words (halo slots to HALO_EMPTY at set-up and after a failed sweep)
__global__ void __launch_bounds__(256) k_fill_u64(unsigned long long *x, size_t n, unsigned long long v)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = v;
}

// DIVV12(DRO, DRO, CP) on interior cells (imps3.f:483)
__global__ void __launch_bounds__(256) k_divcp(Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > b.imax) return;
    int n = cid(b, i, j, k);
    FLD(b, FF_DRO)[n] = FLD(b, FF_DRO)[n] / FLD(b, FF_CP)[n];
}

// Start-of-cycle copies (TOLD ...) on every cell; a cycle of a block with INL/OUT patches
// also counts its ICYCLE (icyc, null otherwise)
__global__ void __launch_bounds__(256) k_save_old(Blk b, int *icyc)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (icyc && n == 0) *icyc = *icyc + 1;
    if (n >= b.ntot) return;
    FLD(b, FF_TOLD)[n] = FLD(b, FF_TEMP)[n];
    FLD(b, FF_UOLD)[n] = FLD(b, FF_U)[n];
    FLD(b, FF_VOLD)[n] = FLD(b, FF_V)[n];
    FLD(b, FF_WOLD)[n] = FLD(b, FF_W)[n];
    FLD(b, FF_POLD)[n] = FLD(b, FF_PDIFF)[n];
    FLD(b, FF_RKOLD)[n] = FLD(b, FF_RK)[n];
    FLD(b, FF_EPSOLD)[n] = FLD(b, FF_REPS)[n];
}

// FIN (res3c.f:1122-1126) on interior cells
__global__ void __launch_bounds__(256) k_fin(Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > b.imax) return;
    int n = cid(b, i, j, k);
    FLD(b, FF_TEMP)[n] = FLD(b, FF_TOLD)[n] + FLD(b, FF_DRO)[n];
    FLD(b, FF_U)[n] = FLD(b, FF_UOLD)[n] + FLD(b, FF_DM)[n];
    FLD(b, FF_V)[n] = FLD(b, FF_VOLD)[n] + FLD(b, FF_DN)[n];
    FLD(b, FF_W)[n] = FLD(b, FF_WOLD)[n] + FLD(b, FF_DW)[n];
    FLD(b, FF_PDIFF)[n] = FLD(b, FF_POLD)[n] + FLD(b, FF_DE)[n];
}

// FIN for two-equation models (res3c.f:1182-1198): the mean flow as above plus the relaxed
// update of rho*k and rho*omega
__global__ void __launch_bounds__(256) k_fin_kw(Blk b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    int k = blockIdx.z + 1;
    if (i > b.imax) return;
    int n = cid(b, i, j, k);
    FLD(b, FF_TEMP)[n] = FLD(b, FF_TOLD)[n] + FLD(b, FF_DRO)[n];
    FLD(b, FF_U)[n] = FLD(b, FF_UOLD)[n] + FLD(b, FF_DM)[n];
    FLD(b, FF_V)[n] = FLD(b, FF_VOLD)[n] + FLD(b, FF_DN)[n];
    FLD(b, FF_W)[n] = FLD(b, FF_WOLD)[n] + FLD(b, FF_DW)[n];
    FLD(b, FF_PDIFF)[n] = FLD(b, FF_POLD)[n] + FLD(b, FF_DE)[n];
    double rko = FLD(b, FF_RKOLD)[n], epo = FLD(b, FF_EPSOLD)[n];
    double drk = FLD(b, FF_DRK)[n], dep = FLD(b, FF_DEPS)[n];
    double rkoldn = .45 * rko;
    double epoldn = .90 * epo;
    double absdrk = fabs(drk);
    double absdep = fabs(dep);
    FLD(b, FF_RK)[n] = rko + drk / (1. + absdrk / rkoldn);
    FLD(b, FF_REPS)[n] = epo + dep / (1. + absdep / epoldn);
}

// Limiters of CYCLES (ns3c.f:1553-1596) on every cell: UPLIM acts on interior K-planes and
// J-rows (all I, res3c.f:3401-3417), LOLIM / PDTOP / LOLIMP on all cells.
__global__ void __launch_bounds__(256) k_limits(Prm p, Blk b)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= b.ntot) return;
    int kk = n / b.il;
    int jj = (n / b.is) % b.js;
    bool up = kk >= 2 && kk < 2 + b.kmax && jj >= 2 && jj < 2 + b.jmax;
    double *x;
    x = FLD(b, FF_TEMP);
    double t = x[n];
    if (up) t = t < p.tuplim ? t : p.tuplim;
    t = t > p.tlolim ? t : p.tlolim;
    x[n] = t;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        x = FLD(b, FF_U + q);
        double s = x[n];
        if (up) s = s < 5000. ? s : 5000.;
        s = s > -5000. ? s : -5000.;
        x[n] = s;
    }
    if (p.iturb == 6) {
        // LOLIM of RK and REPS (FIN res3c.f:1202-1205; CYCLES ns3c.f:1575-1576)
        double rk = FLD(b, FF_RK)[n], re = FLD(b, FF_REPS)[n];
        FLD(b, FF_RK)[n] = rk > p.rklim ? rk : p.rklim;
        FLD(b, FF_REPS)[n] = re > p.epslim ? re : p.epslim;
    }
    double pd = FLD(b, FF_PDIFF)[n];
    double P = pd + p.frspre;
    double y = P - 0.01 * p.frspre;
    double prelim = 0. < y ? 0. : y;
    FLD(b, FF_P)[n] = P - prelim;
    FLD(b, FF_PDIFF)[n] = pd - prelim;
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------

// The solid-wall patches of a block level: the caller's descriptors and host arrays, the
// patches in BOTSKE's frame with device arrays, and the per-face scratch of WallDev
struct WallRec {
    std::vector<ff_wall_patch> host;
    std::vector<WallPatchDev> dev;
    WallPatchDev *dev_tab = nullptr;   // device copy of dev (the force pass)
    double *parr = nullptr;            // device patch arrays
    unsigned char *flag = nullptr;
    double *scr = nullptr;             // flux, corrections, QY/QZ terms
    double *sums = nullptr;            // FF_NFORCE per patch (device)
    double *sums_host = nullptr;       // pinned copy
    WallDev wd{};
    WallEnv env{};
    hipEvent_t fwd = nullptr;          // recorded on the main stream before the force pass
    hipEvent_t done = nullptr;         // recorded on the side stream after it
    bool pending = false;              // a force pass may still read the scratch
};

// One REFLEC / GHOSTEXT / INLET / OUTLET patch in REFLEC's frame (ffbc.hip): ICON(1,L),
// face, ILMO, the face plane, the KX/KY ranges after the J-face swap, the stride normal to
// the face, IHEAT, the viscous-branch flag, TSURF, the device wall velocity arrays (or
// null), the INLOUT type and the device patch arrays BOUNR .. BOUNEP of INL/OUT patches
struct BcPatchDev {
    int ibc, iface, dir, idir, kpl, kx1, kx2, ky1, ky2, kstr, iheat, viscous, n, inlout;
    double tsurf;
    const double *uw, *vw, *ww;
    double *bn[9];
};

struct BlockRec {
    ff_block host{};
    Blk dev{};
    bool mirror = false;
    double *arena = nullptr;   // device arena of a mirrored block
    Sweep sw{};                // LU-SGS tile-wavefront state (halo slots, ticket, error word)
    SweepK swk{};              // the same for the K-marching (I, J) tiles (shares the slots)
    void *sw_mem = nullptr;
    size_t nhalo = 0;          // halo slot words
    size_t ntile_slots = 0;    // words of the (J,K)-tile forms' slots (J faces, then K faces)
    double *kw_scr = nullptr;  // k-omega: per-workgroup maxima of SEPS and their maximum
    int kw_parts = 0;
    int level = 0;             // multigrid level M the slab was registered as
    WallRec walls;
    SweepSt st{};              // staged records of the (J,K)-tile sweeps (allocated on first use)
    double *st_mem = nullptr;
    void *dmx_mem = nullptr;   // DROMAX chunk maxima and candidates (allocated on first use)
    std::vector<BcPatchDev> bcs;   // boundary patches (ff_bc_register)
    double *bc_mem = nullptr;      // their wall velocity and INL/OUT arrays on the device
    int *bc_icyc = nullptr;        // the cycle count ICYCLE INLET reads (INL/OUT patches)
    std::vector<ff_con_patch> cons;   // connection patches (ff_con_register)
    int *con_map = nullptr;           // their ghost <- source cell offset pairs on the device
    std::vector<size_t> con_off;      // start of each patch's pairs in con_map
    bool con_unique = true;           // no ghost cell is the target of two of its pairs
};

struct PendingEv {
    const char *name;
    int level;
    hipEvent_t a, b;
};

// the learned EARSM closure of a context (kernels in ffmlp.hip)
struct MlpRec {
    float *p = nullptr, *m = nullptr, *v = nullptr;    // parameters and Adam moments
    float *nrm = nullptr;                               // normalisation block
    double *nscr = nullptr;                             // normalisation passes: partials, means
    float *part = nullptr;                              // per-wave gradient partials
    unsigned *idx = nullptr;                            // minibatch indices
    float *grad = nullptr;                              // reduced gradient and loss
    int part_waves = 0, idx_cap = 0;
    int step = 0;                                       // Adam steps taken
    bool ready = false;                                 // parameters and normalisation set
};

// candidate tile shapes of the residual autotuner (described with it, below)
struct TileShape { int ri, rj, rk; };
constexpr TileShape kTileShapes[] = {{64, 4, 8}, {64, 4, 16}, {32, 8, 8}, {32, 4, 4},
                                     {128, 2, 8}, {32, 4, 2}, {16, 4, 2}, {64, 2, 2}};
constexpr int NSHAPE = (int)(sizeof(kTileShapes) / sizeof(kTileShapes[0]));

// The halo exchange of one level as a single launch: every connection of every block, one
// workgroup per 256 pairs (ffcon.hip).  Built outside graph capture, rebuilt when the
// registrations change (gen).
struct HaloConnDev { int dst, src; const int *map; int n; };
struct HaloBatch {
    long long gen = -1;
    Blk *blks = nullptr;              // the level's blocks (device copy)
    HaloConnDev *conns = nullptr;     // the connections, in block and patch order
    int2 *wg = nullptr;               // per workgroup: connection, first pair
    int nwg = 0;
};

struct ff_ctx {
    ff_params prm{};
    Prm kp{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::map<long long, BlockRec> blocks;
    std::string err;
    double *red_part = nullptr;   // RESL2 partials
    double *red_out = nullptr;    // device sums
    double *red_host = nullptr;   // pinned host sums
    int red_parts = 0;
    bool profile = false;
    int lusgs_mode = FF_LUSGS_STAGED;
    int residual_mode = FF_RESIDUAL_TILES;
    std::vector<PendingEv> pending;
    // launches and event-timed milliseconds per kernel name, and per "name@level"
    std::map<std::string, std::pair<long long, double>> stats;
    int lev = 0;                  // level of the stage being launched (statistics only)
    hipStream_t wstream = nullptr;   // side stream of the wall force sums (default priority)
    int earsm_mode = FF_EARSM_EXACT;
    int precision = FF_PREC_FP64;
    MlpRec mlp;
    // the captured V-cycle (ff_vcycle_graph) and what it was captured for; gen counts the
    // registrations and option changes that make a capture stale
    hipGraph_t vg_graph = nullptr;
    hipGraphExec_t vg_exec = nullptr;
    long long vg_key[6] = {0, 0, 0, 0, 0, 0};
    long long gen = 0;
    // the captured multi-block cycle (ff_vcycle_blocks_graph): its streams, fork/join events,
    // graph and the key it was captured for
    static constexpr int NPOOL = 32;      // streams of the pool (FF_OPT_BLOCK_STREAMS)
    static constexpr int NBRANCH = 8;     // graph branches of ff_vcycle_blocks_graph
    hipStream_t pool[NPOOL] = {};
    hipEvent_t pool_ev[NPOOL + 1] = {};
    hipGraph_t mb_graph = nullptr;
    hipGraphExec_t mb_exec = nullptr;
    std::vector<long long> mb_key;
    // the concurrent multi-block cycle (FF_OPT_BLOCK_STREAMS): RESL2 partials per pool
    // stream and the sums of every block, read back once after the cycle
    bool block_streams = false;
    bool in_pool = false;         // a pool stream is the context's stream (its V-cycles)
    double *mb_part[NPOOL] = {};
    double *mb_out = nullptr;
    int mb_out_n = 0;
    // the tile-shape autotuner: on, and the chosen shape per (IMAX, JMAX, KMAX, kind)
    bool autotune = false;
    int force_shape = -1;         // FF_OPT_TILE_SHAPE: a candidate shape for every residual
    std::map<std::tuple<int, int, int, int>, std::pair<int, double>> tuned;
    std::map<int, HaloBatch> halo_batch;   // per level (ffcon.hip)
    int device = -1;              // the HIP device the context was created on
    double *scratch = nullptr;    // that device's address of the fixed scratch array
};

static void mlp_free(MlpRec &r);
static int earsm_mlp_pass(ff_ctx *c, const Blk &b);

// The fixed device scratch of ff_residual_noalloc: the RESL2 partials and sums and the
// per-workgroup SEPS maxima of the k-omega residual, 36,680 doubles (about 287 KB) that
// every GPU the library runs on holds until the process ends.  In host code `scratch`
// names a host-side stand-in; the device address comes from hipGetSymbolAddress, taken
// once per device when a context is created there, and only that address is passed on.
constexpr int NA_RED_PARTS = 512;                       // RESL2: 7 partials each
constexpr int NA_KW_PARTS = 32768;                      // SEPS maxima: 256 entries each
constexpr size_t NA_OFF_OUT = 7 * NA_RED_PARTS;         // RESL2 sums (8)
constexpr size_t NA_OFF_KW = NA_OFF_OUT + 8;            // SEPS maxima (NA_KW_PARTS + 1)
constexpr size_t NA_SCRATCH = NA_OFF_KW + NA_KW_PARTS + 8;
static __device__ __align__(256) double scratch[NA_SCRATCH];
constexpr int NA_MAX_DEVICES = 64;
static double *g_scratch_dev[NA_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the current device's address of `scratch` (looked up once per device)
static hipError_t scratch_address(int dev, double **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = (double *)ptr;
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

static long long key(int level, int block) { return ((long long)level << 32) | (unsigned)block; }

static int fail(ff_ctx *c, int code, const char *fmt, const char *a = "")
{
    if (c) {
        char buf[512];
        snprintf(buf, sizeof buf, fmt, a);
        c->err = buf;
    }
    return code;
}

#define HIPCHK(ctx, call)                                                                \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return fail(ctx, FF_E_HIP, "%s", hipGetErrorString(e_));   \
    } while (0)

static Prm make_prm(const ff_params *p, int level)
{
    Prm k{};
    k.gamma = p->gamma; k.rgas = p->rgas; k.pr = p->pr; k.prt = p->prt;
    k.visu0 = p->visu0; k.tsu0 = p->tsu0; k.rmult = p->rmult;
    k.frspre = p->frspre; k.frsden = p->frsden; k.frstem = p->frstem;
    k.cfl = level >= 2 ? p->cfll : p->cfl;
    k.theta = p->theta; k.cdiff = p->cdiff; k.tuplim = p->tuplim; k.tlolim = p->tlolim;
    k.epslim = p->epslim;
    k.rklim = p->rklim; k.frsvis = p->frsvis; k.re = p->re; k.csimps = p->csimps;
    k.turlim = p->turlim;
    k.iturb = p->iturb;
    k.kover = p->kover; k.katol = p->katol; k.istres = p->istres;
    k.fullns = level == 1 ? p->fullns : 0;       // FLUCYC/FLXCYC: FULLNS .AND. M == 1
    k.idiff = p->idiff;
    k.turcor = p->turcor;
    k.cmgk = p->cmgk; k.cmgeps = p->cmgeps;
    for (int d = 0; d < 3; d++) k.inter[d] = level >= 2 ? 55 : p->inter[d];   // res3c RESIDI ns3c.f:2468-2472
    return k;
}

// Launch helper with optional event timing on the context's stream
template <typename K, typename... A>
static int launch_on(ff_ctx *c, hipStream_t st, const char *name, K kern, dim3 g, dim3 t, A... args)
{
    hipEvent_t ea = nullptr, eb = nullptr;
    if (c->profile) {
        HIPCHK(c, hipEventCreate(&ea));
        HIPCHK(c, hipEventCreate(&eb));
        HIPCHK(c, hipEventRecord(ea, st));
    }
    hipLaunchKernelGGL(kern, g, t, 0, st, args...);
    HIPCHK(c, hipGetLastError());
    if (c->profile) {
        HIPCHK(c, hipEventRecord(eb, st));
        c->pending.push_back({name, c->lev, ea, eb});
    }
    return FF_OK;
}

template <typename K, typename... A>
static int launch(ff_ctx *c, const char *name, K kern, dim3 g, dim3 t, A... args)
{
    return launch_on(c, c->stream, name, kern, g, t, args...);
}

static int drain_events(ff_ctx *c)
{
    for (auto &e : c->pending) {
        HIPCHK(c, hipEventSynchronize(e.b));
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, e.a, e.b));
        auto &s = c->stats[e.name];
        s.first += 1;
        s.second += ms;
        auto &sl = c->stats[std::string(e.name) + "@" + std::to_string(e.level)];
        sl.first += 1;
        sl.second += ms;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->pending.clear();
    return FF_OK;
}

static BlockRec *find(ff_ctx *c, int level, int block)
{
    auto it = c->blocks.find(key(level, block));
    return it == c->blocks.end() ? nullptr : &it->second;
}

static Blk make_blk(const ff_block *v)
{
    Blk b{};
    b.imax = v->imax; b.jmax = v->jmax; b.kmax = v->kmax;
    b.is = b.imax + 2 * FF_GHOST; b.js = b.jmax + 2 * FF_GHOST; b.ks = b.kmax + 2 * FF_GHOST;
    b.il = b.is * b.js;
    b.ntot = b.il * b.ks;
    for (int f = 0; f < FF_NFIELDS; f++) b.f[f] = v->f[f];
    return b;
}

static int check_dims(ff_ctx *c, const ff_block *v)
{
    if (!v || v->imax < 1 || v->jmax < 1 || v->kmax < 1)
        return fail(c, FF_E_ARG, "block dimensions must be positive%s");
    long long n = (long long)(v->imax + 4) * (v->jmax + 4) * (v->kmax + 4);
    if (n * FF_NFIELDS >= (1LL << 31))
        return fail(c, FF_E_ARG, "block too large for 32-bit cell offsets%s");
    return FF_OK;
}

// LU-SGS tile-wavefront state of a block: halo slots of the interior tile faces, the
// ticket and the error word, in one allocation; slots start empty.
static int sweep_alloc(ff_ctx *c, BlockRec &r)
{
    const Blk &b = r.dev;
    Sweep &sw = r.sw;
    SweepK &sk = r.swk;
    sw.ntj = (b.jmax + TJ - 1) / TJ;
    sw.ntk = (b.kmax + TK - 1) / TK;
    sk.nti = (b.imax + KI - 1) / KI;
    sk.ntj = (b.jmax + KJ - 1) / KJ;
    // both tile forms leave every slot empty after a sweep, so they share the slots
    const size_t nhj = (size_t)HS * TK * (sw.ntj + 1) * sw.ntk * (b.imax + TK);
    const size_t nhk = (size_t)HS * TJ * (sw.ntk + 1) * sw.ntj * (b.imax + TJ);
    const size_t nhi2 = (size_t)HS * (sk.nti + 1) * b.kmax * b.jmax;
    const size_t nhj2 = (size_t)HS * (sk.ntj + 1) * b.kmax * b.imax;
    r.nhalo = std::max(nhj + nhk, nhi2 + nhj2);
    r.ntile_slots = nhj + nhk;
    // ticket, error word, ticket of the K-marching form, pad to 64 bytes (the slot runs of the
    // tile form are 64-byte aligned)
    const size_t head = 8;
    if (hipMalloc(&r.sw_mem, (r.nhalo + head) * sizeof(unsigned long long)) != hipSuccess)
        return fail(c, FF_E_NOMEM, "LU-SGS halo slots%s");
    unsigned long long *m = (unsigned long long *)r.sw_mem;
    sw.ticket = m;
    sw.err = (unsigned int *)(m + 1);
    sw.hj = m + head;
    sw.hk = sw.hj + nhj;
    sk.ticket = m + 2;
    sk.err = sw.err;
    sk.hi = m + head;
    sk.hj = sk.hi + nhi2;
    HIPCHK(c, hipMemsetAsync(r.sw_mem, 0, head * sizeof(unsigned long long), c->stream));
    if (c->prm.iturb == 6) {
        r.kw_parts = (b.ntot + 255) / 256;
        if (hipMalloc(&r.kw_scr, sizeof(double) * (size_t)(r.kw_parts + 1)) != hipSuccess)
            return fail(c, FF_E_NOMEM, "k-omega reduction scratch%s");
    }
    size_t nh = r.nhalo;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, c->stream,
                       sw.hj, nh, HALO_EMPTY);
    HIPCHK(c, hipGetLastError());
    return FF_OK;
}

static void walls_free(WallRec &w)
{
    if (w.pending && w.done) (void)hipEventSynchronize(w.done);
    if (w.dev_tab) (void)hipFree(w.dev_tab);
    if (w.parr) (void)hipFree(w.parr);
    if (w.flag) (void)hipFree(w.flag);
    if (w.scr) (void)hipFree(w.scr);
    if (w.sums) (void)hipFree(w.sums);
    if (w.sums_host) (void)hipHostFree(w.sums_host);
    if (w.fwd) (void)hipEventDestroy(w.fwd);
    if (w.done) (void)hipEventDestroy(w.done);
    w = WallRec();
}

static void sweep_free(BlockRec &r)
{
    walls_free(r.walls);
    if (r.st_mem) (void)hipFree(r.st_mem);
    r.st_mem = nullptr;
    if (r.sw_mem) (void)hipFree(r.sw_mem);
    r.sw_mem = nullptr;
    if (r.kw_scr) (void)hipFree(r.kw_scr);
    r.kw_scr = nullptr;
    if (r.dmx_mem) (void)hipFree(r.dmx_mem);
    r.dmx_mem = nullptr;
    if (r.bc_mem) (void)hipFree(r.bc_mem);
    r.bc_mem = nullptr;
    if (r.bc_icyc) (void)hipFree(r.bc_icyc);
    r.bc_icyc = nullptr;
    r.bcs.clear();
    if (r.con_map) (void)hipFree(r.con_map);
    r.con_map = nullptr;
    r.cons.clear();
    r.con_off.clear();
}

// Read the error word of the block's sweeps; after a failed halo wait, refill the slots
// (a wait that gave up may leave filled slots behind) and report the failure.
static int sweep_check(ff_ctx *c, BlockRec &r)
{
    if (!r.sw_mem) return FF_OK;
    unsigned int *h = (unsigned int *)(c->red_host + 7);
    HIPCHK(c, hipMemcpyAsync(h, r.sw.err, sizeof(unsigned int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (*h == 0) return FF_OK;
    *h = 0;
    size_t nh = r.nhalo;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, c->stream,
                       r.sw.hj, nh, HALO_EMPTY);
    HIPCHK(c, hipMemsetAsync(r.sw.err, 0, sizeof(unsigned int), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return fail(c, FF_E_HIP, "LU-SGS tile wavefront: a halo wait timed out%s");
}

static int check_all(ff_ctx *c)
{
    for (auto &kv : c->blocks) {
        int rc = sweep_check(c, kv.second);
        if (rc) return rc;
    }
    return FF_OK;
}

// fields a kernel sequence needs
static const int kNeeded[] = {
    FF_PDIFF, FF_TEMP, FF_U, FF_V, FF_W, FF_P, FF_RO, FF_E, FF_C, FF_VIS, FF_CH, FF_CP,
    FF_DRDP, FF_DRDH, FF_RM, FF_RN, FF_RW, FF_EPS2, FF_VIST, FF_DRO, FF_DM, FF_DN, FF_DW,
    FF_DE, FF_DTL, FF_TOLD, FF_UOLD, FF_VOLD, FF_WOLD, FF_POLD, FF_VOL, FF_A1, FF_A1X,
    FF_A1Y, FF_A1Z, FF_A2, FF_A2X, FF_A2Y, FF_A2Z, FF_A3, FF_A3X, FF_A3Y, FF_A3Z};
// and those of the k-omega SST path
static const int kNeededEarsm[] = {FF_TTS, FF_B11, FF_B12, FF_B13, FF_B22, FF_B23, FF_B33};
static const int kNeededKw[] = {
    FF_RK, FF_REPS, FF_DRK, FF_DEPS, FF_RKOLD, FF_EPSOLD, FF_D1, FF_D2, FF_D3, FF_DISTW,
    FF_SRK, FF_SEPS, FF_PTUR, FF_FUN1, FF_OHMI, FF_W12, FF_W13, FF_W23, FF_S11, FF_S12,
    FF_S13, FF_S22, FF_S23, FF_S33, FF_STRAIN, FF_TTS};
// and those of FULLNS (DERFUN, TEMDER, FLUXBL; FLKEBL adds GRADK/GREPS)
static const int kNeededFullns[] = {
    FF_D1, FF_D2, FF_D3, FF_XC, FF_YC, FF_ZC, FF_OHMI, FF_W12, FF_W13, FF_W23, FF_S11, FF_S12,
    FF_S13, FF_S22, FF_S23, FF_S33, FF_STRAIN, FF_DTDX, FF_DTDY, FF_DTDZ};
static const int kNeededFullnsKw[] = {FF_DKDX, FF_DKDY, FF_DKDZ, FF_DEDX, FF_DEDY, FF_DEDZ};

static int check_fields(ff_ctx *c, const ff_block *v, const char *what)
{
    for (int f : kNeeded)
        if (!v->f[f]) return fail(c, FF_E_ARG, what, std::to_string(f).c_str());
    if (c->prm.iturb == 6)
        for (int f : kNeededKw)
            if (!v->f[f]) return fail(c, FF_E_ARG, what, std::to_string(f).c_str());
    if (c->prm.istres >= 1)
        for (int f : kNeededEarsm)
            if (!v->f[f]) return fail(c, FF_E_ARG, what, std::to_string(f).c_str());
    if (c->prm.fullns) {
        for (int f : kNeededFullns)
            if (!v->f[f]) return fail(c, FF_E_ARG, what, std::to_string(f).c_str());
        if (c->prm.iturb == 6)
            for (int f : kNeededFullnsKw)
                if (!v->f[f]) return fail(c, FF_E_ARG, what, std::to_string(f).c_str());
    }
    return FF_OK;
}

extern "C" {

int ff_ctx_create(ff_ctx **out, const ff_params *prm)
{
    if (!out || !prm) return FF_E_ARG;
    *out = nullptr;
    if (prm->iturb != 0 && prm->iturb != 6) return FF_E_UNSUP;   // laminar or k-omega SST
    if (prm->iturb == 6 && (prm->kover != 0 && prm->kover != 5)) return FF_E_UNSUP;
    if (prm->iturb == 6 && prm->iflux != 1) return FF_E_UNSUP;   // FLXCYC: Roe only
    if (prm->istres != 0 && (prm->iturb != 6 || prm->istres != 1)) return FF_E_UNSUP;  // EARSM1 on SST
    if (prm->iflux != 1 && prm->iflux != 0) return FF_E_UNSUP;   // Roe FLUXKE, van Leer FLUXVL
    if (prm->idiff > 3) return FF_E_UNSUP;         // arithmetic FLUXDO/FLUXBL average
    if (prm->fullns != 0 && prm->fullns != 1) return FF_E_UNSUP;
    if (prm->expsu != 1.5) return FF_E_UNSUP;      // Sutherland sqrt branch of PERFGA
    ff_ctx *c = new ff_ctx();
    c->prm = *prm;
    c->kp = make_prm(prm, 1);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return FF_E_HIP;
    }
    c->own_stream = true;
    if (hipGetDevice(&c->device) != hipSuccess || c->device < 0 || c->device >= NA_MAX_DEVICES ||
        scratch_address(c->device, &c->scratch) != hipSuccess) {
        ff_ctx_destroy(c);
        return FF_E_HIP;
    }
    c->red_parts = 512;
    if (hipMalloc(&c->red_part, sizeof(double) * 7 * c->red_parts) != hipSuccess ||
        hipMalloc(&c->red_out, sizeof(double) * 8) != hipSuccess ||
        hipHostMalloc(&c->red_host, sizeof(double) * 8, hipHostMallocDefault) != hipSuccess) {
        ff_ctx_destroy(c);
        return FF_E_NOMEM;
    }
    *out = c;
    return FF_OK;
}

int ff_ctx_destroy(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int q = 0; q < ff_ctx::NPOOL; q++)
        if (c->pool[q]) (void)hipStreamSynchronize(c->pool[q]);
    if (c->wstream) (void)hipStreamSynchronize(c->wstream);
    drain_events(c);
    if (c->mb_exec) (void)hipGraphExecDestroy(c->mb_exec);
    if (c->mb_graph) (void)hipGraphDestroy(c->mb_graph);
    for (int q = 0; q < ff_ctx::NPOOL; q++)
        if (c->pool[q]) (void)hipStreamDestroy(c->pool[q]);
    for (int q = 0; q <= ff_ctx::NPOOL; q++)
        if (c->pool_ev[q]) (void)hipEventDestroy(c->pool_ev[q]);
    for (int q = 0; q < ff_ctx::NPOOL; q++)
        if (c->mb_part[q]) (void)hipFree(c->mb_part[q]);
    if (c->mb_out) (void)hipFree(c->mb_out);
    for (auto &kv : c->halo_batch) {
        if (kv.second.blks) (void)hipFree(kv.second.blks);
        if (kv.second.conns) (void)hipFree(kv.second.conns);
        if (kv.second.wg) (void)hipFree(kv.second.wg);
    }
    for (auto &kv : c->blocks) {
        if (kv.second.mirror && kv.second.arena) (void)hipFree(kv.second.arena);
        sweep_free(kv.second);
    }
    if (c->red_part) (void)hipFree(c->red_part);
    if (c->red_out) (void)hipFree(c->red_out);
    if (c->red_host) (void)hipHostFree(c->red_host);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    if (c->wstream) (void)hipStreamDestroy(c->wstream);
    mlp_free(c->mlp);
    if (c->vg_exec) (void)hipGraphExecDestroy(c->vg_exec);
    if (c->vg_graph) (void)hipGraphDestroy(c->vg_graph);
    delete c;
    return FF_OK;
}

int ff_ctx_set_stream(ff_ctx *c, void *stream)
{
    if (!c) return FF_E_ARG;
    if (stream) {
        if (c->own_stream && c->stream) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipStreamDestroy(c->stream);
        }
        c->stream = (hipStream_t)stream;
        c->own_stream = false;
    }
    return FF_OK;
}

int ff_ctx_set_option(ff_ctx *c, int option, int value)
{
    if (!c) return FF_E_ARG;
    if (option == FF_OPT_LUSGS) {
        if (value != FF_LUSGS_TILES && value != FF_LUSGS_PLANES && value != FF_LUSGS_KMARCH &&
            value != FF_LUSGS_STAGED && value != FF_LUSGS_STAGED1)
            return fail(c, FF_E_ARG, "FF_OPT_LUSGS: unknown value%s");
        c->lusgs_mode = value;
        c->gen++;
        return FF_OK;
    }
    if (option == FF_OPT_EARSM) {
        if (value != FF_EARSM_EXACT && value != FF_EARSM_MLP)
            return fail(c, FF_E_ARG, "FF_OPT_EARSM: unknown value%s");
        c->earsm_mode = value;
        c->gen++;
        return FF_OK;
    }
    if (option == FF_OPT_RESIDUAL) {
        if (value != FF_RESIDUAL_TILES && value != FF_RESIDUAL_CELLS)
            return fail(c, FF_E_ARG, "FF_OPT_RESIDUAL: unknown value%s");
        c->residual_mode = value;
        c->gen++;
        return FF_OK;
    }
    if (option == FF_OPT_TILE_SHAPE) {
        if (value < -1 || value >= NSHAPE) return fail(c, FF_E_ARG, "FF_OPT_TILE_SHAPE: -1 or a candidate%s");
        c->force_shape = value;
        c->gen++;
        return FF_OK;
    }
    if (option == FF_OPT_AUTOTUNE) {
        if (value != 0 && value != 1) return fail(c, FF_E_ARG, "FF_OPT_AUTOTUNE: 0 or 1%s");
        c->autotune = value == 1;
        c->gen++;
        return FF_OK;
    }
    if (option == FF_OPT_BLOCK_STREAMS) {
        if (value != 0 && value != 1) return fail(c, FF_E_ARG, "FF_OPT_BLOCK_STREAMS: 0 or 1%s");
        c->block_streams = value == 1;
        return FF_OK;
    }
    if (option == FF_OPT_PRECISION) {
        if (value != FF_PREC_FP64 && value != FF_PREC_MIXED)
            return fail(c, FF_E_ARG, "FF_OPT_PRECISION: unknown value%s");
        c->precision = value;
        c->gen++;
        return FF_OK;
    }
    return fail(c, FF_E_ARG, "unknown option%s");
}

const char *ff_last_error(const ff_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ff_ctx_profile(ff_ctx *c, int enable)
{
    if (!c) return FF_E_ARG;
    c->profile = enable != 0;
    return FF_OK;
}

int ff_kernel_stats(const ff_ctx *cc, const char *name, long long *launches, double *ms)
{
    ff_ctx *c = const_cast<ff_ctx *>(cc);
    if (!c || !name) return FF_E_ARG;
    int rc = drain_events(c);
    if (rc) return rc;
    auto it = c->stats.find(name);
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.first;
    if (ms) *ms = it == c->stats.end() ? 0. : it->second.second;
    return FF_OK;
}

int ff_kernel_stats_reset(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    int rc = drain_events(c);
    c->stats.clear();
    return rc;
}

int ff_block_register(ff_ctx *c, int level, int block, const ff_block *host)
{
    if (!c) return FF_E_ARG;
    c->gen++;
    int rc = check_dims(c, host);
    if (rc) return rc;
    if ((rc = check_fields(c, host, "host field %s is NULL"))) return rc;
    BlockRec r;
    r.host = *host;
    r.mirror = true;
    r.level = level;
    Blk b = make_blk(host);
    size_t bytes = sizeof(double) * (size_t)b.ntot * FF_NFIELDS;
    if (hipMalloc(&r.arena, bytes) != hipSuccess) return fail(c, FF_E_NOMEM, "device arena%s");
    HIPCHK(c, hipMemsetAsync(r.arena, 0, bytes, c->stream));
    for (int f = 0; f < FF_NFIELDS; f++) b.f[f] = r.arena + (size_t)f * b.ntot;
    r.dev = b;
    if ((rc = sweep_alloc(c, r))) {
        (void)hipFree(r.arena);
        return rc;
    }
    BlockRec *old = find(c, level, block);
    if (old) {
        if (old->mirror && old->arena) (void)hipFree(old->arena);
        (void)hipStreamSynchronize(c->stream);
        sweep_free(*old);
    }
    c->blocks[key(level, block)] = r;
    return FF_OK;
}

int ff_block_register_device(ff_ctx *c, int level, int block, const ff_block *dev)
{
    if (!c) return FF_E_ARG;
    c->gen++;
    int rc = check_dims(c, dev);
    if (rc) return rc;
    if ((rc = check_fields(c, dev, "device field %s is NULL"))) return rc;
    BlockRec r;
    r.mirror = false;
    r.level = level;
    r.dev = make_blk(dev);
    if ((rc = sweep_alloc(c, r))) return rc;
    BlockRec *old = find(c, level, block);
    if (old) {
        if (old->mirror && old->arena) (void)hipFree(old->arena);
        (void)hipStreamSynchronize(c->stream);
        sweep_free(*old);
    }
    c->blocks[key(level, block)] = r;
    return FF_OK;
}

int ff_sync(ff_ctx *c, int level, int block, int direction, const int *fields, int nfields)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    if (direction == 1) {
        int rc = sweep_check(c, *r);
        if (rc) return rc;
    }
    if (!r->mirror) return FF_OK;
    size_t bytes = sizeof(double) * (size_t)r->dev.ntot;
    bool all = fields == nullptr || nfields < 0;
    if (!all)
        for (int t = 0; t < nfields; t++)
            if (fields[t] < 0 || fields[t] >= FF_NFIELDS)
                return fail(c, FF_E_ARG, "ff_sync: field index out of range%s");
    int nf = all ? FF_NFIELDS : nfields;
    for (int t = 0; t < nf; t++) {
        int f = all ? t : fields[t];
        if (!r->host.f[f]) continue;
        if (direction == 0)
            HIPCHK(c, hipMemcpyAsync(r->dev.f[f], r->host.f[f], bytes, hipMemcpyHostToDevice, c->stream));
        else
            HIPCHK(c, hipMemcpyAsync(r->host.f[f], r->dev.f[f], bytes, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_finish(ff_ctx *c)
{
    if (!c) return FF_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->wstream) HIPCHK(c, hipStreamSynchronize(c->wstream));
    return check_all(c);
}

static inline dim3 cells_grid(const Blk &b, int tpb) { return dim3((b.imax + tpb - 1) / tpb, b.jmax, b.kmax); }
// Threads per workgroup of a launch over rows of IMAX cells: of maxtpb, maxtpb/2, ... 64
// the size that leaves the fewest idle lanes in the last workgroup of a row (the larger on
// a tie); IMAX = 192 takes 64 (three full waves) where 128 would leave a quarter idle.
// Used where it measured faster (TIME3 + COPR); the light per-cell kernels ran slower in
// 64-thread workgroups than with the idle lanes of 256.
static inline int row_tpb(const Blk &b, int maxtpb)
{
    int best = maxtpb, idle = (b.imax + maxtpb - 1) / maxtpb * maxtpb - b.imax;
    for (int t = maxtpb / 2; t >= 64; t /= 2) {
        const int w = (b.imax + t - 1) / t * t - b.imax;
        if (w < idle) best = t, idle = w;
    }
    return best;
}
static inline dim3 all_grid(const Blk &b) { return dim3((b.ntot + 255) / 256); }

static int reflec_impl(ff_ctx *c, BlockRec *r, int level);

// STATIM(M,N,INN,.TRUE.): the state, then REFLEC over the block level's boundary patches
// (ns3c.f:10497-10512)
int ff_state(ff_ctx *c, int level, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    c->lev = level;
    int rc = launch(c, "k_state", k_state, all_grid(r->dev), dim3(256), make_prm(&c->prm, level), r->dev);
    if (rc) return rc;
    return reflec_impl(c, r, level);
}

int ff_turb(ff_ctx *c, int level, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    c->lev = level;
    if (c->prm.iturb == 6) {
        const Blk &b = r->dev;
        dim3 g((b.imax + 2 + 255) / 256, b.jmax + 2, b.kmax + 2);
        Prm kp = make_prm(&c->prm, level);
        kp.earsm_mlp = c->prm.istres == 1 && c->earsm_mode == FF_EARSM_MLP;
        if (kp.earsm_mlp && !c->mlp.ready) return fail(c, FF_E_STATE, "FF_EARSM_MLP: learned closure not set%s");
        int rc = launch(c, "k_turb_sst", k_turb_sst, g, dim3(256), kp, b);
        if (rc || !kp.earsm_mlp) return rc;
        return earsm_mlp_pass(c, b);
    }
    return launch(c, "k_turb_lam", k_turb_lam, cells_grid(r->dev, 256), dim3(256), r->dev);
}

}  // extern "C"

// ---- the tile-shape autotuner of the residual kernels ----------------------------------
// Candidate (I, J, K) tile shapes of the face-once residual kernels: I along the wave for
// coalesced rows, J rows sharing I-faces through LDS, K planes marched in registers.  Every
// shape gives the same results bit for bit (each face flux is formed by the same code from
// the same operands); they differ in occupancy, halo re-reads and how many workgroups a
// small (coarse-level) block offers the 256 CUs.

// the k-omega (kind 1) or laminar Roe (kind 0) residual kernel of shape s over block b
template <int S>
static int launch_shape(ff_ctx *c, bool kw, const Prm &kp, const Blk &b, const WallDev &wd)
{
    constexpr TileShape t = kTileShapes[S];
    const TileGrid tg = tile_grid((b.imax + t.ri - 1) / t.ri, (b.jmax + t.rj - 1) / t.rj,
                                  (b.kmax + t.rk - 1) / t.rk);
    const dim3 g(NXCD * tg.per), th(t.ri * t.rj);
    if (kw) return launch(c, "k_residual_kw", k_residual_kw<t.ri, t.rj, t.rk>, g, th, kp, b, wd, tg);
    return launch(c, "k_residual_tile", k_residual_tile<1, false, t.ri, t.rj, t.rk>, g, th, kp, b, wd, tg);
}

static int launch_shape_any(ff_ctx *c, int s, bool kw, const Prm &kp, const Blk &b, const WallDev &wd)
{
    switch (s) {
    case 0: return launch_shape<0>(c, kw, kp, b, wd);
    case 1: return launch_shape<1>(c, kw, kp, b, wd);
    case 2: return launch_shape<2>(c, kw, kp, b, wd);
    case 3: return launch_shape<3>(c, kw, kp, b, wd);
    case 4: return launch_shape<4>(c, kw, kp, b, wd);
    case 5: return launch_shape<5>(c, kw, kp, b, wd);
    case 6: return launch_shape<6>(c, kw, kp, b, wd);
    default: return launch_shape<7>(c, kw, kp, b, wd);
    }
}

// The shape for this block and kind: from the table, or, on the first residual of such a
// block outside a graph capture, by timing every candidate (one launch to warm up, three
// timed with HIP events).  The k-omega kernel reads SEPS (the cross-diffusion CROSKO left
// there) before it overwrites it with the omega source, so SEPS is saved first and put back
// before every launch; every other output is recomputed from unchanged inputs.  -1: the
// built-in shapes.
static int tuned_shape(ff_ctx *c, bool kw, const Prm &kp, const Blk &b, const WallDev &wd)
{
    if (c->force_shape >= 0) return c->force_shape;
    if (!c->autotune || c->profile) return -1;
    const auto key = std::make_tuple(b.imax, b.jmax, b.kmax, kw ? 1 : 0);
    auto it = c->tuned.find(key);
    if (it != c->tuned.end()) return it->second.first;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return -1;
    const size_t bytes = sizeof(double) * (size_t)b.ntot;
    double *seps = kw ? FLD(b, FF_SEPS) : nullptr, *save = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if ((kw && hipMalloc(&save, bytes) != hipSuccess) || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        if (save) (void)hipFree(save);
        if (e0) (void)hipEventDestroy(e0);
        return -1;
    }
    auto restore = [&]() { return !kw || hipMemcpyAsync(seps, save, bytes, hipMemcpyDeviceToDevice, c->stream) == hipSuccess; };
    bool ok = !kw || hipMemcpyAsync(save, seps, bytes, hipMemcpyDeviceToDevice, c->stream) == hipSuccess;
    int best = -1;
    double best_ms = 1e30;
    for (int sh = 0; sh < NSHAPE && ok; sh++) {
        ok = restore() && launch_shape_any(c, sh, kw, kp, b, wd) == FF_OK;
        double tot = 0.;
        for (int rep = 0; rep < 3 && ok; rep++) {
            float ms = 0.f;
            ok = restore() && hipEventRecord(e0, c->stream) == hipSuccess &&
                 launch_shape_any(c, sh, kw, kp, b, wd) == FF_OK &&
                 hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            tot += ms;
        }
        if (ok && tot / 3. < best_ms) {
            best_ms = tot / 3.;
            best = sh;
        }
    }
    ok = restore() && ok;
    (void)hipStreamSynchronize(c->stream);
    if (save) (void)hipFree(save);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!ok) return -1;
    if (best >= 0) c->tuned[key] = {best, best_ms};
    return best;
}

extern "C" {

// where the RESL2 partials and sums of a residual go: the context's buffers, or the fixed
// scratch array of ff_residual_noalloc
struct RedBuf {
    double *part;
    int nparts;
    double *out;
};

static int inlout_impl(ff_ctx *c, BlockRec *r, int level);   // ffbc.hip

static int residual_impl(ff_ctx *c, BlockRec *r, int level, int m2, double *resi,
                         const RedBuf *rb = nullptr, bool host = true)
{
    const Blk &b = r->dev;
    c->lev = level;
    int rc;
    if (c->prm.iturb == 6 && m2 >= 2) {
        // RESIKE's own DERFUN (res3c.f:3749-3769), EXTEND on coarse levels
        const int g0 = level == 1 ? 1 : 0;
        if ((rc = launch(c, "k_grad_zero", k_grad_zero, all_grid(b), dim3(256), b))) return rc;
        dim3 g((b.imax + 2 * g0 + 255) / 256, b.jmax + 2 * g0, b.kmax + 2 * g0);
        if ((rc = launch(c, "k_derfun", k_derfun, g, dim3(256), b, g0))) return rc;
        if (level >= 2 && (rc = launch(c, "k_extend", k_extend, cells_grid(b, 256), dim3(256), b))) return rc;
    }
    const bool kw = c->prm.iturb == 6;
    const int nq = kw ? 7 : 5;
    const bool fns = level == 1 && c->prm.fullns;   // FLUXBL/FLKEBL on M = 1 only
    // FULLNS gradients: RESIDU forms DERFUN itself (res3c.f:250-256; the SST has TURBEQ's);
    // TEMDER after RESIKE's CROSKO rounding of RK/REPS and PRODKE (res3c.f:3989-4007)
    auto gradients = [&]() -> int {
        int e;
        if (!kw) {
            dim3 g((b.imax + 2 + 255) / 256, b.jmax + 2, b.kmax + 2);
            if ((e = launch(c, "k_derfun", k_derfun, g, dim3(256), b, 1))) return e;
        }
        const int gk = b.kmax > 1 ? 1 : 0;
        dim3 g((b.imax + 2 + 255) / 256, b.jmax + 2, b.kmax + 2 * gk);
        return kw ? launch(c, "k_temder", k_temder<true>, g, dim3(256), b, gk)
                  : launch(c, "k_temder", k_temder<false>, g, dim3(256), b, gk);
    };
    // BOTSKE at the solid-wall patches: the wall fluxes, patch arrays, PTUR next to the wall
    // and the near-wall corrections, per direction K, I, J in patch order (the order of the
    // FLXCYC/FLUCYC calls, res3c.f:7075-7157), ahead of the residual kernel that uses them
    WallRec &wr = r->walls;
    auto walls = [&]() -> int {
        if (wr.dev.empty()) return FF_OK;
        int e;
        if (wr.pending) HIPCHK(c, hipStreamWaitEvent(c->stream, wr.done, 0));
        Prm kp = make_prm(&c->prm, level);
        for (int dir : {2, 0, 1})
            for (const WallPatchDev &wp : wr.dev) {
                if (wp.dir != dir) continue;
                dim3 g((wp.n + 255) / 256);
                e = kw ? launch(c, "k_wall", k_wall<true>, g, dim3(256), kp, b, wr.wd, wp, wr.env, level)
                       : launch(c, "k_wall", k_wall<false>, g, dim3(256), kp, b, wr.wd, wp, wr.env, level);
                if (e) return e;
            }
        return FF_OK;
    };
    // the force sums of level 1 (BOTSKE loop 2000) on the side stream, overlapping the
    // stages that follow; the next wall pass of this block waits for them
    auto forces = [&]() -> int {
        if (wr.dev.empty() || level != 1) return FF_OK;
        if (c->in_pool) {
            // blocks side by side on the stream pool: the block's own stream is already one
            // of many running at once, and a single side stream for every block's serial
            // sums would hold up the pool streams that share its hardware queue
            return launch(c, "k_wall_forces", k_wall_forces, dim3((unsigned)wr.dev.size()), dim3(WF_N),
                          make_prm(&c->prm, level), b, wr.wd, (const WallPatchDev *)wr.dev_tab,
                          wr.env, wr.sums);
        }
        HIPCHK(c, hipEventRecord(wr.fwd, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->wstream, wr.fwd, 0));
        int e = launch_on(c, c->wstream, "k_wall_forces", k_wall_forces, dim3((unsigned)wr.dev.size()),
                          dim3(WF_N), make_prm(&c->prm, level), b, wr.wd,
                          (const WallPatchDev *)wr.dev_tab, wr.env, wr.sums);
        if (e) return e;
        HIPCHK(c, hipEventRecord(wr.done, c->wstream));
        wr.pending = true;
        return FF_OK;
    };
    // ZEROZZ of DRO..DE (DRK, DEPS) on every cell (res3c.f:194-198, 3668-3674): ghost
    // residuals stay zero
    if ((rc = launch(c, "k_zero", k_zero, all_grid(b), dim3(256), b, nq))) return rc;
    const bool vl = c->prm.iflux == 0;
    if (kw) {
        if (c->residual_mode != FF_RESIDUAL_TILES)
            return fail(c, FF_E_UNSUP, "the k-omega residual has the tile form only%s");
        Prm kp = make_prm(&c->prm, level);
        // CROSKO, its rounding of RK/REPS, FUN1KO's floor, FUN1KO + PRODKE (res3c.f:3817-3994)
        if ((rc = launch(c, "k_crosko", k_crosko, cells_grid(b, 256), dim3(256), kp, b))) return rc;
        if ((rc = launch(c, "k_kw_round", k_kw_round, dim3(r->kw_parts), dim3(256), b, r->kw_scr))) return rc;
        if ((rc = launch(c, "k_max_final", k_max_final, dim3(1), dim3(256), (const double *)r->kw_scr,
                         r->kw_parts, r->kw_scr + r->kw_parts))) return rc;
        if ((rc = launch(c, "k_fun1_prod", k_fun1_prod, cells_grid(b, 256), dim3(256), kp, b,
                         (const double *)(r->kw_scr + r->kw_parts)))) return rc;
        if (fns && (rc = gradients())) return rc;
        if ((rc = walls())) return rc;
        if (level == 1 && (rc = inlout_impl(c, r, level))) return rc;
        const int sh = fns ? -1 : tuned_shape(c, true, kp, b, wr.wd);
        if (sh >= 0) {
            rc = launch_shape_any(c, sh, true, kp, b, wr.wd);
        } else if (level >= 2) {
            // the coarse levels' blocks are small: 32x4 tiles of two planes keep every CU busy
            const TileGrid tg = tile_grid((b.imax + 31) / 32, (b.jmax + 3) / 4, (b.kmax + 1) / 2);
            rc = launch(c, "k_residual_kw", k_residual_kw<32, 4, 2>, dim3(NXCD * tg.per), dim3(128),
                        kp, b, wr.wd, tg);
        } else {
            const TileGrid tg = tile_grid((b.imax + 63) / 64, (b.jmax + 3) / 4, (b.kmax + 7) / 8);
            dim3 g(NXCD * tg.per);
            rc = fns ? launch(c, "k_residual_kw_ns", k_residual_kw<64, 4, 8, true>, g, dim3(256), kp, b, wr.wd, tg)
                     : launch(c, "k_residual_kw", k_residual_kw<64, 4, 8>, g, dim3(256), kp, b, wr.wd, tg);
        }
    } else {
        if (fns && (rc = gradients())) return rc;
        if ((rc = walls())) return rc;
        if (level == 1 && (rc = inlout_impl(c, r, level))) return rc;
        const int sh = (fns || vl || c->residual_mode != FF_RESIDUAL_TILES) ? -1
                       : tuned_shape(c, false, make_prm(&c->prm, level), b, wr.wd);
        if (sh >= 0) {
            rc = launch_shape_any(c, sh, false, make_prm(&c->prm, level), b, wr.wd);
        } else if (c->residual_mode == FF_RESIDUAL_TILES) {
            const TileGrid tg = tile_grid((b.imax + RI - 1) / RI, (b.jmax + RJ - 1) / RJ,
                                          (b.kmax + RKC - 1) / RKC);
            dim3 g(NXCD * tg.per);
            auto kern = fns ? (vl ? k_residual_tile<0, true> : k_residual_tile<1, true>)
                            : (vl ? k_residual_tile<0> : k_residual_tile<1>);
            rc = launch(c, fns ? "k_residual_tile_ns" : "k_residual_tile", kern, g,
                        dim3(RI * RJ), make_prm(&c->prm, level), b, wr.wd, tg);
        } else {
            auto kern = fns ? (vl ? k_residual<0, true> : k_residual<1, true>)
                            : (vl ? k_residual<0> : k_residual<1>);
            rc = launch(c, fns ? "k_residual_ns" : "k_residual", kern, cells_grid(b, 128),
                        dim3(128), make_prm(&c->prm, level), b, wr.wd);
        }
    }
    if (rc) return rc;
    if ((rc = forces())) return rc;
    if (level == 1 && resi) {
        const RedBuf own{c->red_part, c->red_parts, c->red_out};
        const RedBuf &rd = rb ? *rb : own;
        int ncell = b.imax * b.jmax * b.kmax;
        int parts = (ncell + RED_THREADS - 1) / RED_THREADS;
        if (parts > rd.nparts) parts = rd.nparts;
        rc = kw ? launch(c, "k_resl2_part", k_resl2_part<7>, dim3(parts), dim3(RED_THREADS), b, rd.part)
                : launch(c, "k_resl2_part", k_resl2_part<5>, dim3(parts), dim3(RED_THREADS), b, rd.part);
        if (rc) return rc;
        rc = launch(c, "k_resl2_final", k_resl2_final, dim3(1), dim3(RED_THREADS), (const double *)rd.part,
                    parts, nq, rd.out);
        if (rc) return rc;
        // host = false: the sums stay in rd.out for the caller to read back later
        if (!host) return FF_OK;
        HIPCHK(c, hipMemcpyAsync(c->red_host, rd.out, sizeof(double) * nq, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        for (int q = 0; q < 7; q++) resi[q] = q < nq ? c->red_host[q] : 0.;
    }
    return FF_OK;
}

int ff_residual(ff_ctx *c, int level, int block, int m2, double *resi)
{
    (void)m2;
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    return residual_impl(c, r, level, m2, resi);
}

int ff_tuned_shape(const ff_ctx *c, int level, int block, int *shape, double *ms)
{
    if (!c || !shape) return FF_E_ARG;
    auto it = c->blocks.find(key(level, block));
    if (it == c->blocks.end()) return FF_E_STATE;
    const Blk &b = it->second.dev;
    const auto t = c->tuned.find(std::make_tuple(b.imax, b.jmax, b.kmax, c->prm.iturb == 6 ? 1 : 0));
    if (t == c->tuned.end()) return FF_E_STATE;
    const TileShape &sh = kTileShapes[t->second.first];
    shape[0] = sh.ri; shape[1] = sh.rj; shape[2] = sh.rk;
    if (ms) *ms = t->second.second;
    return FF_OK;
}

// ff_residual on caller-owned device arrays with every scratch buffer in the fixed array:
// no registration, no allocation
int ff_residual_noalloc(ff_ctx *c, int level, const ff_block *dev, int m2, double *resi)
{
    if (!c) return FF_E_ARG;
    if (level < 1) return fail(c, FF_E_ARG, "ff_residual_noalloc: level must be >= 1%s");
    int rc = check_dims(c, dev);
    if (rc) return rc;
    if ((rc = check_fields(c, dev, "device field %s is NULL"))) return rc;
    if (c->profile) return fail(c, FF_E_STATE, "ff_residual_noalloc: per-kernel timing is on%s");
    int cur = -1;
    HIPCHK(c, hipGetDevice(&cur));
    if (cur != c->device)
        return fail(c, FF_E_STATE, "ff_residual_noalloc: the current device is not the context's%s");
    BlockRec r;
    r.mirror = false;
    r.level = level;
    r.dev = make_blk(dev);
    if (c->prm.iturb == 6) {
        r.kw_parts = (r.dev.ntot + 255) / 256;
        if (r.kw_parts > NA_KW_PARTS)
            return fail(c, FF_E_ARG, "ff_residual_noalloc: the k-omega residual takes blocks of at most "
                        "8388608 entries including ghost layers (this one has %s)",
                        std::to_string(r.dev.ntot).c_str());
        r.kw_scr = c->scratch + NA_OFF_KW;
    }
    const RedBuf rb{c->scratch, NA_RED_PARTS, c->scratch + NA_OFF_OUT};
    return residual_impl(c, &r, level, m2, resi, &rb);
}

// the staged LU-SGS records of a block level, allocated on first use and kept
static bool staged_mode(const ff_ctx *c)
{
    return c->lusgs_mode == FF_LUSGS_STAGED || c->lusgs_mode == FF_LUSGS_STAGED1;
}

static int staged_alloc(ff_ctx *c, BlockRec *r)
{
    if (r->st_mem) return FF_OK;
    const Blk &b = r->dev;
    SweepSt &st = r->st;
    st.nstep = b.imax + TJ + TK - 2;
    const size_t nt = (size_t)r->sw.ntj * r->sw.ntk;
    const size_t nrec = nt * st.nstep * NRF * TW, ndn = nt * st.nstep * NDN * TW;
    if (hipMalloc(&r->st_mem, sizeof(double) * (2 * nrec + ndn)) != hipSuccess)
        return fail(c, FF_E_NOMEM, "LU-SGS staged records%s");
    // sized for FP64 records; FP32 records use the first half of each
    st.recp = r->st_mem;
    st.recc = r->st_mem + nrec;
    st.dn = r->st_mem + 2 * nrec;
    return FF_OK;
}

static int ddadi_impl(ff_ctx *c, const Prm &kp, const Blk &b);   // ffddadi.hip

static int implicit_impl(ff_ctx *c, BlockRec *r, int level, bool divcp)
{
    const Blk &b = r->dev;
    c->lev = level;
    Prm kp = make_prm(&c->prm, level);
    int rc = launch(c, "k_time3_copr", k_time3_copr, cells_grid(b, row_tpb(b, 128)), dim3(row_tpb(b, 128)), kp, b);
    if (rc) return rc;
    // BLKS%LUSGS = 0: the DDADI sweeps in place of the LU-SGS
    if (c->prm.ddadi) return ddadi_impl(c, kp, b);
    if (c->lusgs_mode == FF_LUSGS_KMARCH) {
        // DRO/CP is applied inside the forward sweep
        dim3 g(r->swk.nti * r->swk.ntj);
        if (c->prm.iturb == 6) {
            if ((rc = launch(c, "k_pred_km", k_sweep_km<false, true>, g, dim3(4 * TW), kp, b, r->swk))) return rc;
            return launch(c, "k_corr_km", k_sweep_km<true, true>, g, dim3(4 * TW), kp, b, r->swk);
        }
        if ((rc = launch(c, "k_pred_km", k_sweep_km<false, false>, g, dim3(4 * TW), kp, b, r->swk))) return rc;
        return launch(c, "k_corr_km", k_sweep_km<true, false>, g, dim3(4 * TW), kp, b, r->swk);
    }
    if (staged_mode(c)) {
        // records of both sweeps in step order, then the two wavefronts on them
        if ((rc = staged_alloc(c, r))) return rc;
        dim3 g(r->sw.ntj * r->sw.ntk);
        const TileGrid tp = tile_grid((r->st.nstep + PSW - 1) / PSW, r->sw.ntj * r->sw.ntk, 1);
        dim3 gp(NXCD * tp.per);
        const bool ke = c->prm.iturb == 6;
        // FF_PREC_MIXED: the records (diagonal, coefficients, face bases) in FP32
        const bool f32 = c->precision == FF_PREC_MIXED;
        auto prep = ke ? (f32 ? k_sweep_prep<true, float> : k_sweep_prep<true>)
                       : (f32 ? k_sweep_prep<false, float> : k_sweep_prep<false>);
        // FF_LUSGS_STAGED: four waves per tile; FF_LUSGS_STAGED1: one
        const bool w1 = c->lusgs_mode == FF_LUSGS_STAGED1;
        auto pred = w1 ? (ke ? (f32 ? k_pred_st<true, float> : k_pred_st<true>)
                             : (f32 ? k_pred_st<false, float> : k_pred_st<false>))
                       : (ke ? (f32 ? k_pred_wg<true, float> : k_pred_wg<true>)
                             : (f32 ? k_pred_wg<false, float> : k_pred_wg<false>));
        auto corr = w1 ? (ke ? (f32 ? k_corr_st<true, float> : k_corr_st<true>)
                             : (f32 ? k_corr_st<false, float> : k_corr_st<false>))
                       : (ke ? (f32 ? k_corr_wg<true, float> : k_corr_wg<true>)
                             : (f32 ? k_corr_wg<false, float> : k_corr_wg<false>));
        const dim3 ts(w1 ? TW : WGW * TW);
        // (forming the corrector's records on a side stream during the predictor's wavefront
        // was measured slower: the latency-bound sweep loses more than the pass gains)
        rc = launch(c, "k_sweep_prep", prep, gp, dim3(PSW * TW), kp, b, r->sw, r->st, tp);
        if (rc) return rc;
        // each sweep leaves its halo slots filled; empty them for the next one
        const dim3 gf((unsigned)((r->ntile_slots + 255) / 256));
        rc = launch(c, "k_pred_st", pred, g, ts, kp, b, r->sw, r->st);
        if (rc) return rc;
        if ((rc = launch(c, "k_fill_slots", k_fill_u64, gf, dim3(256), r->sw.hj, r->ntile_slots, HALO_EMPTY))) return rc;
        rc = launch(c, "k_corr_st", corr, g, ts, kp, b, r->sw, r->st);
        if (rc) return rc;
        return launch(c, "k_fill_slots", k_fill_u64, gf, dim3(256), r->sw.hj, r->ntile_slots, HALO_EMPTY);
    }
    if (c->lusgs_mode == FF_LUSGS_TILES) {
        // DRO/CP is applied inside the forward sweep
        dim3 g(r->sw.ntj * r->sw.ntk);
        if (c->prm.iturb == 6) {
            if ((rc = launch(c, "k_pred_tiles", k_pred_tiles<true>, g, dim3(TW), kp, b, r->sw))) return rc;
            return launch(c, "k_corr_tiles", k_corr_tiles<true>, g, dim3(TW), kp, b, r->sw);
        }
        if ((rc = launch(c, "k_pred_tiles", k_pred_tiles<false>, g, dim3(TW), kp, b, r->sw))) return rc;
        return launch(c, "k_corr_tiles", k_corr_tiles<false>, g, dim3(TW), kp, b, r->sw);
    }
    if (c->prm.iturb == 6)
        return fail(c, FF_E_UNSUP, "the k-omega LU-SGS has the tile-wavefront form only%s");
    const int mlo = 3, mhi = b.imax + b.jmax + b.kmax;
    // backward sweep: hyperplanes from the far corner down
    for (int m = mhi; m >= mlo; m--) {
        int klo = m - b.imax - b.jmax > 1 ? m - b.imax - b.jmax : 1;
        int khi = m - 2 < b.kmax ? m - 2 : b.kmax;
        int jlo = m - b.imax - b.kmax > 1 ? m - b.imax - b.kmax : 1;
        int jhi = m - 2 < b.jmax ? m - 2 : b.jmax;
        int jn = jhi - jlo + 1, kn = khi - klo + 1;
        int nt = jn * kn;
        rc = launch(c, "k_pred_plane", k_pred_plane, dim3((nt + 127) / 128), dim3(128), kp, b, m, jlo, klo, jn);
        if (rc) return rc;
    }
    for (int m = mlo; m <= mhi; m++) {
        int klo = m - b.imax - b.jmax > 1 ? m - b.imax - b.jmax : 1;
        int khi = m - 2 < b.kmax ? m - 2 : b.kmax;
        int jlo = m - b.imax - b.kmax > 1 ? m - b.imax - b.kmax : 1;
        int jhi = m - 2 < b.jmax ? m - 2 : b.jmax;
        int jn = jhi - jlo + 1, kn = khi - klo + 1;
        int nt = jn * kn;
        rc = launch(c, "k_corr_plane", k_corr_plane, dim3((nt + 127) / 128), dim3(128), kp, b, m, jlo, klo, jn);
        if (rc) return rc;
    }
    if (divcp) {
        rc = launch(c, "k_divcp", k_divcp, cells_grid(b, 256), dim3(256), b);
        if (rc) return rc;
    }
    return FF_OK;
}

int ff_implicit(ff_ctx *c, int level, int block, int m2)
{
    (void)m2;
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    int rc = implicit_impl(c, r, level, true);
    if (rc) return rc;
    return c->lusgs_mode != FF_LUSGS_PLANES ? sweep_check(c, *r) : FF_OK;
}

static int update_impl(ff_ctx *c, BlockRec *r, int level)
{
    const Blk &b = r->dev;
    c->lev = level;
    const bool kw = c->prm.iturb == 6;
    int rc = kw ? launch(c, "k_fin", k_fin_kw, cells_grid(b, 256), dim3(256), b)
                : launch(c, "k_fin", k_fin, cells_grid(b, 256), dim3(256), b);
    if (rc) return rc;
    if (level >= 2)
        return launch(c, "k_limits_mg", k_limits_mg, all_grid(b), dim3(256), make_prm(&c->prm, level),
                      b, 1, kw ? 1 : 0);
    return launch(c, "k_limits", k_limits, all_grid(b), dim3(256), make_prm(&c->prm, level), b);
}

// TRANSF/TRANKE from level m to level m+1 of a block (multi.f:156-620)
static int restrict_impl(ff_ctx *c, BlockRec *fr, BlockRec *cr)
{
    const Blk &f = fr->dev, &cb = cr->dev;
    const bool kw = c->prm.iturb == 6;
    int rc;
    c->lev = cr->level;
    dim3 g((cb.imax + 2 + 255) / 256, cb.jmax + 2, cb.kmax + 2);
    rc = kw ? launch(c, "k_restrict", k_restrict<true>, g, dim3(256), f, cb)
            : launch(c, "k_restrict", k_restrict<false>, g, dim3(256), f, cb);
    if (rc) return rc;
    const int na[3] = {cb.jmax + 4, cb.imax + 4, cb.imax + 4};
    const int nb[3] = {cb.kmax + 4, cb.kmax + 4, cb.jmax + 4};
    for (int face = 0; face < 3; face++) {
        dim3 gg((na[face] + 255) / 256, nb[face]);
        rc = kw ? launch(c, "k_restrict_ghosts", k_restrict_ghosts<true>, gg, dim3(256), cb, face)
                : launch(c, "k_restrict_ghosts", k_restrict_ghosts<false>, gg, dim3(256), cb, face);
        if (rc) return rc;
    }
    return FF_OK;
}

static int forcing_impl(ff_ctx *c, BlockRec *r, int phase)
{
    c->lev = r->level;
    return launch(c, "k_forcing", k_forcing, all_grid(r->dev), dim3(256), r->dev,
                  c->prm.iturb == 6 ? 7 : 5, phase);
}

// INTERP/INTERT from level m to level m-1 and the limiters of CYCLES (ns3c.f:2017-2118)
static int prolong_impl(ff_ctx *c, BlockRec *cr, BlockRec *fr, int level_fine)
{
    const Blk &f = fr->dev, &cb = cr->dev;
    const bool kw = c->prm.iturb == 6;
    Prm kp = make_prm(&c->prm, level_fine);
    c->lev = level_fine;
    int rc = kw ? launch(c, "k_prolong", k_prolong<true>, cells_grid(f, 256), dim3(256), kp, cb, f)
                : launch(c, "k_prolong", k_prolong<false>, cells_grid(f, 256), dim3(256), kp, cb, f);
    if (rc) return rc;
    return launch(c, "k_limits_mg", k_limits_mg, all_grid(f), dim3(256), kp, f, kw ? 1 : 0, kw ? 1 : 0);
}

static int check_pair(ff_ctx *c, BlockRec *fr, BlockRec *cr)
{
    if (!fr || !cr) return fail(c, FF_E_STATE, "multigrid level not registered%s");
    const Blk &f = fr->dev, &cb = cr->dev;
    // a planar block (KMAX = 1) keeps KMAX = 1 on every level: KMAXG = MAX0(1, KMAX/NDIV)
    // (main.f:1756), with TRANSF's planar path
    const bool planar = f.kmax == 1 && cb.kmax == 1;
    if (f.imax != 2 * cb.imax || f.jmax != 2 * cb.jmax || (f.kmax != 2 * cb.kmax && !planar))
        return fail(c, FF_E_ARG, "a coarse level must have half the cells of the finer level in each "
                                 "direction (or KMAX = 1 on both)%s");
    return FF_OK;
}

int ff_update(ff_ctx *c, int level, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    return update_impl(c, r, level);
}

int ff_save_old(ff_ctx *c, int level, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    c->lev = level;
    return launch(c, "k_save_old", k_save_old, all_grid(r->dev), dim3(256), r->dev, (int *)nullptr);
}

int ff_cycle(ff_ctx *c, int block, int ncycles, double *resi)
{
    if (!c || ncycles < 0) return FF_E_ARG;
    BlockRec *r = find(c, 1, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    int rc;
    for (int it = 0; it < ncycles; it++) {
        bool last = it == ncycles - 1;
        c->lev = 1;
        if ((rc = launch(c, "k_save_old", k_save_old, all_grid(r->dev), dim3(256), r->dev, r->bc_icyc))) return rc;
        if ((rc = ff_state(c, 1, block))) return rc;
        if ((rc = ff_turb(c, 1, block))) return rc;
        if ((rc = residual_impl(c, r, 1, 1, last ? resi : nullptr))) return rc;
        if ((rc = implicit_impl(c, r, 1, true))) return rc;
        if ((rc = update_impl(c, r, 1))) return rc;
    }
    return c->lusgs_mode != FF_LUSGS_PLANES ? sweep_check(c, *r) : FF_OK;
}

int ff_restrict(ff_ctx *c, int mfine, int block)
{
    if (!c) return FF_E_ARG;
    BlockRec *fr = find(c, mfine, block), *cr = find(c, mfine + 1, block);
    int rc = check_pair(c, fr, cr);
    if (rc) return rc;
    return restrict_impl(c, fr, cr);
}

int ff_forcing(ff_ctx *c, int level, int block, int phase)
{
    if (!c || (phase != 0 && phase != 1)) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    return forcing_impl(c, r, phase);
}

int ff_prolong(ff_ctx *c, int mcoarse, int block)
{
    if (!c || mcoarse < 2) return FF_E_ARG;
    BlockRec *fr = find(c, mcoarse - 1, block), *cr = find(c, mcoarse, block);
    int rc = check_pair(c, fr, cr);
    if (rc) return rc;
    return prolong_impl(c, cr, fr, mcoarse - 1);
}

// One V-cycle (CYCLES ns3c.f:1450-2137 with MGRID = nlev): the finest-level iteration,
// the FAS down-loop with pre-smoothing, the prolongation; RESL2 on level 1 into rs if given
static int vcycle_once(ff_ctx *c, const std::vector<BlockRec *> &lv, int nlev, int block, double *rs,
                       const RedBuf *rb = nullptr, bool host = true)
{
    int rc;
    BlockRec *r1 = lv[0];
    // level 1: CYCLES ns3c.f:1450-1596
    c->lev = 1;
    if ((rc = launch(c, "k_save_old", k_save_old, all_grid(r1->dev), dim3(256), r1->dev, r1->bc_icyc))) return rc;
    if ((rc = ff_state(c, 1, block))) return rc;
    if ((rc = ff_turb(c, 1, block))) return rc;
    if ((rc = residual_impl(c, r1, 1, 1, rs, rb, host))) return rc;
    if ((rc = implicit_impl(c, r1, 1, true))) return rc;
    if ((rc = update_impl(c, r1, 1))) return rc;
    // FAS down-loop, pre-smoothing only (ns3c.f:1607-1999)
    for (int m = 2; m <= nlev; m++) {
        BlockRec *fr = lv[m - 2], *cr = lv[m - 1];
        if ((rc = ff_state(c, m - 1, block))) return rc;
        if ((rc = residual_impl(c, fr, m - 1, m, m - 1 == 1 ? rs : nullptr, rb, host))) return rc;
        if (m > 2 && (rc = forcing_impl(c, fr, 1))) return rc;
        if ((rc = restrict_impl(c, fr, cr))) return rc;
        if ((rc = ff_state(c, m, block))) return rc;
        if ((rc = residual_impl(c, cr, m, m, nullptr))) return rc;
        if ((rc = forcing_impl(c, cr, 0))) return rc;
        if ((rc = forcing_impl(c, cr, 1))) return rc;
        if ((rc = implicit_impl(c, cr, m, true))) return rc;
        if ((rc = update_impl(c, cr, m))) return rc;
    }
    // prolongation (ns3c.f:2005-2137)
    for (int m = nlev; m >= 2; m--)
        if ((rc = prolong_impl(c, lv[m - 1], lv[m - 2], m - 1))) return rc;
    return FF_OK;
}

static int vcycle_levels(ff_ctx *c, int block, int nlev, std::vector<BlockRec *> &lv)
{
    lv.assign(nlev, nullptr);
    for (int m = 1; m <= nlev; m++) {
        lv[m - 1] = find(c, m, block);
        if (!lv[m - 1]) return fail(c, FF_E_STATE, "multigrid level not registered%s");
        if (m >= 2) {
            int rc = check_pair(c, lv[m - 2], lv[m - 1]);
            if (rc) return rc;
        }
    }
    return FF_OK;
}

int ff_vcycle(ff_ctx *c, int block, int nlev, int ncycles, double *resi)
{
    if (!c || ncycles < 0 || nlev < 1) return FF_E_ARG;
    std::vector<BlockRec *> lv;
    int rc = vcycle_levels(c, block, nlev, lv);
    if (rc) return rc;
    for (int it = 0; it < ncycles; it++)
        if ((rc = vcycle_once(c, lv, nlev, block, it == ncycles - 1 ? resi : nullptr))) return rc;
    if (c->lusgs_mode != FF_LUSGS_PLANES)
        for (int m = 0; m < nlev; m++)
            if ((rc = sweep_check(c, *lv[m]))) return rc;
    return FF_OK;
}

static void vgraph_free(ff_ctx *c)
{
    if (c->vg_exec) (void)hipGraphExecDestroy(c->vg_exec);
    if (c->vg_graph) (void)hipGraphDestroy(c->vg_graph);
    c->vg_exec = nullptr;
    c->vg_graph = nullptr;
}

int ff_vcycle_graph(ff_ctx *c, int block, int nlev, int ncycles)
{
    if (!c || ncycles < 0 || nlev < 1) return FF_E_ARG;
    std::vector<BlockRec *> lv;
    int rc = vcycle_levels(c, block, nlev, lv);
    if (rc) return rc;
    if (c->profile) return fail(c, FF_E_STATE, "ff_vcycle_graph: per-kernel timing is on%s");
    if (!c->stream) return fail(c, FF_E_STATE, "ff_vcycle_graph: the null stream cannot be captured%s");
    const long long key[6] = {block, nlev, c->lusgs_mode, c->residual_mode, c->earsm_mode, c->gen};
    bool same = c->vg_exec != nullptr;
    for (int q = 0; q < 6; q++) same = same && c->vg_key[q] == key[q];
    if (!same) {
        vgraph_free(c);
        if (staged_mode(c))
            for (BlockRec *r : lv)
                if ((rc = staged_alloc(c, r))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        // a force pass of an eager call may still run on the side stream: finish it, so that
        // the capture forks and joins that stream only through events recorded inside it
        for (BlockRec *r : lv)
            if (r->walls.pending) {
                HIPCHK(c, hipEventSynchronize(r->walls.done));
                r->walls.pending = false;
            }
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        rc = vcycle_once(c, lv, nlev, block, nullptr);
        // join the last force pass of each level back into the captured stream: a replay
        // ends only when its force sums are complete
        for (BlockRec *r : lv)
            if (r->walls.pending) {
                if (!rc && hipStreamWaitEvent(c->stream, r->walls.done, 0) != hipSuccess)
                    rc = fail(c, FF_E_HIP, "ff_vcycle_graph: joining the force pass failed%s");
                r->walls.pending = false;
            }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess || !g) return fail(c, FF_E_HIP, "%s", hipGetErrorString(ec));
        c->vg_graph = g;
        if (hipGraphInstantiate(&c->vg_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            vgraph_free(c);
            return fail(c, FF_E_HIP, "ff_vcycle_graph: instantiation failed%s");
        }
        for (int q = 0; q < 6; q++) c->vg_key[q] = key[q];
    }
    // a force pass of an eager call may still read the wall scratch the graph rewrites
    for (BlockRec *r : lv)
        if (r->walls.pending) {
            HIPCHK(c, hipStreamWaitEvent(c->stream, r->walls.done, 0));
            r->walls.pending = false;
        }
    for (int it = 0; it < ncycles; it++) HIPCHK(c, hipGraphLaunch(c->vg_exec, c->stream));
    if (c->lusgs_mode != FF_LUSGS_PLANES)
        for (int m = 0; m < nlev; m++)
            if ((rc = sweep_check(c, *lv[m]))) return rc;
    return FF_OK;
}


// ---- solid-wall patches ---------------------------------------------------------------

static const int kNeededWall[] = {FF_XC, FF_YC, FF_ZC, FF_STRAIN, FF_RK, FF_REPS, FF_PTUR};

// the patch arrays of a descriptor in the order of the WA_* slots
static void wall_arrays(const ff_wall_patch &h, double *a[NWARR])
{
    double *const v[NWARR] = {h.uwall, h.vwall, h.wwall, h.twall, h.hflux, h.rbk, h.xcp, h.ycp,
                              h.zcp, h.cpwall, h.qwall, h.qwfric, h.tauw1, h.tauw2, h.tauwx,
                              h.tauwy, h.tauwz, h.surfx, h.surfy, h.surfz, h.surfpx, h.surfpy,
                              h.surfpz};
    for (int q = 0; q < NWARR; q++) a[q] = v[q];
}

int ff_wall_register(ff_ctx *c, int level, int block, const ff_wall_params *wpar,
                     const ff_wall_patch *patches, int npatch)
{
    if (c) c->gen++;
    if (!c || npatch < 0 || (npatch > 0 && (!patches || !wpar))) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    walls_free(r->walls);
    if (npatch == 0) return FF_OK;
    const Blk &b = r->dev;
    for (int f : kNeededWall)
        if (!b.f[f]) return fail(c, FF_E_ARG, "wall patches need field %s", std::to_string(f).c_str());
    const int nd[3] = {b.imax, b.jmax, b.kmax};
    const int nf[3] = {b.jmax * b.kmax, b.imax * b.kmax, b.imax * b.jmax};
    WallRec w;
    WallDev &wd = w.wd;
    wd.off[0] = 0;
    for (int f = 1; f < 6; f++) wd.off[f] = wd.off[f - 1] + nf[(f - 1) % 3];
    wd.nt = wd.off[5] + nf[2];
    std::vector<unsigned char> flag((size_t)wd.nt, 0);
    const bool corr = c->prm.iturb == 6 && c->prm.turcor;
    bool has_face[6] = {false, false, false, false, false, false};
    size_t nparr = 0;
    for (int l = 0; l < npatch; l++) {
        const ff_wall_patch &h = patches[l];
        if (h.iface < 1 || h.iface > 6) return fail(c, FF_E_ARG, "wall patch face %s", std::to_string(h.iface).c_str());
        if (h.iheat < 0 || h.iheat > 3) return fail(c, FF_E_UNSUP, "IHEAT %s (0-3 supported)", std::to_string(h.iheat).c_str());
        const int face = h.iface - 1, dir = face % 3;
        // the ICON ranges: (J, K) on I-faces, (I, K) on J-faces, (I, J) on K-faces
        const int rx = dir == 0 ? b.jmax : b.imax, ry = dir == 2 ? b.jmax : b.kmax;
        if (h.kx1 < 1 || h.kx2 > rx || h.kx1 > h.kx2 || h.ky1 < 1 || h.ky2 > ry || h.ky1 > h.ky2)
            return fail(c, FF_E_ARG, "wall patch %s range outside its face", std::to_string(l).c_str());
        if (nd[dir] < 3) return fail(c, FF_E_UNSUP, "wall patch on a block under 3 cells thick%s");
        double *ha[NWARR];
        wall_arrays(h, ha);
        for (int q = 0; q < NWARR; q++)
            if (!ha[q]) return fail(c, FF_E_ARG, "wall patch %s array missing", std::to_string(l).c_str());
        WallPatchDev d{};
        d.face = face; d.dir = dir; d.idir = h.iface <= 3 ? 1 : -1;
        d.kbot = h.iface <= 3 ? 1 : nd[dir];
        d.iheat = h.iheat;
        if (dir == 0) { d.sa = b.is; d.sb = b.il; d.sw = 1; d.a1 = h.kx1; d.a2 = h.kx2; d.b1 = h.ky1; d.b2 = h.ky2; }
        else if (dir == 1) { d.sa = b.il; d.sb = 1; d.sw = b.is; d.a1 = h.ky1; d.a2 = h.ky2; d.b1 = h.kx1; d.b2 = h.kx2; }
        else { d.sa = 1; d.sb = b.is; d.sw = b.il; d.a1 = h.kx1; d.a2 = h.kx2; d.b1 = h.ky1; d.b2 = h.ky2; }
        d.na = d.a2 - d.a1 + 1 + 4;
        d.n = (d.a2 - d.a1 + 1) * (d.b2 - d.b1 + 1);
        for (int J = d.b1; J <= d.b2; J++)
            for (int I = d.a1; I <= d.a2; I++) {
                int i, j, k;
                if (dir == 0) { i = d.kbot; j = I; k = J; }
                else if (dir == 1) { i = J; j = d.kbot; k = I; }
                else { i = I; j = J; k = d.kbot; }
                const int idx = dir == 0 ? (k - 1) * b.jmax + (j - 1)
                                         : (dir == 1 ? (k - 1) * b.imax + (i - 1) : (j - 1) * b.imax + (i - 1));
                unsigned char &fl = flag[(size_t)wd.off[face] + idx];
                if (fl) return fail(c, FF_E_ARG, "wall patches %s overlap", std::to_string(l).c_str());
                fl = 1;
            }
        has_face[face] = true;
        nparr += (size_t)d.na * (d.b2 - d.b1 + 1 + 4);
        w.dev.push_back(d);
        w.host.push_back(h);
    }
    if (corr)
        for (int dir = 0; dir < 3; dir++)
            if (has_face[dir] && has_face[dir + 3] && nd[dir] < 5)
                return fail(c, FF_E_UNSUP, "near-wall corrections of both faces meet (under 5 cells)%s");
    wd.on = 1;
    wd.corr_on = corr ? 1 : 0;
    w.env = WallEnv{wpar->t0, wpar->difpre, wpar->xmom, wpar->ymom, wpar->zmom,
                    wpar->gx, wpar->gy, wpar->gz, wpar->ground, wpar->altitude};
    const size_t nscr = (size_t)wd.nt * (7 + 8 + 3);
    if (hipMalloc(&w.parr, sizeof(double) * NWARR * nparr) != hipSuccess ||
        hipMalloc(&w.flag, (size_t)wd.nt) != hipSuccess ||
        hipMalloc(&w.scr, sizeof(double) * nscr) != hipSuccess ||
        hipMalloc(&w.sums, sizeof(double) * FF_NFORCE * npatch) != hipSuccess ||
        hipMalloc(&w.dev_tab, sizeof(WallPatchDev) * npatch) != hipSuccess ||
        hipHostMalloc(&w.sums_host, sizeof(double) * FF_NFORCE * npatch, hipHostMallocDefault) != hipSuccess) {
        walls_free(w);
        return fail(c, FF_E_NOMEM, "wall patch buffers%s");
    }
    if (hipEventCreateWithFlags(&w.fwd, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&w.done, hipEventDisableTiming) != hipSuccess) {
        walls_free(w);
        return fail(c, FF_E_HIP, "wall events%s");
    }
    if (!c->wstream && hipStreamCreateWithFlags(&c->wstream, hipStreamNonBlocking) != hipSuccess) {
        walls_free(w);
        return fail(c, FF_E_HIP, "wall force stream%s");
    }
    wd.flag = w.flag;
    wd.flux = w.scr;
    wd.corr = w.scr + (size_t)wd.nt * 7;
    wd.qterm = w.scr + (size_t)wd.nt * 15;
    size_t o = 0;
    for (auto &d : w.dev) {
        const size_t sz = (size_t)d.na * ((d.b2 - d.b1 + 1) + 4);
        for (int q = 0; q < NWARR; q++) d.arr[q] = w.parr + (size_t)q * nparr + o;
        o += sz;
    }
    HIPCHK(c, hipMemcpyAsync(w.flag, flag.data(), flag.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(w.dev_tab, w.dev.data(), sizeof(WallPatchDev) * npatch,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(w.scr, 0, sizeof(double) * nscr, c->stream));
    HIPCHK(c, hipMemsetAsync(w.sums, 0, sizeof(double) * FF_NFORCE * npatch, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    r->walls = w;
    return ff_wall_sync(c, level, block, 0);
}

int ff_wall_sync(ff_ctx *c, int level, int block, int direction)
{
    if (!c || (direction != 0 && direction != 1)) return FF_E_ARG;
    BlockRec *r = find(c, level, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    WallRec &w = r->walls;
    if (w.pending) HIPCHK(c, hipStreamWaitEvent(c->stream, w.done, 0));
    for (size_t l = 0; l < w.dev.size(); l++) {
        const WallPatchDev &d = w.dev[l];
        const size_t sz = (size_t)d.na * ((d.b2 - d.b1 + 1) + 4) * sizeof(double);
        double *ha[NWARR];
        wall_arrays(w.host[l], ha);
        for (int q = 0; q < NWARR; q++) {
            double *hp = ha[q];
            if (direction == 0)
                HIPCHK(c, hipMemcpyAsync(d.arr[q], hp, sz, hipMemcpyHostToDevice, c->stream));
            else
                HIPCHK(c, hipMemcpyAsync(hp, d.arr[q], sz, hipMemcpyDeviceToHost, c->stream));
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FF_OK;
}

int ff_wall_forces(ff_ctx *c, int block, int patch, double *sums)
{
    if (!c || !sums) return FF_E_ARG;
    BlockRec *r = find(c, 1, block);
    if (!r) return fail(c, FF_E_STATE, "block not registered%s");
    WallRec &w = r->walls;
    if (patch < 0 || patch >= (int)w.dev.size()) return fail(c, FF_E_ARG, "no wall patch %s", std::to_string(patch).c_str());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (w.pending) HIPCHK(c, hipEventSynchronize(w.done));
    HIPCHK(c, hipMemcpy(w.sums_host, w.sums, sizeof(double) * FF_NFORCE * w.dev.size(),
                        hipMemcpyDeviceToHost));
    for (int q = 0; q < FF_NFORCE; q++) sums[q] = w.sums_host[(size_t)patch * FF_NFORCE + q];
    return FF_OK;
}
}  // extern "C"

#include "ffmlp.hip"
#include "ffconv.hip"
#include "ffbc.hip"
#include "ffcon.hip"
#include "ffddadi.hip"
