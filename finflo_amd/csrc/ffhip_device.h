// ffhip_device.h — device arithmetic of the FINFLO hot path for gfx950 (FP64).
//
// Every function reproduces the operation order and literals of the reference routine it
// cites, so that with -ffp-contract=off the results equal the reference's FP64 results bit
// for bit.  Arrays are the reference's flat i-fastest slabs with two ghost layers.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/ffhip.h"

namespace ff {

// Field pointers of one block on the device, passed by value to every kernel.
struct Blk {
    int imax, jmax, kmax;
    int is, js, ks;        // ISTRID, JSTRID, KSTRID
    int il;                // ISTRID*JSTRID
    int ntot;
    double *f[FF_NFIELDS];
};

// Physical constants used by the kernels (subset of ff_params)
struct Prm {
    double gamma, rgas, pr, prt, visu0, tsu0, rmult;
    double frspre, frsden, frstem;
    double cfl, theta, cdiff, tuplim, tlolim, epslim;
    double rklim, frsvis, re, csimps, turlim, cmgk, cmgeps;
    int iturb;
    int inter[3];
    int kover, katol, istres;
    int fullns;     // full viscous fluxes (FULLNS) on this level: level 1 only
    int idiff;
    int turcor;     // BOTSKE's near-wall corrections (the SST path)
    int earsm_mlp;  // EARSM1's coefficients from the learned closure (a separate pass)
};

#define FLD(b, n) ((b).f[n])

__device__ __forceinline__ int cid(const Blk &b, int i, int j, int k)
{
    return (k + 1) * b.il + (j + 1) * b.is + (i + 1);
}

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// MUSCL kappa-scheme, XXTRAP base3c.f:750-850, face between cells L-S and L
__device__ __forceinline__ void xxtrap(const double *__restrict__ q, int L, int S, int intem,
                                       double &qr, double &ql)
{
    int inter = intem < 0 ? -intem : intem;
    int inte2 = inter < 1 ? 1 : (inter > 5 ? 5 : inter);
    // RKK1 = {0., 1., 1.3333333, 2., 1.5} (base3c.f:765)
    double rk1 = inte2 == 1 ? 0. : inte2 == 2 ? 1. : inte2 == 3 ? 1.3333333 : inte2 == 4 ? 2. : 1.5;
    double rk2 = 2. - rk1;
    double r1 = .25 * rk1, r2 = .25 * rk2;
    double q1 = q[L - S], q2 = q[L];
    if (intem <= 0 && intem >= -5) {
        double q0 = q[L - 2 * S], q3 = q[L + S];
        double diff0 = q1 - q0, diff1 = q2 - q1, diff2 = q3 - q2;
        qr = q2 - (diff1 * r1 + diff2 * r2);
        ql = q1 + (diff1 * r1 + diff0 * r2);
    } else if (inter <= 4) {
        double q0 = q[L - 2 * S], q3 = q[L + S];
        const double eps = 1.E-10;
        double diff0 = q1 - q0, diff1 = q2 - q1, diff2 = q3 - q2;
        double diff12 = diff1 * diff1 + eps;
        double phibi = (diff12 + diff1 * diff0) / (diff12 + diff0 * diff0);
        double phifi = (diff12 + diff1 * diff2) / (diff12 + diff2 * diff2);
        qr = q2 - phifi * (diff1 * r1 + diff2 * r2);
        ql = q1 + phibi * (diff1 * r1 + diff0 * r2);
    } else {
        qr = q2;
        ql = q1;
    }
}

// SURFVECTS base3c.f:12245-12295
__device__ __forceinline__ void surfvects(double a11, double a12, double a13,
                                          double &a21, double &a22, double &a23,
                                          double &a31, double &a32, double &a33)
{
    double am = fabs(a11);
    int idir = 1;
    if (fabs(a12) < am) { am = fabs(a12); idir = 2; }
    if (fabs(a13) < am) { idir = 3; }
    double b1, b2, b3;
    if (idir == 1) { b1 = 0.0; b2 = -a13; b3 = a12; }
    else if (idir == 2) { b1 = a13; b2 = 0.0; b3 = -a11; }
    else { b1 = -a12; b2 = a11; b3 = 0.0; }
    double palpo = 1. / sqrt(b1 * b1 + b2 * b2 + b3 * b3);
    b1 = b1 * palpo; b2 = b2 * palpo; b3 = b3 * palpo;
    a21 = b1; a22 = b2; a23 = b3;
    a31 = a12 * b3 - a13 * b2;
    a32 = a13 * b1 - a11 * b3;
    a33 = a11 * b2 - a12 * b1;
}

// EFPRO state.f:636-638 with FRSPRE = 0 and E0REF = 0 as FLUXKE calls it (base3c.f:4882)
__device__ __forceinline__ double efpro(double pp, double ro, double gamma)
{
    double gm1 = gamma - 1.;
    return (pp + 0.) / (gm1 * ro) - 0.;
}

template <int NV> struct FluxN { double f[NV]; };
using Flux5 = FluxN<5>;
using Flux7 = FluxN<7>;

// MUSCL switch of rho*k and rho*omega, INTETI/J/K of RESIKE res3c.f:3634-3641 (INTERTU = 0)
__device__ __forceinline__ int intet_of(int inter) { return (inter > -5 && inter < 5) ? 1 : inter; }

// FLUXKE base3c.f:4753-5125 at one face, with DERIV's I2 = I branch (base3c.f:719) on
// k-direction faces.  NV = 5: the algebraic path of FLUCYC (k and omega zero); NV = 7: the
// two-equation path of FLXCYC with k and omega reconstructed by the INTET switch and floored
// at 0.001*RKLIM, 0.001*EPSLIM (base3c.f:4906-4909).
template <bool SINGLE, int NV = 5>
__device__ __forceinline__ FluxN<NV> roe_face(const Prm &p, const Blk &b, int L, int S, int intem,
                                              const double *__restrict__ an,
                                              const double *__restrict__ anx,
                                              const double *__restrict__ any,
                                              const double *__restrict__ anz)
{
    const double *ro = FLD(b, FF_RO), *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W),
                 *P = FLD(b, FF_P), *c = FLD(b, FF_C), *drdp = FLD(b, FF_DRDP),
                 *drdh = FLD(b, FF_DRDH);
    double roip, roim, rmip, rmim, rnip, rnim, rwip, rwim, ppi, pmi;
    xxtrap(ro, L, S, intem, roip, roim);
    xxtrap(u, L, S, intem, rmip, rmim);
    xxtrap(v, L, S, intem, rnip, rnim);
    xxtrap(w, L, S, intem, rwip, rwim);
    xxtrap(P, L, S, intem, ppi, pmi);
    double fd = 0.001 * p.frsden, fp = 0.001 * p.frspre;
    roip = roip > fd ? roip : fd;
    ppi = ppi > fp ? ppi : fp;
    roim = roim > fd ? roim : fd;
    pmi = pmi > fp ? pmi : fp;
    double eip = efpro(ppi, roip, p.gamma);
    double eim = efpro(pmi, roim, p.gamma);
    double elr = efpro(ppi, roim, p.gamma);
    double erl = efpro(pmi, roip, p.gamma);
    // DERIV base3c.f:707-742
    int i2 = SINGLE ? L : L - S;
    double ypro1 = 1. / ro[L], ypro2 = 1. / ro[i2];
    double dedp1 = -drdp[L] / drdh[L] - ypro1;
    double dedp2 = -drdp[i2] / drdh[i2] - ypro2;
    double dedph = 0.5 * (dedp1 + dedp2);
    double dedr1 = 1. / drdh[L] + P[L] * (ypro1 * ypro1);
    double dedr2 = 1. / drdh[i2] + P[i2] * (ypro2 * ypro2);
    double dedrh = 0.5 * (dedr1 + dedr2);
    {
        double dp = 2. * (ppi - pmi);
        double dro = 2. * (roip - roim);
        if (fabs(dp) >= 1.E-2 * (ppi + pmi)) dedph = (eip + elr - eim - erl) / dp;
        if (fabs(dro) >= 1.E-2 * (roip + roim)) dedrh = (eip + erl - eim - elr) / dro;
    }
    double a11 = anx[L], a12 = any[L], a13 = anz[L];
    double a21, a22, a23, a31, a32, a33;
    surfvects(a11, a12, a13, a21, a22, a23, a31, a32, a33);
    double ur = a11 * rmip + a12 * rnip + a13 * rwip;
    double vr = a21 * rmip + a22 * rnip + a23 * rwip;
    double wr = a31 * rmip + a32 * rnip + a33 * rwip;
    double ul = a11 * rmim + a12 * rnim + a13 * rwim;
    double vl = a21 * rmim + a22 * rnim + a23 * rwim;
    double wl = a31 * rmim + a32 * rnim + a33 * rwim;
    // k and omega: MAX(0.001*RKLIM, RKIP) ...; zero on the algebraic path
    double rkipi = 0., rkimi = 0., reipi = 0., reimi = 0.;
    if (NV == 7) {
        double rkip, rkim, reip, reim;
        const int intet = intet_of(intem);
        xxtrap(FLD(b, FF_RK), L, S, intet, rkip, rkim);
        xxtrap(FLD(b, FF_REPS), L, S, intet, reip, reim);
        double lk = 0.001 * p.rklim, le = 0.001 * p.epslim;
        rkipi = lk > rkip ? lk : rkip;
        reipi = le > reip ? le : reip;
        rkimi = lk > rkim ? lk : rkim;
        reimi = le > reim ? le : reim;
    }
    double ypror = 1. / roip;
    double rkr = rkipi * ypror;
    double yprol = 1. / roim;
    double rkl = rkimi * yprol;
    double eir = eip + .5 * (ur * ur + vr * vr + wr * wr) + rkr;
    double eil = eim + .5 * (ul * ul + vl * vl + wl * wl) + rkl;
    double pr = ppi;
    double pkr = .6667 * rkipi;
    double epsr = reipi * ypror;
    double hr = eir + pr * ypror;
    double er = eip;
    double rohat = sqrt(roip * roim);
    double yproh = 1. / rohat;
    double pl = pmi;
    double pkl = .6667 * rkimi;
    double epsl = reimi * yprol;
    double hl = eil + yprol * pl;
    double el = eim;
    double sqrl = 1. / (roim + rohat);
    double cql = roim * sqrl;
    double cqr = 1. - cql;
    double uhat = cqr * ur + cql * ul;
    double vhat = cqr * vr + cql * vl;
    double what = cqr * wr + cql * wl;
    double hhat = cqr * hr + cql * hl;
    double ehat = cqr * er + cql * el;
    double rkhat = cqr * rkr + cql * rkl;
    double epshat = cqr * epsr + cql * epsl;
    double phat = rohat * (hhat - ehat - .5 * (uhat * uhat + vhat * vhat + what * what) - rkhat);
    double pmin = pr < pl ? pr : pl;
    double pmax = pr > pl ? pr : pl;
    phat = pmax < phat ? pmax : phat;
    phat = pmin > phat ? pmin : phat;
    double drdhh = 1. / (dedrh - phat * (yproh * yproh));
    double drdph = -drdhh * (dedph + yproh);
    double xturb = 1. + .6667 * rkhat * yproh * drdph;
    double chat2 = xturb / (drdhh * yproh + drdph);
    double cL = c[L], cM = c[L - S];
    double cmin = cL < cM ? cL : cM;
    double cmax = cL > cM ? cL : cM;
    chat2 = (cmax * cmax) < chat2 ? (cmax * cmax) : chat2;
    chat2 = (cmin * cmin) > chat2 ? (cmin * cmin) : chat2;
    double chat = sqrt(chat2);
    double pdpde = dedph;
    double dro = roip - roim;
    double du = ur - ul;
    double dk = rkr - rkl;
    double dp = pr - pl + .6667 * (rkhat * dro + rohat * dk);
    double pchat2 = 1. / (chat * chat);
    double alfa1 = .5 * (dp + rohat * chat * du) * pchat2;
    double alfa2 = .5 * (dp - rohat * chat * du) * pchat2;
    double alfa3 = dro - dp * pchat2;
    double alfa4 = rohat * (vr - vl);
    double alfa5 = rohat * (wr - wl);
    double alfa6 = rohat * dk;
    double alfa7 = rohat * (epsr - epsl);
    double rl1 = fabs(uhat + chat);
    double rl2 = fabs(uhat - chat);
    double rl3 = fabs(uhat);
    double rl4 = rl3, rl5 = rl3, rl6 = rl3, rl7 = rl3;
    double roucr = roip * ur;
    double roucl = roim * ul;
    double damp = rl1 * alfa1 + rl2 * alfa2 + rl3 * alfa3;
    double ha = .5 * an[L];
    double f1 = ha * (roucr + roucl - damp);
    double f2 = ha * (roucr * ur + pr + pkr + roucl * ul + pl + pkl -
                      (damp * uhat + (rl1 * alfa1 - rl2 * alfa2) * chat));
    double f3 = ha * (roucr * vr + roucl * vl - damp * vhat - rl4 * alfa4);
    double f4 = ha * (roucr * wr + roucl * wl - damp * what - rl5 * alfa5);
    double f5 = ha * (roucr * (hr + pkr * ypror) + roucl * (hl + pkl * yprol)
                      - damp * hhat - ((rl1 * alfa1 - rl2 * alfa2) * uhat * chat
                      - rl3 * alfa3 * rohat * (chat * chat) * pdpde + rl4 * alfa4 * vhat
                      + rl5 * alfa5 * what) + rl6 * alfa6 * (1. - .6667 * rohat * pdpde));
    FluxN<NV> out;
    out.f[0] = f1;
    out.f[1] = a11 * f2 + a21 * f3 + a31 * f4;
    out.f[2] = a12 * f2 + a22 * f3 + a32 * f4;
    out.f[3] = a13 * f2 + a23 * f3 + a33 * f4;
    out.f[4] = f5;
    if (NV == 7) {
        out.f[NV == 7 ? 5 : 0] = ha * (roucr * rkr + roucl * rkl - damp * rkhat - rl6 * alfa6);
        out.f[NV == 7 ? 6 : 0] = ha * (roucr * epsr + roucl * epsl - damp * epshat - rl7 * alfa7);
    }
    return out;
}

// FLUXDO base3c.f:3677-3869 (IDIFF <= 3) at one face: momentum fluxes and the energy term
__device__ __forceinline__ void visc_face(const Prm &p, const Blk &b, int L, int S,
                                          const double *__restrict__ an,
                                          const double *__restrict__ anx,
                                          const double *__restrict__ any,
                                          const double *__restrict__ anz, double &fm,
                                          double &fn, double &fw, double &fe)
{
    const double *vol = FLD(b, FF_VOL), *eps2 = FLD(b, FF_EPS2), *vis = FLD(b, FF_VIS),
                 *cp = FLD(b, FF_CP), *vist = FLD(b, FF_VIST), *ch = FLD(b, FF_CH),
                 *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W),
                 *t = FLD(b, FF_TEMP);
    int m = L - S;
    double ypd2 = an[L] / (vol[L] + vol[m]);
    double surg = -an[L] * ypd2;
    double surf = surg * (eps2[L] * vis[L] + eps2[m] * vis[m]);
    double surfe = surg * (cp[L] * vist[L] / p.prt + ch[L] + cp[m] * vist[m] / p.prt + ch[m]);
    double duc = .333333 * (anx[L] * (u[L] - u[m]) + any[L] * (v[L] - v[m]) + anz[L] * (w[L] - w[m]));
    double f2 = surf * (u[L] - u[m] + anx[L] * duc);
    double f3 = surf * (v[L] - v[m] + any[L] * duc);
    double f4 = surf * (w[L] - w[m] + anz[L] * duc);
    double f5 = .5 * ((u[L] + u[m]) * f2 + (v[L] + v[m]) * f3 + (w[L] + w[m]) * f4);
    fm = f2; fn = f3; fw = f4;
    fe = f5 + surfe * (t[L] - t[m]);
}

// FLUXDO with the two energy terms it adds to FE returned apart (FE = FE + F5VIG + SURFE*dT)
__device__ __forceinline__ void visc_face2(const Prm &p, const Blk &b, int L, int S,
                                           const double *__restrict__ an,
                                           const double *__restrict__ anx,
                                           const double *__restrict__ any,
                                           const double *__restrict__ anz, double &fm,
                                           double &fn, double &fw, double &f5v, double &fh)
{
    const double *vol = FLD(b, FF_VOL), *eps2 = FLD(b, FF_EPS2), *vis = FLD(b, FF_VIS),
                 *cp = FLD(b, FF_CP), *vist = FLD(b, FF_VIST), *ch = FLD(b, FF_CH),
                 *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W),
                 *t = FLD(b, FF_TEMP);
    int m = L - S;
    double ypd2 = an[L] / (vol[L] + vol[m]);
    double surg = -an[L] * ypd2;
    double surf = surg * (eps2[L] * vis[L] + eps2[m] * vis[m]);
    double surfe = surg * (cp[L] * vist[L] / p.prt + ch[L] + cp[m] * vist[m] / p.prt + ch[m]);
    double duc = .333333 * (anx[L] * (u[L] - u[m]) + any[L] * (v[L] - v[m]) + anz[L] * (w[L] - w[m]));
    double f2 = surf * (u[L] - u[m] + anx[L] * duc);
    double f3 = surf * (v[L] - v[m] + any[L] * duc);
    double f4 = surf * (w[L] - w[m] + anz[L] * duc);
    f5v = .5 * ((u[L] + u[m]) * f2 + (v[L] + v[m]) * f3 + (w[L] + w[m]) * f4);
    fm = f2; fn = f3; fw = f4;
    fh = surfe * (t[L] - t[m]);
}

// FLUXBL base3c.f:4085-4242 (IDIFF <= 3) at one face: the full viscous flux from the
// distance-weighted face average of the cell gradients SIJ, W12 and GRADT, with the Blazek
// correction along the cell-centre connection when ALFA = SXI.n > 0 and the lambda term
// unless IDIFF = 2.  dd is the direction's D1/D2/D3.  fm, fn, fw are what FLUXBL stores
// (-F2VIB..-F4VIB); f5 is the F5VIB it subtracts from FE.
__device__ __forceinline__ void visc_face_bl(const Prm &p, const Blk &b, int L, int S,
                                             const double *__restrict__ dd,
                                             const double *__restrict__ an,
                                             const double *__restrict__ anx,
                                             const double *__restrict__ any,
                                             const double *__restrict__ anz, double &fm,
                                             double &fn, double &fw, double &f5)
{
    const double *eps2 = FLD(b, FF_EPS2), *vis = FLD(b, FF_VIS), *cp = FLD(b, FF_CP),
                 *vist = FLD(b, FF_VIST), *ch = FLD(b, FF_CH), *u = FLD(b, FF_U),
                 *v = FLD(b, FF_V), *w = FLD(b, FF_W), *t = FLD(b, FF_TEMP),
                 *xc = FLD(b, FF_XC), *yc = FLD(b, FF_YC), *zc = FLD(b, FF_ZC),
                 *s11 = FLD(b, FF_S11), *s12 = FLD(b, FF_S12), *s13 = FLD(b, FF_S13),
                 *s22 = FLD(b, FF_S22), *s23 = FLD(b, FF_S23), *s33 = FLD(b, FF_S33),
                 *w12 = FLD(b, FF_W12), *w13 = FLD(b, FF_W13), *w23 = FLD(b, FF_W23),
                 *gx = FLD(b, FF_DTDX), *gy = FLD(b, FF_DTDY), *gz = FLD(b, FF_DTDZ);
    const double r23 = .6666667;
    const int m = L - S;
    double w1 = 2.0 * (dd[L] / (dd[L] + dd[m]));
    double w2 = 2.0 - w1;
    double w1p = .5 * w1, w2p = .5 * w2;
    double dudx = w2p * s11[L] + w1p * s11[m];
    double dvdy = w2p * s22[L] + w1p * s22[m];
    double dwdz = w2p * s33[L] + w1p * s33[m];
    double dudy = w2p * (s12[L] + w12[L]) + w1p * (s12[m] + w12[m]);
    double dudz = w2p * (s13[L] + w13[L]) + w1p * (s13[m] + w13[m]);
    double dvdx = w2p * (s12[L] - w12[L]) + w1p * (s12[m] - w12[m]);
    double dvdz = w2p * (s23[L] + w23[L]) + w1p * (s23[m] + w23[m]);
    double dwdx = w2p * (s13[L] - w13[L]) + w1p * (s13[m] - w13[m]);
    double dwdy = w2p * (s23[L] - w23[L]) + w1p * (s23[m] - w23[m]);
    double s11i = w2 * s11[L] + w1 * s11[m], s12i = w2 * s12[L] + w1 * s12[m];
    double s13i = w2 * s13[L] + w1 * s13[m], s22i = w2 * s22[L] + w1 * s22[m];
    double s23i = w2 * s23[L] + w1 * s23[m], s33i = w2 * s33[L] + w1 * s33[m];
    double dtdxi = w2p * gx[L] + w1p * gx[m];
    double dtdyi = w2p * gy[L] + w1p * gy[m];
    double dtdzi = w2p * gz[L] + w1p * gz[m];
    double sxi = xc[L] - xc[m], syi = yc[L] - yc[m], szi = zc[L] - zc[m];
    double slen = sqrt(sxi * sxi + syi * syi + szi * szi) + 1.E-21;
    sxi = sxi / slen; syi = syi / slen; szi = szi / slen;
    const double nx = anx[L], ny = any[L], nz = anz[L];
    double alfa = sxi * nx + syi * ny + szi * nz;
    double surf = 0.5 * an[L] * (eps2[L] * vis[L] + eps2[m] * vis[m]);
    double surfe = 0.5 * an[L] * (cp[L] * vist[L] / p.prt + ch[L] + cp[m] * vist[m] / p.prt + ch[m]);
    double dudxsj = dudx * sxi + dudy * syi + dudz * szi;
    double dvdxsj = dvdx * sxi + dvdy * syi + dvdz * szi;
    double dwdxsj = dwdx * sxi + dwdy * syi + dwdz * szi;
    double duvwsj = nx * dudxsj + ny * dvdxsj + nz * dwdxsj;
    double adotu = nx * (u[L] - u[m]) + ny * (v[L] - v[m]) + nz * (w[L] - w[m]);
    double f2 = nx * s11i + ny * s12i + nz * s13i;
    double f3 = nx * s12i + ny * s22i + nz * s23i;
    double f4 = nx * s13i + ny * s23i + nz * s33i;
    if (alfa > 0.) {
        double sdotdu = sxi * dudx + syi * dudy + szi * dudz;
        double sdotdv = sxi * dvdx + syi * dvdy + szi * dvdz;
        double sdotdw = sxi * dwdx + syi * dwdy + szi * dwdz;
        f2 = alfa * ((u[L] - u[m]) / slen - sdotdu) + sxi * (adotu / slen - duvwsj) + f2;
        f3 = alfa * ((v[L] - v[m]) / slen - sdotdv) + syi * (adotu / slen - duvwsj) + f3;
        f4 = alfa * ((w[L] - w[m]) / slen - sdotdw) + szi * (adotu / slen - duvwsj) + f4;
    }
    if (p.idiff != 2 && p.idiff != 5) {
        f2 = f2 - r23 * nx * (dudx + dvdy + dwdz);
        f3 = f3 - r23 * ny * (dudx + dvdy + dwdz);
        f4 = f4 - r23 * nz * (dudx + dvdy + dwdz);
    }
    f2 = surf * f2;
    f3 = surf * f3;
    f4 = surf * f4;
    double e5 = .5 * ((u[L] + u[m]) * f2 + (v[L] + v[m]) * f3 + (w[L] + w[m]) * f4)
              + surfe * (dtdxi * nx + dtdyi * ny + dtdzi * nz);
    if (alfa > 0.)
        e5 = surfe * alfa * ((t[L] - t[m]) / slen - sxi * dtdxi - syi * dtdyi - szi * dtdzi) + e5;
    fm = -f2; fn = -f3; fw = -f4;
    f5 = e5;
}

// SST diffusion constants of COEFKO turb3.f:8068-8108 (ITURB = 6)
struct SstC {
    double a1kleb, bstar, ckappa, beta1, sigrk1, sigom1, beta2, sigrk2, sigom2;
};
__device__ __forceinline__ SstC coefko(int kover)
{
    SstC s;
    s.a1kleb = 0.31; s.bstar = 0.09; s.ckappa = 0.41; s.beta1 = 0.075;
    s.sigrk1 = kover < 5 ? 0.85 : 0.5;
    s.sigom1 = 0.5; s.beta2 = 0.0828; s.sigrk2 = 1.0; s.sigom2 = 0.856;
    return s;
}

// FLUDKE base3c.f:5547-5651 at one face: thin-layer k and omega diffusion with the SST
// sigmas blended by FUN1 of the face's upper cell; returns F6VIG, F7VIG
__device__ __forceinline__ void kw_diff_face(const Blk &b, int L, int S,
                                             const double *__restrict__ an, const SstC &s,
                                             double &f6, double &f7)
{
    const double *vol = FLD(b, FF_VOL), *eps2 = FLD(b, FF_EPS2), *vis = FLD(b, FF_VIS),
                 *ro = FLD(b, FF_RO), *rk = FLD(b, FF_RK), *reps = FLD(b, FF_REPS),
                 *fun1 = FLD(b, FF_FUN1);
    int m = L - S;
    double f1 = fun1[L];
    double onemf1 = 0.0 > (1.0 - f1) ? 0.0 : (1.0 - f1);
    double psigk = f1 * s.sigrk1 + onemf1 * s.sigrk2;
    double psige = f1 * s.sigom1 + onemf1 * s.sigom2;
    double ypd2 = an[L] / (vol[L] + vol[m]);
    double surg = -an[L] * ypd2;
    double rmut2 = (eps2[L] - 1.) * vis[L] + (eps2[m] - 1.) * vis[m];
    double surfk = surg * (vis[L] + vis[m] + rmut2 * psigk);
    double surfep = surg * (vis[L] + vis[m] + rmut2 * psige);
    double ypro1 = 1. / ro[L];
    double ypro2 = 1. / ro[m];
    f6 = surfk * (rk[L] * ypro1 - rk[m] * ypro2);
    f7 = surfep * (reps[L] * ypro1 - reps[m] * ypro2);
}

// FLKEBL base3c.f:5727-5830 at one face: k and omega diffusion from the distance-weighted face
// average of the gradients of k/rho and omega/rho, corrected along the cell-centre connection
// with ALFA = MAX(SXI.n, 0); returns FRK and FEPS
__device__ __forceinline__ void kw_diff_face_bl(const Blk &b, int L, int S,
                                                const double *__restrict__ dd,
                                                const double *__restrict__ an,
                                                const double *__restrict__ anx,
                                                const double *__restrict__ any,
                                                const double *__restrict__ anz, const SstC &s,
                                                double &f6, double &f7)
{
    const double *vis = FLD(b, FF_VIS), *vist = FLD(b, FF_VIST), *ro = FLD(b, FF_RO),
                 *rk = FLD(b, FF_RK), *reps = FLD(b, FF_REPS), *fun1 = FLD(b, FF_FUN1),
                 *xc = FLD(b, FF_XC), *yc = FLD(b, FF_YC), *zc = FLD(b, FF_ZC),
                 *kx = FLD(b, FF_DKDX), *ky = FLD(b, FF_DKDY), *kz = FLD(b, FF_DKDZ),
                 *ex = FLD(b, FF_DEDX), *ey = FLD(b, FF_DEDY), *ez = FLD(b, FF_DEDZ);
    const int m = L - S;
    double f1 = fun1[L];
    double onemf1 = 0.0 > (1.0 - f1) ? 0.0 : (1.0 - f1);
    double psigk = f1 * s.sigrk1 + onemf1 * s.sigrk2;
    double psige = f1 * s.sigom1 + onemf1 * s.sigom2;
    double w1 = 2.0 * (dd[L] / (dd[L] + dd[m]));
    double w2 = 2.0 - w1;
    double dkdxi = 0.5 * (w2 * kx[L] + w1 * kx[m]);
    double dkdyi = 0.5 * (w2 * ky[L] + w1 * ky[m]);
    double dkdzi = 0.5 * (w2 * kz[L] + w1 * kz[m]);
    double dedxi = 0.5 * (w2 * ex[L] + w1 * ex[m]);
    double dedyi = 0.5 * (w2 * ey[L] + w1 * ey[m]);
    double dedzi = 0.5 * (w2 * ez[L] + w1 * ez[m]);
    double sxi = xc[L] - xc[m], syi = yc[L] - yc[m], szi = zc[L] - zc[m];
    double slen = sqrt(sxi * sxi + syi * syi + szi * szi) + 1.E-21;
    sxi = sxi / slen; syi = syi / slen; szi = szi / slen;
    const double nx = anx[L], ny = any[L], nz = anz[L];
    double alfa = sxi * nx + syi * ny + szi * nz;
    alfa = dmax(alfa, 0.);
    double rmut2 = vist[L] + vist[m];
    double surfk = 0.5 * an[L] * (vis[L] + vis[m] + rmut2 * psigk);
    double surfep = 0.5 * an[L] * (vis[L] + vis[m] + rmut2 * psige);
    double ypro1 = 1. / ro[L];
    double ypro2 = 1. / ro[m];
    double tx = nx - alfa * sxi, ty = ny - alfa * syi, tz = nz - alfa * szi;
    f6 = -surfk * (alfa * (rk[L] * ypro1 - rk[m] * ypro2) / slen + dkdxi * tx + dkdyi * ty
                   + dkdzi * tz);
    f7 = -surfep * (alfa * (reps[L] * ypro1 - reps[m] * ypro2) / slen + dedxi * tx
                    + dedyi * ty + dedzi * tz);
}

// FLXCYC face flux (ITURB = 6, IFLUX = 1): FLUDKE (FE = FRK = F6VIG, FEPS = F7VIG), SETV12,
// FLUXDO (FE = FE + F5VIG + heat), then FLUXKE added (res3c.f:6817-6946).  FNS: FULLNS on
// level 1, FLKEBL and FLUXBL (FE = FRK - F5VIB) in place of FLUDKE and FLUXDO
// (res3c.f:6802-6848).
template <bool SINGLE, bool FNS = false>
__device__ __forceinline__ Flux7 face_flux_kw(const Prm &p, const SstC &s, const Blk &b, int L,
                                              int S, int intem,
                                              const double *__restrict__ an,
                                              const double *__restrict__ anx,
                                              const double *__restrict__ any,
                                              const double *__restrict__ anz,
                                              const double *__restrict__ dd = nullptr)
{
    double f6, f7, fm, fn, fw, fe;
    if (FNS) {
        double f5;
        kw_diff_face_bl(b, L, S, dd, an, anx, any, anz, s, f6, f7);
        visc_face_bl(p, b, L, S, dd, an, anx, any, anz, fm, fn, fw, f5);
        fe = f6 - f5;
    } else {
        double f5v, fh;
        kw_diff_face(b, L, S, an, s, f6, f7);
        visc_face2(p, b, L, S, an, anx, any, anz, fm, fn, fw, f5v, fh);
        fe = f6 + f5v + fh;
    }
    if (p.istres >= 1) {
        // ANISFLUX earsm.f:278-307 (FLXCYC res3c.f:6933-6937): the EARSM Reynolds stresses
        // added to the momentum and energy fluxes ahead of the inviscid flux
        const int M = L - S;
        const double *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W), *rk = FLD(b, FF_RK);
        const double *b1 = FLD(b, FF_B11), *b2 = FLD(b, FF_B12), *b3 = FLD(b, FF_B13),
                     *b4 = FLD(b, FF_B22), *b5 = FLD(b, FF_B23);
        double ucm = 0.5 * (u[L] + u[M]), vcm = 0.5 * (v[L] + v[M]), wcm = 0.5 * (w[L] + w[M]);
        double b11 = b1[L] + b1[M], b12 = b2[L] + b2[M], b13 = b3[L] + b3[M];
        double b22 = b4[L] + b4[M], b23 = b5[L] + b5[M];
        double b33 = -b1[L] - b4[L] - b1[M] - b4[M];
        double rka = 0.5 * (rk[L] + rk[M]) * an[L];
        double f2ex = rka * (anx[L] * b11 + any[L] * b12 + anz[L] * b13);
        double f3ex = rka * (anx[L] * b12 + any[L] * b22 + anz[L] * b23);
        double f4ex = rka * (anx[L] * b13 + any[L] * b23 + anz[L] * b33);
        double f5ex = ucm * f2ex + vcm * f3ex + wcm * f4ex;
        fm = fm + f2ex;
        fn = fn + f3ex;
        fw = fw + f4ex;
        fe = fe + f5ex;
    }
    Flux7 r = roe_face<SINGLE, 7>(p, b, L, S, intem, an, anx, any, anz);
    Flux7 o;
    o.f[0] = r.f[0] + 0.;
    o.f[1] = r.f[1] + fm;
    o.f[2] = r.f[2] + fn;
    o.f[3] = r.f[3] + fw;
    o.f[4] = r.f[4] + fe;
    o.f[5] = r.f[5] + f6;
    o.f[6] = r.f[6] + f7;
    return o;
}

// FLUXVL base3c.f:1733-1961 at one face: van Leer flux-vector splitting of the MUSCL states
// of RO, RM, RN, RW, E (ideal gas), as the sums FLUCYC adds to the viscous flux.
__device__ __forceinline__ Flux5 vl_face(const Prm &p, const Blk &b, int L, int S, int intem,
                                         const double *__restrict__ an,
                                         const double *__restrict__ anx,
                                         const double *__restrict__ any,
                                         const double *__restrict__ anz)
{
    const double gamma = p.gamma;
    const double ygm = 1. / gamma;
    const double gm1 = gamma - 1.;
    double roip, roim, rmip, rmim, rnip, rnim, rwip, rwim, eip, eim;
    xxtrap(FLD(b, FF_RO), L, S, intem, roip, roim);
    xxtrap(FLD(b, FF_RM), L, S, intem, rmip, rmim);
    xxtrap(FLD(b, FF_RN), L, S, intem, rnip, rnim);
    xxtrap(FLD(b, FF_RW), L, S, intem, rwip, rwim);
    xxtrap(FLD(b, FF_E), L, S, intem, eip, eim);
    // pressures from the unfloored states, then the floors (base3c.f:1810-1818)
    double ppi = (gamma - 1.) * (eip - .5 / roip * (rmip * rmip + rnip * rnip + rwip * rwip));
    double pmi = (gamma - 1.) * (eim - .5 / roim * (rmim * rmim + rnim * rnim + rwim * rwim));
    double fd = 0.001 * p.frsden, fp = 0.001 * p.frspre;
    roip = roip > fd ? roip : fd;
    ppi = ppi > fp ? ppi : fp;
    roim = roim > fd ? roim : fd;
    pmi = pmi > fp ? pmi : fp;
    double a11 = anx[L], a12 = any[L], a13 = anz[L];
    double a21, a22, a23, a31, a32, a33;
    surfvects(a11, a12, a13, a21, a22, a23, a31, a32, a33);
    double cmip = a11 * rmip + a12 * rnip + a13 * rwip;     // contravariant momenta
    double cnip = a21 * rmip + a22 * rnip + a23 * rwip;
    double cwip = a31 * rmip + a32 * rnip + a33 * rwip;
    double cmim = a11 * rmim + a12 * rnim + a13 * rwim;
    double cnim = a21 * rmim + a22 * rnim + a23 * rwim;
    double cwim = a31 * rmim + a32 * rnim + a33 * rwim;
    const double A = an[L];
    double f1m, f2m, f3m, f4m, f5m;                          // right state, base3c.f:1866-1893
    {
        double ypro = 1. / roip;
        double uip = ypro * cmip, vip = ypro * cnip, wip = ypro * cwip;
        double pip = 0.01 > ppi ? 0.01 : ppi;
        double cip = sqrt(gamma * pip * ypro);
        double rmaip = uip / cip;
        if (rmaip < -1.) {
            f1m = A * cmip;
            f2m = A * (cmip * uip + pip);
            f3m = A * cmip * vip;
            f4m = A * cmip * wip;
            f5m = A * uip * (eip + pip);
        } else if (rmaip < 1.) {
            double h = .5 * (rmaip - 1.);
            double f1 = -A * roip * cip * (h * h);
            double f2mi = gm1 * uip - 2. * cip;
            f2m = f1 * f2mi * ygm;
            f3m = f1 * vip;
            f4m = f1 * wip;
            f5m = f1 * (eip + pip) * ypro;
            f1m = f1;
        } else {
            f1m = 0.; f2m = 0.; f3m = 0.; f4m = 0.; f5m = 0.;
        }
    }
    double f1p, f2p, f3p, f4p, f5p;                          // left state, base3c.f:1897-1928
    {
        double ypro = 1. / roim;
        double uim = ypro * cmim, vim = ypro * cnim, wim = ypro * cwim;
        double pim = 0.01 > pmi ? 0.01 : pmi;
        double cim = sqrt(gamma * pim * ypro);
        double rmaim = uim / cim;
        if (rmaim > 1.) {
            f1p = A * cmim;
            f2p = A * (cmim * uim + pim);
            f3p = A * cmim * vim;
            f4p = A * cmim * wim;
            f5p = A * uim * (eim + pim);
        } else if (rmaim > -1.) {
            double h = .5 * (rmaim + 1.);
            double f1 = +A * roim * cim * (h * h);
            double f2pi = gm1 * uim + 2. * cim;
            f2p = f1 * f2pi * ygm;
            f3p = f1 * vim;
            f4p = f1 * wim;
            f5p = f1 * (eim + pim) * ypro;
            f1p = f1;
        } else {
            f1p = 0.; f2p = 0.; f3p = 0.; f4p = 0.; f5p = 0.;
        }
    }
    double f2 = f2p + f2m, f3 = f3p + f3m, f4 = f4p + f4m;   // base3c.f:1934-1950
    Flux5 out;
    out.f[0] = f1m + f1p;
    out.f[1] = a11 * f2 + a21 * f3 + a31 * f4;
    out.f[2] = a12 * f2 + a22 * f3 + a32 * f4;
    out.f[3] = a13 * f2 + a23 * f3 + a33 * f4;
    out.f[4] = f5m + f5p;
    return out;
}

// FLUCYC face flux: FLUXDO (FNS: FLUXBL, FULLNS on level 1, res3c.f:7839), then the
// inviscid flux of the scheme IFLUX selects added (res3c.f:7856, 7892 Roe FLUXKE, 7907 van
// Leer FLUXVL)
template <bool SINGLE, int IFLUX = 1, bool FNS = false>
__device__ __forceinline__ Flux5 face_flux(const Prm &p, const Blk &b, int L, int S, int intem,
                                           const double *__restrict__ an,
                                           const double *__restrict__ anx,
                                           const double *__restrict__ any,
                                           const double *__restrict__ anz,
                                           const double *__restrict__ dd = nullptr)
{
    double fm, fn, fw, fe;
    if (FNS) {
        double f5;
        visc_face_bl(p, b, L, S, dd, an, anx, any, anz, fm, fn, fw, f5);
        fe = 0. - f5;                                   // FE = FE - F5VIB on the zeroed FE
    } else {
        visc_face(p, b, L, S, an, anx, any, anz, fm, fn, fw, fe);
    }
    Flux5 r = IFLUX == 0 ? vl_face(p, b, L, S, intem, an, anx, any, anz)
                         : roe_face<SINGLE>(p, b, L, S, intem, an, anx, any, anz);
    Flux5 o;
    o.f[0] = r.f[0] + 0.;
    o.f[1] = r.f[1] + fm;
    o.f[2] = r.f[2] + fn;
    o.f[3] = r.f[3] + fw;
    o.f[4] = r.f[4] + (0. + fe);
    return o;
}

// SYI imps4.f:10-40
__device__ __forceinline__ void syi(double &w1, double &w2, double &w3, double &w4, double &w5,
                                    double x1, double x2, double x3, double x4, double x5,
                                    double c, double ro, double a2x, double a2y, double a2z,
                                    double &a21, double &a22, double &a23, double &a31,
                                    double &a32, double &a33)
{
    double z1i = x1, z5i = x5;
    surfvects(a2x, a2y, a2z, a21, a22, a23, a31, a32, a33);
    double z2i = a2x * x2 + a2y * x3 + a2z * x4;
    double z3i = a21 * x2 + a22 * x3 + a23 * x4;
    double z4i = a31 * x2 + a32 * x3 + a33 * x4;
    double roc = ro * c;
    w1 = z1i - z5i / ro;
    w2 = roc * z2i + z5i;
    w3 = z3i;
    w4 = z4i;
    w5 = -roc * z2i + z5i;
}

// SYM1I imps4.f:46-70
__device__ __forceinline__ void sym1i(double &x1, double &x2, double &x3, double &x4,
                                      double &x5, double w1, double w2, double w3, double w4,
                                      double w5, double c, double ro, double a2x, double a2y,
                                      double a2z, double a21, double a22, double a23,
                                      double a31, double a32, double a33)
{
    double ypro = 1. / ro;
    double pc2 = .5 * ypro;
    double proc = pc2 / c;
    x1 = w1 + pc2 * (w2 + w5);
    double x2i = proc * (w2 - w5);
    double x3i = w3, x4i = w4;
    x5 = .5 * (w2 + w5);
    x2 = a2x * x2i + a21 * x3i + a31 * x4i;
    x3 = a2y * x2i + a22 * x3i + a32 * x4i;
    x4 = a2z * x2i + a23 * x3i + a33 * x4i;
}

// SYM1I(XI * SYI(dW)) across one face (imps4.f:252-290, 513-552, 316-328)
__device__ __forceinline__ void face_contrib(const double dw[5], double c, double ro,
                                             double nx, double ny, double nz, double x1,
                                             double x2, double x5, double out[5])
{
    double a21, a22, a23, a31, a32, a33, w1, w2, w3, w4, w5;
    syi(w1, w2, w3, w4, w5, dw[0], dw[1], dw[2], dw[3], dw[4], c, ro, nx, ny, nz, a21, a22,
        a23, a31, a32, a33);
    w1 = x1 * w1;
    w2 = x2 * w2;
    w3 = x1 * w3;
    w4 = x1 * w4;
    w5 = x5 * w5;
    sym1i(out[0], out[1], out[2], out[3], out[4], w1, w2, w3, w4, w5, c, ro, nx, ny, nz, a21,
          a22, a23, a31, a32, a33);
}

// A face's unit normal with the tangential basis SURFVECTS builds from it.  It depends on
// the geometry only, so the LU-SGS sweeps form it before the recurrence reaches the cell.
struct FaceBasis { double nx, ny, nz, a21, a22, a23, a31, a32, a33; };

__device__ __forceinline__ FaceBasis face_basis(double nx, double ny, double nz)
{
    FaceBasis f;
    f.nx = nx; f.ny = ny; f.nz = nz;
    surfvects(nx, ny, nz, f.a21, f.a22, f.a23, f.a31, f.a32, f.a33);
    return f;
}

// face_contrib with the basis and the cell scalars RO*C (SYI), .5/RO and (.5/RO)/C (SYM1I)
// formed beforehand; every operation that involves the increments is the same as there.
__device__ __forceinline__ void face_contrib_b(const double dw[5], double ro, double roc,
                                               double pc2, double proc, const FaceBasis &f,
                                               double x1, double x2, double x5, double out[5])
{
    double z1i = dw[0], z5i = dw[4];
    double z2i = f.nx * dw[1] + f.ny * dw[2] + f.nz * dw[3];
    double z3i = f.a21 * dw[1] + f.a22 * dw[2] + f.a23 * dw[3];
    double z4i = f.a31 * dw[1] + f.a32 * dw[2] + f.a33 * dw[3];
    double w1 = z1i - z5i / ro;
    double w2 = roc * z2i + z5i;
    double w3 = z3i;
    double w4 = z4i;
    double w5 = -roc * z2i + z5i;
    w1 = x1 * w1;
    w2 = x2 * w2;
    w3 = x1 * w3;
    w4 = x1 * w4;
    w5 = x5 * w5;
    out[0] = w1 + pc2 * (w2 + w5);
    double x2i = proc * (w2 - w5);
    out[4] = .5 * (w2 + w5);
    out[1] = f.nx * x2i + f.a21 * w3 + f.a31 * w4;
    out[2] = f.ny * x2i + f.a22 * w3 + f.a32 * w4;
    out[3] = f.nz * x2i + f.a23 * w3 + f.a33 * w4;
}

// Split-eigenvalue coefficients of cell n across one face.
//   PREDGS (imps4.f:405-454): face = own lower face n, neighbour n-S, sign -1.
//   CORRGS (imps4.f:150-195): face = own upper face n+S, neighbour n+S, viscous area A(n).
struct Coef { double x1, x2, x5, rmax, xa; };

// The same coefficients from values: (nx,ny,nz) the face normal, an = A(n), the cell's
// and the neighbour's VOL, RO, VIS, VIST.
template <bool PRED>
__device__ __forceinline__ Coef coef_vals(double nx, double ny, double nz, double u, double v,
                                          double w, double c, double an, double vol,
                                          double volb, double ro, double rob, double vis,
                                          double vist, double visb, double vistb,
                                          double cdiffd)
{
    double rl1 = nx * u + ny * v + nz * w - 0.;
    double rl2 = rl1 + c;
    double rl5 = rl1 - c;
    double rnum = (vol + volb) * (ro + rob);
    double t = cdiffd * an / rnum * (vist + vis + vistb + visb);
    double rmax = 1.00 * (fabs(rl1) + c);
    Coef k;
    if (PRED) {
        k.x1 = (rl1 - rmax) * .5 - t;
        k.x2 = (rl2 - rmax) * .5 - t;
        k.x5 = (rl5 - rmax) * .5 - t;
    } else {
        k.x1 = (rl1 + rmax) * .5 + t;
        k.x2 = (rl2 + rmax) * .5 + t;
        k.x5 = (rl5 + rmax) * .5 + t;
    }
    k.rmax = rmax;
    k.xa = rmax + t;
    return k;
}

template <bool PRED>
__device__ __forceinline__ Coef lusgs_coef(const Blk &b, int n, int S, double cdiffd,
                                           const double *__restrict__ an,
                                           const double *__restrict__ anx,
                                           const double *__restrict__ any,
                                           const double *__restrict__ anz)
{
    const double *u = FLD(b, FF_U), *v = FLD(b, FF_V), *w = FLD(b, FF_W), *c = FLD(b, FF_C),
                 *vol = FLD(b, FF_VOL), *ro = FLD(b, FF_RO), *vis = FLD(b, FF_VIS),
                 *vist = FLD(b, FF_VIST);
    int nf = PRED ? n : n + S;          // face whose normal is used
    int nb = PRED ? n - S : n + S;      // neighbour across it
    double rl1 = anx[nf] * u[n] + any[nf] * v[n] + anz[nf] * w[n] - 0.;
    double rl2 = rl1 + c[n];
    double rl5 = rl1 - c[n];
    double rnum = (vol[n] + vol[nb]) * (ro[n] + ro[nb]);
    double t = cdiffd * an[n] / rnum * (vist[n] + vis[n] + vist[nb] + vis[nb]);
    double rmax = 1.00 * (fabs(rl1) + c[n]);
    Coef k;
    if (PRED) {
        k.x1 = (rl1 - rmax) * .5 - t;
        k.x2 = (rl2 - rmax) * .5 - t;
        k.x5 = (rl5 - rmax) * .5 - t;
    } else {
        k.x1 = (rl1 + rmax) * .5 + t;
        k.x2 = (rl2 + rmax) * .5 + t;
        k.x5 = (rl5 + rmax) * .5 + t;
    }
    k.rmax = rmax;
    k.xa = rmax + t;
    return k;
}

}  // namespace ff
