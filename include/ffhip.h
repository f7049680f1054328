/*
 * ffhip.h — C-ABI of the MI355X (gfx950) implementation of FINFLO's per-iteration hot path.
 *
 * The entry points replace FINFLO's per-block, per-level call surface used by CYCLES
 * (ns3c.f:1354-2349).  Every array is the reference's own layout: one flat FP64 array per
 * field, i-fastest, IN=JN=KN=2 ghost layers (charac.f:179), the cell (I,J,K) of a block with
 * IMAX x JMAX x KMAX interior cells at the 0-based offset
 *     (K+1)*ISTRID*JSTRID + (J+1)*ISTRID + (I+1),   ISTRID=IMAX+4, JSTRID=JMAX+4,
 * which is the reference's L = (KN+K-1)*ISTRID*JSTRID + (JN+J-1)*ISTRID + IN + I minus one
 * (res3c.f:127-136, base3c.f:1784-1790).  Face quantities of direction d are stored at the
 * upper cell of the face (face I-1/2 at cell I), as the reference does (res3c.f:7686-7703).
 *
 * No HIP or torch types appear here: streams are passed as void*.
 * Return codes: 0 on success, a negative FF_E* code on failure; ff_last_error() explains.
 */
#ifndef FFHIP_H
#define FFHIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FF_GHOST 2

/* Field slots of a block.  Names follow the reference's MAIN_ARRAYS (charac.f:391-574). */
enum ff_field {
    /* solved (primitive) variables: ns3c.f:1517 FIN outputs */
    FF_PDIFF = 0, FF_TEMP, FF_U, FF_V, FF_W, FF_RK, FF_REPS,
    /* derived by STATIM (state.f:159 PERFGA, base3c.f:489 CONSVA) and TURBDI */
    FF_P, FF_RO, FF_E, FF_C, FF_VIS, FF_CH, FF_CP, FF_DRDP, FF_DRDH,
    FF_RM, FF_RN, FF_RW, FF_EPS2, FF_VIST,
    /* residuals / increments (RESIDU res3c.f:194-198), local time step (TIME3) */
    FF_DRO, FF_DM, FF_DN, FF_DW, FF_DE, FF_DRK, FF_DEPS, FF_DTL,
    /* values at the start of the cycle, read by FIN (res3c.f:1122-1126) */
    FF_TOLD, FF_UOLD, FF_VOLD, FF_WOLD, FF_POLD, FF_RKOLD, FF_EPSOLD,
    /* geometry (gluebm.f AREVC1/VOLINC): volume, face areas and unit normals of the
       i-, j- and k-faces, cell spacings, cell centres, wall distance */
    FF_VOL, FF_A1, FF_A1X, FF_A1Y, FF_A1Z, FF_A2, FF_A2X, FF_A2Y, FF_A2Z,
    FF_A3, FF_A3X, FF_A3Y, FF_A3Z, FF_D1, FF_D2, FF_D3, FF_XC, FF_YC, FF_ZC, FF_DISTW,
    /* k-omega SST work fields (RESIKE res3c.f:3423, TURBEQ res3c.f:5845): the sources
       of rho*k and rho*omega, production, blending function F1, vorticity magnitude */
    FF_SRK, FF_SEPS, FF_PTUR, FF_FUN1, FF_OHMI,
    /* velocity gradients of DERFUN (turb3.f:2476): vorticity W12 W13 W23, strain-rate
       tensor S11 S12 S13 S22 S23 S33 (the reference's W12(:,1..3), SIJ(:,1..6)), strain
       magnitude STRAIN, turbulence time scale TTS (TTIMES earsm.f:629) */
    FF_W12, FF_W13, FF_W23, FF_S11, FF_S12, FF_S13, FF_S22, FF_S23, FF_S33,
    FF_STRAIN, FF_TTS,
    /* EARSM extra anisotropy BIJ(:,1..6) (earsm.f:209-219) */
    FF_B11, FF_B12, FF_B13, FF_B22, FF_B23, FF_B33,
    /* multigrid forcing of a coarse level, the reference's ROP2H..EP2H, RKP2H, EPSP2H at
       IH(M,N) (TRANSF multi.f:245-270, FORCFU multi.f:1630-1648, ADDRES multi.f:1709-1725) */
    FF_ROP2H, FF_RMP2H, FF_RNP2H, FF_RWP2H, FF_EP2H, FF_RKP2H, FF_EPSP2H,
    /* cell gradients of TEMDER (turb3.f:3456) for the full viscous fluxes (FULLNS): of T
       (GRADT, RESIDU res3c.f:257), of k/rho (GRADK) and omega/rho (GREPS, RESIKE
       res3c.f:3991-4005) */
    FF_DTDX, FF_DTDY, FF_DTDZ, FF_DKDX, FF_DKDY, FF_DKDZ, FF_DEDX, FF_DEDY, FF_DEDZ,
    FF_NFIELDS
};

/* Physical and numerical parameters of a block (BLKS(NGL) and NS3CO scalars). */
typedef struct ff_params {
    double gamma;    /* GAMMA, main.f:5306 default 1.4 */
    double rgas;     /* RGAS 287.05287 */
    double pr;       /* laminar Prandtl number PR 0.72 */
    double prt;      /* turbulent Prandtl number PRT 0.9 */
    double visu0;    /* Sutherland VISU0 1.458e-6 (main.f:5315) */
    double expsu;    /* Sutherland exponent EXPSU 1.5 */
    double tsu0;     /* Sutherland TSU0 110.4 */
    double frspre;   /* free-stream pressure FRSPRE */
    double frsden;   /* free-stream density FRSDEN */
    double frstem;   /* free-stream temperature FRSTEM */
    double frsvel;   /* free-stream speed FRSVEL */
    double frsvis;   /* free-stream viscosity FRSVIS */
    double cfl;      /* CFL on the finest level */
    double cfll;     /* CFLL on the coarse levels */
    double theta;    /* LU-SGS time-step factor THETA, hard-coded 1.5 (main.f:3499) */
    double cdiff;    /* TIME3 viscous divisor: 1/CDIFF of the input (main.f:4886), 0.5 */
    double tuplim;   /* upper temperature limit TUPLIM (6*FRSTEM, main.f:3828) */
    double tlolim;   /* lower temperature limit TLOLIM (0.05*FRSTEM, main.f:3827) */
    double rklim;    /* RKLIM floor of rho*k */
    double epslim;   /* EPSLIM floor of rho*omega */
    double turlim;   /* TURLIM, upper limit of mu_t/mu */
    double frsmut;   /* FRSMUT, lower limit of mu_t */
    double rmult;    /* RMULTV viscosity multiplier, 1 */
    double re;       /* Reynolds number RE of the reference length (FUN1KO's wall limit) */
    double csimps;   /* CSIMPS of the k-omega point-implicit sources (main.f:5346), 5 */
    double cmgk;     /* CMGK, damping of the prolonged rho*k correction (main.f:5344), 0.1 */
    double cmgeps;   /* CMGEPS, the same for rho*omega (main.f:5345), 0.2 */
    double reserved[3];
    int iturb;       /* 0 laminar, 6 k-omega SST */
    int iflux;       /* 1 Roe (FLUXKE), 0 van Leer (FLUXVL) */
    int idiff;       /* FLUXDO averaging, 1 = arithmetic (main.f:5754) */
    int inter[3];    /* MUSCL switch per direction (INTERI/J/K, main.f:5648-5650), -3 */
    int kover;       /* KOVER of the SST variant */
    int katol;       /* Kato-Launder production switch */
    int istres;      /* ISTRES: 1 = Wallin-Johansson EARSM on top of the SST (EARSM1, ANISFLUX,
                        PEEKOO), 0 = eddy-viscosity SST */
    int fullns;      /* FULLNS: 1 = full viscous fluxes from cell gradients on level 1 (FLUXBL,
                        FLKEBL with TEMDER), 0 = the thin-layer FLUXDO/FLUDKE */
    int turcor;      /* TURCOR (main.f:5257, default YES): the near-wall blending and flux
                        corrections of BOTSKE with the SST; the laminar FLUCYC passes .FALSE. */
    int ddadi;       /* 1: the implicit stage is DDADI (BLKS%LUSGS = 0, imps3.f:321-373): a K,
                        an I and a J sweep of PRED/DIAG/CORR(KE); 0 (default): LU-SGS
                        (LUSGS = 1, PREDGS/DIAGGS/CORRGS) */
    int reserved_i[4];
} ff_params;

/* ---- solid-wall patches (BOTSKE base3c.f:6147-7095) ---------------------------------- */
/* Constants BOTSKE reads besides ff_params (NS3CO, main.f:5367-5369) */
typedef struct ff_wall_params {
    double t0;       /* T0: wall temperature of IHEAT = 0 and 2, and the TEMPB floor 0.01*T0 */
    double difpre;   /* DIFPRE: the pressure subtracted in the force sums (YPG) */
    double xmom, ymom, zmom;   /* XMOM, YMOM, ZMOM: moment reference point */
    double gx, gy, gz;         /* GX, GY, GZ: gravity of the hydrostatic force part */
    double ground, altitude;   /* GROUND, ALTITUDE: ALTITUDE >= GROUND switches gravity off */
    double reserved[6];
} ff_wall_params;

/* One solid-wall patch of a block level (ICON(1,L) = 8..10 or 15), in the reference's terms:
 * face ICON(3,L) (1, 2, 3 = the I-, J-, K-low face; 4, 5, 6 = the high faces) and the cell
 * ranges ICON(4..7,L) on it (KX1..KX2, KY1..KY2: J and K on an I-face, I and K on a J-face,
 * I and J on a K-face).  The patch arrays are the reference's IHF(ICON(2,L)) slabs with two
 * ghost rows each side, (NA+4) x (NB+4) doubles, fastest along J on an I-face, K on a J-face
 * (FLXCYC swaps KX and KY there, res3c.f:7090-7098) and I on a K-face. */
typedef struct ff_wall_patch {
    int iface;                  /* ICON(3,L) */
    int kx1, kx2, ky1, ky2;     /* ICON(4..7,L) */
    int iheat;                  /* ICON(20,L): 0 adiabatic, 1 TWALL given, 2 HFLUX given,
                                   3 wall temperature extrapolated */
    int reserved[2];
    /* inputs */
    double *uwall, *vwall, *wwall;   /* wall velocity */
    double *twall;                   /* wall temperature (IHEAT = 1; written otherwise) */
    double *hflux;                   /* wall heat flux (IHEAT = 2) */
    double *rbk;                     /* roughness of the k+ wall condition of omega */
    double *xcp, *ycp, *zcp;         /* face centres of the moment arms */
    /* outputs */
    double *cpwall, *qwall, *qwfric, *tauw1, *tauw2, *tauwx, *tauwy, *tauwz;
    double *surfx, *surfy, *surfz, *surfpx, *surfpy, *surfpz;
} ff_wall_patch;

/* The force sums of one patch (CX(L) .. of FLXCYC, zeroed per level-1 RESIDI ns3c.f:2376-2396) */
enum ff_force {
    FF_CX, FF_CY, FF_CZ, FF_CMX, FF_CMY, FF_CMZ, FF_TOM, FF_DX, FF_DY, FF_DZ,
    FF_QX, FF_QY, FF_QZ, FF_VMX, FF_VMY, FF_VMZ, FF_BUX, FF_BUY, FF_BUZ, FF_NFORCE
};

/* One (level M, block N) slab: pointers to the first element of each field, i.e. the
 * reference's X(IG(M,N)).  Unused slots may be NULL. */
typedef struct ff_block {
    int imax, jmax, kmax, reserved;
    double *f[FF_NFIELDS];
} ff_block;

typedef struct ff_ctx ff_ctx;

/* Error codes */
#define FF_OK          0
#define FF_E_ARG      -1   /* bad argument (NULL pointer, out-of-range index, size) */
#define FF_E_HIP      -2   /* a HIP runtime call failed */
#define FF_E_NOMEM    -3   /* device allocation failed */
#define FF_E_STATE    -4   /* call sequence error (block not registered, ...) */
#define FF_E_UNSUP    -5   /* configuration outside the implemented path */

/* ---- context ------------------------------------------------------------------------ */
/* Create a context on the current HIP device.  One stream, one device arena. */
int  ff_ctx_create(ff_ctx **ctx, const ff_params *prm);
int  ff_ctx_destroy(ff_ctx *ctx);
/* Use the caller's HIP stream (a hipStream_t passed as void*); NULL = the context's own. */
int  ff_ctx_set_stream(ff_ctx *ctx, void *stream);
const char *ff_last_error(const ff_ctx *ctx);
/* Context options (no counterpart in the reference, which has one form of each stage):
 *   FF_OPT_LUSGS: FF_LUSGS_STAGED (default, below); FF_LUSGS_TILES runs each LU-SGS sweep
 *   as one persistent launch of (J,K) tiles marching along I that form their coefficients
 *   on the way; FF_LUSGS_KMARCH as one persistent launch of (I,J) tiles marching along K;
 *   FF_LUSGS_PLANES launches one kernel per I+J+K hyperplane (laminar only).  All give the
 *   same results bit for bit. */
#define FF_OPT_LUSGS     1
#define FF_LUSGS_TILES   0
#define FF_LUSGS_PLANES  1
#define FF_LUSGS_KMARCH  2
/*   FF_LUSGS_STAGED: the (J,K) tiles of FF_LUSGS_TILES on cell records that a parallel pass
 *   forms beforehand in the sweeps' step order (coalesced, the recurrence-independent work
 *   off the wavefront), four waves per tile (three form one face's outgoing contribution
 *   each, one streams the records into LDS);
 *   needs 2 x 56 + 8 doubles per cell and step of device memory per level (about 1 KB per
 *   cell), allocated on the block's first implicit stage.  FF_LUSGS_STAGED1: the same with
 *   one wave per tile. */
#define FF_LUSGS_STAGED  3
#define FF_LUSGS_STAGED1 4
/*   FF_OPT_RESIDUAL: FF_RESIDUAL_TILES (default) forms every face flux once in (I,J) tiles
 *   marching along K; FF_RESIDUAL_CELLS forms both faces of every direction per cell.
 *   Both give the same results bit for bit. */
#define FF_OPT_RESIDUAL    2
#define FF_RESIDUAL_TILES  0
#define FF_RESIDUAL_CELLS  1
/*   FF_OPT_EARSM: FF_EARSM_EXACT (default) evaluates EARSM1 as the reference does;
 *   FF_EARSM_MLP takes its coefficients from the learned closure (ff_mlp_*, below) in the
 *   cells whose invariants lie in the training box and the exact EARSM1 everywhere else. */
#define FF_OPT_EARSM       3
#define FF_EARSM_EXACT     0
#define FF_EARSM_MLP       1
/*   FF_OPT_PRECISION: FF_PREC_FP64 (default) computes and stores everything in FP64, as the
 *   reference does; FF_PREC_MIXED stores the staged LU-SGS records -- the DIAGGS diagonal,
 *   the split-eigenvalue coefficients, the face bases and the SYI/SYM1I scalars, i.e. the
 *   implicit operator -- in FP32 (half the record traffic), while the residual (fluxes,
 *   sources), its RESL2 norms, the increments of the sweeps, the FIN update and the
 *   multigrid transfers and corrections stay FP64.  The implicit operator only steers the
 *   iteration, so the converged state is the FP64 one; iterates differ from it at the
 *   FP32 rounding of the operator.  Takes effect with FF_LUSGS_STAGED. */
#define FF_OPT_PRECISION   4
#define FF_PREC_FP64       0
#define FF_PREC_MIXED      1
/*   FF_OPT_AUTOTUNE: 1 lets the library pick the tile shape of the residual kernels (the
 *   k-omega RESIKE kernel and the laminar Roe RESIDU kernel) per block size: on the first
 *   residual of a block of new dimensions outside a graph capture it times each of eight
 *   (I, J, K) tile shapes and keeps the fastest for every later residual of such a block.
 *   All shapes give the same results bit for bit.  0 (default): the built-in shapes. */
#define FF_OPT_AUTOTUNE    5
/*   FF_OPT_TILE_SHAPE: 0-7 runs every residual in that candidate shape of the autotuner
 *   (for tests and measurements), -1 (default) leaves the choice to FF_OPT_AUTOTUNE. */
#define FF_OPT_TILE_SHAPE  6
/*   FF_OPT_BLOCK_STREAMS: 1 makes ff_vcycle_blocks run the blocks' V-cycles of a cycle side
 *   by side, block n on the n mod 8-th stream of a pool forked from the context's stream and
 *   joined back before the level-1 exchange, so that the small blocks' launches and sweep
 *   wavefronts overlap instead of leaving most of the chip idle.  The blocks of a cycle are
 *   independent (each reads only its own ghost cells, which the exchange refreshes after
 *   the cycle), so the results are those of 0 (default, one block after the other) bit for
 *   bit; the RESL2 sums are read back once after the last cycle. */
#define FF_OPT_BLOCK_STREAMS 7
int  ff_ctx_set_option(ff_ctx *ctx, int option, int value);
/* The tile shape (I, J, K) the autotuner chose for block N on level M and its measured mean
 * launch time in ms (ms may be NULL); FF_E_STATE if that block size has not been tuned. */
int  ff_tuned_shape(const ff_ctx *ctx, int level, int block, int *shape, double *ms);
/* Per-kernel event timing (0 off, 1 on); results via ff_kernel_stats. */
int  ff_ctx_profile(ff_ctx *ctx, int enable);
/* Copy the accumulated timing of kernel `name`: launches and total milliseconds.  `name`
 * is a kernel name (all levels) or "name@M" for the launches on multigrid level M. */
int  ff_kernel_stats(const ff_ctx *ctx, const char *name, long long *launches, double *ms);
int  ff_kernel_stats_reset(ff_ctx *ctx);

/* ---- blocks ------------------------------------------------------------------------- */
/* Register host arrays of block N on level M (the reference's ALLOCA'd slabs).  The
 * library allocates device mirrors; ff_sync moves data between them. */
int  ff_block_register(ff_ctx *ctx, int level, int block, const ff_block *host);
/* Register device-resident arrays (caller-owned device memory, no mirror). */
int  ff_block_register_device(ff_ctx *ctx, int level, int block, const ff_block *dev);
/* direction 0: host -> device, 1: device -> host, for the nfields field slots listed in
 * fields (enum ff_field values); fields == NULL or nfields < 0 moves every registered field. */
int  ff_sync(ff_ctx *ctx, int level, int block, int direction, const int *fields, int nfields);
/* Synchronise the context's stream. */
int  ff_finish(ff_ctx *ctx);

/* ---- the hot path (each call works on the device copy of block N, level M) ---------- */
/* STATIM (ns3c.f:10318): P = PDIFF + FRSPRE (PDTOP state.f:1690), ideal-gas EOS with
 * Sutherland viscosity (PERFGA state.f:159), sound speed and conservative variables
 * (CONSVA base3c.f:489), on every cell including ghosts. */
int  ff_state(ff_ctx *ctx, int level, int block);
/* TURBDI (ns3c.f:2874).  ITURB=0: EPS2 = 1, VIST = 0 (ADJUIN ns3c.f:2987-2991).  ITURB=6:
 * TURBEQ (res3c.f:5845) = DERFUN velocity gradients (turb3.f:2476), TTIMES (earsm.f:629),
 * with ISTRES=1 the EARSM anisotropy BIJ (EARSM1 earsm.f:1), and the SST eddy viscosity
 * EPS2KO (turb3.f:7454). */
int  ff_turb(ff_ctx *ctx, int level, int block);
/* RESIDI (ns3c.f:2355) -> RESIDU (res3c.f:6): viscous FLUXDO plus the inviscid flux of
 * params.iflux (1: Roe FLUXKE base3c.f:4753; 0: van Leer FLUXVL base3c.f:1733) in the k, i,
 * j directions, SWEEP differences, division by VOL.  ITURB=6: RESIKE (res3c.f:3423) with
 * CROSKO, FUN1KO, PRODKE (PEEKOO and the ANISFLUX stress fluxes with ISTRES=1), FLXCYC's
 * FLUDKE/FLUXDO/FLUXKE, SWEEPK, SOURKO and ADDSKE.  With level == 1 the L2
 * sums of RESL2 (res3c.f:8920) are returned in resi[0..6] (DRO, DM, DN, DW, DE, DRK, DEPS
 * squared sums; the caller takes sqrt()/N as SUBIN does); resi may be NULL. */
int  ff_residual(ff_ctx *ctx, int level, int block, int m2, double *resi);
/* IMPSDI (ns3c.f:2683) -> IMPS (imps3.f:6) with LUSGS=1: TIME3, LOCDT, COPR, the
 * backward PREDGS, DIAGGS and forward CORRGS sweeps, DRO/CP.  ITURB=6: LOCDTK, IMSOKO,
 * COPRKE, PRKEGS/DIKEGS/COKEGS and D(rho k), D(rho omega) times RO. */
int  ff_implicit(ff_ctx *ctx, int level, int block, int m2);
/* FIN (res3c.f:1077) + the limiters of CYCLES (ns3c.f:1553-1596): UPLIM/LOLIM of TEMP,
 * U, V, W, PDTOP, LOLIMP. */
int  ff_update(ff_ctx *ctx, int level, int block);
/* Copy the solved variables into their *OLD slots (the start-of-cycle values FIN reads). */
int  ff_save_old(ff_ctx *ctx, int level, int block);
/* Multigrid transfers between level M and M+1 of block N; the coarse level has half the cells
 * of the finer one in each direction (a planar block, KMAX = 1, keeps KMAX = 1 on every level
 * and takes TRANSF's KIND = 1 path, multi.f:185-191, which counts each of the four fine cells
 * twice) and carries the FAS forcing in its *P2H slots.
 *   ff_restrict:  TRANSF multi.f:156 (+ TRANKE multi.f:434 for k-omega) from Mfine to Mfine+1:
 *                 state, EPS2, VIST and the residual (into the forcing slots), OLD copies.
 *   ff_forcing:   phase 0 FORCFU multi.f:1609 (P = P - D), phase 1 ADDRES multi.f:1686
 *                 (D = P + D), level M.
 *   ff_prolong:   INTERP multi.f:1763 (+ INTERT multi.f:2051) from Mcoarse to Mcoarse-1 with the
 *                 limiters CYCLES applies after it (ns3c.f:2073-2118).
 * ff_update on a level M >= 2 applies the coarse-level limiters (ns3c.f:1936-1985) and
 * ff_residual with M2 >= 2 forms the k-omega velocity gradients itself, as RESIKE does. */
int  ff_restrict(ff_ctx *ctx, int mfine, int block);
int  ff_forcing(ff_ctx *ctx, int level, int block, int phase);
int  ff_prolong(ff_ctx *ctx, int mcoarse, int block);
/* ncycles device-resident CYCLES iterations (ns3c.f:1400-2137) of block N with nlev
 * multigrid levels 1..nlev registered: the level-1 stages, the pre-smoothing-only FAS
 * down-loop and the prolongation.  resi receives the RESL2 sums of the last level-1
 * residual (the one evaluated after the update when nlev >= 2). */
int  ff_vcycle(ff_ctx *ctx, int block, int nlev, int ncycles, double *resi);
/* ff_vcycle without residual sums, as one HIP graph: the first call captures one V-cycle
 * of the context's stream (every kernel of ff_vcycle, same launches, same results bit for
 * bit) and each call replays it ncycles times; nlev = 1 is the iteration of ff_cycle.  A registration or option change makes the
 * next call capture again.  The force pass of wall patches stays on its side stream inside
 * the graph (forked and joined through events).  Not with per-kernel timing on, or on the
 * null stream (FF_E_STATE).  No counterpart in the reference: CYCLES
 * (ns3c.f:1400-2137) launches nothing; this replaces the per-stage launches of ff_vcycle. */
int  ff_vcycle_graph(ff_ctx *ctx, int block, int nlev, int ncycles);
/* ncycles device-resident iterations of the finest-level sequence of CYCLES for block N:
 * save_old, STATIM, TURBDI, RESIDI, IMPSDI, FIN + limits.  Ghost cells are left as given
 * (the caller's boundary routines own them).  resi (may be NULL) receives the RESL2 sums
 * of the last cycle. */
int  ff_cycle(ff_ctx *ctx, int block, int ncycles, double *resi);

/* ---- boundary patches: REFLEC (turb3.f:5294-5878) and GHOSTEXT (ns3c.f:10979-11226) ---- */
/* One boundary patch of a block level, in the reference's ICON terms. */
typedef struct ff_bc_patch {
    int ibc;        /* ICON(1,L): 2 free stream (EXT: GHOSTEXT), 3 inlet (INL: INLET), 5
                       outlet (OUT: OUTLET), 7 symmetry plane or singularity, 8-10 solid
                       wall (REFLEC) */
    int iface;      /* ICON(3,L): 1-3 the I-, J-, K-low face, 4-6 the high faces */
    int kx1, kx2, ky1, ky2;   /* ICON(4..7,L): J and K on an I-face, I and K on a J-face,
                                 I and J on a K-face */
    int iheat;      /* ICON(20,L): 1 takes TSURF = tsurf for the ghost temperature */
    int viscous;    /* IBC 8-10: 1 takes REFLEC's viscous-surface branch, which the
                       reference selects with (IDI /= 0 .AND. IFACE <= 3) .OR. (IDI > K
                       .AND. IFACE >= 4) from the block's IDI1-3 (turb3.f:5540-5541); 0 the
                       inviscid one */
    double tsurf;   /* RCON(5,L): the wall temperature of IHEAT = 1 */
    double reserved[3];
    const double *uwall, *vwall, *wwall;   /* viscous walls: the wall velocity per face,
                       (NA+4) x (NB+4) as the patch arrays of ff_wall_patch (host arrays,
                       copied at registration); NULL: a wall at rest */
    /* INL and OUT patches (IBC 3 and 5) */
    int inlout;     /* ICON(8,L), the INLOUT type (11-20 and 21-30 reduce by 10 and 20, as
                       FLXCYC does, res3c.f:6647-6648).  Inlets (INLET base3c.f:9223-9589):
                       4 subsonic (velocity, density and temperature given, pressure
                       extrapolated), 5 supersonic (everything given).  Outlets (OUTLET
                       base3c.f:9595-10216): 2 and 3 pressure given, velocity extrapolated
                       on outflow and on inflow relaxed towards the given direction (0.7 of
                       the old ghost value with 2, none with 3); 4 every variable
                       extrapolated; 5 the x-momentum given (the face value BOUNR BOUNU),
                       the rest extrapolated */
    int reserved_i[3];
    /* the patch arrays BOUNR, BOUNU, BOUNV, BOUNW, BOUNT, BOUNP, BOUNPD, BOUNRK, BOUNEP,
       (NA+4) x (NB+4) as uwall (host arrays, copied at registration).  Each INLOUT type
       needs the ones it reads: inlet 4 R U V W T, inlet 5 R U V W T P PD, outlets 2 and 3
       R U V W T P PD, outlet 4 none, outlet 5 R U P PD; inlets with the k-omega model RK
       EP.  The others may be NULL.  INLET 4 writes BOUNP and BOUNPD (ff_bc_read). */
    const double *bounr, *bounu, *bounv, *bounw, *bount, *bounp, *bounpd, *bounrk, *bounep;
} ff_bc_patch;
/* Register (or replace, or with n = 0 remove) the boundary patches of block N on level M.
 * From then on ff_state ends with REFLEC over its IBC 7-10 patches, as every STATIM of
 * CYCLES does (STATIM(M,N,INN,.TRUE.), ns3c.f:1450, 1622, 1848 -> ns3c.f:10497-10512):
 * the ghost temperature, VIS, CH, C, DRDP, DRDH (loop 1200), and on viscous walls the
 * reflected velocities with the wall velocity, the extrapolated RO, P, PDIFF, EPS2 = 1,
 * VIST = 0, RM, RN, RW, E and the k-omega RK = -RK, REPS = MAX(0, 2 REPS - REPS') (loop
 * 500); on symmetry planes and inviscid walls the velocities mirrored in the face normal
 * (loop 1150; as in the reference a K-face takes the I-face normal there, turb3.f:5687-
 * 5707).  REFLECL = .FALSE., a perfect gas (ISTATE /= 10), no DDEPS. */
int  ff_bc_register(ff_ctx *ctx, int level, int block, const ff_bc_patch *p, int n);
/* With IBC 3 and 5 patches registered on level 1, every level-1 residual (ff_residual,
 * and the ones inside ff_cycle, ff_vcycle, the graphs and the multi-block cycles) first
 * updates their ghost cells by INLET/OUTLET, as FLXCYC and FLUCYC do at M = 1 before a
 * direction's fluxes (res3c.f:6639-6735, 7724-7830): after the gradients, CROSKO and
 * FUN1KO, before the face fluxes.  Single phase, ISTATE = 1, a grid at rest (WROT = 0), no
 * scalars, no transition model; not with the EARSM (ISTRES = 1: FF_E_UNSUP).  The laminar
 * residual passes TURLIM = 1e-6 as FLUCYC does (res3c.f:7750), the k-omega one
 * ff_params.turlim.  INLET's ICYCLE = 1 branch (the ghost density from BOUNR) is taken
 * while the block's cycle count is 1: registration sets it to 0, and the cycles above add
 * one at their start (ff_bc_cycle sets it, e.g. for a restart). */
/* INLET/OUTLET alone over the registered IBC 3 and 5 patches of block N on level M, with
 * the block's current cycle count (FLXCYC's update, res3c.f:6639-6735). */
int  ff_inlout(ff_ctx *ctx, int level, int block);
/* Set block N's cycle count ICYCLE on level M (0 after ff_bc_register); a cycle adds one at
 * its start, so the next cycle runs as ICYCLE = icycle + 1.  No counterpart in the
 * reference: CYCLES keeps ICYCLE itself (ns3c.f:1400). */
int  ff_bc_cycle(ff_ctx *ctx, int level, int block, int icycle);
/* Copy patch `patch` (0-based, the order of ff_bc_register) of block N on level M back:
 * its device BOUNP and BOUNPD ((NA+4) x (NB+4) each; INLET 4 writes the extrapolated
 * pressure there, base3c.f:9404-9405).  Either pointer may be NULL. */
int  ff_bc_read(ff_ctx *ctx, int level, int block, int patch, double *bounp, double *bounpd);
/* REFLEC alone over the registered IBC 7-10 patches (CYCLES' own call, ns3c.f:2168). */
int  ff_reflec(ff_ctx *ctx, int level, int block);
/* GHOSTEXT over the registered IBC 2 patches: both ghost layers take the free stream at
 * flow angles alpha, beta (radians; U = FRSVEL cos(alpha) cos(beta), V = FRSVEL sin(alpha),
 * W = FRSVEL cos(alpha) sin(beta)), RK = RKLIM, REPS = EPSLIM, VIST = FRSMUT with the
 * k-omega model; no atmosphere model (ALTITUDE < GROUND).  The reference calls it once at
 * start-up (ns3c.f:294); the ghost cells then stay as set. */
int  ff_ghostext(ff_ctx *ctx, int level, int block, double alpha, double beta);

/* ---- one-to-one block connections (CON patches, ICON(1,L) = 1) ------------------------ */
/* A connection patch of block N: the patch on this block's face, the connected block and
 * face, and how the connected face's in-plane indices follow this face's.  The in-plane
 * indices (a, b) of a face are (J, K) on an I-face, (I, K) on a J-face and (I, J) on a
 * K-face, as in ICON(4..7).  Cell (ka1 + p, kb1 + q) of the patch faces the connected
 * face's cell (na0 + da*p, nb0 + db*q), or (na0 + da*q, nb0 + db*p) with swap = 1; that
 * covers the eight orientations ICON(14,L) and PATCHE (gluebm.f:12850-13087) describe.  Both
 * blocks must be registered on the same level; each block lists its own side. */
typedef struct ff_con_patch {
    int iface, ka1, ka2, kb1, kb2;   /* ICON(3..7,L) */
    int nblock;                      /* ICON(8,L): the connected block */
    int nface;                       /* ICON(9,L): its face */
    int na0, nb0;                    /* its in-plane indices of the patch's first cell */
    int swap, da, db;                /* the orientation (da, db = +1 or -1) */
    int reserved[4];
} ff_con_patch;
/* Register (or replace, n = 0 removes) the connection patches of block N on level M. */
int  ff_con_register(ff_ctx *ctx, int level, int block, const ff_con_patch *p, int n);
/* The halo exchange of BOUNDI's connections (EXPTOT/IMPVO2 gluebm.f:6819-7262 with the
 * conversions of CONNEC ns3c.f:8605-8857) on the device, for every connection registered on
 * level M: both ghost layers behind each connected face take the connected block's two cell
 * layers next to its face.  what = 0: the flow fields (PDIFF TEMP U V W RK REPS, P RO E C
 * VIS CH CP DRDP DRDH RM RN RW EPS2 VIST); what = 1: the cell geometry (VOL, XC YC ZC,
 * DISTW, D1 D2 D3), once after registration, as IMPVOL does for the volumes (gluebm.f:6166).
 * The values are copied, not recomputed from exchanged conserved variables.  The level's
 * connections go in one launch (a table built, outside graph capture, after registrations
 * change; the first exchange after a change allocates it), or one launch per connection
 * when two patches of a block name the same ghost cell. */
int  ff_halo_exchange(ff_ctx *ctx, int level, int what);
/* ncycles multi-block cycles: per cycle every listed block's V-cycle (levels 1..nlev, as
 * ff_vcycle), then the level-1 halo exchange, as CYCLES' block loop and BOUNDI(1) do
 * (ns3c.f:1400-2195, 556); resi receives the RESL2 sums of the last cycle summed over the
 * blocks (SUBIN's sums, ns3c.f:10242-10248). */
int  ff_vcycle_blocks(ff_ctx *ctx, const int *blocks, int nblocks, int nlev, int ncycles,
                      double *resi);

/* ff_vcycle_blocks without residual sums, as one HIP graph: the first call captures one
 * multi-block cycle with every block's V-cycle on a stream of a pool of eight (independent
 * branches the device runs side by side) and the halo exchange after them; each call
 * replays it ncycles times, bit for bit the eager cycles.  A registration or option change
 * makes the next call capture again.  No counterpart in the reference (CYCLES launches
 * nothing).  Not with per-kernel timing or on the null stream (FF_E_STATE); at most 9
 * block levels per branch, i.e. ceil(nblocks/8) * nlev <= 9 (FF_E_UNSUP beyond: a 64-block,
 * 4-level capture crashed the process and its cause is not known). */
int  ff_vcycle_blocks_graph(ff_ctx *ctx, const int *blocks, int nblocks, int nlev, int ncycles);

/* SUBIN's convergence monitor for block N (ns3c.f:9877-9905): the change of the cycle that
 * IDRXX selects, over every entry of the level-1 slab including the ghost layers (SUBIN
 * ns3c.f:9836-9844: 2 DM = RO*(U-UOLD), 3 DN, 4 DW, 5 TEMP-TOLD, 6 DE = PDIFF-POLD, the
 * default, main.f:5280; 8 DEPS = REPS-EPSOLD), scanned as DROMXX (res3c.f:9293-9324) does:
 * *dromax, *ixerr (1-based slab entry) and *inerr (block) are the running maximum, its
 * entry and its block in and out, as SUBIN carries them across its block loop (DROMAX =
 * 1E-20, IXERR = 1 before the first block, ns3c.f:9799-9800).  On return *dromax is the
 * signed, ghost-weighted change when this block holds the maximum.  Reads the state after
 * ff_update and the *OLD copies of ff_save_old; writes no field.  IDRXX 1 and 7 (which need
 * ROLD and SUBIN's k limiter) are FF_E_UNSUP. */
int  ff_dromax(ff_ctx *ctx, int block, int idrxx, double *dromax, int *ixerr, int *inerr);

/* ff_residual of a block given directly as caller-owned device arrays (the field pointers
 * of `dev` are device addresses), with no registration and no device allocation: every
 * scratch buffer (the RESL2 partials and sums, the k-omega SEPS maxima) lives in a fixed
 * device array of the library, 36,680 doubles (about 287 KB) that each GPU the library
 * runs on holds until the process ends.  No counterpart in the reference (RESIDI
 * ns3c.f:2355 is the computation); it is for callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * The context comes from ff_ctx_create on the current device (which looks the array up).
 * Limits: no wall patches; with ITURB = 6 at most 8,388,608 entries per field including the
 * ghost layers (FF_E_ARG beyond); per-kernel timing off (FF_E_STATE).  Every call on one GPU
 * shares that GPU's array and leaves its contents behind (each call rewrites what it
 * reads): calls that overlap on one GPU, from other streams or contexts, race. */
int  ff_residual_noalloc(ff_ctx *ctx, int level, const ff_block *dev, int m2, double *resi);

/* Register (or replace) the solid-wall patches of block N on level M, in the reference's
 * patch order.  The patch arrays are host arrays mirrored on the device (ff_wall_sync moves
 * them).  From then on ff_residual applies BOTSKE (base3c.f:6147-7095) at the patch faces
 * in each direction's FLUCYC/FLXCYC pass (res3c.f:8092, 7157): the wall fluxes replace the
 * face fluxes there, with TURCOR the SST's near-wall flux corrections and production, and
 * on level 1 the force sums.  Heat options IHEAT 0-3; no porosity, wall functions or
 * rotating frame.  npatch = 0 removes the patches. */
int  ff_wall_register(ff_ctx *ctx, int level, int block, const ff_wall_params *wp,
                      const ff_wall_patch *patches, int npatch);
/* direction 0: host -> device, 1: device -> host, for every patch array of the block level */
int  ff_wall_sync(ff_ctx *ctx, int level, int block, int direction);
/* The FF_NFORCE force sums of patch `patch` (0-based, registration order) of block N from the
 * last level-1 ff_residual. */
int  ff_wall_forces(ff_ctx *ctx, int block, int patch, double *sums);

/* ---- the learned EARSM closure ------------------------------------------------------ */
/* No counterpart in the reference: a 5 -> 64 -> 64 -> 3 SiLU network that replaces EARSM1's
 * cubic solve, N correction and coefficients (earsm.f:170-205).  Inputs: the invariants
 * II_S, II_W, III_S, IV, V of the T-scaled tensors (SII, WII, tr S^3, SWWIV, SSWWV of
 * earsm.f:108-124); outputs: beta4, beta6, beta9 (beta3 = 2 IV beta9^2 / beta6).  BF16
 * MFMA products with FP32 accumulation on the device.
 * Parameters (FF_MLP_NPARAM floats): W1[64][5], b1[64], W2[64][64], b2[64], W3[3][64],
 * b3[3], row = output.  Normalisation (FF_MLP_NNORM floats): input means (5) and scales
 * (5), output means (3) and scales (3), then the training box of the normalised inputs:
 * low (5), high (5).  Sample arrays are device memory: x[n][5], y[n][3] (float). */
#define FF_MLP_NPARAM  4739
#define FF_MLP_NNORM   26
/* Host copies in and out; either pointer may be NULL.  Setting parameters resets Adam. */
int  ff_mlp_set(ff_ctx *ctx, const float *params, const float *norm);
int  ff_mlp_get(ff_ctx *ctx, float *params, float *norm);
/* Exact-EARSM samples of every interior cell of a registered block (after ff_turb):
 * cell (I,J,K) at (K-1)*JMAX*IMAX + (J-1)*IMAX + I-1.  Needs ITURB = 6, ISTRES = 1. */
int  ff_earsm_samples(ff_ctx *ctx, int level, int block, float *x, float *y);
/* Normalisation and training box from n samples. */
int  ff_mlp_fit_norm(ff_ctx *ctx, const float *x, const float *y, int n);
/* Physical outputs of n samples. */
int  ff_mlp_infer(ff_ctx *ctx, const float *x, int n, float *y);
/* Mean squared error of the normalised outputs over n samples and its gradient: host
 * array of FF_MLP_NPARAM + 1 floats, the loss last. */
int  ff_mlp_grad(ff_ctx *ctx, const float *x, const float *y, int n, float *grad_and_loss);
/* `steps` Adam steps (learning rate lr) on minibatches of `batch` samples drawn from the
 * n samples by a hash of (seed, step, slot); loss[s] (host, may be NULL) is the
 * minibatch loss before step s. */
int  ff_mlp_train(ff_ctx *ctx, const float *x, const float *y, int n, int steps, int batch,
                  float lr, unsigned seed, float *loss);

#ifdef __cplusplus
}
#endif
#endif /* FFHIP_H */
