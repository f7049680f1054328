C     ffhip_mod.f -- ISO_C_BINDING interface of libffhip.so (ffhip.h)
C     for FINFLO's fixed-form Fortran driver.  Every entry point
C     returns 0 on success and a negative FF_E* code on failure;
C     FF_LAST_ERROR gives the message.  Compiles with the reference's
C     own fixed-form flags (72 columns, -fdefault-real-8); the kinds
C     below are the C ones, so -r8 does not change the layouts.
      MODULE FFHIP
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTEGER, PARAMETER :: FF_NFIELDS = 94, FF_NFORCE = 19
      TYPE, BIND(C) :: FF_PARAMS
         REAL(C_DOUBLE) :: GAMMA, RGAS, PR, PRT, VISU0, EXPSU, TSU0,
     +        FRSPRE, FRSDEN, FRSTEM, FRSVEL, FRSVIS, CFL, CFLL,
     +        THETA, CDIFF, TUPLIM, TLOLIM, RKLIM, EPSLIM, TURLIM,
     +        FRSMUT, RMULT, RE, CSIMPS, CMGK, CMGEPS
         REAL(C_DOUBLE) :: RESERVED(3)
         INTEGER(C_INT) :: ITURB, IFLUX, IDIFF, INTER(3), KOVER,
     +        KATOL, ISTRES, FULLNS, TURCOR
         INTEGER(C_INT) :: RESERVED_I(5)
      END TYPE
      TYPE, BIND(C) :: FF_WALL_PARAMS
         REAL(C_DOUBLE) :: T0, DIFPRE, XMOM, YMOM, ZMOM, GX, GY, GZ,
     +        GROUND, ALTITUDE, RESERVED(6)
      END TYPE
      TYPE, BIND(C) :: FF_WALL_PATCH
         INTEGER(C_INT) :: IFACE, KX1, KX2, KY1, KY2, IHEAT,
     +        RESERVED(2)
         TYPE(C_PTR)    :: UWALL, VWALL, WWALL, TWALL, HFLUX, RBK,
     +        XCP, YCP, ZCP, CPWALL, QWALL, QWFRIC, TAUW1, TAUW2,
     +        TAUWX, TAUWY, TAUWZ, SURFX, SURFY, SURFZ, SURFPX,
     +        SURFPY, SURFPZ
      END TYPE
      TYPE, BIND(C) :: FF_BLOCK
         INTEGER(C_INT) :: IMAX, JMAX, KMAX, RESERVED
         TYPE(C_PTR)    :: F(FF_NFIELDS)
      END TYPE
      INTERFACE
       INTEGER(C_INT) FUNCTION FF_CTX_CREATE(CTX, PRM) BIND(C)
         IMPORT
         TYPE(C_PTR) :: CTX
         TYPE(FF_PARAMS) :: PRM
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_CTX_DESTROY(CTX) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_BLOCK_REGISTER(CTX, M, N, BLK)
     +   BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N
         TYPE(FF_BLOCK) :: BLK
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_SYNC(CTX, M, N, DIR, FIELDS, NF)
     +   BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N, DIR, NF
         INTEGER(C_INT) :: FIELDS(*)
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_FINISH(CTX) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_STATE(CTX, M, N) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_TURB(CTX, M, N) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_RESIDUAL(CTX, M, N, M2, RESI)
     +   BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N, M2
         REAL(C_DOUBLE) :: RESI(7)
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_IMPLICIT(CTX, M, N, M2) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N, M2
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_UPDATE(CTX, M, N) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_SAVE_OLD(CTX, M, N) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_RESTRICT(CTX, MFINE, N) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: MFINE, N
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_FORCING(CTX, M, N, PHASE) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N, PHASE
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_PROLONG(CTX, MCOARSE, N) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: MCOARSE, N
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_CYCLE(CTX, N, NCYC, RESI) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: N, NCYC
         REAL(C_DOUBLE) :: RESI(7)
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_VCYCLE(CTX, N, NLEV, NCYC, RESI)
     +   BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: N, NLEV, NCYC
         REAL(C_DOUBLE) :: RESI(7)
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_WALL_REGISTER(CTX, M, N, WP, P, NP)
     +   BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N, NP
         TYPE(FF_WALL_PARAMS) :: WP
         TYPE(FF_WALL_PATCH) :: P(*)
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_WALL_SYNC(CTX, M, N, DIR) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: M, N, DIR
       END FUNCTION
       INTEGER(C_INT) FUNCTION FF_WALL_FORCES(CTX, N, L, SUMS)
     +   BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: N, L
         REAL(C_DOUBLE) :: SUMS(FF_NFORCE)
       END FUNCTION
       TYPE(C_PTR) FUNCTION FF_LAST_ERROR(CTX) BIND(C)
         IMPORT
         TYPE(C_PTR), VALUE :: CTX
       END FUNCTION
      END INTERFACE
      TYPE(C_PTR), SAVE :: FFCTX = C_NULL_PTR
      END MODULE FFHIP
