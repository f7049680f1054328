! ref_driver.f90 — test-infrastructure driver that calls FINFLO's own, unmodified hot-path
! routines (compiled from the reference's sources by oracle/Makefile into oracle/_ref/).
!
! TEST INFRASTRUCTURE ONLY: used to pin the C restatement in ff_oracle.c.  Nothing in the
! product links it.  Each entry point follows the reference's call sequence for one stage:
!   ref_state     STATIM ns3c.f:10318-10447: PDTOP, STATE->PERFGA (state.f:159), CONSVA
!   ref_turb      TURBDI ns3c.f:2987-2991 (ITURB=0): EPS2=1, VIST=0; ITURB=6: TURBEQ
!                 res3c.f:5951-6371: DERFUN, TTIMES, (ISTRES=1) EARSM1, EPS2KO
!   ref_residual  RESIDU res3c.f:194-559: ZEROZZ, then per direction k,i,j the FLUCYC
!                 sequence res3c.f:7705-7907 (ZEROZZ of F1*, FLUXDO, FLUXKE or FLUXVL)
!                 and SWEEP;
!                 RESL2 res3c.f:8920 on level 1; ITURB=6: the RESIKE sequence
!                 res3c.f:3700-5068 (CROSKO, FUN1KO, MAKE_EPS, PRODKE, FLXCYC with FLUDKE,
!                 SETV12, FLUXDO, FLUXKE; SWEEPK; SOURKO, ADDSKE); with ISTRES >= 1 PEEKOO
!                 in place of PRODKE and ANISFLUX after FLUXDO.  FULLNS: RESIDU adds DERFUN
!                 and TEMDER of T (res3c.f:250-261), RESIKE TEMDER of T, k/rho, omega/rho
!                 (res3c.f:3989-4007); on level 1 FLUXBL replaces FLUXDO and FLKEBL
!                 replaces FLUDKE (res3c.f:6802-6848, 7839-7846)
!   ref_implicit  IMPS imps3.f:111-483 with LUSGS=1: TIME3, LOCDT, COPR, PREDGS, DIAGGS,
!                 CORRGS, DIVV12; ITURB=6: LOCDTK, IMSOKO, COPRKE, PRKEGS, DIKEGS,
!                 COKEGS, DIVV12, MULV12.  With LUSGS=0 (IPR(QDDADI) /= 0) the DDADI
!                 sweeps of imps3.f:321-373 in place of the LU-SGS: IMPSWE's PRED, DIAG,
!                 CORR (PREDKE, DIAGKE, CORRKE) along K, I and J.  The reference sorts its
!                 arrays for the I and J sweeps with SORT38/SORTZZ of multi.f, which does
!                 not build here; ref_ddadi_dir makes the same JKI and KIJ orders with its
!                 own transposed copies (dd_sort) and copies the increments back
!   ref_update    CYCLES ns3c.f:1517-1596: FIN, UPLIM/LOLIM, PDTOP, LOLIMP
!
! All fields of a block are packed as F(NTOT, NF) in the order of enum ff_field
! (include/ffhip.h); the parameter vector PRM and integer vector IPR are filled by the
! Python test harness (oracle/oracle_py.py, RefBlock).
module ffref_ids
  implicit none
  integer, parameter :: IPDIFF=1, ITEMP=2, IU=3, IV=4, IW=5, IRK=6, IREPS=7
  integer, parameter :: IP=8, IRO=9, IE=10, IC=11, IVIS=12, ICH=13, ICP=14, IDRDP=15, IDRDH=16
  integer, parameter :: IRM=17, IRN=18, IRW=19, IEPS2=20, IVIST=21
  integer, parameter :: IDRO=22, IDM=23, IDN=24, IDW=25, IDE=26, IDRK=27, IDEPS=28, IDTL=29
  integer, parameter :: ITOLD=30, IUOLD=31, IVOLD=32, IWOLD=33, IPOLD=34, IRKOLD=35, IEPSOLD=36
  integer, parameter :: IVOL=37, IA1=38, IA1X=39, IA1Y=40, IA1Z=41, IA2=42, IA2X=43, IA2Y=44
  integer, parameter :: IA2Z=45, IA3=46, IA3X=47, IA3Y=48, IA3Z=49, ID1=50, ID2=51, ID3=52
  integer, parameter :: IXC=53, IYC=54, IZC=55, IDISTW=56
  integer, parameter :: ISRK=57, ISEPS=58, IPTUR=59, IFUN1=60, IOHMI=61
  integer, parameter :: IW12=62, IW13=63, IW23=64, IS11=65, IS12=66, IS13=67, IS22=68
  integer, parameter :: IS23=69, IS33=70, ISTRAIN=71, ITTS=72
  integer, parameter :: IB11=73, IB12=74, IB13=75, IB22=76, IB23=77, IB33=78
  integer, parameter :: IROP2H=79, IRMP2H=80, IRNP2H=81, IRWP2H=82, IEP2H=83, IRKP2H=84
  integer, parameter :: IEPSP2H=85, IDTDX=86, IDTDY=87, IDTDZ=88, IDKDX=89, IDKDY=90
  integer, parameter :: IDKDZ=91, IDEDX=92, IDEDY=93, IDEDZ=94
  ! parameter vector PRM (same order as the doubles of ff_params)
  integer, parameter :: PGAMMA=1, PRGAS=2, PPR=3, PPRT=4, PVISU0=5, PEXPSU=6, PTSU0=7
  integer, parameter :: PFRSPRE=8, PFRSDEN=9, PFRSTEM=10, PFRSVEL=11, PFRSVIS=12, PCFL=13
  integer, parameter :: PCFLL=14, PTHETA=15, PCDIFF=16, PTUPLIM=17, PTLOLIM=18, PRKLIM=19
  integer, parameter :: PEPSLIM=20, PTURLIM=21, PFRSMUT=22, PRMULT=23, PRE=24, PCSIMPS=25
  ! integer vector IPR
  integer, parameter :: QIMAX=1, QJMAX=2, QKMAX=3, QITURB=4, QIFLUX=5, QIDIFF=6
  integer, parameter :: QINTI=7, QINTJ=8, QINTK=9, QLEVEL=10, QKOVER=11, QKATOL=12, QM2=13
  integer, parameter :: QISTRES=14, QFULLNS=15, QTURCOR=16, QDDADI=17, QICYCLE=18
end module ffref_ids

! Solid-wall patches per level, set by ref_wall_set: ICON(3..7,L) and ICON(20,L), the
! constants of ff_wall_params and the patch arrays (in the order of ff_wall_patch), plus the
! force sums of the last level-1 residual
module ffref_walls
  use iso_c_binding
  implicit none
  integer, parameter :: MAXP = 64, NWA = 23, MAXL = 8
  integer :: nwall(MAXL) = 0
  integer :: wint(8, MAXP, MAXL) = 0
  type(c_ptr) :: wptr(NWA, MAXP, MAXL)
  real(c_double) :: wenv(16, MAXL) = 0.
  real(c_double) :: wforce(19, MAXP, MAXL) = 0.
end module ffref_walls

! INL and OUT patches per level, set by ref_bc_set: ICON(1,L), ICON(3..7,L), ICON(8,L) and
! the patch arrays BOUNR, BOUNU, BOUNV, BOUNW, BOUNT, BOUNP, BOUNPD, BOUNRK, BOUNEP
module ffref_bcs
  use iso_c_binding
  implicit none
  integer, parameter :: MAXB = 64, NBA = 9, MAXBL = 8
  integer :: nbc(MAXBL) = 0
  integer :: bint(8, MAXB, MAXBL) = 0
  type(c_ptr) :: bptr(NBA, MAXB, MAXBL)
end module ffref_bcs

subroutine ref_bc_set(level, n, ints, ptrs) bind(C, name="ref_bc_set")
  use iso_c_binding
  use ffref_bcs
  implicit none
  integer(c_int), value :: level, n
  integer(c_int) :: ints(8, *)
  type(c_ptr) :: ptrs(NBA, *)
  integer :: l
  if (level < 1 .or. level > MAXBL .or. n < 0 .or. n > MAXB) return
  nbc(level) = n
  do l = 1, n
     bint(:, l, level) = ints(:, l)
     bptr(:, l, level) = ptrs(:, l)
  end do
end subroutine ref_bc_set

! The update of FLXCYC (res3c.f:6639-6735) or FLUCYC (res3c.f:7724-7830) at M = 1 for
! direction IDIR: INLET (ICON(1,L) = 3, IBTYPE <= 6) and OUTLET (5, IBTYPE <= 5) on the
! direction's patches in patch order, with FLXCYC's rotated strides (res3c.f:4018, 4274,
! 4499), KX/KY swapped on J-faces, KBOTOP = 1 or IMAXRE/JMAXRE/KMAXRE, ILMO = +-1,
! ICYCLE = IPR(QICYCLE); FLUCYC's TURLIM = 1E-6 and ISTRES = 0 (res3c.f:7750, 7797) for
! the laminar residual.  A grid at rest (WROT = 0), one phase, ISTATE = 1, E0REF = T0REF = 0,
! no scalars, FRESUL = TRANSL = .FALSE., SOLUTION_TYPE 'FLUID'.
subroutine ref_inlout_dir(f, ntot, prm, ipr, idir)
  use iso_c_binding
  use ffref_ids
  use ffref_bcs
  use TYPE_ARRAYS
  implicit none
  integer :: ntot, idir
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(INTERMITTENCY) :: trm(1)
  type(PRE_COR) :: prc(1)
  type pa
     real(c_double), pointer :: a(:)
  end type pa
  type(pa) :: w(NBA)
  real(c_double), allocatable :: wrot(:), dum(:)
  real(c_double) :: s11(1), fi(1), bij(1), turlim, bounfi(1), bounbi(1)
  integer :: jstate(1), iwaveb(1)
  integer :: lev, l, q, iface, pdir, kx1, kx2, ky1, ky2, ilmo, kbotop, npa, ityp, ibtype
  integer :: imax, jmax, kmax, is, js, istr, jstr, kstr, ia, iax, iay, iaz, nd(3)
  lev = ipr(QLEVEL)
  if (lev /= 1) return
  if (nbc(lev) == 0) return
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  nd = (/ imax, jmax, kmax /)
  is = imax + 4; js = jmax + 4
  select case (idir)
  case (1)
     istr = is; jstr = is*js; kstr = 1; ia = IA1
  case (2)
     istr = is*js; jstr = 1; kstr = is; ia = IA2
  case default
     istr = 1; jstr = is; kstr = is*js; ia = IA3
  end select
  iax = ia + 1; iay = ia + 2; iaz = ia + 3
  if (ipr(QITURB) == 6) then
     turlim = prm(PTURLIM)
  else
     turlim = 1.E-6_c_double
  end if
  allocate(wrot(ntot))
  wrot = 0.
  s11 = 0.; fi = 0.; bij = 0.; bounfi = 0.; bounbi = 0.; jstate = 0; iwaveb = 0
  do l = 1, nbc(lev)
     ityp = bint(1, l, lev)
     iface = bint(2, l, lev)
     ibtype = bint(7, l, lev)
     if (ibtype > 10 .and. ibtype <= 20) ibtype = ibtype - 10
     if (ibtype > 20 .and. ibtype <= 30) ibtype = ibtype - 20
     pdir = iface
     if (iface > 3) pdir = iface - 3
     if (pdir /= idir) cycle
     if (.not. ((ityp == 3 .and. ibtype <= 6) .or. (ityp == 5 .and. ibtype <= 5))) cycle
     kx1 = bint(3, l, lev); kx2 = bint(4, l, lev); ky1 = bint(5, l, lev); ky2 = bint(6, l, lev)
     if (idir == 2) then
        ky1 = bint(3, l, lev); ky2 = bint(4, l, lev); kx1 = bint(5, l, lev); kx2 = bint(6, l, lev)
     end if
     if (iface <= 3) then
        ilmo = 1; kbotop = 1
     else
        ilmo = -1; kbotop = nd(idir)
     end if
     npa = (kx2 - kx1 + 5)*(ky2 - ky1 + 5)
     do q = 1, NBA
        call c_f_pointer(bptr(q, l, lev), w(q)%a, [npa])
     end do
     allocate(dum(npa))
     dum = 0.
     if (ityp == 3) then
        call INLET(f(1,IP), f(1,IPDIFF), f(1,IU), f(1,IV), f(1,IW), f(1,IC), f(1,IE), &
             f(1,IVOL), f(1,ia), f(1,iax), f(1,iay), f(1,iaz), f(1,IXC), f(1,IYC), f(1,IZC), &
             f(1,IOHMI), s11, f(1,IVIS), f(1,IEPS2), f(1,IVIST), is, js, kmax + 4, 0, kbotop, &
             kx1, kx2, ky1, ky2, f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IRK), f(1,IREPS), &
             wrot, f(1,ITEMP), f(1,ICH), f(1,ICP), prm(PPRT), ipr(QIFLUX), 1, ipr(QM2), 2, 2, &
             2, prm(PFRSDEN), prm(PFRSPRE), prm(PFRSVEL), prm(PFRSVIS), 0._c_double, &
             0._c_double, ipr(QITURB), istr, jstr, kstr, 0, ipr(QICYCLE), 1, ilmo, 1, 1, 1, &
             1, ibtype, prm(PRGAS), prm(PGAMMA), 0._c_double, 0._c_double, fi, bij, &
             w(1)%a, w(2)%a, w(3)%a, w(4)%a, dum, w(6)%a, w(7)%a, w(5)%a, w(8)%a, w(9)%a, &
             bounfi, bounbi, dum, dum, 1, 1, 1, 1, 0, turlim, 0, jstate, pro, var, 1, &
             iwaveb, .false., 0._c_double, dum, .false., trm, dum, dum, dum, dum, dum, &
             dum, dum, dum, dum, 'FLUID', prc)
     else
        call OUTLET(f(1,IP), f(1,IPDIFF), f(1,IU), f(1,IV), f(1,IW), f(1,IC), f(1,IE), &
             f(1,IVOL), f(1,ia), f(1,iax), f(1,iay), f(1,iaz), f(1,IXC), f(1,IYC), f(1,IZC), &
             f(1,IOHMI), s11, f(1,IVIS), f(1,IEPS2), f(1,IVIST), is, js, kmax + 4, 0, kbotop, &
             kx1, kx2, ky1, ky2, f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IRK), f(1,IREPS), &
             wrot, f(1,ITEMP), f(1,ICH), f(1,ICP), prm(PPRT), ipr(QIFLUX), 1, ipr(QM2), 2, 2, &
             2, prm(PFRSDEN), prm(PFRSPRE), prm(PFRSVEL), prm(PFRSVIS), 0._c_double, &
             0._c_double, ipr(QITURB), istr, jstr, kstr, 0, ipr(QICYCLE), 1, ilmo, 1, 1, 1, &
             1, ibtype, prm(PRGAS), prm(PGAMMA), 0._c_double, 0._c_double, fi, bij, &
             w(1)%a, w(2)%a, w(3)%a, w(4)%a, dum, w(6)%a, w(7)%a, w(5)%a, w(8)%a, w(9)%a, &
             bounfi, bounbi, dum, dum, 1, 1, 1, 1, 0, turlim, 0, jstate, pro, var, 1, &
             iwaveb, .false., dum, .false., trm, dum, dum, dum, dum, dum, &
             dum, dum, dum, dum, 'FLUID', 1, prc)
     end if
     deallocate(dum)
  end do
  deallocate(wrot)
end subroutine ref_inlout_dir

! INLET/OUTLET alone over the directions K, I, J (the order of RESIDU's flux cycles)
subroutine ref_inlout(f, ntot, prm, ipr) bind(C, name="ref_inlout")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ntot
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  call ref_inlout_dir(f, ntot, prm, ipr, 3)
  call ref_inlout_dir(f, ntot, prm, ipr, 1)
  call ref_inlout_dir(f, ntot, prm, ipr, 2)
end subroutine ref_inlout

subroutine ref_wall_set(level, n, ints, env, ptrs) bind(C, name="ref_wall_set")
  use iso_c_binding
  use ffref_walls
  implicit none
  integer(c_int), value :: level, n
  integer(c_int) :: ints(8, *)
  real(c_double) :: env(16)
  type(c_ptr) :: ptrs(NWA, *)
  integer :: l
  if (level < 1 .or. level > MAXL .or. n < 0 .or. n > MAXP) return
  nwall(level) = n
  wenv(:, level) = env
  do l = 1, n
     wint(:, l, level) = ints(:, l)
     wptr(:, l, level) = ptrs(:, l)
  end do
  wforce(:, :, level) = 0.
end subroutine ref_wall_set

subroutine ref_wall_forces(level, patch, out) bind(C, name="ref_wall_forces")
  use iso_c_binding
  use ffref_walls
  implicit none
  integer(c_int), value :: level, patch
  real(c_double) :: out(19)
  out = wforce(:, patch + 1, level)
end subroutine ref_wall_forces

! BOTSKE base3c.f:6147 for every solid-wall patch of the level in direction IDIR, in patch
! order, as FLUCYC (res3c.f:8092, TURCOR = .FALSE., IDIS = 0, PSIG = CMU = 0) or FLXCYC
! (res3c.f:7157, IDIS = 2, SST sigmas, CMU = 0.09) call it after the direction's fluxes:
! FLXCYC's rotated strides (res3c.f:4018, 4274, 4499), KX/KY swapped on J-faces
! (res3c.f:7090-7098), no wall functions, UROT = 0
subroutine ref_walls(f, ntot, prm, ipr, idir, fl, zero, sig, kw)
  use iso_c_binding
  use ffref_ids
  use ffref_walls
  use TYPE_ARRAYS
  use NS3CO, only : WALLFUNL, GX, GY, GZ, GROUND, ALTITUDE
  implicit none
  integer :: ntot, idir
  real(c_double) :: f(ntot, *), prm(*), fl(ntot, 7), zero(ntot), sig(4)
  integer(c_int) :: ipr(*)
  logical :: kw
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(INTERMITTENCY) :: trm(1)
  type(PRE_COR) :: prc(1)
  type pa
     real(c_double), pointer :: a(:)
  end type pa
  type(pa) :: w(NWA)
  real(c_double), allocatable :: dum(:)
  real(c_double) :: rcon(1), xco(1), f1h(1), wave(1), utaum(1), fvw(1)
  real(c_double) :: psk1, pse1, psk2, pse2, cmu
  integer :: iwaveb(1)
  integer :: lev, l, q, iface, pdir, kx1, kx2, ky1, ky2, ilmo, kbotop, ndi, npa
  integer :: imax, jmax, kmax, is, js, istr, jstr, kstr, ia, iax, iay, iaz, inter, intet, idis
  logical :: turcor
  lev = ipr(QLEVEL)
  if (lev < 1 .or. lev > MAXL) return
  if (nwall(lev) == 0) return
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  is = imax + 4; js = jmax + 4
  WALLFUNL = .false.
  GX = wenv(6, lev); GY = wenv(7, lev); GZ = wenv(8, lev)
  GROUND = wenv(9, lev); ALTITUDE = wenv(10, lev)
  select case (idir)
  case (1)
     istr = is; jstr = is*js; kstr = 1; ia = IA1; inter = ipr(QINTI); ndi = imax + 1
  case (2)
     istr = is*js; jstr = 1; kstr = is; ia = IA2; inter = ipr(QINTJ); ndi = jmax + 1
  case default
     istr = 1; jstr = is; kstr = is*js; ia = IA3; inter = ipr(QINTK); ndi = kmax + 1
  end select
  iax = ia + 1; iay = ia + 2; iaz = ia + 3
  intet = inter
  if (inter > -5 .and. inter < 5) intet = +1
  if (kw) then
     idis = 2; psk1 = sig(1); pse1 = sig(2); psk2 = sig(3); pse2 = sig(4); cmu = 0.09_c_double
     turcor = ipr(QTURCOR) /= 0
  else
     idis = 0; psk1 = 0.; pse1 = 0.; psk2 = 0.; pse2 = 0.; cmu = 0.; turcor = .false.
  end if
  do l = 1, nwall(lev)
     iface = wint(1, l, lev)
     pdir = iface
     if (iface > 3) pdir = iface - 3
     if (pdir /= idir) cycle
     kx1 = wint(2, l, lev); kx2 = wint(3, l, lev); ky1 = wint(4, l, lev); ky2 = wint(5, l, lev)
     if (idir == 2) then
        kx1 = wint(4, l, lev); kx2 = wint(5, l, lev); ky1 = wint(2, l, lev); ky2 = wint(3, l, lev)
     end if
     if (iface <= 3) then
        ilmo = 1; kbotop = 1
     else
        ilmo = -1; kbotop = ndi - 1
     end if
     npa = (kx2 - kx1 + 5)*(ky2 - ky1 + 5)
     do q = 1, NWA
        call c_f_pointer(wptr(q, l, lev), w(q)%a, [npa])
     end do
     allocate(dum(npa))
     dum = 0.
     call BOTSKE(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), f(1,IP), f(1,IU), f(1,IV), &
          f(1,IW), f(1,IC), f(1,IE), f(1,IVOL), f(1,ia), f(1,iax), f(1,iay), f(1,iaz), &
          f(1,IOHMI), f(1,IVIS), f(1,IEPS2), f(1,IVIST), is, js, ipr(QKMAX) + 4, ndi, kbotop, &
          f(1,IPDIFF), pro, var, kx1, kx2, ky1, ky2, f(1,ICP), prm(PPR), prm(PPRT), f(1,IRO), &
          wforce(7,l,lev), ipr(QIFLUX), wforce(1,l,lev), wforce(2,l,lev), wforce(3,l,lev), &
          wforce(4,l,lev), wforce(5,l,lev), wforce(6,l,lev), lev, 2, 2, 2, f(1,IXC), &
          f(1,IYC), f(1,IZC), prm(PRKLIM), wenv(3,lev), wenv(4,lev), wenv(5,lev), ilmo, &
          wforce(8,l,lev), wforce(9,l,lev), wforce(10,l,lev), fl(1,6), fl(1,7), f(1,IRK), &
          f(1,IREPS), f(1,IPTUR), zero, f(1,ITEMP), f(1,ICH), wforce(11,l,lev), &
          wforce(12,l,lev), wforce(13,l,lev), prm(PFRSDEN), prm(PFRSPRE), prm(PFRSVEL), &
          prm(PFRSVIS), wenv(1,lev), wenv(2,lev), ipr(QITURB), istr, jstr, kstr, &
          wint(6,l,lev), rcon, idis, psk1, pse1, psk2, pse2, w(1)%a, w(2)%a, w(3)%a, &
          w(10)%a, w(4)%a, w(11)%a, w(12)%a, w(5)%a, w(13)%a, w(14)%a, w(15)%a, w(16)%a, &
          w(17)%a, dum, dum, dum, dum, dum, dum, w(6)%a, dum, dum, dum, w(18)%a, w(19)%a, &
          w(20)%a, w(7)%a, w(8)%a, w(9)%a, ntot, cmu, xco, xco, xco, inter, intet, &
          .false., 'FLUID', 1, turcor, 1, .false., f1h, iwaveb, wforce(17,l,lev), &
          wforce(18,l,lev), wforce(19,l,lev), wforce(14,l,lev), wforce(15,l,lev), &
          wforce(16,l,lev), f(1,ISTRAIN), .false., trm, wave, w(21)%a, w(22)%a, w(23)%a, &
          prc, utaum, fvw, fvw, fvw, prm(PTURLIM))
     deallocate(dum)
  end do
end subroutine ref_walls

subroutine ref_state(f, ntot, prm, ipr) bind(C, name="ref_state")
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  use NS3CO, only : GX, GY, GZ, GROUND, ALTITUDE
  implicit none
  integer(c_int), value :: ntot
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  real(c_double), allocatable :: trgs(:), fi(:), rnut(:)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  integer :: ercode, iturb
  character(len=10) :: mulphc
  ALTITUDE = 0.; GROUND = 0.; GX = 0.; GY = 0.; GZ = 0.
  allocate(trgs(ntot), fi(ntot), rnut(ntot))
  fi = 0.; rnut = 0.
  iturb = ipr(QITURB)
  ercode = 0
  mulphc = 'FLUID'
  call PDTOP(f(1,IP), f(1,IPDIFF), f(1,IXC), f(1,IYC), f(1,IZC), prm(PFRSDEN), prm(PFRSPRE), ntot)
  call PERFGA(f(1,IVIS), f(1,IP), f(1,IRO), f(1,IE), f(1,ITEMP), f(1,ICP), f(1,ICH), &
       f(1,IDRDP), f(1,IDRDH), prm(PRGAS), prm(PGAMMA), prm(PPR), prm(PVISU0), prm(PEXPSU), &
       prm(PTSU0), ntot, ercode, 0._c_double, 0._c_double, prm(PRMULT), trgs, prm(PFRSPRE), prm(PFRSTEM), 0)
  call CONSVA(f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IE), f(1,IC), f(1,IU), f(1,IV), &
       f(1,IW), f(1,IRK), f(1,IP), f(1,IDRDP), f(1,IDRDH), f(1,IXC), f(1,IYC), f(1,IZC), &
       pro, var, 0._c_double, prm(PFRSVEL), ntot, iturb, 0, fi, ntot, 0, 0, 0._c_double, mulphc, 1, 1, &
       rnut, f(1,IREPS))
  deallocate(trgs, fi, rnut)
end subroutine ref_state

subroutine ref_turb(f, ntot, prm, ipr) bind(C, name="ref_turb")
  use iso_c_binding
  use ffref_ids
  implicit none
  integer(c_int), value :: ntot
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  real(c_double), allocatable :: vellap(:), rksi(:)
  real(c_double) :: a1kleb, bstar, ckappa, beta1, sigrk1, sigom1, beta2, sigrk2, sigom2
  real(c_double) :: sigphi, zeta2, frsmut, eps4
  real(c_double), allocatable :: wir(:)
  integer :: kover
  if (ipr(QITURB) == 6) then
     ! TURBEQ (res3c.f:5922, 5951-5954, 5984, 6362-6371), KOVER = 0: EPS2KO with STRAIN
     allocate(vellap(ntot), rksi(ntot))
     vellap = 0.; rksi = 0.
     kover = ipr(QKOVER)
     frsmut = 1.E-3_c_double*prm(PFRSVIS)
     call DERFUN(f(1,IOHMI), f(1,IW12), f(1,IW13), f(1,IW23), f(1,IS11), f(1,IS12), &
          f(1,IS13), f(1,IS22), f(1,IS23), f(1,IS33), f(1,IU), f(1,IV), f(1,IW), f(1,IA1), &
          f(1,IA2), f(1,IA3), f(1,IVOL), f(1,ID1), f(1,ID2), f(1,ID3), f(1,IA1X), f(1,IA1Y), &
          f(1,IA1Z), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3X), f(1,IA3Y), f(1,IA3Z), &
          f(1,IXC), f(1,IYC), f(1,IZC), ipr(QIMAX), ipr(QJMAX), ipr(QKMAX), 2, 2, 2, 1, 1, &
          f(1,ISTRAIN), vellap)
     call TTIMES(f(1,IRO), f(1,IVIS), f(1,IRK), f(1,IREPS), 6, ipr(QIMAX), ipr(QJMAX), &
          ipr(QKMAX), f(1,ITTS))
     if (ipr(QISTRES) == 1) then
        ! res3c.f:5987-5990; WIR is never given a value (PA0 = 0, earsm.f:47)
        allocate(wir(3*ntot))
        wir = 0.
        eps4 = 0.
        call EARSM1(f(1,IVIS), f(1,IVIST), f(1,IEPS2), eps4, f(1,IRK), wir, f(1,IW12), &
             f(1,IS11), f(1,IB11), ipr(QIMAX), ipr(QJMAX), ipr(QKMAX), ntot, ntot, f(1,ITTS), &
             prm(PPR), prm(PPRT), prm(PTURLIM), frsmut)
        deallocate(wir)
     end if
     call COEFKO(6, a1kleb, bstar, ckappa, beta1, sigrk1, sigom1, beta2, sigrk2, sigom2, &
          sigphi, zeta2, kover)
     call EPS2KO(f(1,IEPS2), f(1,IVIST), f(1,IVIS), f(1,ISTRAIN), f(1,IRO), f(1,IRK), &
          f(1,IREPS), f(1,IDISTW), prm(PTURLIM), frsmut, ipr(QIMAX), ipr(QJMAX), ipr(QKMAX), &
          2, 2, 2, prm(PPR), prm(PPRT), bstar, a1kleb, kover, rksi)
     deallocate(vellap, rksi)
     return
  end if
  ! TURBDI ITURB = 0 (ns3c.f:2987-2991)
  call ADJUIN(f(1,IEPS2), 1._c_double, ipr(QIMAX), ipr(QJMAX), ipr(QKMAX), 2, 2, 2)
  call ADJUIN(f(1,IVIST), 0._c_double, ipr(QIMAX), ipr(QJMAX), ipr(QKMAX), 2, 2, 2)
end subroutine ref_turb

! One direction of RESIDU: the FLUCYC sequence for a block with no boundary patches
! (NPATCH = 0), FULLNS = .FALSE., IFLUX = 1 (Roe) or 0 (van Leer), followed by SWEEP.
subroutine ref_direction(f, ntot, prm, ipr, idir, fl, zzz, maw, maxw, rksi, hat, zero, trm)
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  implicit none
  integer :: ntot, idir, maw, maxw
  real(c_double) :: f(ntot, *), prm(*), fl(ntot, 7), zzz(maxw), rksi(ntot), hat(ntot, 4)
  real(c_double) :: zero(ntot)
  integer(c_int) :: ipr(*)
  type(INTERMITTENCY) :: trm(1)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  real(c_double) :: f1h(1), wave(1)
  integer :: imax, jmax, kmax, is, js, kstr, kmaxp, ia, iax, iay, iaz, inter, ndi, istr, ido
  integer :: id
  real(c_double) :: sig(4)
  logical :: xxtral, inchiml, entropy_fix, transl, mulphl, fresul
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  is = imax + 4; js = jmax + 4
  xxtral = .false.; inchiml = .false.; entropy_fix = .false.; transl = .false.
  mulphl = .false.; fresul = .false.
  select case (idir)
  case (1)
     kstr = 1; kmaxp = kmax; ia = IA1; inter = ipr(QINTI); ndi = imax + 1
     istr = 1; id = ID1
  case (2)
     kstr = is; kmaxp = kmax; ia = IA2; inter = ipr(QINTJ); ndi = jmax + 1
     istr = is; id = ID2
  case default
     kstr = is*js; kmaxp = kmax + 1; ia = IA3; inter = ipr(QINTK); ndi = kmax + 1
     istr = is*js; id = ID3
  end select
  iax = ia + 1; iay = ia + 2; iaz = ia + 3
  ! INLET/OUTLET at M = 1 (FLUCYC res3c.f:7724-7830)
  call ref_inlout_dir(f, ntot, prm, ipr, idir)
  fl = 0.
  if (ipr(QFULLNS) /= 0 .and. ipr(QLEVEL) == 1) then
  ! FLUXBL (res3c.f:7839): full viscous flux from the cell gradients
  call FLUXBL(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), f(1,IVOL), f(1,ia), f(1,IXC), &
       f(1,IYC), f(1,IZC), f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IP), f(1,IU), f(1,IV), &
       f(1,IW), f(1,IE), f(1,IC), imax, jmax, kmax, 1, inter, ndi, f(1,IVIS), f(1,IEPS2), &
       f(1,IVIST), f(1,IW12), f(1,IW13), f(1,IW23), f(1,IS11), f(1,IS12), f(1,IS13), &
       f(1,IS22), f(1,IS23), f(1,IS33), f(1,IDTDX), f(1,IDTDY), f(1,IDTDZ), f(1,id), &
       f(1,iax), f(1,iay), f(1,iaz), f(1,ICP), prm(PPR), prm(PPRT), zero, pro, var, &
       f(1,ITEMP), f(1,ICH), f(1,IDRDH), f(1,IDRDP), 1, prm(PGAMMA), prm(PFRSDEN), &
       prm(PFRSPRE), hat(1,1), hat(1,2), hat(1,3), hat(1,4), kstr, idir, ipr(QITURB), mulphl, &
       ipr(QIDIFF))
  else
  ! FLUXDO (res3c.f:7856): thin-layer viscous flux, arithmetic average
  call FLUXDO(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), f(1,IVOL), f(1,ia), f(1,IRO), &
       f(1,IRM), f(1,IRN), f(1,IRW), f(1,IP), f(1,IU), f(1,IV), f(1,IW), f(1,IE), f(1,IC), &
       imax, jmax, kmax, 1, inter, ndi, f(1,IVIS), f(1,IEPS2), f(1,IVIST), f(1,iax), &
       f(1,iay), f(1,iaz), f(1,ICP), prm(PPR), prm(PPRT), zero, pro, var, f(1,ITEMP), &
       f(1,ICH), f(1,IDRDH), f(1,IDRDP), 1, prm(PGAMMA), prm(PFRSDEN), prm(PFRSPRE), &
       hat(1,1), hat(1,2), hat(1,3), hat(1,4), kstr, idir, mulphl, 1, fresul, 0._c_double, f1h, &
       wave, ipr(QLEVEL), ipr(QIDIFF))
  end if
  if (ipr(QIFLUX) == 0) then
  ! FLUXVL (res3c.f:7907): van Leer flux-vector splitting
  call FLUXVL(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), fl(1,6), fl(1,7), f(1,IVOL), &
       f(1,ia), f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IRK), f(1,IREPS), f(1,IP), &
       f(1,IU), f(1,IV), f(1,IW), f(1,IE), f(1,IC), imax, jmax, kmax, 1, inter, ndi, &
       f(1,IVIS), f(1,IEPS2), f(1,IVIST), f(1,IPDIFF), f(1,iax), f(1,iay), f(1,iaz), &
       prm(PPR), prm(PPRT), zero, f(1,ITEMP), f(1,ICH), f(1,IDRDH), f(1,IDRDP), 1, &
       prm(PGAMMA), 0._c_double, 0._c_double, prm(PFRSDEN), prm(PFRSPRE), 0._c_double, &
       0._c_double, hat(1,1), hat(1,2), hat(1,3), hat(1,4), kmaxp, kstr, idir, ipr(QITURB), &
       zzz, maw, maxw, rksi, inchiml)
  else
  ! FLUXKE (res3c.f:7892): Roe flux; FLUCYC passes RKLIM = EPSLIM = 0 (res3c.f:7680-7681)
  call FLUXKE(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), fl(1,6), fl(1,7), f(1,IVOL), &
       f(1,ia), f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IRK), f(1,IREPS), f(1,IP), &
       f(1,IU), f(1,IV), f(1,IW), f(1,IE), f(1,IC), imax, jmax, kmax, 1, inter, 9, ndi, &
       f(1,IVIS), f(1,IEPS2), f(1,IVIST), f(1,IPDIFF), f(1,iax), f(1,iay), f(1,iaz), &
       f(1,ID1), prm(PPR), prm(PPRT), zero, f(1,ITEMP), f(1,ICH), f(1,IDRDH), f(1,IDRDP), &
       1, prm(PRGAS), prm(PGAMMA), 0._c_double, 0._c_double, prm(PFRSDEN), prm(PFRSPRE), 0._c_double, 0._c_double, &
       hat(1,1), hat(1,2), hat(1,3), hat(1,4), kmaxp, kstr, idir, ipr(QITURB), zzz, maw, &
       maxw, xxtral, rksi, inchiml, entropy_fix, transl, trm)
  end if
  ! BOTSKE on the solid-wall patches of the direction (res3c.f:8092)
  sig = 0.
  call ref_walls(f, ntot, prm, ipr, idir, fl, zero, sig, .false.)
  ! SWEEP (res3c.f:330, 448, 559); the j-direction is last and divides by VOL
  ido = 0
  if (idir == 2) ido = 1
  call SWEEP(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), fl(1,1), fl(1,2), fl(1,3), &
       fl(1,4), fl(1,5), f(1,IVOL), imax, jmax, kmax, istr, 2, 2, 2, 0._c_double, ido)
end subroutine ref_direction

! TEMDER turb3.f:3456-3703 of the scalar Q into the three fields from IOUT on, with
! IDERI = IDER = 1 (main.f:5653) and the level of IPR
subroutine ref_temder(f, ntot, q, iout, ipr)
  use iso_c_binding
  use ffref_ids
  implicit none
  integer :: ntot, iout
  real(c_double) :: f(ntot, *), q(ntot)
  integer(c_int) :: ipr(*)
  call TEMDER(q, f(1,iout), f(1,iout+1), f(1,iout+2), f(1,IA1), f(1,IA2), f(1,IA3), f(1,IVOL), &
       f(1,ID1), f(1,ID2), f(1,ID3), f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2X), f(1,IA2Y), &
       f(1,IA2Z), f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IXC), f(1,IYC), f(1,IZC), ipr(QIMAX), &
       ipr(QJMAX), ipr(QKMAX), 2, 2, 2, 1, ipr(QLEVEL))
end subroutine ref_temder

! RESIDI zeroes the force sums of the patches on level 1 (ns3c.f:2376-2396)
subroutine ref_wall_zero()
  use ffref_walls
  implicit none
  wforce(:, :, 1) = 0.
end subroutine ref_wall_zero

subroutine ref_residual(f, ntot, prm, ipr, resi) bind(C, name="ref_residual")
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  use NS3CO, only : TWO_FLUIDL
  implicit none
  integer(c_int), value :: ntot
  real(c_double) :: f(ntot, *), prm(*), resi(*)
  integer(c_int) :: ipr(*)
  real(c_double), allocatable :: fl(:,:), zzz(:), rksi(:), hat(:,:), zero(:), dfi(:)
  real(c_double), allocatable :: vellap(:)
  type(INTERMITTENCY) :: trm(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(PRE_COR) :: prc(1)
  real(c_double) :: r(67)
  integer :: maw, maxw, imax, jmax, kmax
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  if (ipr(QLEVEL) == 1) call ref_wall_zero()
  if (ipr(QITURB) == 6) then
     call ref_residual_kw(f, ntot, prm, ipr, resi)
     return
  end if
  ! FLUXKE's slab work arrays hold ISTRID*(JMAX+2) entries (base3c.f:4812-4824)
  maw = (imax + 4)*(jmax + 4)
  maxw = 27*maw
  allocate(fl(ntot,7), zzz(maxw), rksi(ntot), hat(ntot,4), zero(ntot), dfi(ntot))
  rksi = 0.; zero = 0.; zzz = 0.; dfi = 0.
  ! ZEROZZ of the residuals (res3c.f:194-199)
  call ZEROZZ(f(1,IDRO), ntot); call ZEROZZ(f(1,IDM), ntot); call ZEROZZ(f(1,IDN), ntot)
  call ZEROZZ(f(1,IDW), ntot); call ZEROZZ(f(1,IDE), ntot); call ZEROZZ(f(1,IDEPS), ntot)
  if (ipr(QFULLNS) /= 0) then
     ! RESIDU res3c.f:250-261: DERFUN and TEMDER of T, on every level
     allocate(vellap(ntot))
     vellap = 0.
     call DERFUN(f(1,IOHMI), f(1,IW12), f(1,IW13), f(1,IW23), f(1,IS11), f(1,IS12), &
          f(1,IS13), f(1,IS22), f(1,IS23), f(1,IS33), f(1,IU), f(1,IV), f(1,IW), f(1,IA1), &
          f(1,IA2), f(1,IA3), f(1,IVOL), f(1,ID1), f(1,ID2), f(1,ID3), f(1,IA1X), f(1,IA1Y), &
          f(1,IA1Z), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3X), f(1,IA3Y), f(1,IA3Z), &
          f(1,IXC), f(1,IYC), f(1,IZC), imax, jmax, kmax, 2, 2, 2, 1, ipr(QLEVEL), &
          f(1,ISTRAIN), vellap)
     call ref_temder(f, ntot, f(1,ITEMP), IDTDX, ipr)
     deallocate(vellap)
  end if
  call ref_direction(f, ntot, prm, ipr, 3, fl, zzz, maw, maxw, rksi, hat, zero, trm)
  call ref_direction(f, ntot, prm, ipr, 1, fl, zzz, maw, maxw, rksi, hat, zero, trm)
  call ref_direction(f, ntot, prm, ipr, 2, fl, zzz, maw, maxw, rksi, hat, zero, trm)
  if (ipr(QLEVEL) == 1) then
     TWO_FLUIDL = .false.
     r = 0.
     call RESL2(f(1,IVOL), f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), dfi, &
          f(1,IDRK), f(1,IDEPS), var, r, rksi, imax, jmax, kmax, ipr(QITURB), 0, ntot, 67, &
          prm(PRKLIM), prm(PEPSLIM), .false., 1, 1, 1, prc)
     resi(1:7) = r(31:37)
  end if
  deallocate(fl, zzz, rksi, hat, zero, dfi)
end subroutine ref_residual

! One direction of RESIKE for ITURB = 6: the FLXCYC sequence res3c.f:6740-6946 for a block
! with no boundary patches, FULLNS = .FALSE., IFLUX = 1 (zero F1*, FLUDKE, SETV12, FLUXDO,
! FLUXKE with the turbulence MUSCL switch INTET and RKLIM/EPSLIM), followed by SWEEPK
! (res3c.f:4098, 4325, 4551).
subroutine ref_direction_kw(f, ntot, prm, ipr, idir, fl, zzz, maw, maxw, rksi, hat, zero, &
     trm, sig)
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  implicit none
  integer :: ntot, idir, maw, maxw
  real(c_double) :: f(ntot, *), prm(*), fl(ntot, 7), zzz(maxw), rksi(ntot), hat(ntot, 4)
  real(c_double) :: zero(ntot), sig(4)
  integer(c_int) :: ipr(*)
  type(INTERMITTENCY) :: trm(1)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(PRE_COR) :: prc(1)
  real(c_double) :: f1h(1), wave(1)
  integer :: imax, jmax, kmax, is, js, kstr, kmaxp, ia, iax, iay, iaz, inter, intet, ndi, ido
  integer :: jadd, kadd, id
  logical :: fullns
  logical :: xxtral, inchiml, entropy_fix, transl, mulphl, fresul
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  is = imax + 4; js = jmax + 4
  xxtral = .false.; inchiml = .false.; entropy_fix = .false.; transl = .false.
  mulphl = .false.; fresul = .false.
  select case (idir)
  case (1)
     kstr = 1; kmaxp = kmax; ia = IA1; inter = ipr(QINTI); ndi = imax + 1; id = ID1
  case (2)
     kstr = is; kmaxp = kmax; ia = IA2; inter = ipr(QINTJ); ndi = jmax + 1; id = ID2
  case default
     kstr = is*js; kmaxp = kmax + 1; ia = IA3; inter = ipr(QINTK); ndi = kmax + 1; id = ID3
  end select
  fullns = ipr(QFULLNS) /= 0 .and. ipr(QLEVEL) == 1
  ! INTETI/J/K of RESIKE (res3c.f:3634-3641, INTERTU = 0)
  intet = inter
  if (inter > -5 .and. inter < 5) intet = +1
  iax = ia + 1; iay = ia + 2; iaz = ia + 3
  ! INLET/OUTLET at M = 1 (FLXCYC res3c.f:6639-6735)
  call ref_inlout_dir(f, ntot, prm, ipr, idir)
  fl = 0.
  if (fullns) then
  ! FLKEBL (res3c.f:6808): k and omega diffusion from the cell gradients of k/rho, omega/rho
  call FLKEBL(fl(1,5), fl(1,6), fl(1,7), f(1,IVOL), f(1,ia), f(1,iax), f(1,iay), f(1,iaz), &
       f(1,id), f(1,IRO), f(1,IRK), f(1,IREPS), imax, jmax, kmax, ndi, f(1,IVIS), f(1,IEPS2), &
       f(1,IVIST), f(1,IFUN1), f(1,IDKDX), f(1,IDKDY), f(1,IDKDZ), f(1,IDEDX), f(1,IDEDY), &
       f(1,IDEDZ), f(1,IXC), f(1,IYC), f(1,IZC), sig(1), sig(2), sig(3), sig(4), kstr, idir, &
       ipr(QIDIFF))
  else
  call FLUDKE(fl(1,5), fl(1,6), fl(1,7), f(1,IVOL), f(1,ia), f(1,IRO), f(1,IRK), f(1,IREPS), &
       imax, jmax, kmax, ndi, f(1,IVIS), f(1,IEPS2), f(1,IVIST), f(1,IFUN1), sig(1), sig(2), &
       sig(3), sig(4), kstr, idir, ipr(QIDIFF), transl, trm)
  end if
  call SETV12(fl(1,5), fl(1,6), ntot)
  if (fullns) then
  ! FLUXBL (res3c.f:6845)
  call FLUXBL(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), f(1,IVOL), f(1,ia), f(1,IXC), &
       f(1,IYC), f(1,IZC), f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IP), f(1,IU), f(1,IV), &
       f(1,IW), f(1,IE), f(1,IC), imax, jmax, kmax, 1, inter, ndi, f(1,IVIS), f(1,IEPS2), &
       f(1,IVIST), f(1,IW12), f(1,IW13), f(1,IW23), f(1,IS11), f(1,IS12), f(1,IS13), &
       f(1,IS22), f(1,IS23), f(1,IS33), f(1,IDTDX), f(1,IDTDY), f(1,IDTDZ), f(1,id), &
       f(1,iax), f(1,iay), f(1,iaz), f(1,ICP), prm(PPR), prm(PPRT), zero, pro, var, &
       f(1,ITEMP), f(1,ICH), f(1,IDRDH), f(1,IDRDP), 1, prm(PGAMMA), prm(PFRSDEN), &
       prm(PFRSPRE), hat(1,1), hat(1,2), hat(1,3), hat(1,4), kstr, idir, ipr(QITURB), mulphl, &
       ipr(QIDIFF))
  else
  call FLUXDO(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), f(1,IVOL), f(1,ia), f(1,IRO), &
       f(1,IRM), f(1,IRN), f(1,IRW), f(1,IP), f(1,IU), f(1,IV), f(1,IW), f(1,IE), f(1,IC), &
       imax, jmax, kmax, 1, inter, ndi, f(1,IVIS), f(1,IEPS2), f(1,IVIST), f(1,iax), &
       f(1,iay), f(1,iaz), f(1,ICP), prm(PPR), prm(PPRT), zero, pro, var, f(1,ITEMP), &
       f(1,ICH), f(1,IDRDH), f(1,IDRDP), 1, prm(PGAMMA), prm(PFRSDEN), prm(PFRSPRE), &
       hat(1,1), hat(1,2), hat(1,3), hat(1,4), kstr, idir, mulphl, 1, fresul, 0._c_double, f1h, &
       wave, ipr(QLEVEL), ipr(QIDIFF))
  end if
  if (ipr(QISTRES) >= 1) then
     ! FLXCYC res3c.f:6933-6937 (JADD/KADD of res3c.f:6565-6579)
     jadd = 0; kadd = 0
     if (idir == 2) jadd = 1
     if (idir == 3) kadd = 1
     call ANISFLUX(imax, jmax, kmax, jadd, kadd, kstr, f(1,IU), f(1,IV), f(1,IW), f(1,IRK), &
          f(1,IB11), f(1,ia), f(1,iax), f(1,iay), f(1,iaz), fl(1,2), fl(1,3), fl(1,4), &
          fl(1,5), ntot)
  end if
  hat = 0.
  call FLUXKE(fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), fl(1,6), fl(1,7), f(1,IVOL), &
       f(1,ia), f(1,IRO), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IRK), f(1,IREPS), f(1,IP), &
       f(1,IU), f(1,IV), f(1,IW), f(1,IE), f(1,IC), imax, jmax, kmax, 1, inter, intet, ndi, &
       f(1,IVIS), f(1,IEPS2), f(1,IVIST), f(1,IPDIFF), f(1,iax), f(1,iay), f(1,iaz), &
       f(1,ID1), prm(PPR), prm(PPRT), zero, f(1,ITEMP), f(1,ICH), f(1,IDRDH), f(1,IDRDP), &
       1, prm(PRGAS), prm(PGAMMA), 0._c_double, 0._c_double, prm(PFRSDEN), prm(PFRSPRE), &
       prm(PRKLIM), prm(PEPSLIM), hat(1,1), hat(1,2), hat(1,3), hat(1,4), kmaxp, kstr, idir, &
       ipr(QITURB), zzz, maw, maxw, xxtral, rksi, inchiml, entropy_fix, transl, trm)
  ! BOTSKE on the solid-wall patches of the direction (res3c.f:7157)
  call ref_walls(f, ntot, prm, ipr, idir, fl, zero, sig, .true.)
  ido = 0
  if (idir == 2) ido = 1
  call SWEEPK(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IDRK), f(1,IDEPS), &
       fl(1,1), fl(1,2), fl(1,3), fl(1,4), fl(1,5), fl(1,6), fl(1,7), f(1,IVOL), var, trm, &
       imax, jmax, kmax, kstr, 2, 2, 2, 0._c_double, ido, mulphl, transl, 'FLUID', prc)
end subroutine ref_direction_kw

! RESIKE res3c.f:3423-5387 for ITURB = 6 (SST, IDIS = 2), M = M2 = 1, no patches
subroutine ref_residual_kw(f, ntot, prm, ipr, resi)
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  use MAIN_ARRAYS, only : BLKS
  use NS3CO, only : TWO_FLUIDL
  implicit none
  integer(c_int) :: ntot
  real(c_double) :: f(ntot, *), prm(*), resi(*)
  integer(c_int) :: ipr(*)
  real(c_double), allocatable :: fl(:,:), zzz(:), rksi(:), hat(:,:), zero(:), dfi(:)
  real(c_double), allocatable :: t2(:), trueps(:), ddeps(:), vtran(:), rich(:), vellap(:)
  real(c_double), allocatable :: qsas(:), rst(:)
  type(INTERMITTENCY) :: trm(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(PRE_COR) :: prc(1)
  real(c_double) :: r(67), sig(4)
  real(c_double) :: a1kleb, bstar, ckappa, beta1, sigrk1, sigom1, beta2, sigrk2, sigom2
  real(c_double) :: sigphi, zeta2
  integer :: maw, maxw, imax, jmax, kmax, kover
  logical :: katol
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  kover = ipr(QKOVER)
  katol = ipr(QKATOL) /= 0
  if (.not. allocated(BLKS)) allocate(BLKS(1))
  BLKS(1)%RKLIM = prm(PRKLIM); BLKS(1)%EPSLIM = prm(PEPSLIM)
  maw = (imax + 4)*(jmax + 4)
  maxw = 27*maw
  allocate(fl(ntot,7), zzz(maxw), rksi(ntot), hat(ntot,4), zero(ntot), dfi(ntot))
  allocate(t2(ntot), trueps(ntot), ddeps(ntot), vtran(ntot), rich(ntot), vellap(ntot))
  allocate(qsas(ntot), rst(ntot))
  rksi = 0.; zero = 0.; zzz = 0.; dfi = 0.; t2 = 0.; ddeps = 0.; vtran = 1.; vellap = 0.
  qsas = 0.; rst = 0.
  ! res3c.f:3668-3675, 3730
  call ZEROZZ(f(1,IDRO), ntot); call ZEROZZ(f(1,IDM), ntot); call ZEROZZ(f(1,IDN), ntot)
  call ZEROZZ(f(1,IDW), ntot); call ZEROZZ(f(1,IDE), ntot); call ZEROZZ(f(1,IDRK), ntot)
  call ZEROZZ(f(1,IDEPS), ntot); call ZEROZZ(f(1,ISRK), ntot); call ZEROZZ(f(1,IFUN1), ntot)
  ! with M2 >= 2 RESIKE forms the velocity gradients itself (res3c.f:3749-3769)
  if (ipr(QM2) >= 2) then
     call DERFUN(f(1,IOHMI), f(1,IW12), f(1,IW13), f(1,IW23), f(1,IS11), f(1,IS12), &
          f(1,IS13), f(1,IS22), f(1,IS23), f(1,IS33), f(1,IU), f(1,IV), f(1,IW), f(1,IA1), &
          f(1,IA2), f(1,IA3), f(1,IVOL), f(1,ID1), f(1,ID2), f(1,ID3), f(1,IA1X), f(1,IA1Y), &
          f(1,IA1Z), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3X), f(1,IA3Y), f(1,IA3Z), &
          f(1,IXC), f(1,IYC), f(1,IZC), imax, jmax, kmax, 2, 2, 2, 1, ipr(QLEVEL), &
          f(1,ISTRAIN), vellap)
  end if
  ! COEFKO and the diffusion sigmas (res3c.f:3788-3806)
  call COEFKO(6, a1kleb, bstar, ckappa, beta1, sigrk1, sigom1, beta2, sigrk2, sigom2, &
       sigphi, zeta2, kover)
  sig(1) = sigrk1; sig(2) = sigom1; sig(3) = sigrk2; sig(4) = sigom2
  call CROSKO(f(1,ISEPS), f(1,IRO), f(1,IRK), f(1,IREPS), f(1,IVOL), f(1,IA1), f(1,IA1X), &
       f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), f(1,IA3X), &
       f(1,IA3Y), f(1,IA3Z), imax, jmax, kmax, 2, 2, 2, sigom2, sigphi, t2, kover)
  call FUN1KO(f(1,IFUN1), f(1,ISEPS), f(1,IVIS), f(1,IRO), f(1,IRK), f(1,IREPS), &
       f(1,IDISTW), .false., imax, jmax, kmax, 2, 2, 2, bstar, sigom2, prm(PRKLIM), &
       1._c_double, prm(PRE), 1)
  ! MAKE_EPS, PRODKE, ZEROZZ(DDEPS) (res3c.f:3949-3994)
  trueps = 0.
  call MAKE_EPS(f(1,IRO), f(1,IRK), f(1,IREPS), ddeps, trueps, 6, ntot)
  if (ipr(QISTRES) >= 1) then
     ! res3c.f:3976-3978
     call PEEKOO(f(1,IVIS), f(1,IEPS2), f(1,IS11), f(1,IB11), vtran, imax, jmax, kmax, &
          f(1,IRK), f(1,IPTUR), ntot, ntot)
  else
     call PRODKE(f(1,IOHMI), f(1,IW12), f(1,IW13), f(1,IW23), f(1,IS11), f(1,IS12), &
          f(1,IS13), f(1,IS22), f(1,IS23), f(1,IS33), f(1,IPTUR), f(1,IRK), f(1,IREPS), ddeps, &
          trueps, f(1,IEPS2), f(1,IVIST), f(1,IVIS), rst, rst, rst, rst, rst, rst, vtran, &
          imax, jmax, kmax, 2, 2, 2, 1, 6, katol)
  end if
  if (ipr(QFULLNS) /= 0) then
     ! res3c.f:3989-4007: TEMDER of T, of DDEPS = RK/RO and of DDEPS = REPS/RO
     call ref_temder(f, ntot, f(1,ITEMP), IDTDX, ipr)
     call DIVV12(ddeps, f(1,IRK), f(1,IRO), ntot)
     call ref_temder(f, ntot, ddeps, IDKDX, ipr)
     call DIVV12(ddeps, f(1,IREPS), f(1,IRO), ntot)
     call ref_temder(f, ntot, ddeps, IDEDX, ipr)
  end if
  call ZEROZZ(ddeps, ntot)
  call ref_direction_kw(f, ntot, prm, ipr, 3, fl, zzz, maw, maxw, rksi, hat, zero, trm, sig)
  call ref_direction_kw(f, ntot, prm, ipr, 1, fl, zzz, maw, maxw, rksi, hat, zero, trm, sig)
  call ref_direction_kw(f, ntot, prm, ipr, 2, fl, zzz, maw, maxw, rksi, hat, zero, trm, sig)
  ! MAKE_EPS, RICH = 0 (LRICH = 0), SOURKO, ADDSKE (res3c.f:4850-5068)
  trueps = 0.
  call MAKE_EPS(f(1,IRO), f(1,IRK), f(1,IREPS), ddeps, trueps, 6, ntot)
  rich = 0.
  call SOURKO(f(1,ISRK), f(1,ISEPS), f(1,IPTUR), f(1,IEPS2), f(1,IVIS), f(1,IRO), f(1,IRK), &
       f(1,IREPS), f(1,IFUN1), rich, f(1,ID1), f(1,ID2), f(1,ID3), 1, imax, jmax, kmax, &
       2, 2, 2, kover, bstar, ckappa, beta1, sigom1, beta2, sigom2, zeta2, 0.09_c_double, &
       .false., f(1,ITTS), .false., .false., trm, t2, f(1,ISTRAIN), vellap, qsas, &
       f(1,IOHMI), f(1,IDISTW))
  call ADDSKE(f(1,IDRK), f(1,IDEPS), f(1,ISRK), f(1,IRK), f(1,IREPS), f(1,IPTUR), &
       f(1,ISEPS), f(1,IVOL), imax, jmax, kmax, (imax+4)*(jmax+4), 2, 2, 2, 0._c_double, trm, &
       .false.)
  if (ipr(QLEVEL) == 1) then
     TWO_FLUIDL = .false.
     r = 0.
     call RESL2(f(1,IVOL), f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), dfi, &
          f(1,IDRK), f(1,IDEPS), var, r, rksi, imax, jmax, kmax, ipr(QITURB), 0, ntot, 67, &
          prm(PRKLIM), prm(PEPSLIM), .false., 1, 1, 1, prc)
     resi(1:7) = r(31:37)
  end if
  deallocate(fl, zzz, rksi, hat, zero, dfi, t2, trueps, ddeps, vtran, rich, vellap, qsas, rst)
end subroutine ref_residual_kw

subroutine ref_implicit(f, ntot, prm, ipr) bind(C, name="ref_implicit")
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  implicit none
  integer(c_int), value :: ntot
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  real(c_double), allocatable :: zzz(:), rksi(:), zero(:), abv(:)
  real(c_double), allocatable :: dxi(:), dgi(:), dreti(:), gii(:), reti(:), pdfor(:)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  integer :: maw, maxw, imax, jmax, kmax, idm1, idm2, idm3
  real(c_double) :: cfm, theta
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  ! the LU-SGS slab work arrays hold JMAX*ISTRID entries (imps4.f:96-106)
  maw = (imax + 4)*(jmax + 4)
  maxw = 21*maw
  allocate(zzz(maxw), rksi(ntot), zero(ntot), abv(ntot))
  zzz = 0.; rksi = 0.; zero = 0.; abv = 0.
  cfm = prm(PCFL)
  if (ipr(QLEVEL) >= 2) cfm = prm(PCFLL)
  theta = prm(PTHETA)
  idm1 = imax + 1; idm2 = jmax + 1; idm3 = kmax + 1
  ! TIME3 (imps3.f:111) with PVISC = 1.8, IVISC = 1
  call TIME3(f(1,IA1), f(1,IA2), f(1,IA3), f(1,ID1), f(1,ID2), f(1,ID3), f(1,IVOL), &
       f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3X), &
       f(1,IA3Y), f(1,IA3Z), f(1,IDTL), 0._c_double, f(1,IRO), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
       f(1,IP), imax, jmax, kmax, cfm, f(1,IVIS), f(1,IVIST), zero, zero, zero, &
       prm(PFRSDEN), prm(PFRSVEL), prm(PPR), prm(PPRT), idm1, idm2, idm3, ipr(QLEVEL), &
       .false., 1.8_c_double, abv, 1, prm(PFRSPRE), prm(PCDIFF))
  if (ipr(QITURB) == 6) then
     ! LOCDTK (DTURB = 1, imps3.f:61, 123), IMSOKO (imps3.f:233), COPRKE (imps3.f:276),
     ! PRKEGS/DIKEGS/COKEGS (imps3.f:398-430), DIVV12, MULV12 (imps3.f:483, 497-500)
     allocate(dxi(ntot), dgi(1), dreti(1), gii(1), reti(1), pdfor(ntot))
     dxi = 0.; dgi = 0.; dreti = 0.; gii = 0.; reti = 0.; pdfor = 0.
     call LOCDTK(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IDRK), f(1,IDEPS), &
          f(1,IDTL), ntot, 0._c_double, 1._c_double, 'FLUID     ', dxi, .false., dgi, dreti)
     call IMSOKO(f(1,IDRK), f(1,IDEPS), f(1,IRK), f(1,IREPS), f(1,IPTUR), f(1,IEPS2), &
          f(1,IVIS), f(1,IDTL), f(1,IRO), f(1,IU), f(1,IV), f(1,IW), f(1,IE), prm(PGAMMA), &
          imax, jmax, kmax, 1, imax + 4, (imax + 4)*(jmax + 4), 2, 2, 2, prm(PCSIMPS), &
          ipr(QKOVER))
     call COPRKE(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IDRK), f(1,IDEPS), &
          f(1,IRO), f(1,IE), f(1,IP), f(1,IPDIFF), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRK), f(1,IREPS), f(1,IDRDP), f(1,IDRDH), prm(PFRSVEL), 'FLUID     ', pro, var, &
          dxi, f(1,IDTL), ntot, .false., gii, reti, dgi, dreti, pdfor, rksi, f(1,IVOL), 1, &
          imax, jmax, kmax)
     if (ipr(QDDADI) /= 0) then
        call ddadi_dir(f, ntot, prm, ipr, 0, 3)
        call ddadi_dir(f, ntot, prm, ipr, 1, 3)
        call ddadi_dir(f, ntot, prm, ipr, 2, 3)
     else
     call PRKEGS(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IVOL), f(1,IA1), &
          f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), &
          f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IDTL), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRO), f(1,IP), f(1,IDRDP), f(1,IDRDH), imax, jmax, kmax, theta, 0._c_double, &
          idm1, idm2, idm3, f(1,IVIS), f(1,IVIST), f(1,IDRK), f(1,IDEPS), f(1,IRK), f(1,IREPS), &
          f(1,ISRK), f(1,ISEPS), prm(PPR), prm(PPRT), 2, 2, 2, zero, zero, zero, rksi, zzz, &
          maw, maxw, .false., dgi, dreti)
     call DIKEGS(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IVOL), f(1,IA1), &
          f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), &
          f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IDTL), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRO), f(1,IP), f(1,IDRDP), f(1,IDRDH), imax, jmax, kmax, theta, f(1,IVIS), &
          f(1,IVIST), f(1,IDRK), f(1,IDEPS), f(1,IRK), f(1,IREPS), f(1,ISRK), f(1,ISEPS), &
          2, 2, 2, zero, zero, zero, rksi, .false., dgi, dreti)
     call COKEGS(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IVOL), f(1,IA1), &
          f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), &
          f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IDTL), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRO), f(1,IP), f(1,IDRDP), f(1,IDRDH), imax, jmax, kmax, theta, 0._c_double, &
          idm1, idm2, idm3, f(1,IVIS), f(1,IVIST), f(1,IDRK), f(1,IDEPS), f(1,IRK), f(1,IREPS), &
          f(1,ISRK), f(1,ISEPS), prm(PPR), prm(PPRT), 2, 2, 2, zero, zero, zero, rksi, zzz, &
          maw, maxw, .false., dgi, dreti)
     end if
     call DIVV12(f(1,IDRO), f(1,IDRO), f(1,ICP), ntot)
     call MULV12(f(1,IDRK), f(1,IRO), ntot)
     call MULV12(f(1,IDEPS), f(1,IRO), ntot)
     deallocate(dxi, dgi, dreti, gii, reti, pdfor)
  else
  ! LOCDT (imps3.f:120) and COPR (imps3.f:272)
     call LOCDT(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IDTL), ntot, 0._c_double)
     call COPR(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IRO), f(1,IE), f(1,IP), &
          f(1,IPDIFF), f(1,IU), f(1,IV), f(1,IW), f(1,IC), f(1,IDRDP), f(1,IDRDH), &
          prm(PFRSVEL), ntot)
     if (ipr(QDDADI) /= 0) then
        call ddadi_dir(f, ntot, prm, ipr, 0, 3)
        call ddadi_dir(f, ntot, prm, ipr, 1, 3)
        call ddadi_dir(f, ntot, prm, ipr, 2, 3)
     else
     ! LU-SGS (imps3.f:383-391)
     call PREDGS(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IVOL), f(1,IA1), &
          f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), &
          f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IDTL), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRO), f(1,IP), imax, jmax, kmax, theta, 0._c_double, idm1, idm2, idm3, f(1,IVIS), &
          f(1,IVIST), prm(PPR), prm(PPRT), 2, 2, 2, zero, zero, zero, rksi, zzz, maw, maxw)
     call DIAGGS(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IVOL), f(1,IA1), &
          f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), &
          f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IDTL), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRO), f(1,IP), imax, jmax, kmax, theta, 0._c_double, idm1, idm2, idm3, f(1,IVIS), &
          f(1,IVIST), prm(PPR), prm(PPRT), 2, 2, 2, zero, zero, zero, rksi, zzz, maw, maxw)
     call CORRGS(f(1,IDRO), f(1,IDM), f(1,IDN), f(1,IDW), f(1,IDE), f(1,IVOL), f(1,IA1), &
          f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), f(1,IA3), &
          f(1,IA3X), f(1,IA3Y), f(1,IA3Z), f(1,IDTL), f(1,IU), f(1,IV), f(1,IW), f(1,IC), &
          f(1,IRO), f(1,IP), imax, jmax, kmax, theta, 0._c_double, idm1, idm2, idm3, f(1,IVIS), &
          f(1,IVIST), prm(PPR), prm(PPRT), 2, 2, 2, zero, zero, zero, rksi, zzz, maw, maxw)
     end if
     ! dh -> dT (imps3.f:483)
     call DIVV12(f(1,IDRO), f(1,IDRO), f(1,ICP), ntot)
  end if
  deallocate(zzz, rksi, zero, abv)
end subroutine ref_implicit

! The sorted copy (dir 1: JKI, dir 2: KIJ) of one IJK field of an (imax+4, jmax+4, kmax+4)
! block, or (back = .true.) the IJK field from its sorted copy
subroutine dd_sort(x, xs, imax, jmax, kmax, dir, back)
  use iso_c_binding
  implicit none
  integer :: imax, jmax, kmax, dir
  logical :: back
  real(c_double) :: x(imax+4, jmax+4, kmax+4)
  real(c_double) :: xs(*)
  integer :: i, j, k, n, ni, nj, nk
  ni = imax + 4; nj = jmax + 4; nk = kmax + 4
  do k = 1, nk
     do j = 1, nj
        do i = 1, ni
           if (dir == 1) then
              n = j + (k - 1)*nj + (i - 1)*nj*nk
           else
              n = k + (i - 1)*nk + (j - 1)*nk*ni
           end if
           if (back) then
              x(i, j, k) = xs(n)
           else
              xs(n) = x(i, j, k)
           end if
        end do
     end do
  end do
end subroutine dd_sort

! One DDADI direction of IMPS (imps3.f:321-373): dir 0 the K sweep on the IJK arrays, 1 the
! I sweep on JKI copies, 2 the J sweep on KIJ copies; phase 0 PRED, 1 DIAG, 2 CORR, 3 all
! three (IMPSWE imps3.f:604-637; the k-omega PREDKE/DIAGKE/CORRKE with DTURB = 1)
subroutine ref_ddadi_dir(f, ntot, prm, ipr, dir, phase) bind(C, name="ref_ddadi_dir")
  use iso_c_binding
  implicit none
  integer(c_int), value :: ntot, dir, phase
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  call ddadi_dir(f, ntot, prm, ipr, dir, phase)
end subroutine ref_ddadi_dir

subroutine ddadi_dir(f, ntot, prm, ipr, dir, phase)
  use iso_c_binding
  use ffref_ids
  implicit none
  integer(c_int) :: ntot, dir, phase
  real(c_double) :: f(ntot, *)
  real(c_double) :: prm(*)
  integer(c_int) :: ipr(*)
  integer, parameter :: NSRT = 27
  integer :: ids(NSRT), q, imax, jmax, kmax, i1, j1, k1, idi, maw, maxw, nq
  integer :: kcp(1)
  real(c_double), allocatable :: g(:,:)
  real(c_double), allocatable :: zzz(:), urot(:)
  real(c_double) :: theta, dturb
  integer :: ia, iax, iay, iaz
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  theta = prm(PTHETA); dturb = 1.
  if (dir == 0) then
     ia = IA3; iax = IA3X; iay = IA3Y; iaz = IA3Z
     i1 = imax; j1 = jmax; k1 = kmax; idi = kmax + 1
  else if (dir == 1) then
     ia = IA1; iax = IA1X; iay = IA1Y; iaz = IA1Z
     i1 = jmax; j1 = kmax; k1 = imax; idi = imax + 1
  else
     ia = IA2; iax = IA2X; iay = IA2Y; iaz = IA2Z
     i1 = kmax; j1 = imax; k1 = jmax; idi = jmax + 1
  end if
  ! the fields the sweeps read or write, in a fixed order
  ids = (/ IDRO, IDM, IDN, IDW, IDE, IDRK, IDEPS, ia, IVOL, iax, iay, iaz, IDTL, IU, IV, IW, &
           IC, IRO, IP, IVIS, IVIST, IDRDP, IDRDH, IRK, IREPS, ISRK, ISEPS /)
  nq = NSRT
  allocate(g(ntot, NSRT), urot(ntot))
  urot = 0.
  do q = 1, nq
     if (dir == 0) then
        g(:, q) = f(:, ids(q))
     else
        call dd_sort(f(1, ids(q)), g(1, q), imax, jmax, kmax, dir, .false.)
     end if
  end do
  maw = (i1 + 4)*(j1 + 4)
  maxw = 8*maw
  allocate(zzz(maxw))
  zzz = 0.
  if (ipr(QITURB) == 6) then
     if (phase == 0 .or. phase == 3) &
        call PREDKE(g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,8), g(1,9), g(1,10), g(1,11), &
             g(1,12), g(1,13), g(1,14), g(1,15), g(1,16), g(1,17), g(1,18), g(1,19), g(1,22), &
             g(1,23), i1, j1, k1, theta, 0._c_double, idi, g(1,20), g(1,21), g(1,6), g(1,7), &
             g(1,24), g(1,25), g(1,26), g(1,27), prm(PPR), prm(PPRT), kcp, 2, 2, 2, urot, dturb, &
             zzz, maw, maxw)
     if (phase == 1 .or. phase == 3) &
        call DIAGKE(g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,8), g(1,9), g(1,10), g(1,11), &
             g(1,12), g(1,13), g(1,14), g(1,15), g(1,16), g(1,17), g(1,18), g(1,19), g(1,22), &
             g(1,23), i1, j1, k1, theta, 0._c_double, idi, g(1,20), g(1,21), g(1,6), g(1,7), &
             g(1,24), g(1,25), g(1,26), g(1,27), prm(PPR), prm(PPRT), kcp, 2, 2, 2, urot, dturb, &
             zzz, maw, maxw)
     if (phase == 2 .or. phase == 3) &
        call CORRKE(g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,8), g(1,9), g(1,10), g(1,11), &
             g(1,12), g(1,13), g(1,14), g(1,15), g(1,16), g(1,17), g(1,18), g(1,19), g(1,22), &
             g(1,23), i1, j1, k1, theta, 0._c_double, idi, g(1,20), g(1,21), g(1,6), g(1,7), &
             g(1,24), g(1,25), g(1,26), g(1,27), prm(PPR), prm(PPRT), kcp, 2, 2, 2, urot, dturb, &
             zzz, maw, maxw)
  else
     if (phase == 0 .or. phase == 3) &
        call PRED(g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,8), g(1,9), g(1,10), g(1,11), &
             g(1,12), g(1,13), g(1,14), g(1,15), g(1,16), g(1,17), g(1,18), g(1,19), i1, j1, k1, &
             theta, 0._c_double, idi, g(1,20), g(1,21), prm(PPR), prm(PPRT), 2, 2, 2, urot, zzz, &
             maw, maxw)
     if (phase == 1 .or. phase == 3) &
        call DIAG(g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,8), g(1,9), g(1,10), g(1,11), &
             g(1,12), g(1,13), g(1,14), g(1,15), g(1,16), g(1,17), g(1,18), g(1,19), i1, j1, k1, &
             theta, 0._c_double, idi, g(1,20), g(1,21), prm(PPR), prm(PPRT), 2, 2, 2, urot)
     if (phase == 2 .or. phase == 3) &
        call CORR(g(1,1), g(1,2), g(1,3), g(1,4), g(1,5), g(1,8), g(1,9), g(1,10), g(1,11), &
             g(1,12), g(1,13), g(1,14), g(1,15), g(1,16), g(1,17), g(1,18), g(1,19), i1, j1, k1, &
             theta, 0._c_double, idi, g(1,20), g(1,21), prm(PPR), prm(PPRT), 2, 2, 2, urot, zzz, &
             maw, maxw)
  end if
  ! the increments back
  do q = 1, 7
     if (dir == 0) then
        f(:, ids(q)) = g(:, q)
     else
        call dd_sort(f(1, ids(q)), g(1, q), imax, jmax, kmax, dir, .true.)
     end if
  end do
  deallocate(g, urot, zzz)
end subroutine ddadi_dir

subroutine ref_update(f, ntot, prm, ipr) bind(C, name="ref_update")
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  use NS3CO, only : GX, GY, GZ, GROUND, ALTITUDE
  implicit none
  integer(c_int), value :: ntot
  real(c_double) :: f(ntot, *), prm(*)
  integer(c_int) :: ipr(*)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(INTERMITTENCY) :: trm(1)
  type(BLOCKS) :: blks(1)
  real(c_double), allocatable :: fi(:), dfi(:), fiold(:)
  real(c_double) :: rlolim(1), upplim(1)
  integer :: imax, jmax, kmax
  ALTITUDE = 0.; GROUND = 0.; GX = 0.; GY = 0.; GZ = 0.
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  allocate(fi(ntot), dfi(ntot), fiold(ntot))
  fi = 0.; dfi = 0.; fiold = 0.
  blks(1)%SOLUTION_TYPE = 'FLUID'
  blks(1)%ISTATE = 1
  call FIN(f(1,ITEMP), f(1,IU), f(1,IV), f(1,IW), f(1,IPDIFF), f(1,IDRO), f(1,IDM), &
       f(1,IDN), f(1,IDW), f(1,IDE), f(1,ITOLD), f(1,IUOLD), f(1,IVOLD), f(1,IWOLD), &
       f(1,IPOLD), f(1,IRK), f(1,IREPS), f(1,IDRK), f(1,IDEPS), f(1,IRKOLD), f(1,IEPSOLD), &
       prm(PRKLIM), prm(PEPSLIM), pro, var, fi, dfi, fiold, rlolim, upplim, imax, jmax, &
       kmax, ntot, 0, 0, ipr(QITURB), 0, prm(PTURLIM), f(1,IVIS), .false., trm, blks, 1, 1)
  call UPLIM(f(1,ITEMP), imax, jmax, kmax, 2, 2, 2, prm(PTUPLIM))
  call LOLIM(f(1,ITEMP), imax, jmax, kmax, 2, 2, 2, prm(PTLOLIM))
  if (ipr(QLEVEL) >= 2) then
     ! coarse levels (ns3c.f:1945-1985): TEMP limits and LOLIMP with the level's pressure
     call LOLIMP(f(1,IP), f(1,IPDIFF), prm(PFRSDEN), prm(PFRSPRE), 0._c_double, 0._c_double, &
          f(1,IXC), f(1,IYC), f(1,IZC), imax, jmax, kmax, 2, 2, 2)
     deallocate(fi, dfi, fiold)
     return
  end if
  call UPLIM(f(1,IU), imax, jmax, kmax, 2, 2, 2, 5000._c_double)
  call LOLIM(f(1,IU), imax, jmax, kmax, 2, 2, 2, -5000._c_double)
  call UPLIM(f(1,IV), imax, jmax, kmax, 2, 2, 2, 5000._c_double)
  call LOLIM(f(1,IV), imax, jmax, kmax, 2, 2, 2, -5000._c_double)
  call UPLIM(f(1,IW), imax, jmax, kmax, 2, 2, 2, 5000._c_double)
  call LOLIM(f(1,IW), imax, jmax, kmax, 2, 2, 2, -5000._c_double)
  call PDTOP(f(1,IP), f(1,IPDIFF), f(1,IXC), f(1,IYC), f(1,IZC), prm(PFRSDEN), prm(PFRSPRE), &
       ntot)
  if (ipr(QITURB) >= 3) call LOLIM(f(1,IREPS), imax, jmax, kmax, 2, 2, 2, prm(PEPSLIM))
  call LOLIMP(f(1,IP), f(1,IPDIFF), prm(PFRSDEN), prm(PFRSPRE), 0._c_double, 0._c_double, f(1,IXC), &
       f(1,IYC), f(1,IZC), imax, jmax, kmax, 2, 2, 2)
  deallocate(fi, dfi, fiold)
end subroutine ref_update

! REFLEC (turb3.f:5294-5878) over npatch boundary patches, as STATIM calls it with
! REFLEL = .TRUE. (ns3c.f:10497-10512): ICON(1,L) = ibc, ICON(2,L) = L, ICON(3..7,L) =
! iface, kx1..ky2, ICON(20,L) = iheat, RCON(5,L) = tsurf; the wall velocity arrays of the
! patches are concatenated at IHF offsets.  The viscous branch is selected, per direction,
! by IDI = IMAX+1 (every wall of that direction viscous) or 0 (none), with IBOT = 1 and
! ITOP = IMAX; SOLUTION_TYPE = 'FLUID', ISTATE = 1, REFLECL = MULPHL = TRANSL = .FALSE.,
! IPRESC = 0.
subroutine ref_reflec(f, ntot, prm, ipr, npatch, ints, tsurf, uw, vw, ww, nwarr) &
     bind(C, name="ref_reflec")
  use iso_c_binding
  use ffref_ids
  use TYPE_ARRAYS
  use NS3CO, only : IC9
  implicit none
  integer(c_int), value :: ntot, npatch, nwarr
  real(c_double) :: f(ntot, *), prm(*), tsurf(*), uw(*), vw(*), ww(*)
  integer(c_int) :: ipr(*), ints(8, *)
  integer, allocatable :: icon(:,:), ihf(:)
  real(c_double), allocatable :: rcon(:,:), ddeps(:), twall(:)
  type(PROPERTIES) :: pro(1)
  type(MPHASE_VARIABLES) :: var(1)
  type(INTERMITTENCY) :: trm(1)
  type(PRE_COR) :: prc(1)
  integer :: l, idi(3), imax, jmax, kmax, dir, nd(3), off, na, nb
  character(len=10) :: soltype
  imax = ipr(QIMAX); jmax = ipr(QJMAX); kmax = ipr(QKMAX)
  nd = (/ imax, jmax, kmax /)
  allocate(icon(IC9, max(npatch, 1)), rcon(IC9, max(npatch, 1)), ihf(max(npatch, 1)))
  allocate(ddeps(ntot), twall(max(nwarr, 1)))
  icon = 0; rcon = 0.; ddeps = 0.; twall = 0.
  idi = 0
  off = 0
  do l = 1, npatch
     icon(1, l) = ints(1, l)
     icon(2, l) = l
     icon(3:7, l) = ints(2:6, l)
     icon(20, l) = ints(7, l)
     rcon(5, l) = tsurf(l)
     dir = mod(ints(2, l) - 1, 3) + 1
     if (ints(8, l) /= 0) idi(dir) = nd(dir) + 1
     ! IHF(IPL): the 1-based start of the patch's arrays, each (NA+4)*(NB+4) with LN = 2
     ! ghost rows each side (NR = ... + IHF(IPL), turb3.f:5545; BOTSKE's patch-local form
     ! has + 1 in its place, base3c.f:6359)
     ihf(l) = off + 1
     na = ints(4, l) - ints(3, l) + 1
     nb = ints(6, l) - ints(5, l) + 1
     off = off + (na + 4) * (nb + 4)
  end do
  soltype = 'FLUID'
  call REFLEC(f(1,IU), f(1,IV), f(1,IW), f(1,IXC), f(1,IYC), f(1,IZC), imax, jmax, kmax, &
       1, imax, 1, jmax, 1, kmax, idi(1), idi(2), idi(3), 2, 2, 2, &
       prm(PGAMMA), 0._c_double, 0._c_double, prm(PFRSDEN), prm(PFRSPRE), 1, npatch, icon, &
       f(1,IRO), f(1,IP), f(1,IPDIFF), f(1,IRM), f(1,IRN), f(1,IRW), f(1,IE), &
       f(1,IVIS), f(1,IVIST), f(1,ICH), f(1,IEPS2), f(1,ITEMP), f(1,IC), f(1,IDRDP), &
       f(1,IDRDH), f(1,IRK), f(1,IREPS), ddeps, pro, var, prm(PRGAS), prm(PFRSTEM), &
       ipr(QITURB), 1, 1, 1, rcon, uw, vw, ww, twall, ihf, prm(PTLOLIM), prm(PTUPLIM), &
       soltype, .false., .false., .false., trm, &
       f(1,IA1X), f(1,IA1Y), f(1,IA1Z), f(1,IA2X), f(1,IA2Y), f(1,IA2Z), &
       f(1,IA2X), f(1,IA2Y), f(1,IA2Z), prc, 0)
  deallocate(icon, rcon, ihf, ddeps, twall)
end subroutine ref_reflec
