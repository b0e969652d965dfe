!-----------------------------------------------------------------------
! gci_rci.f90 -- Fortran host side of the rci90 hot path over libgraspci.
!
! Mirrors the reference routines so that an rci90 build calls the GPU
! engine where it calls its own code, with the same argument meanings:
!
!   GCI_LOAD_HMAT  loads a block the way MANEIG fills module hmat_C
!                  (rci90/maneig.f90:312-332): rows as SETHAM writes them
!                  (IENDC(0:NCF) cumulative ends, IROW, EMT, diagonal last),
!                  diagonal shifted by -EAV, kept on the device
!   GCI_SPICMV2    SPICMV2(N, M, B, C) (lib9290/spicmv2.f90:2-56), the GDVD
!                  operator: CALL GDVD (GCI_SPICMV2, NCF, ...) replaces
!                  CALL GDVD (SPICMV2, NCF, ...) at rci90/maneig.f90:451
!   GCI_MANEIG_SOLVE  MANEIG's eigensolve of the loaded block (maneig.f90:157-
!                  555: NCF <= 2000 dense, else INIEST2 + GDVD with LIM,
!                  CRITE/CRITC/CRITR, MAXITR as MANEIG sets them), eigenvalues
!                  of H - EAV, eigenvectors sign-fixed
!   GCI_GENINTRK   GENINTRK (rci90/genintrk.f90:3-109): the R^k list into
!                  the arrays of cteilsrk_C / kkstart_C (INDTEIRK, VALTEIRK,
!                  KSTART(0:KMAX)) with the reference's order and packing
!
! and, for the SCF program rmcdhf90 (the second caller of the same pieces):
!
!   GCI_MCDHF_MANEIG  rmcdhf90's MANEIG solve of the loaded block
!                  (rmcdhf90/maneig.f90:78-200): levels ICCMIN of the lowest
!                  min(NVECT, NCF), cold (INIEST2 over 4000 CSFs) or warm
!                  (previous EVAL/EVEC) start, GDVD above 4000 CSFs
!   GCI_YZK_BATCH  YZK(K, I, J) (lib9290/yzk.f90) for all pairs of an SCF
!                  iteration's potentials (rmcdhf90/ypot.f90:75,
!                  xpot.f90:80): TB(1:N) and MTP per pair
!
! Like the reference, these routines STOP on an error (here with the
! engine's message); the handle plays the role of hmat_C's module state.
!-----------------------------------------------------------------------
module gci_rci
   use, intrinsic :: iso_c_binding
   use gci_bind
   implicit none
   private
   public :: gci_start, gci_load_hmat, gci_free_hmat, gci_set_op_mode, gci_spicmv2, gci_maneig_solve, &
             gci_genintrk, gci_check, gci_mcdhf_maneig, gci_yzk_batch

   type(c_ptr), save :: hmat = c_null_ptr
   integer, save :: hmat_n = 0
   integer(c_int), save :: op_mode = GCI_OP_FP64

contains

   subroutine gci_check(st, where)
      integer(c_int), intent(in) :: st
      character(len=*), intent(in) :: where
      if (st /= GCI_OK) then
         write (0, '(a, a, i0, a, a)') where, ': graspci error ', st, ': ', gci_message()
         error stop 'libgraspci call failed'
      end if
   end subroutine gci_check

   subroutine gci_start(idev)
      integer, intent(in) :: idev
      call gci_check(gci_init(int(idev, c_int)), 'GCI_START')
   end subroutine gci_start

   ! MANEIG's load of hmat_C: NCF rows, IENDC(0:NCF), IROW/EMT with the
   ! diagonal last in each row; iprec = GCI_FP64 (plain drop-in) or
   ! GCI_SPLIT32 (also keeps the fp32 hi/lo split for the mixed form).
   subroutine gci_load_hmat(ncf, iendc, irow, emt, eav, iprec)
      integer, intent(in) :: ncf
      integer, intent(in) :: iendc(0:ncf)
      integer, intent(in) :: irow(*)
      real(c_double), intent(in) :: emt(*)
      real(c_double), intent(in) :: eav
      integer, intent(in) :: iprec
      integer(c_int64_t), allocatable :: ie(:)
      integer(c_int32_t), allocatable :: ir(:)
      real(c_double), allocatable :: em(:)
      integer(c_int64_t) :: nel
      call gci_free_hmat()
      allocate (ie(0:ncf))
      ie = int(iendc, c_int64_t)
      nel = ie(ncf)
      allocate (ir(max(nel, 1_c_int64_t)), em(max(nel, 1_c_int64_t)))
      ir(1:nel) = int(irow(1:nel), c_int32_t)
      em(1:nel) = emt(1:nel)
      call gci_check(gci_hmat_upload(int(ncf, c_int), ie, ir, em, 1_c_int, -eav, int(iprec, c_int), hmat), &
                     'GCI_LOAD_HMAT')
      hmat_n = ncf
      op_mode = GCI_OP_FP64
   end subroutine gci_load_hmat

   subroutine gci_free_hmat()
      if (c_associated(hmat)) call gci_check(gci_hmat_free(hmat), 'GCI_FREE_HMAT')
      hmat = c_null_ptr
      hmat_n = 0
   end subroutine gci_free_hmat

   ! Operator used by GCI_SPICMV2: GCI_OP_FP64 (default, the exact drop-in),
   ! GCI_OP_SPLIT_BULK or GCI_OP_SPLIT_FULL for a block loaded as GCI_SPLIT32.
   subroutine gci_set_op_mode(mode)
      integer, intent(in) :: mode
      op_mode = int(mode, c_int)
   end subroutine gci_set_op_mode

   ! SPICMV2(N, M, B, C): C := (H - EAV) B for the loaded block; C is
   ! overwritten, B and C are N x M column-major (DVDSON's WORK slices).
   subroutine gci_spicmv2(n, m, b, c)
      integer, intent(in) :: n, m
      real(c_double), intent(in) :: b(n, m)
      real(c_double), intent(inout) :: c(n, m)
      if (n /= hmat_n) error stop 'GCI_SPICMV2: N does not match the loaded block'
      call gci_check(gci_spmm(hmat, int(m, c_int), b, c, op_mode), 'GCI_SPICMV2')
   end subroutine gci_spicmv2

   ! MANEIG's eigensolve for the loaded block: the NVEC lowest eigenpairs of
   ! H - EAV; EVAL(NVEC), EVEC(NCF, NVEC); NLOOPS, NMV and IERR as GDVD
   ! reports them (0 for the dense branch).
   subroutine gci_maneig_solve(nvec, eval, evec, nloops, nmv, ierr)
      integer, intent(in) :: nvec
      real(c_double), intent(out) :: eval(nvec)
      real(c_double), intent(out) :: evec(hmat_n, nvec)
      integer, intent(out) :: nloops, nmv, ierr
      type(gci_dvd_stats) :: st
      call gci_check(gci_maneig(hmat, int(nvec, c_int), eval, evec, st), 'GCI_MANEIG')
      nloops = st%nloops
      nmv = st%nmv
      ierr = st%ierr
   end subroutine gci_maneig_solve

   ! rmcdhf90's MANEIG for the loaded block: EVAL(NLEV), EVEC(NCF, NLEV) of
   ! the levels ICCMIN(1:NLEV) (H - EAV); on entry, with WARM, the previous
   ! SCF iteration's values for those levels (the start vectors).
   subroutine gci_mcdhf_maneig(nvect, nlev, iccmin, warm, eval, evec, nloops, nmv, ierr)
      integer, intent(in) :: nvect, nlev
      integer, intent(in) :: iccmin(nlev)
      logical, intent(in) :: warm
      real(c_double), intent(inout) :: eval(nlev)
      real(c_double), intent(inout) :: evec(hmat_n, nlev)
      integer, intent(out) :: nloops, nmv, ierr
      type(gci_dvd_stats) :: st
      integer(c_int32_t) :: lev(nlev)
      integer(c_int) :: iw
      lev = int(iccmin, c_int32_t)
      iw = 0
      if (warm) iw = 1
      call gci_check(gci_maneig_mcdhf(hmat, int(nvect, c_int), int(nlev, c_int), lev, iw, eval, evec, st), &
                     'GCI_MCDHF_MANEIG')
      nloops = st%nloops
      nmv = st%nmv
      ierr = st%ierr
   end subroutine gci_mcdhf_maneig

   ! YZK for NPAIRS pairs at once: YK(1:NGRID, q) and MTP(q) as YZK(K(q),
   ! I(q), J(q)) leaves TB and MTP in module tatb_C.  Grid and orbitals as
   ! for GCI_GENINTRK.
   subroutine gci_yzk_batch(ngrid, h, r, rp, rpor, nw, ld, pf, qf, mf, nak, nkl, nkj, npairs, k, i, j, yk, mtp)
      integer, intent(in) :: ngrid, nw, ld, npairs
      real(c_double), intent(in) :: h
      real(c_double), intent(in), target :: r(*), rp(*), rpor(*)
      real(c_double), intent(in), target :: pf(ld, nw), qf(ld, nw)
      integer, intent(in) :: mf(nw), nak(nw), nkl(nw), nkj(nw)
      integer, intent(in) :: k(npairs), i(npairs), j(npairs)
      real(c_double), intent(out) :: yk(ngrid, npairs)
      integer, intent(out) :: mtp(npairs)
      type(gci_grid) :: g
      type(gci_orbitals) :: o
      integer(c_int32_t), allocatable, target :: imf(:), inak(:), inkl(:), inkj(:)
      integer(c_int32_t) :: kk(npairs), ii(npairs), jj(npairs), mm(npairs)
      allocate (imf(nw), inak(nw), inkl(nw), inkj(nw))
      imf = int(mf, c_int32_t)
      inak = int(nak, c_int32_t)
      inkl = int(nkl, c_int32_t)
      inkj = int(nkj, c_int32_t)
      g%n = int(ngrid, c_int)
      g%h = h
      g%accy = h**6
      g%r = c_loc(r)
      g%rp = c_loc(rp)
      g%rpor = c_loc(rpor)
      o%nw = int(nw, c_int)
      o%ld = int(ld, c_int)
      o%p = c_loc(pf)
      o%q = c_loc(qf)
      o%mf = c_loc(imf)
      o%nak = c_loc(inak)
      o%nkl = c_loc(inkl)
      o%nkj = c_loc(inkj)
      kk = int(k, c_int32_t)
      ii = int(i, c_int32_t)
      jj = int(j, c_int32_t)
      call gci_check(gci_yzk(g, o, int(npairs, c_int), kk, ii, jj, yk, mm), 'GCI_YZK_BATCH')
      mtp = int(mm)
   end subroutine gci_yzk_batch

   ! GENINTRK: every R^k(ab;cd) the orbitals can form, in the reference's
   ! order, as packed keys INDTEIRK and values VALTEIRK (1..NINT), with the
   ! per-k starts KSTART(0:J2MAX+1).  Grid: R, RP, RPOR on NGRID + 10
   ! points, step H (ACCY = H**6 as GETCID sets it); orbitals PF/QF(LD, NW),
   ! MF, NAK, NKL, NKJ as in module orb_C / wave_C.
   subroutine gci_genintrk(ngrid, h, r, rp, rpor, nw, ld, pf, qf, mf, nak, nkl, nkj, nint, j2max, kstart, &
                           indteirk, valteirk)
      integer, intent(in) :: ngrid, nw, ld
      real(c_double), intent(in) :: h
      real(c_double), intent(in), target :: r(*), rp(*), rpor(*)
      real(c_double), intent(in), target :: pf(ld, nw), qf(ld, nw)
      integer, intent(in) :: mf(nw), nak(nw), nkl(nw), nkj(nw)
      integer, intent(out) :: nint, j2max
      integer, intent(out) :: kstart(0:)
      integer, allocatable, intent(out) :: indteirk(:)
      real(c_double), allocatable, intent(out) :: valteirk(:)
      type(gci_grid) :: g
      type(gci_orbitals) :: o
      integer(c_int32_t), allocatable, target :: imf(:), inak(:), inkl(:), inkj(:), keys(:)
      real(c_double), allocatable, target :: vals(:)
      integer(c_int32_t) :: j2, ks(0:31)
      integer(c_int64_t) :: cnt
      allocate (imf(nw), inak(nw), inkl(nw), inkj(nw))
      imf = int(mf, c_int32_t)
      inak = int(nak, c_int32_t)
      inkl = int(nkl, c_int32_t)
      inkj = int(nkj, c_int32_t)
      g%n = int(ngrid, c_int)
      g%h = h
      g%accy = h**6
      g%r = c_loc(r(1))
      g%rp = c_loc(rp(1))
      g%rpor = c_loc(rpor(1))
      o%nw = int(nw, c_int)
      o%ld = int(ld, c_int)
      o%p = c_loc(pf(1, 1))
      o%q = c_loc(qf(1, 1))
      o%mf = c_loc(imf(1))
      o%nak = c_loc(inak(1))
      o%nkl = c_loc(inkl(1))
      o%nkj = c_loc(inkj(1))
      ks = 0
      cnt = 0
      call gci_check(gci_rk_table(g, o, j2, ks, c_null_ptr, c_null_ptr, 0_c_int64_t, cnt), 'GCI_GENINTRK')
      allocate (keys(max(cnt, 1_c_int64_t)), vals(max(cnt, 1_c_int64_t)))
      call gci_check(gci_rk_table(g, o, j2, ks, c_loc(keys(1)), c_loc(vals(1)), cnt, cnt), 'GCI_GENINTRK')
      nint = int(cnt)
      j2max = int(j2)
      if (ubound(kstart, 1) < j2max + 1) error stop 'GCI_GENINTRK: KSTART too short'
      kstart = 0
      kstart(0:j2max + 1) = int(ks(0:j2max + 1))
      allocate (indteirk(nint), valteirk(nint))
      indteirk = int(keys(1:nint))
      valteirk = vals(1:nint)
   end subroutine gci_genintrk

end module gci_rci
