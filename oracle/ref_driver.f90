! ref_driver.f90 -- C-callable harness around the reference's own routines.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_ref.sh together with
! the GRASP2018 sources where they lie (radial grid, quadrature, YZK,
! SLATER, RINTI, NUCPOT, GENINTRK, RKINTC, SPICMV2), into oracle/_ref/.
! It only loads module state from C arrays and calls those routines; it
! contains no numerics of its own.  Used by tests to pin the C oracle.
module ref_driver
   use iso_c_binding
   implicit none
contains

   subroutine ref_init(n, rnt, h, cl, zin, nparm_in, parm_in) bind(C, name="ref_init")
      use vast_kind_param, only: DOUBLE
      use grid_C, only: gn => N, grnt => RNT, gh => H, ghp => HP
      use def_C, only: ACCY, C, Z, PI
      use npar_C, only: NPARM, PARM
      integer(c_int), value :: n, nparm_in
      real(c_double), value :: rnt, h, cl, zin
      real(c_double), intent(in) :: parm_in(2)
      interface
         subroutine setqic()
         end subroutine
         subroutine radgrd()
         end subroutine
         subroutine nucpot()
         end subroutine
      end interface
      gn = n
      grnt = rnt
      gh = h
      ghp = 0.0d0
      ACCY = h**6
      C = cl
      Z = zin
      PI = 4.0d0*atan(1.0d0)
      NPARM = nparm_in
      PARM(1) = parm_in(1)
      PARM(2) = parm_in(2)
      call setqic()
      call radgrd()
      call nucpot()
   end subroutine ref_init

   subroutine ref_grid(r_out, rp_out, rpor_out, zz_out) bind(C, name="ref_grid")
      use grid_C, only: N, R, RP, RPOR
      use npot_C, only: ZZ
      real(c_double), intent(out) :: r_out(*), rp_out(*), rpor_out(*), zz_out(*)
      r_out(1:N+10) = R(1:N+10)
      rp_out(1:N+10) = RP(1:N+10)
      rpor_out(1:N+10) = RPOR(1:N+10)
      zz_out(1:N) = ZZ(1:N)
   end subroutine ref_grid

   subroutine ref_set_orbitals(nw_in, ld, p, q, mf_in, nak_in, nkl_in, nkj_in) &
         bind(C, name="ref_set_orbitals")
      use parameter_def, only: NNNP
      use wave_C, only: MF, PF, QF
      use orb_C, only: NW, NAK, NKL, NKJ
      integer(c_int), value :: nw_in, ld
      real(c_double), intent(in) :: p(ld, nw_in), q(ld, nw_in)
      integer(c_int), intent(in) :: mf_in(nw_in), nak_in(nw_in), nkl_in(nw_in), nkj_in(nw_in)
      integer :: j, m
      if (associated(PF)) deallocate(PF)
      if (associated(QF)) deallocate(QF)
      allocate(PF(NNNP, nw_in), QF(NNNP, nw_in))
      PF = 0.0d0
      QF = 0.0d0
      m = min(ld, NNNP)
      do j = 1, nw_in
         PF(1:m, j) = p(1:m, j)
         QF(1:m, j) = q(1:m, j)
         MF(j) = mf_in(j)
         NAK(j) = nak_in(j)
         NKL(j) = nkl_in(j)
         NKJ(j) = nkj_in(j)
      end do
      NW = nw_in
   end subroutine ref_set_orbitals

   function ref_quad(ta_in, mtp_in) result(res) bind(C, name="ref_quad")
      use tatb_C, only: TA, MTP
      real(c_double), intent(in) :: ta_in(*)
      integer(c_int), value :: mtp_in
      real(c_double) :: res
      real(c_double) :: out
      interface
         subroutine quad(result)
            real(kind(0.0d0)), intent(out) :: result
         end subroutine
      end interface
      TA(:) = 0.0d0
      TA(1:mtp_in) = ta_in(1:mtp_in)
      MTP = mtp_in
      call quad(out)
      res = out
   end function ref_quad

   function ref_yzk(k, i, j, yk_out) result(mtp_out) bind(C, name="ref_yzk")
      use grid_C, only: N
      use tatb_C, only: TB, MTP
      integer(c_int), value :: k, i, j
      real(c_double), intent(out) :: yk_out(*)
      integer(c_int) :: mtp_out
      integer :: ii, jj
      interface
         subroutine yzk(k, i, j)
            integer, intent(in) :: k
            integer :: i, j
         end subroutine
      end interface
      ii = i
      jj = j
      call yzk(k, ii, jj)
      yk_out(1:N) = TB(1:N)
      mtp_out = MTP
   end function ref_yzk

   function ref_slater(ia, ib, ic, id, k) result(res) bind(C, name="ref_slater")
      integer(c_int), value :: ia, ib, ic, id, k
      real(c_double) :: res
      integer :: a, b, c, d, kk
      interface
         real(kind(0.0d0)) function slater(ia, ib, ic, id, k)
            integer :: ia, ib, ic, id, k
         end function
      end interface
      a = ia; b = ib; c = ic; d = id; kk = k
      res = slater(a, b, c, d, kk)
   end function ref_slater

   function ref_rinti(j, k, mode) result(res) bind(C, name="ref_rinti")
      integer(c_int), value :: j, k, mode
      real(c_double) :: res
      integer :: jj, kk
      interface
         real(kind(0.0d0)) function rinti(j, k, mode)
            integer :: j, k
            integer, intent(in) :: mode
         end function
      end interface
      jj = j; kk = k
      res = rinti(jj, kk, mode)
   end function ref_rinti

   function ref_genintrk(kstart_out, keys_out, vals_out, cap) result(nout) bind(C, name="ref_genintrk")
      use cteilsrk_C, only: INDTEIRK, VALTEIRK
      use kkstart_C, only: KSTART
      integer(c_int), intent(out) :: kstart_out(*)
      integer(c_int), intent(out) :: keys_out(*)
      real(c_double), intent(out) :: vals_out(*)
      integer(c_int), value :: cap
      integer(c_int) :: nout
      integer :: nn, j2max
      interface
         subroutine genintrk(myid, nprocs, n, j2max)
            integer, intent(in) :: myid, nprocs
            integer, intent(out) :: n, j2max
         end subroutine
      end interface
      call genintrk(0, 1, nn, j2max)
      nout = nn
      kstart_out(1:j2max+2) = KSTART(0:j2max+1)
      if (nn <= cap) then
         keys_out(1:nn) = INDTEIRK(1:nn)
         vals_out(1:nn) = VALTEIRK(1:nn)
      end if
      deallocate(INDTEIRK)
      deallocate(VALTEIRK)
   end function ref_genintrk

   subroutine ref_spicmv2(n, m, iendc_in, irow_in, emt_in, nnz, b, c) bind(C, name="ref_spicmv2")
      use hmat_C, only: EMT, IENDC, IROW
      integer(c_int), value :: n, m
      integer(c_int64_t), value :: nnz
      integer(c_int), intent(in) :: iendc_in(0:n), irow_in(nnz)
      real(c_double), intent(in) :: emt_in(nnz), b(n, m)
      real(c_double), intent(out) :: c(n, m)
      interface
         subroutine spicmv2(n, m, b, c)
            integer, intent(in) :: n, m
            real(kind(0.0d0)), dimension(n, m), intent(in) :: b
            real(kind(0.0d0)), dimension(n, m), intent(inout) :: c
         end subroutine
      end interface
      allocate(EMT(nnz), IROW(nnz), IENDC(0:n))
      EMT = emt_in
      IROW = irow_in
      IENDC = iendc_in
      call spicmv2(n, m, b, c)
      deallocate(EMT, IROW, IENDC)
   end subroutine ref_spicmv2
   ! transverse Breit state: orbital energies E (orb_C) and WFACT (wfac_C);
   ! BESSEL's cache and the BRINT3..6 memo start empty
   subroutine ref_breit_setup(nw_in, e_in, wfact_in) bind(C, name="ref_breit_setup")
      use orb_C, only: E
      use wfac_C, only: WFACT
      use stor_C, only: KEEP
      use bess_C, only: WIJ, BESSJ, BESSN
      use bilst_C, only: FIRST, NTPI, INDTP1, INDTP2, INDTP3, INDTP4, INDTP5, INDTP6, &
                         VALTP1, VALTP2, VALTP3, VALTP4, VALTP5, VALTP6
      integer(c_int), value :: nw_in
      real(c_double), intent(in) :: e_in(nw_in)
      real(c_double), value :: wfact_in
      E(1:nw_in) = e_in(1:nw_in)
      WFACT = wfact_in
      KEEP = 0
      WIJ = 0.0d0
      BESSJ = 0.0d0
      BESSN = 0.0d0
      FIRST = .true.
      NTPI = 0
      if (associated(INDTP1)) deallocate(INDTP1, VALTP1)
      if (associated(INDTP2)) deallocate(INDTP2, VALTP2)
      if (associated(INDTP3)) deallocate(INDTP3, VALTP3)
      if (associated(INDTP4)) deallocate(INDTP4, VALTP4)
      if (associated(INDTP5)) deallocate(INDTP5, VALTP5)
      if (associated(INDTP6)) deallocate(INDTP6, VALTP6)
      nullify(INDTP1, VALTP1, INDTP2, VALTP2, INDTP3, VALTP3, INDTP4, VALTP4, INDTP5, VALTP5, INDTP6, VALTP6)
   end subroutine ref_breit_setup

   ! GENINTBREIT1 (type 1) or GENINTBREIT2 (type 2): the table and its
   ! KSTARTBREIT; the table stays in bilst_C for BRINT1/2
   function ref_genintbreit(itype, kstart_out, keys_out, vals_out, cap) result(nout) &
         bind(C, name="ref_genintbreit")
      use bilst_C, only: INDTP1, VALTP1, INDTP2, VALTP2
      use kkstartbreit_C, only: KSTARTBREIT1, KSTARTBREIT2
      integer(c_int), value :: itype, cap
      integer(c_int), intent(out) :: kstart_out(*)
      integer(c_int), intent(out) :: keys_out(*)
      real(c_double), intent(out) :: vals_out(*)
      integer(c_int) :: nout
      integer :: nn, j2max
      interface
         subroutine genintbreit1(myid, nprocs, nb, j2max)
            integer, intent(in) :: myid, nprocs
            integer :: nb, j2max
         end subroutine
         subroutine genintbreit2(myid, nprocs, nb, j2max)
            integer, intent(in) :: myid, nprocs
            integer :: nb, j2max
         end subroutine
      end interface
      if (itype == 1) then
         call genintbreit1(0, 1, nn, j2max)
         kstart_out(1:j2max+2) = KSTARTBREIT1(0:j2max+1)
         if (nn <= cap) then
            keys_out(1:nn) = INDTP1(1:nn)
            vals_out(1:nn) = VALTP1(1:nn)
         end if
      else
         call genintbreit2(0, 1, nn, j2max)
         kstart_out(1:j2max+2) = KSTARTBREIT2(0:j2max+1)
         if (nn <= cap) then
            keys_out(1:nn) = INDTP2(1:nn)
            vals_out(1:nn) = VALTP2(1:nn)
         end if
      end if
      nout = nn
   end function ref_genintbreit

   ! BRINT1..6 (rci90/brint1.f90 .. brint6.f90) as SETHAM calls them
   function ref_brint(itype, ia, ib, ic, id, k) result(res) bind(C, name="ref_brint")
      integer(c_int), value :: itype, ia, ib, ic, id, k
      real(c_double) :: res
      real(kind(0.0d0)) :: t
      interface
         subroutine brint1(ia, ib, ic, id, k, tegral)
            integer, intent(in) :: ia, ib, ic, id, k
            real(kind(0.0d0)), intent(out) :: tegral
         end subroutine
         subroutine brint2(ia, ib, ic, id, k, tegral)
            integer, intent(in) :: ia, ib, ic, id, k
            real(kind(0.0d0)), intent(out) :: tegral
         end subroutine
         subroutine brint3(ia, ib, ic, id, nu, tegral)
            integer, intent(in) :: ia, ib, ic, id, nu
            real(kind(0.0d0)), intent(out) :: tegral
         end subroutine
         subroutine brint4(ia, ib, ic, id, nu, tegral)
            integer, intent(in) :: ia, ib, ic, id, nu
            real(kind(0.0d0)), intent(out) :: tegral
         end subroutine
         subroutine brint5(ia, ib, ic, id, nu, tegral)
            integer, intent(in) :: ia, ib, ic, id, nu
            real(kind(0.0d0)), intent(out) :: tegral
         end subroutine
         subroutine brint6(ia, ib, ic, id, nu, tegral)
            integer, intent(in) :: ia, ib, ic, id, nu
            real(kind(0.0d0)), intent(out) :: tegral
         end subroutine
      end interface
      select case (itype)
      case (1)
         call brint1(ia, ib, ic, id, k, t)
      case (2)
         call brint2(ia, ib, ic, id, k, t)
      case (3)
         call brint3(ia, ib, ic, id, k, t)
      case (4)
         call brint4(ia, ib, ic, id, k, t)
      case (5)
         call brint5(ia, ib, ic, id, k, t)
      case default
         call brint6(ia, ib, ic, id, k, t)
      end select
      res = t
   end function ref_brint
end module ref_driver
