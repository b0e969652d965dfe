!-----------------------------------------------------------------------
! gci_drive.f90 -- a Fortran host calling the engine through gci_rci.
!
! Usage: gci_drive <input> <output>
! Input (list-directed):  NCF NVEC M
!                         EAV
!                         IENDC(0:NCF)
!                         IROW(1:NELMNT)
!                         EMT(1:NELMNT)
!                         NGRID NW LD H          (NW = 0: no orbital section)
!                         R, RP, RPOR            (NGRID + 10 each)
!                         PF(LD, NW), QF(LD, NW)
!                         MF, NAK, NKL, NKJ      (NW each)
! Output: C = (H - EAV) B for B(i, k) = sin(0.37 i + 1.3 k), k = 1..M
!         (column-major), the NVEC lowest eigenvalues of H - EAV with
!         NLOOPS NMV IERR, then NINT J2MAX, KSTART(0:J2MAX+1) and the
!         (INDTEIRK, VALTEIRK) pairs.
!-----------------------------------------------------------------------
program gci_drive
   use, intrinsic :: iso_c_binding
   use gci_bind, only: GCI_FP64
   use gci_rci
   implicit none
   character(len=4096) :: fin, fout
   integer :: ncf, nvec, m, i, k, nloops, nmv, ierr, u, v
   integer(c_int64_t) :: nel
   real(c_double) :: eav
   integer, allocatable :: iendc(:), irow(:)
   real(c_double), allocatable :: emt(:), b(:, :), c(:, :), eval(:), evec(:, :)
   integer :: ngrid, nw, ld, nint, j2max
   real(c_double) :: h
   real(c_double), allocatable :: r(:), rp(:), rpor(:), pf(:, :), qf(:, :), valteirk(:)
   integer, allocatable :: mf(:), nak(:), nkl(:), nkj(:), indteirk(:)
   integer :: kstart(0:20)

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (newunit=u, file=trim(fin), status='old', action='read')
   read (u, *) ncf, nvec, m
   read (u, *) eav
   allocate (iendc(0:ncf))
   read (u, *) iendc
   nel = iendc(ncf)
   allocate (irow(nel), emt(nel))
   read (u, *) irow
   read (u, *) emt
   read (u, *) ngrid, nw, ld, h
   if (nw > 0) then
      allocate (r(ngrid + 10), rp(ngrid + 10), rpor(ngrid + 10), pf(ld, nw), qf(ld, nw))
      allocate (mf(nw), nak(nw), nkl(nw), nkj(nw))
      read (u, *) r
      read (u, *) rp
      read (u, *) rpor
      read (u, *) pf
      read (u, *) qf
      read (u, *) mf
      read (u, *) nak
      read (u, *) nkl
      read (u, *) nkj
   end if
   close (u)

   call gci_start(0)
   call gci_load_hmat(ncf, iendc, irow, emt, eav, GCI_FP64)

   allocate (b(ncf, m), c(ncf, m))
   do k = 1, m
      do i = 1, ncf
         b(i, k) = sin(0.37d0 * i + 1.3d0 * k)
      end do
   end do
   call gci_spicmv2(ncf, m, b, c)

   allocate (eval(nvec), evec(ncf, nvec))
   call gci_maneig_solve(nvec, eval, evec, nloops, nmv, ierr)

   open (newunit=v, file=trim(fout), status='replace', action='write')
   write (v, '(es25.17e3)') c
   write (v, '(es25.17e3)') eval
   write (v, '(3i12)') nloops, nmv, ierr
   if (nw > 0) then
      call gci_genintrk(ngrid, h, r, rp, rpor, nw, ld, pf, qf, mf, nak, nkl, nkj, nint, j2max, kstart, &
                        indteirk, valteirk)
      write (v, '(2i12)') nint, j2max
      write (v, '(i12)') kstart(0:j2max + 1)
      do i = 1, nint
         write (v, '(i12, 1x, es25.17e3)') indteirk(i), valteirk(i)
      end do
   end if
   close (v)
   call gci_free_hmat()
end program gci_drive
