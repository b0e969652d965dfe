! rang_driver.f90 -- writes the angular-coefficient lists the rci90 element
! loop consumes, produced by the reference's own librang90.
!
! TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_rang.sh together with
! the GRASP2018 sources where they lie (libmod, lib9290, librang90 and
! rci90's TALK/CXK) into oracle/_ref/rang_lists.  It contains no angular
! algebra of its own: it loads a CSF list the way rci90 does (SETCSL,
! rci90/setcsl.f90; FACTT, rci90/rci92.f90:189; LODCSH2 per block,
! rci90/matrix.f90:84-88) and, for every pair IR <= IC of one block, makes
! the two calls of the SETHAM element loop (rci90/setham_gg.f90:213-292):
! ONESCALAR(IC, IR, IA, IB, TSHELL) and RKCO_GG(IC, IR, CORD, INCOR=1, 1),
! whose TALK/SPEAK callbacks fill LABEL/COEFF (lib9290/speak.f90:46-59).
! With a fourth argument 1 it also makes SETHAM's transverse Breit call
! RKCO_GG(IC, IR, BREID, 1, 2) (setham_gg.f90:298-302), after setting
! ICORE for the block as SETHAM does (setham_gg.f90:145-156).
!
! Usage: rang_lists <file.c> <block> <out.bin> [breit]
! Output (stream, native byte order):
!   int32 NW, NCF, NCORE, BREIT
!   NW x int32 (NP, NAK, NKL, NKJ)
!   per pair (IC ascending, IR = 1..IC):
!     int32 IC, IR, IA, IB, NT   (NT = NW if IA == IB /= 0, 1 if IA /= IB, 0 if IA == 0)
!     NT x real64 TSHELL(1:NT)
!     int32 NVCOEF, then NVCOEF x (5 x int32 LABEL(1:5, i), real64 COEFF(i))
!     if BREIT: int32 NVCOEF, then NVCOEF x (6 x int32 LABEL(1:6, i), real64 COEFF(i))
program rang_lists
   use vast_kind_param, only: DOUBLE
   use parameter_def, only: NNNW
   use memory_man
   use debug_C, only: IBUG1, IBUG2, IBUG3, IBUG4, IBUG5, IBUG6
   use hblock_C, only: NBLOCK, NCFBLK
   use orb_C, only: NCF, NW, IQA, NP, NAK, NKL, NKJ
   use stat_C, only: JQSA, JCUPA
   use buffer_C, only: NVCOEF, LABEL, COEFF
   use setcon_I
   use setcsll_I
   use lodcsh_I
   use lodcsh2_I
   use factt_I
   use alcbuf_I
   use onescalar_I
   use rkco_gg_I
   use cord_I
   use breid_I
   use bcore_C, only: ICORE
   use ichop_I
   implicit none
   integer, parameter :: NBLKIN = 100
   character(len=512) :: cfile, ofile, arg
   character(len=8) :: idblk(NBLKIN)
   integer :: jb, ncore, ncftot, ic, ir, ia, ib, i, nt, u, iblk, ibreit, j
   integer :: ncf_of(NBLKIN)
   real(DOUBLE), allocatable :: tshell(:)

   call get_command_argument(1, cfile)
   call get_command_argument(2, arg)
   read (arg, *) jb
   call get_command_argument(3, ofile)
   ibreit = 0
   if (command_argument_count() >= 4) then
      call get_command_argument(4, arg)
      read (arg, *) ibreit
   end if
   IBUG1 = 0; IBUG2 = 0; IBUG3 = 0; IBUG4 = 0; IBUG5 = 0; IBUG6 = 0

   ! SETMC (machine constants from LAPACK's DLAMCH, absent from the image)
   ! is not called: no routine on the angular path reads its constants
   ! (only CGAMMA and RADGRD do)
   call SETCON
   call ALLOC (NCFBLK, 0, NBLKIN, 'NCFBLK', 'RANG_LISTS')
   ! block sizes into a plain array (NCFBLK is a pointer array, which may not
   ! be passed element-wise to SETCSLL's assumed-size argument)
   call SETCSLL (21, cfile, NBLKIN, NBLOCK, ncf_of, ncftot, idblk)
   NCFBLK(1:NBLOCK) = ncf_of(1:NBLOCK)
   rewind (21)
   read (21, *)
   call LODCSH (21, ncore)
   call FACTT
   if (jb < 1 .or. jb > NBLOCK) stop 'rang_lists: block out of range'
   do iblk = 1, jb
      NCF = NCFBLK(iblk)
      call ALLOC (IQA, NNNW, NCF, 'IQA', 'RANG_LISTS')
      call ALLOC (JQSA, NNNW, 3, NCF, 'JQSA', 'RANG_LISTS')
      call ALLOC (JCUPA, NNNW, NCF, 'JCUPA', 'RANG_LISTS')
      call LODCSH2 (21, ncore, iblk)
      if (iblk < jb) then
         call DALLOC (IQA, 'IQA', 'RANG_LISTS')
         call DALLOC (JQSA, 'JQSA', 'RANG_LISTS')
         call DALLOC (JCUPA, 'JCUPA', 'RANG_LISTS')
      end if
   end do
   call ALCBUF (1)
   allocate (tshell(NW))
   if (ibreit /= 0) then
      ! setham_gg.f90:145-156
      do i = 1, NW
         ICORE(i) = 0
         do j = 1, NCF
            if (ICHOP(i, j) <= 0) goto 2
         end do
         ICORE(i) = 1
2        continue
      end do
   end if

   open (newunit=u, file=ofile, access='stream', form='unformatted', status='replace')
   write (u) NW, NCF, ncore, ibreit
   write (u) (NP(i), NAK(i), NKL(i), NKJ(i), i = 1, NW)
   do ic = 1, NCF
      do ir = 1, ic
         tshell = 0.0d0
         call ONESCALAR (ic, ir, ia, ib, tshell)
         nt = 0
         if (ia /= 0) then
            if (ia == ib) then
               nt = NW
            else
               nt = 1
            end if
         end if
         write (u) ic, ir, ia, ib, nt
         if (nt > 0) write (u) tshell(1:nt)
         NVCOEF = 0
         call RKCO_GG (ic, ir, CORD, 1, 1)
         write (u) NVCOEF
         do i = 1, NVCOEF
            write (u) LABEL(1:5, i), COEFF(i)
         end do
         if (ibreit /= 0) then
            NVCOEF = 0
            call RKCO_GG (ic, ir, BREID, 1, 2)
            write (u) NVCOEF
            do i = 1, NVCOEF
               write (u) LABEL(1:6, i), COEFF(i)
            end do
         end if
      end do
   end do
   close (u)
end program rang_lists
