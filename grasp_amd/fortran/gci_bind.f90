!-----------------------------------------------------------------------
! gci_bind.f90 -- ISO_C_BINDING interfaces to libgraspci.so (graspci.h).
!
! One interface per C entry point, with the C argument types spelled out
! (value scalars, assumed-size arrays, c_ptr for opaque handles), and
! bind(C) derived types laid out like the C structs.  Status codes are
! returned as integer(c_int); gci_message() turns the thread's last error
! string into a Fortran character value.
!-----------------------------------------------------------------------
module gci_bind
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: GCI_OK = 0, GCI_ERR_ARG = -1, GCI_ERR_DEVICE = -2, &
                                GCI_ERR_NOMEM = -3, GCI_ERR_NOTFOUND = -4, GCI_ERR_NOCONV = -5, &
                                GCI_ERR_BUSY = -6
   integer(c_int), parameter :: GCI_FP64 = 0, GCI_SPLIT32 = 1
   integer(c_int), parameter :: GCI_OP_FP64 = 0, GCI_OP_SPLIT_BULK = 1, GCI_OP_SPLIT_FULL = 2

   type, bind(C) :: gci_grid
      integer(c_int) :: n
      real(c_double) :: h, accy
      type(c_ptr) :: r, rp, rpor
   end type gci_grid

   type, bind(C) :: gci_orbitals
      integer(c_int) :: nw, ld
      type(c_ptr) :: p, q, mf, nak, nkl, nkj
   end type gci_orbitals

   type, bind(C) :: gci_pair_lists
      integer(c_int64_t) :: npairs
      type(c_ptr) :: pair_ic, pair_ir, one_ia, one_ib, tsh_off, tshell
      integer(c_int64_t) :: ntshell
      type(c_ptr) :: tb_beg, tb_end, label, coeff
      integer(c_int64_t) :: nterms
   end type gci_pair_lists

   type, bind(C) :: gci_breit_lists
      type(c_ptr) :: bt_beg, bt_end, label, coeff
      integer(c_int64_t) :: nterms
   end type gci_breit_lists

   type, bind(C) :: gci_dvd_params
      integer(c_int) :: nvec, lim, mblock, maxiter
      real(c_double) :: crite, critc, critr
      integer(c_int) :: niv, mode, refine
   end type gci_dvd_params

   type, bind(C) :: gci_dvd_stats
      integer(c_int) :: nloops, nmv, ierr, refine_loops
      real(c_double) :: seconds
   end type gci_dvd_stats

   type, bind(C) :: gci_spmm_shape
      integer(c_int) :: r, block, nslice, depth
      integer(c_int64_t) :: target
   end type gci_spmm_shape

   interface
      function gci_last_error() bind(C, name="gci_last_error") result(p)
         import :: c_ptr
         type(c_ptr) :: p
      end function gci_last_error

      function gci_init(device) bind(C, name="gci_init") result(st)
         import :: c_int
         integer(c_int), value :: device
         integer(c_int) :: st
      end function gci_init

      ! page-locked host memory for result buffers (hipHostMalloc / hipHostFree)
      function gci_host_alloc(bytes, ptr) bind(C, name="gci_host_alloc") result(st)
         import :: c_int, c_size_t, c_ptr
         integer(c_size_t), value :: bytes
         type(c_ptr) :: ptr
         integer(c_int) :: st
      end function gci_host_alloc

      function gci_host_free(ptr) bind(C, name="gci_host_free") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: ptr
         integer(c_int) :: st
      end function gci_host_free

      function gci_radgrd(n, rnt, h, r, rp, rpor) bind(C, name="gci_radgrd") result(st)
         import :: c_int, c_double
         integer(c_int), value :: n
         real(c_double), value :: rnt, h
         real(c_double) :: r(*), rp(*), rpor(*)
         integer(c_int) :: st
      end function gci_radgrd

      function gci_rk_table(grid, orb, j2max, kstart, keys, vals, cap, n) &
         bind(C, name="gci_rk_table") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_ptr, gci_grid, gci_orbitals
         type(gci_grid), intent(in) :: grid
         type(gci_orbitals), intent(in) :: orb
         integer(c_int32_t) :: j2max
         integer(c_int32_t) :: kstart(*)
         type(c_ptr), value :: keys, vals
         integer(c_int64_t), value :: cap
         integer(c_int64_t) :: n
         integer(c_int) :: st
      end function gci_rk_table

      function gci_yzk(grid, orb, npairs, k, b, d, yk, mtp) bind(C, name="gci_yzk") result(st)
         import :: c_int, c_int32_t, c_double, gci_grid, gci_orbitals
         type(gci_grid), intent(in) :: grid
         type(gci_orbitals), intent(in) :: orb
         integer(c_int), value :: npairs
         integer(c_int32_t), intent(in) :: k(*), b(*), d(*)
         real(c_double) :: yk(*)
         integer(c_int32_t) :: mtp(*)
         integer(c_int) :: st
      end function gci_yzk

      function gci_one_body(grid, orb, zz, c, iab) bind(C, name="gci_one_body") result(st)
         import :: c_int, c_double, gci_grid, gci_orbitals
         type(gci_grid), intent(in) :: grid
         type(gci_orbitals), intent(in) :: orb
         real(c_double) :: zz(*)
         real(c_double), value :: c
         real(c_double) :: iab(*)
         integer(c_int) :: st
      end function gci_one_body

      function gci_h_assemble(ncf, nw, lists, iab, kstart, keys, vals, nkeys, lfordr, iccut, &
                              precision, out, eav, nelmnt) bind(C, name="gci_h_assemble") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, gci_pair_lists
         integer(c_int), value :: ncf, nw
         type(gci_pair_lists), intent(in) :: lists
         real(c_double) :: iab(*)
         integer(c_int32_t) :: kstart(*), keys(*)
         real(c_double) :: vals(*)
         integer(c_int64_t), value :: nkeys
         integer(c_int), value :: lfordr, iccut, precision
         type(c_ptr) :: out
         real(c_double) :: eav
         integer(c_int64_t) :: nelmnt
         integer(c_int) :: st
      end function gci_h_assemble

      function gci_breit_create(grid, orb, e, c, wfact, out) bind(C, name="gci_breit_create") result(st)
         import :: c_int, c_double, c_ptr, gci_grid, gci_orbitals
         type(gci_grid), intent(in) :: grid
         type(gci_orbitals), intent(in) :: orb
         real(c_double), intent(in) :: e(*)
         real(c_double), value :: c, wfact
         type(c_ptr) :: out
         integer(c_int) :: st
      end function gci_breit_create

      function gci_breit_table(b, itype, j2max, kstart, keys, vals, cap, n) &
         bind(C, name="gci_breit_table") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_ptr
         type(c_ptr), value :: b
         integer(c_int), value :: itype
         integer(c_int32_t) :: j2max
         integer(c_int32_t) :: kstart(*)
         type(c_ptr), value :: keys, vals
         integer(c_int64_t), value :: cap
         integer(c_int64_t) :: n
         integer(c_int) :: st
      end function gci_breit_table

      function gci_breit_integral(b, itype, ia, ib, ic, id, k, out) bind(C, name="gci_breit_integral") result(st)
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: b
         integer(c_int), value :: itype, ia, ib, ic, id, k
         real(c_double) :: out
         integer(c_int) :: st
      end function gci_breit_integral

      function gci_breit_free(b) bind(C, name="gci_breit_free") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: b
         integer(c_int) :: st
      end function gci_breit_free

      function gci_h_assemble_tb(ncf, nw, lists, breit, br, iab, kstart, keys, vals, nkeys, lfordr, iccut, &
                                 precision, out, eav, elsto, nelmnt) bind(C, name="gci_h_assemble_tb") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, gci_pair_lists, gci_breit_lists
         integer(c_int), value :: ncf, nw
         type(gci_pair_lists), intent(in) :: lists
         type(gci_breit_lists), intent(in) :: breit
         type(c_ptr), value :: br
         real(c_double) :: iab(*)
         integer(c_int32_t) :: kstart(*), keys(*)
         real(c_double) :: vals(*)
         integer(c_int64_t), value :: nkeys
         integer(c_int), value :: lfordr, iccut, precision
         type(c_ptr) :: out
         real(c_double) :: eav, elsto
         integer(c_int64_t) :: nelmnt
         integer(c_int) :: st
      end function gci_h_assemble_tb

      function gci_hstream_begin(ncf, nw, iab, kstart, keys, vals, nkeys, lfordr, iccut, precision, br, out) &
         bind(C, name="gci_hstream_begin") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         integer(c_int), value :: ncf, nw
         real(c_double) :: iab(*)
         integer(c_int32_t) :: kstart(*), keys(*)
         real(c_double) :: vals(*)
         integer(c_int64_t), value :: nkeys
         integer(c_int), value :: lfordr, iccut, precision
         type(c_ptr), value :: br
         type(c_ptr) :: out
         integer(c_int) :: st
      end function gci_hstream_begin

      function gci_hstream_push(s, chunk, breit_chunk) bind(C, name="gci_hstream_push") result(st)
         import :: c_int, c_ptr, gci_pair_lists
         type(c_ptr), value :: s
         type(gci_pair_lists), intent(in) :: chunk
         type(c_ptr), value :: breit_chunk
         integer(c_int) :: st
      end function gci_hstream_push

      function gci_hstream_end(s, out, eav, elsto, nelmnt) bind(C, name="gci_hstream_end") result(st)
         import :: c_int, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: s
         type(c_ptr) :: out
         real(c_double) :: eav, elsto
         integer(c_int64_t) :: nelmnt
         integer(c_int) :: st
      end function gci_hstream_end

      function gci_hstream_free(s) bind(C, name="gci_hstream_free") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: st
      end function gci_hstream_free

      function gci_hmat_upload(n, iendc, irow, emt, index_base, shift, precision, out) &
         bind(C, name="gci_hmat_upload") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         integer(c_int), value :: n
         integer(c_int64_t) :: iendc(*)
         integer(c_int32_t) :: irow(*)
         real(c_double) :: emt(*)
         integer(c_int), value :: index_base
         real(c_double), value :: shift
         integer(c_int), value :: precision
         type(c_ptr) :: out
         integer(c_int) :: st
      end function gci_hmat_upload

      function gci_hmat_free(h) bind(C, name="gci_hmat_free") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int) :: st
      end function gci_hmat_free

      function gci_hmat_info(h, n, nnz_offdiag, precision) bind(C, name="gci_hmat_info") result(st)
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: h
         integer(c_int64_t) :: n, nnz_offdiag
         integer(c_int) :: precision
         integer(c_int) :: st
      end function gci_hmat_info

      function gci_hmat_download(h, iendc, irow, emt, index_base) bind(C, name="gci_hmat_download") result(st)
         import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: h
         integer(c_int64_t) :: iendc(*)
         integer(c_int32_t) :: irow(*)
         real(c_double) :: emt(*)
         integer(c_int), value :: index_base
         integer(c_int) :: st
      end function gci_hmat_download

      function gci_spmm(h, m, b, c, mode) bind(C, name="gci_spmm") result(st)
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: m
         real(c_double) :: b(*), c(*)
         integer(c_int), value :: mode
         integer(c_int) :: st
      end function gci_spmm

      function gci_spmm_device(h, m, db, ldb, dc, ldc, mode) bind(C, name="gci_spmm_device") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: m
         type(c_ptr), value :: db
         integer(c_int), value :: ldb
         type(c_ptr), value :: dc
         integer(c_int), value :: ldc, mode
         integer(c_int) :: st
      end function gci_spmm_device

      ! the same operator with no device allocation (library-owned scratch)
      function gci_spmm_device_noalloc(h, m, db, ldb, dc, ldc, mode) &
            bind(C, name="gci_spmm_device_noalloc") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: h
         integer(c_int), value :: m
         type(c_ptr), value :: db
         integer(c_int), value :: ldb
         type(c_ptr), value :: dc
         integer(c_int), value :: ldc, mode
         integer(c_int) :: st
      end function gci_spmm_device_noalloc

      function gci_davidson(h, p, eval, evec, resid, stats) bind(C, name="gci_davidson") result(st)
         import :: c_int, c_double, c_ptr, gci_dvd_params, gci_dvd_stats
         type(c_ptr), value :: h
         type(gci_dvd_params), intent(in) :: p
         real(c_double) :: eval(*), evec(*), resid(*)
         type(gci_dvd_stats) :: stats
         integer(c_int) :: st
      end function gci_davidson

      function gci_maneig(h, nvec, eval, evec, stats) bind(C, name="gci_maneig") result(st)
         import :: c_int, c_double, c_ptr, gci_dvd_stats
         type(c_ptr), value :: h
         integer(c_int), value :: nvec
         real(c_double) :: eval(*), evec(*)
         type(gci_dvd_stats) :: stats
         integer(c_int) :: st
      end function gci_maneig

      function gci_maneig_mcdhf(h, nvect, nlev, iccmin, warm, eval, evec, stats) &
         bind(C, name="gci_maneig_mcdhf") result(st)
         import :: c_int, c_int32_t, c_double, c_ptr, gci_dvd_stats
         type(c_ptr), value :: h
         integer(c_int), value :: nvect, nlev, warm
         integer(c_int32_t), intent(in) :: iccmin(*)
         real(c_double) :: eval(*), evec(*)
         type(gci_dvd_stats) :: stats
         integer(c_int) :: st
      end function gci_maneig_mcdhf

      function gci_batch_solve(nblocks, blocks, nev, tol, eav, eval, eval_off, evec, evec_off, stats) &
         bind(C, name="gci_batch_solve") result(st)
         import :: c_int, c_int64_t, c_double, c_ptr, gci_dvd_stats
         integer(c_int), value :: nblocks
         type(c_ptr) :: blocks(*)
         integer(c_int) :: nev(*)
         real(c_double) :: tol(*), eav(*), eval(*)
         integer(c_int64_t) :: eval_off(*)
         real(c_double) :: evec(*)
         integer(c_int64_t) :: evec_off(*)
         type(gci_dvd_stats) :: stats(*)
         integer(c_int) :: st
      end function gci_batch_solve

      function gci_server_create(hbm_budget_bytes, max_active, out) bind(C, name="gci_server_create") result(st)
         import :: c_int, c_size_t, c_ptr
         integer(c_size_t), value :: hbm_budget_bytes
         integer(c_int), value :: max_active
         type(c_ptr) :: out
         integer(c_int) :: st
      end function gci_server_create

      function gci_server_destroy(s) bind(C, name="gci_server_destroy") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: st
      end function gci_server_destroy

      function gci_server_submit(s, h, nev, tol, eav, req) bind(C, name="gci_server_submit") result(st)
         import :: c_int, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: s, h
         integer(c_int), value :: nev
         real(c_double), value :: tol, eav
         integer(c_int64_t) :: req
         integer(c_int) :: st
      end function gci_server_submit

      function gci_server_submit_shared(s, h, nev, tol, eav, req) bind(C, name="gci_server_submit_shared") result(st)
         import :: c_int, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: s, h
         integer(c_int), value :: nev
         real(c_double), value :: tol, eav
         integer(c_int64_t) :: req
         integer(c_int) :: st
      end function gci_server_submit_shared

      function gci_server_poll(s, req, eval, evec, latency_s) bind(C, name="gci_server_poll") result(st)
         import :: c_int, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: s
         integer(c_int64_t), value :: req
         real(c_double) :: eval(*), evec(*)
         real(c_double) :: latency_s
         integer(c_int) :: st
      end function gci_server_poll

      function gci_server_drain(s) bind(C, name="gci_server_drain") result(st)
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: st
      end function gci_server_drain

      function gci_profile_enable(on) bind(C, name="gci_profile_enable") result(st)
         import :: c_int
         integer(c_int), value :: on
         integer(c_int) :: st
      end function gci_profile_enable

      function gci_profile_read(stats) bind(C, name="gci_profile_read") result(st)
         import :: c_int, c_double
         real(c_double) :: stats(*)
         integer(c_int) :: st
      end function gci_profile_read

      function gci_spmm_get_shape(out) bind(C, name="gci_spmm_get_shape") result(st)
         import :: c_int, gci_spmm_shape
         type(gci_spmm_shape) :: out
         integer(c_int) :: st
      end function gci_spmm_get_shape

      function gci_spmm_set_shape(shape) bind(C, name="gci_spmm_set_shape") result(st)
         import :: c_int, gci_spmm_shape
         type(gci_spmm_shape), intent(in) :: shape
         integer(c_int) :: st
      end function gci_spmm_set_shape

      function gci_spmm_layout(h, nseg, nslice, stored) bind(C, name="gci_spmm_layout") result(st)
         import :: c_int, c_int64_t, c_ptr
         type(c_ptr), value :: h
         integer(c_int64_t) :: nseg, nslice, stored
         integer(c_int) :: st
      end function gci_spmm_layout

      function gci_spmm_time(mats, nmats, m, mode, reps, ms) bind(C, name="gci_spmm_time") result(st)
         import :: c_int, c_double, c_ptr
         type(c_ptr) :: mats(*)
         integer(c_int), value :: nmats, m, mode, reps
         real(c_double) :: ms
         integer(c_int) :: st
      end function gci_spmm_time

      function gci_spmm_autotune(mats, nmats, m, mode, reps, best, best_ms) &
         bind(C, name="gci_spmm_autotune") result(st)
         import :: c_int, c_double, c_ptr, gci_spmm_shape
         type(c_ptr) :: mats(*)
         integer(c_int), value :: nmats, m, mode, reps
         type(gci_spmm_shape) :: best
         real(c_double) :: best_ms
         integer(c_int) :: st
      end function gci_spmm_autotune

      function c_strlen(s) bind(C, name="strlen") result(n)
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
         integer(c_size_t) :: n
      end function c_strlen
   end interface

contains

   ! The calling thread's last error message.
   function gci_message() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = gci_last_error()
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate (character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function gci_message

end module gci_bind
