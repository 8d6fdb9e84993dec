! dropin_check -- a Fortran caller of the device library built only from this
! repository's sources: the kinds module beside it, the ISO_C_BINDING module
! (mixsolver_c.f90) and the drop-in solver modules (mod_uns_linsolver.f90,
! mod_uns_linsolver_mpi.f90).  The GPU tests drive it with the committed
! fixtures:
!
!   dropin_check solve <bicg|cgj|cgic> <0|1> in.bin out.bin
!       one solve through the serial drop-in (0: bicgstab_ilu0, cg_jacobi,
!       cg_ic0) or the one-partition _mpi drop-in (1: bicgstab_ilu0_mpi,
!       cg_jacobi_mpi, cg_ic0_mpi).  in.bin: n nnz (int32), row_ptr col_idx
!       (int32, 1-based), val b x0 (real64), tol (real64), maxit (int32).
!       out.bin: it ierr (int32), res (real64), x (real64).
!   dropin_check piso in.bin out.bin
!       one loop-level run: ms_system_create, initial u p T, ms_piso_run,
!       fields back.  in.bin: nnodes nfaces ncells nbc nsteps (int32), x
!       (3,nnodes), fptr (int64, nfaces+1, 0-based), fnodes c0 c1 zone
!       (int32), the ms_ctrl_t and nbc ms_bcgroup_t records, u (3,ncells), p,
!       T.  out.bin: ier (int32), u p flux T (real64).
program dropin_check
   use iso_c_binding
   use mod_precision, only: dp
   use mixsolver_c
   use mod_uns_linsolver, only: csr_t, bicgstab_ilu0, cg_jacobi, cg_ic0
   use mod_uns_linsolver_mpi, only: bicgstab_ilu0_mpi, cg_jacobi_mpi, cg_ic0_mpi
   implicit none
   character(len=512) :: mode, a2, a3, a4, a5
   call get_command_argument(1, mode)
   if ( ms_init(0_c_int) /= MS_OK ) then
      print '(a)', 'dropin_check: ms_init failed'
      error stop 2
   end if
   select case ( trim(mode) )
   case ( 'solve' )
      call get_command_argument(2, a2)
      call get_command_argument(3, a3)
      call get_command_argument(4, a4)
      call get_command_argument(5, a5)
      call run_solve( trim(a2), trim(a3) == '1', trim(a4), trim(a5) )
   case ( 'piso' )
      call get_command_argument(2, a2)
      call get_command_argument(3, a3)
      call run_piso( trim(a2), trim(a3) )
   case default
      print '(a)', 'usage: dropin_check solve <bicg|cgj|cgic> <0|1> in out | piso in out'
      error stop 2
   end select

contains

   subroutine run_solve( which, mpi, fin, fout )
      character(len=*), intent(in) :: which, fin, fout
      logical, intent(in) :: mpi
      type(csr_t) :: A
      integer(c_int) :: n, nnz, maxit
      real(dp) :: tol, res
      real(dp), allocatable :: b(:), x(:)
      integer :: it, ierr, u
      integer :: halo   ! the halo argument: never read at one partition
      open(newunit=u, file=fin, access='stream', form='unformatted', status='old')
      read(u) n, nnz
      A%nrows = n
      allocate(A%row_ptr(n+1), A%col_idx(nnz), A%val(nnz), b(n), x(n))
      read(u) A%row_ptr, A%col_idx, A%val, b, x, tol, maxit
      close(u)
      halo = 0
      select case ( which )
      case ( 'bicg' )
         if ( mpi ) then
            call bicgstab_ilu0_mpi( A, b, x, n, halo, tol, maxit, it, res, ierr )
         else
            call bicgstab_ilu0( A, b, x, tol, maxit, it, res, ierr )
         end if
      case ( 'cgj' )
         if ( mpi ) then
            call cg_jacobi_mpi( A, b, x, n, halo, tol, maxit, it, res, ierr )
         else
            call cg_jacobi( A, b, x, tol, maxit, it, res, ierr )
         end if
      case default
         if ( mpi ) then
            call cg_ic0_mpi( A, b, x, n, halo, tol, maxit, it, res, ierr )
         else
            call cg_ic0( A, b, x, tol, maxit, it, res, ierr )
         end if
      end select
      open(newunit=u, file=fout, access='stream', form='unformatted', status='replace')
      write(u) int(it, c_int), int(ierr, c_int), res, x
      close(u)
   end subroutine run_solve

   subroutine run_piso( fin, fout )
      character(len=*), intent(in) :: fin, fout
      integer(c_int) :: nnodes, nfaces, ncells, nbc, nsteps, ier, rc
      real(c_double), allocatable :: x(:), u(:), p(:), T(:), flux(:)
      integer(c_long_long), allocatable :: fptr(:)
      integer(c_int), allocatable :: fnodes(:), c0(:), c1(:), zone(:)
      type(ms_ctrl_t) :: ctrl
      type(ms_bcgroup_t), allocatable :: groups(:)
      type(c_ptr) :: s
      integer :: un
      integer(c_int8_t), allocatable :: raw(:)
      open(newunit=un, file=fin, access='stream', form='unformatted', status='old')
      read(un) nnodes, nfaces, ncells, nbc, nsteps
      allocate(x(3*nnodes), fptr(nfaces+1), c0(nfaces), c1(nfaces), zone(nfaces))
      read(un) x, fptr
      allocate(fnodes(fptr(nfaces+1)), groups(max(nbc, 1)))
      read(un) fnodes, c0, c1, zone
      ! The records are C struct images, padding included.  Unformatted I/O of a
      ! derived type transfers its components only (no padding), so the bytes
      ! are read as they lie and copied over the interoperable types.
      allocate(raw(c_sizeof(ctrl)))
      read(un) raw
      ctrl = transfer(raw, ctrl)
      deallocate(raw)
      if ( nbc > 0 ) then
         allocate(raw(c_sizeof(groups(1)) * nbc))
         read(un) raw
         groups(1:nbc) = transfer(raw, groups(1:nbc), nbc)
      end if
      allocate(u(3*ncells), p(ncells), T(ncells), flux(nfaces))
      read(un) u, p, T
      close(un)
      rc = ms_system_create(s, nnodes, x, nfaces, fptr, fnodes, c0, c1, zone, ncells, ctrl, groups)
      call check(rc, 'ms_system_create')
      call check(ms_set_field(s, c_str('u'), u, int(size(u), c_long_long)), 'set u')
      call check(ms_set_field(s, c_str('p'), p, int(size(p), c_long_long)), 'set p')
      call check(ms_set_field(s, c_str('T'), T, int(size(T), c_long_long)), 'set T')
      call check(ms_piso_run(s, nsteps, ier), 'ms_piso_run')
      call check(ms_get_field(s, c_str('u'), u, int(size(u), c_long_long)), 'get u')
      call check(ms_get_field(s, c_str('p'), p, int(size(p), c_long_long)), 'get p')
      call check(ms_get_field(s, c_str('flux'), flux, int(size(flux), c_long_long)), 'get flux')
      call check(ms_get_field(s, c_str('T'), T, int(size(T), c_long_long)), 'get T')
      call check(ms_system_destroy(s), 'ms_system_destroy')
      open(newunit=un, file=fout, access='stream', form='unformatted', status='replace')
      write(un) ier, u, p, flux, T
      close(un)
   end subroutine run_piso

   subroutine check( rc, what )
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if ( rc /= MS_OK ) then
         print '(a,a,i0,a,a)', what, ' failed, code ', rc, ': ', last_error()
         error stop 3
      end if
   end subroutine check

end program dropin_check
