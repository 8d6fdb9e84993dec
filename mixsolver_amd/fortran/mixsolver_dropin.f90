! mixsolver_dropin.f90 -- the device solve shared by the drop-in solver
! modules (mod_uns_linsolver.f90 and mod_uns_linsolver_mpi.f90 of this
! directory).  One call hands a 1-based CSR system to libmixsolver_hip.so
! and returns the reference's it / res / ierr.  Arrays go through contiguous
! copies, since callers may pass strided sections.
module mixsolver_dropin
   use iso_c_binding, only: c_int, c_double
   use mixsolver_c, only: ms_cg_jacobi, ms_cg_ic0, ms_bicgstab_ilu0, MS_OK
   implicit none
   private
   public :: dropin_solve, KIND_BICG, KIND_CGJ, KIND_CGIC

   integer, parameter :: KIND_BICG = 1, KIND_CGJ = 2, KIND_CGIC = 3
   integer, parameter :: wp = kind(1.0d0)

contains

   ! Solve rows 1..n of the CSR (row_ptr, col_idx, val) for b; x holds the
   ! initial guess on entry.  kind: KIND_BICG (BiCGSTAB + ILU(0)), KIND_CGJ
   ! (PCG + Jacobi), KIND_CGIC (PCG + IC(0)).  A library error (no device,
   ! bad matrix) stops the program: the reference solvers have no status code
   ! for it.
   subroutine dropin_solve( kind, n, row_ptr, col_idx, val, b, x, tol, maxit, it, res, ierr )
      integer,  intent(in)    :: kind, n
      integer,  intent(in)    :: row_ptr(:), col_idx(:)
      real(wp), intent(in)    :: val(:), b(:)
      real(wp), intent(inout) :: x(:)
      real(wp), intent(in)    :: tol
      integer,  intent(in)    :: maxit
      integer,  intent(out)   :: it, ierr
      real(wp), intent(out)   :: res
      real(c_double), allocatable :: xc(:), bc(:), vc(:)
      integer(c_int), allocatable :: rp(:), ci(:)
      integer(c_int) :: rc, it_c, ierr_c
      real(c_double) :: res_c
      integer :: nnz
      nnz = row_ptr(n+1) - 1
      xc = x(1:n)
      bc = b(1:n)
      vc = val(1:nnz)
      rp = row_ptr(1:n+1)
      ci = col_idx(1:nnz)
      select case ( kind )
      case ( KIND_BICG )
         rc = ms_bicgstab_ilu0( int(n, c_int), rp, ci, vc, bc, xc, real(tol, c_double), &
                                int(maxit, c_int), it_c, res_c, ierr_c )
      case ( KIND_CGJ )
         rc = ms_cg_jacobi( int(n, c_int), rp, ci, vc, bc, xc, real(tol, c_double), &
                            int(maxit, c_int), it_c, res_c, ierr_c )
      case default
         rc = ms_cg_ic0( int(n, c_int), rp, ci, vc, bc, xc, real(tol, c_double), &
                         int(maxit, c_int), it_c, res_c, ierr_c )
      end select
      if ( rc /= MS_OK ) then
         write(*,'(a,i0)') 'mixsolver: device solver call failed, code ', rc
         error stop 1
      end if
      x(1:n) = xc
      it = it_c
      res = res_c
      ierr = ierr_c
   end subroutine dropin_solve

end module mixsolver_dropin
