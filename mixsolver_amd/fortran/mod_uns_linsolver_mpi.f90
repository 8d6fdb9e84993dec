!===============================================================================
! mod_uns_linsolver_mpi.f90 -- drop-in replacement of the reference module of
! the same name (src/unstructured/mod_uns_linsolver_mpi.f90, publics at
! :32-33) for a run on ONE partition, the configuration in which the whole
! mesh, its matrices and fields live on one GPU.
!
! Compiled in place of the reference file, it keeps every public name and
! signature, so mod_uns_simple_mpi (piso_run_mpi, simple_run_mpi,
! pimple_run_mpi) builds unchanged, while its linear solves run on the
! MI355X through libmixsolver_hip.so:
!
!   bicgstab_ilu0_mpi (:155) -> ms_bicgstab_ilu0  (BiCGSTAB, natural ILU(0))
!   cg_jacobi_mpi     (:235) -> ms_cg_jacobi      (PCG, Jacobi)
!   cg_ic0_mpi        (:306) -> ms_cg_ic0         (PCG, natural IC(0))
!
! At one partition the owned block is the whole matrix, the halo exchange
! inside the reference's matrix products moves nothing and its global dot
! products are local sums, so these solves are the serial ones on rows
! 1..nowned; it / res / ierr keep the reference's meaning (CG reports
! maxit+1 on exhaustion, ierr 0/1/2 and 11/12 for a failed factorization).
!
! The halo argument is accepted as any type and never read.  A matrix with
! halo rows or halo columns (nrows > nowned, or a column above nowned) comes
! from a partitioned run, which the single-partition device library does not
! serve: the call stops with a message instead of returning wrong numbers.
!===============================================================================
module mod_uns_linsolver_mpi
   use mod_precision, only: dp
   use mod_uns_linsolver, only: csr_t
   use mixsolver_dropin, only: dropin_solve, KIND_BICG, KIND_CGJ, KIND_CGIC
   implicit none
   private
   public :: matvec_owned, owned_block, &
             bicgstab_ilu0_mpi, cg_jacobi_mpi, cg_ic0_mpi

contains

   ! stop unless A is a one-partition system: nowned rows, columns 1..nowned
   subroutine require_one_partition( A, nowned, who )
      type(csr_t),      intent(in) :: A
      integer,          intent(in) :: nowned
      character(len=*), intent(in) :: who
      integer :: nnz
      nnz = A%row_ptr(nowned+1) - 1
      if ( A%nrows /= nowned .or. any(A%col_idx(1:nnz) > nowned) .or. any(A%col_idx(1:nnz) < 1) ) then
         write(*,'(a,a,i0,a,i0)') who, ': the device library solves one partition only; nrows=', &
            A%nrows, ' nowned=', nowned
         error stop 1
      end if
   end subroutine require_one_partition

   ! y(1:nowned) = A x over the owned rows (matvec_owned, :56-73); with one
   ! partition there are no halo entries of x to refresh.
   subroutine matvec_owned( A, x, y, hi, nowned )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(inout) :: x(:)
      real(dp),    intent(out)   :: y(:)
      type(*),     intent(in)    :: hi
      integer,     intent(in)    :: nowned
      integer :: r, k
      real(dp) :: acc
      do r = 1, nowned
         acc = 0.0_dp
         do k = A%row_ptr(r), A%row_ptr(r+1) - 1
            acc = acc + A%val(k) * x(A%col_idx(k))
         end do
         y(r) = acc
      end do
   end subroutine matvec_owned

   ! The owned-owned block of A (owned_block, :105-149): rows 1..nowned,
   ! columns <= nowned in their stored order.
   subroutine owned_block( A, nowned, B )
      type(csr_t), intent(in)  :: A
      integer,     intent(in)  :: nowned
      type(csr_t), intent(out) :: B
      integer :: r, k, m
      B%nrows = nowned
      allocate( B%row_ptr(nowned+1) )
      B%row_ptr(1) = 1
      do r = 1, nowned
         B%row_ptr(r+1) = B%row_ptr(r) + count( A%col_idx(A%row_ptr(r):A%row_ptr(r+1)-1) <= nowned )
      end do
      allocate( B%col_idx(B%row_ptr(nowned+1)-1), B%val(B%row_ptr(nowned+1)-1) )
      m = 0
      do r = 1, nowned
         do k = A%row_ptr(r), A%row_ptr(r+1) - 1
            if ( A%col_idx(k) <= nowned ) then
               m = m + 1
               B%col_idx(m) = A%col_idx(k)
               B%val(m) = A%val(k)
            end if
         end do
      end do
   end subroutine owned_block

   subroutine bicgstab_ilu0_mpi( A, b, x, nowned, hi, tol, maxit, it, res, ierr )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(in)    :: b(:)
      real(dp),    intent(inout) :: x(:)
      integer,     intent(in)    :: nowned
      type(*),     intent(in)    :: hi
      real(dp),    intent(in)    :: tol
      integer,     intent(in)    :: maxit
      integer,     intent(out)   :: it
      real(dp),    intent(out)   :: res
      integer,     intent(out)   :: ierr
      call require_one_partition( A, nowned, 'bicgstab_ilu0_mpi' )
      call dropin_solve( KIND_BICG, nowned, A%row_ptr, A%col_idx, A%val, b, x, tol, maxit, it, res, ierr )
   end subroutine bicgstab_ilu0_mpi

   subroutine cg_jacobi_mpi( A, b, x, nowned, hi, tol, maxit, it, res, ierr )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(in)    :: b(:)
      real(dp),    intent(inout) :: x(:)
      integer,     intent(in)    :: nowned
      type(*),     intent(in)    :: hi
      real(dp),    intent(in)    :: tol
      integer,     intent(in)    :: maxit
      integer,     intent(out)   :: it
      real(dp),    intent(out)   :: res
      integer,     intent(out)   :: ierr
      call require_one_partition( A, nowned, 'cg_jacobi_mpi' )
      call dropin_solve( KIND_CGJ, nowned, A%row_ptr, A%col_idx, A%val, b, x, tol, maxit, it, res, ierr )
   end subroutine cg_jacobi_mpi

   subroutine cg_ic0_mpi( A, b, x, nowned, hi, tol, maxit, it, res, ierr )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(in)    :: b(:)
      real(dp),    intent(inout) :: x(:)
      integer,     intent(in)    :: nowned
      type(*),     intent(in)    :: hi
      real(dp),    intent(in)    :: tol
      integer,     intent(in)    :: maxit
      integer,     intent(out)   :: it
      real(dp),    intent(out)   :: res
      integer,     intent(out)   :: ierr
      call require_one_partition( A, nowned, 'cg_ic0_mpi' )
      call dropin_solve( KIND_CGIC, nowned, A%row_ptr, A%col_idx, A%val, b, x, tol, maxit, it, res, ierr )
   end subroutine cg_ic0_mpi

end module mod_uns_linsolver_mpi
