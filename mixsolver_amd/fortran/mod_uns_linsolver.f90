!===============================================================================
! mod_uns_linsolver.f90 -- drop-in replacement of the reference module of the
! same name (src/unstructured/mod_uns_linsolver.f90, publics at :24-25).
!
! Compiled in place of the reference file, it keeps every public name and
! signature, so mod_uns_simple and the other callers build unchanged, while
! the three Krylov solvers run on the MI355X through libmixsolver_hip.so:
!
!   bicgstab_ilu0 (:148)  -> ms_bicgstab_ilu0  (BiCGSTAB, natural-order ILU(0))
!   cg_jacobi     (:230)  -> ms_cg_jacobi      (PCG, Jacobi)
!   cg_ic0        (:405)  -> ms_cg_ic0         (PCG, natural-order IC(0))
!
! Argument meaning, status codes (ierr 0/1/2/11/12), the returned iteration
! count and the relative residual are the reference's.  csr_t keeps its
! components; the host helpers (csr_matvec, ilu0/ic0 factor and apply) keep
! the reference's storage conventions: ILU(0) in place with a unit lower
! factor and the upper factor including the diagonal, IC(0) overwriting the
! lower triangle and diagonal with L (A ~ L L^T), both recording diag_pos.
!===============================================================================
module mod_uns_linsolver
   use mod_precision, only: dp
   use mixsolver_dropin, only: dropin_solve, KIND_BICG, KIND_CGJ, KIND_CGIC
   implicit none
   private
   public :: csr_t, csr_matvec, ilu0_factor, ilu0_apply, bicgstab_ilu0, &
             cg_jacobi, ic0_factor, ic0_apply, cg_ic0

   type :: csr_t
      integer              :: nrows = 0
      integer, allocatable :: row_ptr(:)    ! nrows+1, 1-based
      integer, allocatable :: col_idx(:)    ! nnz, ascending per row
      real(dp), allocatable :: val(:)       ! nnz
      integer, allocatable :: diag_pos(:)   ! slot of A(i,i), set by the factorizations
   end type csr_t

contains

   ! y = A x
   subroutine csr_matvec( A, x, y )
      type(csr_t), intent(in)  :: A
      real(dp),    intent(in)  :: x(:)
      real(dp),    intent(out) :: y(:)
      integer :: r, k
      real(dp) :: acc
      do r = 1, A%nrows
         acc = 0.0_dp
         do k = A%row_ptr(r), A%row_ptr(r+1) - 1
            acc = acc + A%val(k) * x(A%col_idx(k))
         end do
         y(r) = acc
      end do
   end subroutine csr_matvec

   ! diagonal slots; .false. when a row has no diagonal entry
   logical function find_diagonals( A, what )
      type(csr_t), intent(inout) :: A
      character(len=*), intent(in) :: what
      integer :: r, k
      if ( allocated(A%diag_pos) ) deallocate( A%diag_pos )
      allocate( A%diag_pos(A%nrows), source = 0 )
      do r = 1, A%nrows
         do k = A%row_ptr(r), A%row_ptr(r+1) - 1
            if ( A%col_idx(k) == r ) A%diag_pos(r) = k
         end do
         if ( A%diag_pos(r) == 0 ) then
            write(*,'(a,a,i0)') what, ' ERROR: missing diagonal in row ', r
            find_diagonals = .false.
            return
         end if
      end do
      find_diagonals = .true.
   end function find_diagonals

   ! ILU(0), row-wise elimination restricted to the pattern (slot(j) maps a
   ! column of the current row to its storage position, 0 outside the row)
   subroutine ilu0_factor( A, ierr )
      type(csr_t), intent(inout) :: A
      integer,     intent(out)   :: ierr
      integer, allocatable :: slot(:)
      integer :: r, kk, piv, jj, col
      real(dp) :: m
      ierr = 0
      if ( .not. find_diagonals( A, 'ILU0' ) ) then
         ierr = 1
         return
      end if
      allocate( slot(A%nrows), source = 0 )
      do r = 1, A%nrows
         do kk = A%row_ptr(r), A%row_ptr(r+1) - 1
            slot(A%col_idx(kk)) = kk
         end do
         do kk = A%row_ptr(r), A%diag_pos(r) - 1          ! eliminate with earlier rows
            piv = A%col_idx(kk)
            if ( abs( A%val(A%diag_pos(piv)) ) < tiny(1.0_dp) ) then
               write(*,'(a,i0)') 'ILU0 ERROR: zero pivot in row ', piv
               ierr = 2
               return
            end if
            m = A%val(kk) / A%val(A%diag_pos(piv))
            A%val(kk) = m
            do jj = A%diag_pos(piv) + 1, A%row_ptr(piv+1) - 1
               col = A%col_idx(jj)
               if ( slot(col) > 0 ) A%val(slot(col)) = A%val(slot(col)) - m * A%val(jj)
            end do
         end do
         do kk = A%row_ptr(r), A%row_ptr(r+1) - 1
            slot(A%col_idx(kk)) = 0
         end do
      end do
   end subroutine ilu0_factor

   ! z = (L U)^-1 r
   subroutine ilu0_apply( A, r, z )
      type(csr_t), intent(in)  :: A
      real(dp),    intent(in)  :: r(:)
      real(dp),    intent(out) :: z(:)
      integer :: i, k
      real(dp) :: acc
      do i = 1, A%nrows
         acc = r(i)
         do k = A%row_ptr(i), A%diag_pos(i) - 1
            acc = acc - A%val(k) * z(A%col_idx(k))
         end do
         z(i) = acc
      end do
      do i = A%nrows, 1, -1
         acc = z(i)
         do k = A%diag_pos(i) + 1, A%row_ptr(i+1) - 1
            acc = acc - A%val(k) * z(A%col_idx(k))
         end do
         z(i) = acc / A%val(A%diag_pos(i))
      end do
   end subroutine ilu0_apply

   ! IC(0): L(i,k) = (A(i,k) - sum_j<k L(i,j) L(k,j)) / L(k,k) over the common
   ! pattern, L(i,i) = sqrt(A(i,i) - sum_k<i L(i,k)^2)
   subroutine ic0_factor( A, ierr )
      type(csr_t), intent(inout) :: A
      integer,     intent(out)   :: ierr
      integer, allocatable :: slot(:)
      integer :: i, kp, k, q
      real(dp) :: s
      ierr = 0
      if ( .not. find_diagonals( A, 'ICC0' ) ) then
         ierr = 1
         return
      end if
      allocate( slot(A%nrows), source = 0 )
      do i = 1, A%nrows
         do kp = A%row_ptr(i), A%diag_pos(i) - 1
            k = A%col_idx(kp)
            ! sum over columns j < k present in both row i (slots before kp) and row k
            s = 0.0_dp
            do q = A%row_ptr(k), A%diag_pos(k) - 1
               if ( slot(A%col_idx(q)) > 0 ) s = s + A%val(slot(A%col_idx(q))) * A%val(q)
            end do
            A%val(kp) = ( A%val(kp) - s ) / A%val(A%diag_pos(k))
            slot(k) = kp                                   ! row i's L(i,k) now usable
         end do
         s = A%val(A%diag_pos(i))
         do kp = A%row_ptr(i), A%diag_pos(i) - 1
            s = s - A%val(kp)**2
         end do
         do kp = A%row_ptr(i), A%diag_pos(i) - 1
            slot(A%col_idx(kp)) = 0
         end do
         if ( s <= 0.0_dp ) then
            write(*,'(a,i0)') 'ICC0 ERROR: non-positive pivot in row ', i
            ierr = 2
            return
         end if
         A%val(A%diag_pos(i)) = sqrt( s )
      end do
   end subroutine ic0_factor

   ! z = (L L^T)^-1 r
   subroutine ic0_apply( L, r, z )
      type(csr_t), intent(in)  :: L
      real(dp),    intent(in)  :: r(:)
      real(dp),    intent(out) :: z(:)
      real(dp), allocatable :: w(:)
      integer :: i, kp
      real(dp) :: acc
      do i = 1, L%nrows
         acc = r(i)
         do kp = L%row_ptr(i), L%diag_pos(i) - 1
            acc = acc - L%val(kp) * z(L%col_idx(kp))
         end do
         z(i) = acc / L%val(L%diag_pos(i))
      end do
      w = z
      do i = L%nrows, 1, -1
         z(i) = w(i) / L%val(L%diag_pos(i))
         do kp = L%row_ptr(i), L%diag_pos(i) - 1
            w(L%col_idx(kp)) = w(L%col_idx(kp)) - L%val(kp) * z(i)
         end do
      end do
   end subroutine ic0_apply

   subroutine bicgstab_ilu0( A, b, x, tol, maxit, it, res, ierr, verb )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(in)    :: b(:)
      real(dp),    intent(inout) :: x(:)
      real(dp),    intent(in)    :: tol
      integer,     intent(in)    :: maxit
      integer,     intent(out)   :: it
      real(dp),    intent(out)   :: res
      integer,     intent(out)   :: ierr
      integer,     intent(in), optional :: verb
      call dropin_solve( KIND_BICG, A%nrows, A%row_ptr, A%col_idx, A%val, b, x, tol, maxit, it, res, ierr )
      if ( present(verb) ) then
         if ( verb > 0 ) write(*,'(a,i6,a,es12.4)') '  bicgstab_ilu0 (device): it=', it, '  res=', res
      end if
   end subroutine bicgstab_ilu0

   subroutine cg_jacobi( A, b, x, tol, maxit, it, res, ierr )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(in)    :: b(:)
      real(dp),    intent(inout) :: x(:)
      real(dp),    intent(in)    :: tol
      integer,     intent(in)    :: maxit
      integer,     intent(out)   :: it
      real(dp),    intent(out)   :: res
      integer,     intent(out)   :: ierr
      call dropin_solve( KIND_CGJ, A%nrows, A%row_ptr, A%col_idx, A%val, b, x, tol, maxit, it, res, ierr )
   end subroutine cg_jacobi

   subroutine cg_ic0( A, b, x, tol, maxit, it, res, ierr )
      type(csr_t), intent(in)    :: A
      real(dp),    intent(in)    :: b(:)
      real(dp),    intent(inout) :: x(:)
      real(dp),    intent(in)    :: tol
      integer,     intent(in)    :: maxit
      integer,     intent(out)   :: it
      real(dp),    intent(out)   :: res
      integer,     intent(out)   :: ierr
      call dropin_solve( KIND_CGIC, A%nrows, A%row_ptr, A%col_idx, A%val, b, x, tol, maxit, it, res, ierr )
   end subroutine cg_ic0

end module mod_uns_linsolver
