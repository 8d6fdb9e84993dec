! mixsolver_c.f90 -- ISO_C_BINDING interface to libmixsolver_hip.so.
!
! Interoperable copies of the C structs of include/mixsolver.h and interface
! blocks for every entry point, so Fortran callers of the unstructured
! solver (mod_uns_simple_mpi.f90, mod_uns_linsolver.f90) can call the device
! path directly.  Derived types with allocatable components (csr_t, fields_t,
! mesh_t) are not interoperable: callers pass their arrays (A%row_ptr,
! A%col_idx, A%val, fld%u, ...) as plain buffers.
module mixsolver_c
   use iso_c_binding, only: c_int, c_long_long, c_double, c_char, c_ptr, c_null_ptr, c_null_char, &
                            c_size_t, c_f_pointer, c_associated
   implicit none
   private
   public :: ms_ctrl_t, ms_bcgroup_t
   public :: MS_OK, MS_EINVAL, MS_EDEVICE, MS_ENOMEM
   public :: MS_PC_JACOBI, MS_PC_NATURAL, MS_PC_DILU, MS_PC_MC, MS_FP64, MS_MIXED
   public :: ms_init, ms_version, ms_system_create, ms_system_destroy, ms_system_info
   public :: ms_coarse_info, ms_coarse_aggregates
   public :: ms_set_field, ms_get_field, ms_set_ts_order, ms_get_topology
   public :: ms_simple_run, ms_piso_run, ms_pimple_run, ms_piso_continue, ms_pimple_continue, ms_get_history
   public :: ms_set_cell_types, ms_get_ts_order, ms_vtu_write, ms_nan_scan
   public :: ms_compute_gradients, ms_momentum_assembly, ms_temperature_assembly
   public :: ms_flux_rhiechow, ms_outflow_mass_rescale, ms_ppe_assembly, ms_correct_fields
   public :: ms_cg_jacobi, ms_cg_ic0, ms_bicgstab_ilu0, ms_solve, ms_time_kernel, ms_time_copy
   public :: c_str, last_error

   integer(c_int), parameter :: MS_OK = 0, MS_EINVAL = -1, MS_EDEVICE = -2, MS_ENOMEM = -3
   integer(c_int), parameter :: MS_PC_JACOBI = 0, MS_PC_NATURAL = 1, MS_PC_DILU = 2, MS_PC_MC = 3
   integer(c_int), parameter :: MS_FP64 = 0, MS_MIXED = 1

   type, bind(c) :: ms_ctrl_t
      real(c_double) :: rho, mu, alpha_u, alpha_p
      real(c_double) :: outer_tol, lin_tol, conv_blend, nonorth_corr
      real(c_double) :: dt, cp, k_cond, tref, beta
      real(c_double) :: gravity(3), body_force(3)
      real(c_double) :: init_T, t_pert
      integer(c_int) :: outer_max, lin_max, ppe_precond, transient
      integer(c_int) :: time_scheme, n_time_max, n_correct, n_out_every
      integer(c_int) :: pimple, n_outer_iter, boussinesq, thermal_model
      integer(c_int) :: inlet_ramp, nbc
      integer(c_int) :: gpu_ppe_precond, gpu_mom_precond, gpu_precision, gpu_ppe_solver
      real(c_double) :: gpu_inner_tol
      integer(c_int) :: gpu_reorder, gpu_coarse
      integer(c_int) :: nperiodic
      integer(c_int) :: periodic_zone(2, 4)
      integer(c_int) :: periodic_auto(4)
      real(c_double) :: periodic_shift(3, 4)
      integer(c_int) :: porous_unified
   end type ms_ctrl_t

   type, bind(c) :: ms_bcgroup_t
      integer(c_int) :: zone, btype
      real(c_double) :: uvel(3), pval, mdot, uspeed
      real(c_double) :: prfl_mean, prfl_span, prfl_origin
      integer(c_int) :: prfl_axis, has_lid
      integer(c_int) :: lid_dir, ttype
      real(c_double) :: lid_coord, lid_tol, lid_vel(3)
      real(c_double) :: tval, qval
      integer(c_int) :: ntbc_plane, tbc_pdir(4), tbc_pttype(4)
      real(c_double) :: tbc_pcoord(4), tbc_tol, tbc_ptval(4), tbc_pqval(4)
   end type ms_bcgroup_t

   interface
      integer(c_int) function ms_init(device) bind(c, name='ms_init')
         import :: c_int
         integer(c_int), value :: device
      end function
      integer(c_int) function ms_version() bind(c, name='ms_version')
         import :: c_int
      end function
      type(c_ptr) function ms_last_error() bind(c, name='ms_last_error')
         import :: c_ptr
      end function
      integer(c_size_t) function strlen(p) bind(c, name='strlen')
         import :: c_size_t, c_ptr
         type(c_ptr), value :: p
      end function
      integer(c_int) function ms_system_create(s, nnodes, x, nfaces, fptr, fnodes, c0, c1, zone, ncells, &
                                               ctrl, groups) bind(c, name='ms_system_create')
         import :: c_int, c_long_long, c_double, c_ptr, ms_ctrl_t, ms_bcgroup_t
         type(c_ptr) :: s
         integer(c_int), value :: nnodes, nfaces, ncells
         real(c_double) :: x(*)
         integer(c_long_long) :: fptr(*)
         integer(c_int) :: fnodes(*), c0(*), c1(*), zone(*)
         type(ms_ctrl_t) :: ctrl
         type(ms_bcgroup_t) :: groups(*)
      end function
      integer(c_int) function ms_system_destroy(s) bind(c, name='ms_system_destroy')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
      end function
      integer(c_int) function ms_system_info(s, out8) bind(c, name='ms_system_info')
         import :: c_int, c_ptr, c_long_long
         type(c_ptr), value :: s
         integer(c_long_long) :: out8(8)
      end function
      integer(c_int) function ms_coarse_info(s, k, nchunks, agg) bind(c, name='ms_coarse_info')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: k, nchunks, agg(*)
      end function
      integer(c_int) function ms_coarse_aggregates(n, xc, k, agg) bind(c, name='ms_coarse_aggregates')
         import :: c_int, c_double
         integer(c_int), value :: n, k
         real(c_double) :: xc(*)
         integer(c_int) :: agg(*)
      end function
      integer(c_int) function ms_set_field(s, name, host, n) bind(c, name='ms_set_field')
         import :: c_int, c_ptr, c_char, c_double, c_long_long
         type(c_ptr), value :: s
         character(kind=c_char) :: name(*)
         real(c_double) :: host(*)
         integer(c_long_long), value :: n
      end function
      integer(c_int) function ms_get_field(s, name, host, n) bind(c, name='ms_get_field')
         import :: c_int, c_ptr, c_char, c_double, c_long_long
         type(c_ptr), value :: s
         character(kind=c_char) :: name(*)
         real(c_double) :: host(*)
         integer(c_long_long), value :: n
      end function
      integer(c_int) function ms_set_ts_order(s, t) bind(c, name='ms_set_ts_order')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int), value :: t
      end function
      integer(c_int) function ms_get_topology(s, row_ptr, col_idx, fgrp) bind(c, name='ms_get_topology')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: row_ptr(*), col_idx(*), fgrp(*)
      end function
      integer(c_int) function ms_simple_run(s, nsteps, it_done, ier) bind(c, name='ms_simple_run')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int), value :: nsteps
         integer(c_int) :: it_done, ier
      end function
      integer(c_int) function ms_piso_run(s, nsteps, ier) bind(c, name='ms_piso_run')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int), value :: nsteps
         integer(c_int) :: ier
      end function
      integer(c_int) function ms_pimple_run(s, nsteps, ier) bind(c, name='ms_pimple_run')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int), value :: nsteps
         integer(c_int) :: ier
      end function
      integer(c_int) function ms_piso_continue(s, nsteps, ier) bind(c, name='ms_piso_continue')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int), value :: nsteps
         integer(c_int) :: ier
      end function
      integer(c_int) function ms_pimple_continue(s, nsteps, ier) bind(c, name='ms_pimple_continue')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int), value :: nsteps
         integer(c_int) :: ier
      end function
      integer(c_int) function ms_set_cell_types(s, cztype) bind(c, name='ms_set_cell_types')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: cztype(*)
      end function
      integer(c_int) function ms_get_ts_order(s, t) bind(c, name='ms_get_ts_order')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
         integer(c_int) :: t
      end function
      integer(c_int) function ms_nan_scan(s, field, cell, val, kind, ifin) bind(c, name='ms_nan_scan')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         integer(c_int) :: field, cell, kind, ifin
         real(c_double) :: val
      end function
      integer(c_int) function ms_vtu_write(path, binary, nnodes, x, nfaces, fptr, fnodes, c0, c1, ncells, xc, p, u, &
                                           T, T_s, q_src) bind(c, name='ms_vtu_write')
         import :: c_int, c_char, c_double, c_long_long, c_ptr
         character(kind=c_char) :: path(*)
         integer(c_int), value :: binary, nnodes, nfaces, ncells
         real(c_double) :: x(*), xc(*), p(*), u(*), T(*)
         integer(c_long_long) :: fptr(*)
         integer(c_int) :: fnodes(*), c0(*), c1(*)
         type(c_ptr), value :: T_s, q_src
      end function
      integer(c_int) function ms_get_history(s, out, maxrows, nrows) bind(c, name='ms_get_history')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         real(c_double) :: out(*)
         integer(c_int), value :: maxrows
         integer(c_int) :: nrows
      end function
      integer(c_int) function ms_compute_gradients(s) bind(c, name='ms_compute_gradients')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
      end function
      integer(c_int) function ms_momentum_assembly(s, comp, val, rhs, ap) bind(c, name='ms_momentum_assembly')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         integer(c_int), value :: comp
         real(c_double) :: val(*), rhs(*), ap(*)
      end function
      integer(c_int) function ms_temperature_assembly(s, val, rhs, ap) bind(c, name='ms_temperature_assembly')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         real(c_double) :: val(*), rhs(*), ap(*)
      end function
      integer(c_int) function ms_flux_rhiechow(s) bind(c, name='ms_flux_rhiechow')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
      end function
      integer(c_int) function ms_outflow_mass_rescale(s) bind(c, name='ms_outflow_mass_rescale')
         import :: c_int, c_ptr
         type(c_ptr), value :: s
      end function
      integer(c_int) function ms_ppe_assembly(s, val, rhs) bind(c, name='ms_ppe_assembly')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         real(c_double) :: val(*), rhs(*)
      end function
      integer(c_int) function ms_correct_fields(s, pp, dpp) bind(c, name='ms_correct_fields')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         real(c_double) :: pp(*), dpp(*)
      end function
      integer(c_int) function ms_cg_jacobi(n, row_ptr, col_idx, val, b, x, tol, maxit, it, res, ierr) &
            bind(c, name='ms_cg_jacobi')
         import :: c_int, c_double
         integer(c_int), value :: n, maxit
         integer(c_int) :: row_ptr(*), col_idx(*), it, ierr
         real(c_double) :: val(*), b(*), x(*), res
         real(c_double), value :: tol
      end function
      integer(c_int) function ms_cg_ic0(n, row_ptr, col_idx, val, b, x, tol, maxit, it, res, ierr) &
            bind(c, name='ms_cg_ic0')
         import :: c_int, c_double
         integer(c_int), value :: n, maxit
         integer(c_int) :: row_ptr(*), col_idx(*), it, ierr
         real(c_double) :: val(*), b(*), x(*), res
         real(c_double), value :: tol
      end function
      integer(c_int) function ms_bicgstab_ilu0(n, row_ptr, col_idx, val, b, x, tol, maxit, it, res, ierr) &
            bind(c, name='ms_bicgstab_ilu0')
         import :: c_int, c_double
         integer(c_int), value :: n, maxit
         integer(c_int) :: row_ptr(*), col_idx(*), it, ierr
         real(c_double) :: val(*), b(*), x(*), res
         real(c_double), value :: tol
      end function
      integer(c_int) function ms_solve(n, row_ptr, col_idx, val, b, x, tol, maxit, solver, precond, precision, &
                                       inner_tol, it, res, ierr) bind(c, name='ms_solve')
         import :: c_int, c_double
         integer(c_int), value :: n, maxit, solver, precond, precision
         integer(c_int) :: row_ptr(*), col_idx(*), it, ierr
         real(c_double) :: val(*), b(*), x(*), res
         real(c_double), value :: tol, inner_tol
      end function
      integer(c_int) function ms_time_kernel(s, which, reps, avg_ms, alg_bytes) bind(c, name='ms_time_kernel')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: s
         integer(c_int), value :: which, reps
         real(c_double) :: avg_ms, alg_bytes
      end function
      integer(c_int) function ms_time_copy(bytes, reps, avg_ms) bind(c, name='ms_time_copy')
         import :: c_int, c_long_long, c_double
         integer(c_long_long), value :: bytes
         integer(c_int), value :: reps
         real(c_double) :: avg_ms
      end function
   end interface

contains

   ! NUL-terminated copy of a Fortran string for the name arguments
   pure function c_str(s) result(r)
      character(len=*), intent(in) :: s
      character(kind=c_char) :: r(len_trim(s) + 1)
      integer :: i
      do i = 1, len_trim(s)
         r(i) = s(i:i)
      end do
      r(len_trim(s) + 1) = c_null_char
   end function c_str

   ! the message of the library's last error (ms_last_error) as a Fortran string
   function last_error() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: i, n
      p = ms_last_error()
      if ( .not. c_associated(p) ) then
         msg = ''
         return
      end if
      n = int(strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function last_error

end module mixsolver_c
