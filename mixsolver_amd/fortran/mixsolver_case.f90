!===============================================================================
! mixsolver_case.f90 -- the reference's unstructured case, as its drivers hold
! it (mesh_t, ctrl_t, bc_t, fields_t), handed to libmixsolver_hip.so: one
! device system per run, kept resident while the loop runs.
!
! Used by the loop-level drop-in mod_uns_simple_mpi.f90 of this directory.
! It converts
!   mesh_t   (mod_uns_mesh.f90:38-54)  -> node coordinates, face-node lists,
!            owner / neighbour / zone per face (ms_system_create)
!   ctrl_t   (mod_uns_control.f90:148-280) -> ms_ctrl_t
!   bc_t     (mod_uns_bc.f90:80-130)   -> ms_bcgroup_t, control order
!   fields_t (mod_uns_fields.f90:29-80) -> the device fields, the per-cell
!            porous coefficients and the block type of every cell
! and moves the fields back.  The reference has no device options, so they
! come from the environment (all optional, the defaults give the reference's
! FP64 path with its preconditioners):
!   MIXSOLVER_PPE_PRECOND   jacobi | ic0 | ilu0 | dilu | mc_ilu0 | mc_ic0
!                           (default: the control's ppe_precond)
!   MIXSOLVER_MOM_PRECOND   ilu0 | jacobi | dilu | mc_ilu0   (default ilu0)
!   MIXSOLVER_PPE_SOLVER    pcg | bicgstab                   (default pcg)
!   MIXSOLVER_PRECISION     fp64 | mixed                     (default fp64)
!   MIXSOLVER_INNER_TOL     mixed PCG reliable-update trigger (default 0.1)
!   MIXSOLVER_REORDER       none | rcm                        (default none)
!   MIXSOLVER_COARSE        p' aggregates, 0..1024            (default 0)
!   MIXSOLVER_DEVICE        HIP device index                  (default 0)
!===============================================================================
module mixsolver_case
   use iso_c_binding
   use mod_precision, only: dp
   use mod_uns_mesh, only: mesh_t
   use mod_uns_connectivity, only: conn_t
   use mod_uns_geometry, only: geom_t
   use mod_uns_control, only: ctrl_t, out_is_binary, out_native_cells, BC_INTERFACE
   use mod_uns_bc, only: bc_t
   use mod_uns_fields, only: fields_t
   use mod_uns_output, only: vtk_write
   use mod_uns_simple, only: make_step_filename, nan_scan, nan_report, NAN_IER
   use mixsolver_c
   implicit none
   private
   public :: case_convert, case_create, case_push, case_pull, case_destroy, case_check
   public :: case_run_transient, case_run_steady, case_nan_guard, RUN_PISO, RUN_PIMPLE

   integer, parameter :: RUN_PISO = 1, RUN_PIMPLE = 2

contains

   ! stop with the library's message when a call fails: the reference's
   ! drivers have no status code for a device or API error
   subroutine case_check( rc, what )
      integer(c_int),   intent(in) :: rc
      character(len=*), intent(in) :: what
      if ( rc /= MS_OK ) then
         write(*,'(a,a,i0,a,a)') 'mixsolver: ', what // ' failed, code ', rc, ': ', last_error()
         error stop 1
      end if
   end subroutine case_check

   ! the value of environment variable `name` (lower case), or `def`
   function env_opt( name, def ) result( v )
      character(len=*), intent(in) :: name, def
      character(len=:), allocatable :: v
      character(len=64) :: buf
      integer :: n, st, i
      call get_environment_variable( name, buf, n, st )
      if ( st /= 0 .or. n == 0 ) then
         v = def
         return
      end if
      v = trim(adjustl(buf))
      do i = 1, len(v)
         if ( v(i:i) >= 'A' .and. v(i:i) <= 'Z' ) v(i:i) = achar( iachar(v(i:i)) + 32 )
      end do
   end function env_opt

   integer function precond_of( name, what )
      character(len=*), intent(in) :: name, what
      select case ( name )
      case ( 'jacobi' );            precond_of = MS_PC_JACOBI
      case ( 'ic0', 'ilu0' );       precond_of = MS_PC_NATURAL
      case ( 'dilu' );              precond_of = MS_PC_DILU
      case ( 'mc_ilu0', 'mc_ic0' ); precond_of = MS_PC_MC
      case default
         write(*,'(a,a,a,a)') 'mixsolver: unknown ', what, ' value ', name
         error stop 1
      end select
   end function precond_of

   ! ctrl_t + bc_t -> ms_ctrl_t + ms_bcgroup_t(1:bcs%nb); mesh_t -> the
   ! arrays of ms_system_create (x(3*nnodes), fptr(nfaces+1) 0-based,
   ! fnodes, c0, c1, zone)
   subroutine case_convert( m, ctrl, bcs, cc, groups, x, fptr, fnodes, c0, c1, zone )
      type(mesh_t),  intent(in)  :: m
      type(ctrl_t),  intent(in)  :: ctrl
      type(bc_t),    intent(in)  :: bcs
      type(ms_ctrl_t), intent(out) :: cc
      type(ms_bcgroup_t), allocatable, intent(out) :: groups(:)
      real(c_double), allocatable, intent(out) :: x(:)
      integer(c_long_long), allocatable, intent(out) :: fptr(:)
      integer(c_int), allocatable, intent(out) :: fnodes(:), c0(:), c1(:), zone(:)
      integer :: i, j, f, k, ios, ng, nz
      integer, allocatable :: zs(:)
      character(len=:), allocatable :: opt

      cc%rho = ctrl%rho;             cc%mu = ctrl%mu
      cc%alpha_u = ctrl%alpha_u;     cc%alpha_p = ctrl%alpha_p
      cc%outer_tol = ctrl%outer_tol; cc%lin_tol = ctrl%lin_tol
      cc%conv_blend = ctrl%conv_blend
      cc%nonorth_corr = ctrl%nonorth_corr
      cc%dt = ctrl%dt;               cc%cp = ctrl%cp
      cc%k_cond = ctrl%k_cond;       cc%tref = ctrl%tref
      cc%beta = ctrl%beta
      cc%gravity = ctrl%gravity;     cc%body_force = ctrl%body_force
      cc%init_T = ctrl%init_T;       cc%t_pert = ctrl%t_pert
      cc%outer_max = ctrl%outer_max; cc%lin_max = ctrl%lin_max
      ! the drivers' preconditioner switch: 'ic0', anything else Jacobi
      ! (mod_uns_simple_mpi.f90:437-444)
      cc%ppe_precond = merge( 1, 0, trim(ctrl%ppe_precond) == 'ic0' )
      cc%transient = merge( 1, 0, ctrl%transient )
      cc%time_scheme = ctrl%time_scheme
      cc%n_time_max = ctrl%n_time_max
      cc%n_correct = ctrl%n_correct
      cc%n_out_every = ctrl%n_out_every
      cc%pimple = merge( 1, 0, ctrl%pimple )
      cc%n_outer_iter = ctrl%n_outer_iter
      cc%boussinesq = merge( 1, 0, ctrl%boussinesq )
      cc%thermal_model = merge( 1, 0, trim(ctrl%thermal_model) == 'ltne' )
      cc%inlet_ramp = ctrl%inlet_ramp
      cc%porous_unified = merge( 1, 0, trim(ctrl%porous_model) == 'unified' )
      cc%nperiodic = 0
      cc%periodic_zone = 0
      cc%periodic_auto = 0
      cc%periodic_shift = 0.0_c_double

      ! device options
      if ( cc%ppe_precond == 1 ) then
         opt = env_opt( 'MIXSOLVER_PPE_PRECOND', 'ic0' )
      else
         opt = env_opt( 'MIXSOLVER_PPE_PRECOND', 'jacobi' )
      end if
      cc%gpu_ppe_precond = precond_of( opt, 'MIXSOLVER_PPE_PRECOND' )
      cc%gpu_mom_precond = precond_of( env_opt( 'MIXSOLVER_MOM_PRECOND', 'ilu0' ), 'MIXSOLVER_MOM_PRECOND' )
      opt = env_opt( 'MIXSOLVER_PPE_SOLVER', 'pcg' )
      cc%gpu_ppe_solver = merge( 1, 0, opt == 'bicgstab' )
      opt = env_opt( 'MIXSOLVER_PRECISION', 'fp64' )
      cc%gpu_precision = merge( MS_MIXED, MS_FP64, opt == 'mixed' )
      opt = env_opt( 'MIXSOLVER_INNER_TOL', '0.1' )
      read( opt, *, iostat=ios ) cc%gpu_inner_tol
      if ( ios /= 0 ) cc%gpu_inner_tol = 0.1_c_double
      cc%gpu_reorder = merge( 1, 0, env_opt( 'MIXSOLVER_REORDER', 'none' ) == 'rcm' )
      opt = env_opt( 'MIXSOLVER_COARSE', '0' )
      read( opt, *, iostat=ios ) cc%gpu_coarse
      if ( ios /= 0 ) cc%gpu_coarse = 0

      ! boundary groups in control order (build_bc, mod_uns_bc.f90:155-323).
      ! build_bc gathers the faces of every coupling-interface zone into one
      ! BC_INTERFACE group with zone 0 (:282-323); the library takes a group
      ! per face zone, so that group becomes one group per interface zone,
      ! in ascending zone order (the faces and their treatment are the same).
      ng = 0
      do i = 1, bcs%nb
         if ( bcs%gb(i)%btype == BC_INTERFACE .and. bcs%gb(i)%zone == 0 ) then
            call group_zones( m, bcs%gb(i)%faces, bcs%gb(i)%nf, zs, nz )
            ng = ng + nz
         else
            ng = ng + 1
         end if
      end do
      cc%nbc = ng
      allocate( groups(max(1, ng)) )
      j = 0
      do i = 1, bcs%nb
         if ( bcs%gb(i)%btype == BC_INTERFACE .and. bcs%gb(i)%zone == 0 ) then
            call group_zones( m, bcs%gb(i)%faces, bcs%gb(i)%nf, zs, nz )
            do k = 1, nz
               j = j + 1
               groups(j)%zone = zs(k)
               groups(j)%btype = BC_INTERFACE
               groups(j)%prfl_span = 1.0_c_double
               groups(j)%prfl_axis = 2
            end do
            cycle
         end if
         j = j + 1
         associate( g => bcs%gb(i), o => groups(j) )
            o%zone = g%zone;               o%btype = g%btype
            o%uvel = g%uvel;               o%pval = g%pval
            o%mdot = g%mdot;               o%uspeed = g%uspeed
            o%prfl_mean = g%prfl_mean;     o%prfl_span = g%prfl_span
            o%prfl_origin = g%prfl_origin; o%prfl_axis = g%prfl_axis
            o%has_lid = merge( 1, 0, g%has_lid )
            o%lid_dir = g%lid_dir;         o%ttype = g%ttype
            o%lid_coord = g%lid_coord;     o%lid_tol = g%lid_tol
            o%lid_vel = g%lid_vel
            o%tval = g%tval;               o%qval = g%qval
            o%ntbc_plane = g%ntbc_plane
            o%tbc_pdir = g%tbc_pdir;       o%tbc_pttype = g%tbc_pttype
            o%tbc_pcoord = g%tbc_pcoord;   o%tbc_tol = g%tbc_tol
            o%tbc_ptval = g%tbc_ptval;     o%tbc_pqval = g%tbc_pqval
         end associate
      end do

      ! mesh: node coordinates and the face list (mod_uns_mesh.f90:17-54)
      allocate( x(3 * m%nnodes) )
      x = reshape( m%x(1:3, 1:m%nnodes), [3 * m%nnodes] )
      allocate( fptr(m%nfaces + 1), c0(m%nfaces), c1(m%nfaces), zone(m%nfaces) )
      fptr(1) = 0
      do f = 1, m%nfaces
         fptr(f + 1) = fptr(f) + m%f(f)%nn
         c0(f) = m%f(f)%c0
         c1(f) = m%f(f)%c1
         zone(f) = m%f(f)%zone
      end do
      allocate( fnodes(fptr(m%nfaces + 1)) )
      do f = 1, m%nfaces
         do k = 1, m%f(f)%nn
            fnodes(fptr(f) + k) = m%f(f)%nodes(k)
         end do
      end do
   end subroutine case_convert

   ! the distinct face zones among a group's faces, ascending
   subroutine group_zones( m, faces, nf, zs, nz )
      type(mesh_t), intent(in)  :: m
      integer,      intent(in)  :: faces(:), nf
      integer, allocatable, intent(out) :: zs(:)
      integer,      intent(out) :: nz
      integer :: k, zlast, znext, z
      allocate( zs(max(1, nf)) )
      nz = 0
      zlast = -huge(1)
      do
         znext = huge(1)
         do k = 1, nf
            z = m%f(faces(k))%zone
            if ( z > zlast .and. z < znext ) znext = z
         end do
         if ( znext == huge(1) ) exit
         nz = nz + 1
         zs(nz) = znext
         zlast = znext
      end do
   end subroutine group_zones

   ! one device system for the case, with the fields of fld on it
   subroutine case_create( m, ctrl, bcs, fld, s )
      type(mesh_t),   intent(in)  :: m
      type(ctrl_t),   intent(in)  :: ctrl
      type(bc_t),     intent(in)  :: bcs
      type(fields_t), intent(in)  :: fld
      type(c_ptr),    intent(out) :: s
      type(ms_ctrl_t) :: cc
      type(ms_bcgroup_t), allocatable :: groups(:)
      real(c_double), allocatable :: x(:)
      integer(c_long_long), allocatable :: fptr(:)
      integer(c_int), allocatable :: fnodes(:), c0(:), c1(:), zone(:), cz(:)
      integer :: dev, ios
      call case_convert( m, ctrl, bcs, cc, groups, x, fptr, fnodes, c0, c1, zone )
      read( env_opt( 'MIXSOLVER_DEVICE', '0' ), *, iostat=ios ) dev
      if ( ios /= 0 ) dev = 0
      call case_check( ms_init( int(dev, c_int) ), 'ms_init' )
      call case_check( ms_system_create( s, int(m%nnodes, c_int), x, int(m%nfaces, c_int), fptr, fnodes, c0, c1, &
                                         zone, int(m%ncells, c_int), cc, groups ), 'ms_system_create' )
      ! block types after resolve_cell_zones: 1 fluid (also unset), 2 porous, 3 solid
      if ( allocated( m%cztype ) ) then
         allocate( cz(m%ncells) )
         cz = max( 1, m%cztype(1:m%ncells) )
         call case_check( ms_set_cell_types( s, cz ), 'ms_set_cell_types' )
      end if
      call set_cell( s, 'porosity', fld%porosity, m%ncells, 1 )
      call set_cell( s, 'inertial', fld%inertial, m%ncells, 1 )
      call set_cell( s, 'disp_l', fld%disp_l, m%ncells, 1 )
      call set_cell( s, 'disp_t', fld%disp_t, m%ncells, 1 )
      call set_cell( s, 'bj_alpha', fld%bj_alpha, m%ncells, 1 )
      call set_cell( s, 'k_s', fld%k_s, m%ncells, 1 )
      call set_cell( s, 'rho_s', fld%rho_s, m%ncells, 1 )
      call set_cell( s, 'cp_s', fld%cp_s, m%ncells, 1 )
      call set_cell( s, 'h_sf', fld%h_sf, m%ncells, 1 )
      call set_cell( s, 'a_sf', fld%a_sf, m%ncells, 1 )
      call set_cell( s, 'q_src', fld%q_src, m%ncells, 1 )
      if ( allocated( fld%perm_dir ) ) &
         call case_check( ms_set_field( s, c_str('perm_dir'), fld%perm_dir, int(3 * m%ncells, c_long_long) ), &
                          'set perm_dir' )
      call case_push( s, fld, m%ncells )
   end subroutine case_create

   subroutine set_cell( s, name, a, n, width )
      type(c_ptr),      intent(in) :: s
      character(len=*), intent(in) :: name
      real(dp), allocatable, intent(in) :: a(:)
      integer,          intent(in) :: n, width
      if ( .not. allocated( a ) ) return
      call case_check( ms_set_field( s, c_str(name), a, int(width * n, c_long_long) ), 'set ' // name )
   end subroutine set_cell

   ! fields of fld -> the device system
   subroutine case_push( s, fld, nc )
      type(c_ptr),    intent(in) :: s
      type(fields_t), intent(in) :: fld
      integer,        intent(in) :: nc
      call case_check( ms_set_field( s, c_str('u'), fld%u, int(3 * nc, c_long_long) ), 'set u' )
      call case_check( ms_set_field( s, c_str('p'), fld%p, int(nc, c_long_long) ), 'set p' )
      if ( allocated( fld%u_old ) ) &
         call case_check( ms_set_field( s, c_str('u_old'), fld%u_old, int(3 * nc, c_long_long) ), 'set u_old' )
      if ( allocated( fld%u_old_old ) ) &
         call case_check( ms_set_field( s, c_str('u_old_old'), fld%u_old_old, int(3 * nc, c_long_long) ), &
                          'set u_old_old' )
      call set_cell( s, 'T', fld%T, nc, 1 )
      call set_cell( s, 'T_old', fld%T_old, nc, 1 )
      call set_cell( s, 'T_old_old', fld%T_old_old, nc, 1 )
      call set_cell( s, 'T_s', fld%T_s, nc, 1 )
      call set_cell( s, 'T_s_old', fld%T_s_old, nc, 1 )
      call set_cell( s, 'T_s_old_old', fld%T_s_old_old, nc, 1 )
      call case_check( ms_set_ts_order( s, int(fld%ts_order, c_int) ), 'ms_set_ts_order' )
   end subroutine case_push

   ! the device system's fields -> fld (every component the drivers leave)
   subroutine case_pull( s, fld, nc, nf )
      type(c_ptr),    intent(in)    :: s
      type(fields_t), intent(inout) :: fld
      integer,        intent(in)    :: nc, nf
      integer(c_int) :: t
      call case_check( ms_get_field( s, c_str('u'), fld%u, int(3 * nc, c_long_long) ), 'get u' )
      call case_check( ms_get_field( s, c_str('p'), fld%p, int(nc, c_long_long) ), 'get p' )
      call case_check( ms_get_field( s, c_str('flux'), fld%flux, int(nf, c_long_long) ), 'get flux' )
      call case_check( ms_get_field( s, c_str('apc'), fld%apc, int(nc, c_long_long) ), 'get apc' )
      call case_check( ms_get_field( s, c_str('gp'), fld%gp, int(3 * nc, c_long_long) ), 'get gp' )
      call case_check( ms_get_field( s, c_str('gu'), fld%gu, int(9 * nc, c_long_long) ), 'get gu' )
      if ( allocated( fld%u_old ) ) &
         call case_check( ms_get_field( s, c_str('u_old'), fld%u_old, int(3 * nc, c_long_long) ), 'get u_old' )
      if ( allocated( fld%u_old_old ) ) &
         call case_check( ms_get_field( s, c_str('u_old_old'), fld%u_old_old, int(3 * nc, c_long_long) ), &
                          'get u_old_old' )
      call get_cell( s, 'T', fld%T, nc )
      call get_cell( s, 'T_old', fld%T_old, nc )
      call get_cell( s, 'T_old_old', fld%T_old_old, nc )
      call get_cell( s, 'T_s', fld%T_s, nc )
      call get_cell( s, 'T_s_old', fld%T_s_old, nc )
      call get_cell( s, 'T_s_old_old', fld%T_s_old_old, nc )
      if ( allocated( fld%gt ) ) &
         call case_check( ms_get_field( s, c_str('gt'), fld%gt, int(3 * nc, c_long_long) ), 'get gt' )
      if ( allocated( fld%gts ) ) &
         call case_check( ms_get_field( s, c_str('gts'), fld%gts, int(3 * nc, c_long_long) ), 'get gts' )
      call case_check( ms_get_ts_order( s, t ), 'ms_get_ts_order' )
      fld%ts_order = t
   end subroutine case_pull

   subroutine get_cell( s, name, a, n )
      type(c_ptr),      intent(in)    :: s
      character(len=*), intent(in)    :: name
      real(dp), allocatable, intent(inout) :: a(:)
      integer,          intent(in)    :: n
      if ( .not. allocated( a ) ) return
      call case_check( ms_get_field( s, c_str(name), a, int(n, c_long_long) ), 'get ' // name )
   end subroutine get_cell

   ! uscale and flux_ref of the drivers' reports (mod_uns_simple_mpi.f90:349-364)
   subroutine ref_scales( g, bcs, rho, uscale, flux_ref )
      type(geom_t), intent(in) :: g
      type(bc_t),   intent(in) :: bcs
      real(dp),     intent(in) :: rho
      real(dp),     intent(out) :: uscale, flux_ref
      real(dp) :: area_btot
      integer :: i, k
      uscale = 0.0_dp
      area_btot = 0.0_dp
      do i = 1, bcs%nb
         uscale = max( uscale, norm2( bcs%gb(i)%lid_vel ), norm2( bcs%gb(i)%uvel ), bcs%gb(i)%uspeed )
         do k = 1, bcs%gb(i)%nf
            area_btot = area_btot + g%area( bcs%gb(i)%faces(k) )
         end do
      end do
      if ( uscale <= 0.0_dp ) uscale = 1.0_dp
      if ( area_btot <= 0.0_dp ) area_btot = 1.0_dp
      flux_ref = rho * uscale * area_btot
   end subroutine ref_scales

   ! the non-finite report of the drivers' guard on the host fields
   ! (nan_guard_mpi at one partition, mod_uns_simple_mpi.f90:1024-1074)
   subroutine case_nan_guard( m, g, fld, nown, stage, it, ier, t_val )
      type(mesh_t),     intent(in)  :: m
      type(geom_t),     intent(in)  :: g
      type(fields_t),   intent(in)  :: fld
      integer,          intent(in)  :: nown
      character(len=*), intent(in)  :: stage
      integer,          intent(in)  :: it
      integer,          intent(out) :: ier
      real(dp), optional, intent(in) :: t_val
      character(len=32) :: bfld
      character(len=8)  :: bkind
      real(dp) :: bval
      integer  :: bcell, ifin
      call nan_scan( m, fld, nown, bfld, bcell, bval, bkind, ifin )
      ier = 0
      if ( ifin == 0 ) return
      if ( present(t_val) ) then
         call nan_report( m, g, fld, bfld, bcell, bval, bkind, stage, it, 'mpi rank 0', t_val, bcell )
      else
         call nan_report( m, g, fld, bfld, bcell, bval, bkind, stage, it, 'mpi rank 0', glb_cell = bcell )
      end if
      ier = NAN_IER
   end subroutine case_nan_guard

   ! piso_run_mpi / pimple_run_mpi at one partition (mod_uns_simple_mpi.f90:
   ! 298-571, 580-850) on the device: the fields stay on the GPU for the whole
   ! run and come back to fld every n_out_every steps (where the reference
   ! writes its snapshot, the same file through vtk_write) and at the end.
   ! Prints the reference's header and per-step lines (the Tmin / Tmax line
   ! is left out: the temperature stays on the device between snapshots).
   subroutine case_run_transient( kind, m, c, g, ctrl, bcs, fld, vtu_prefix, ier )
      integer,          intent(in)    :: kind
      type(mesh_t),     intent(in)    :: m
      type(conn_t),     intent(in)    :: c
      type(geom_t),     intent(in)    :: g
      type(ctrl_t),     intent(in)    :: ctrl
      type(bc_t),       intent(in)    :: bcs
      type(fields_t),   intent(inout) :: fld
      character(len=*), intent(in)    :: vtu_prefix
      integer,          intent(out)   :: ier
      type(c_ptr) :: s
      integer :: n_steps, n_out, it, nchunk, k, nr, ios
      integer(c_int) :: ierc, nrows
      real(dp) :: uscale, flux_ref, t_now
      real(c_double), allocatable :: h(:)
      character(len=512) :: vtufile
      character(len=6) :: tag

      ier = 0
      tag = merge( 'PISO  ', 'PIMPLE', kind == RUN_PISO )
      if ( .not. ctrl%transient ) then
         write(*,'(a,a,a)') 'FATAL: ', merge( 'piso_run_mpi  ', 'pimple_run_mpi', kind == RUN_PISO ), &
                            ' invoked with transient=false'
         ier = merge( 30, 50, kind == RUN_PISO )
         return
      end if
      if ( ctrl%dt <= 0.0_dp ) then
         write(*,'(a,es12.4)') 'FATAL: dt must be > 0, got ', ctrl%dt
         ier = merge( 31, 51, kind == RUN_PISO )
         return
      end if
      n_steps = max( 1, ctrl%n_time_max )
      n_out = max( 0, ctrl%n_out_every )
      call ref_scales( g, bcs, ctrl%rho, uscale, flux_ref )
      write(*,'(a)') ''
      if ( kind == RUN_PISO ) then
         write(*,'(a)') '--- PISO (MPI) time stepping ---'
         write(*,'(a,es10.3,a,i0,a,i0)') '  dt=', ctrl%dt, '  n_steps=', n_steps, '  n_correct=', ctrl%n_correct
         write(*,'(a,es10.3,a,es10.3)') '  reference: uscale =', uscale, '   flux_ref =', flux_ref
      else
         write(*,'(a)') '--- PIMPLE (MPI) time stepping ---'
         write(*,'(a,es10.3,a,i0,a,i0,a,i0)') '  dt=', ctrl%dt, '  n_steps=', n_steps, '  n_outer_iter=', &
            max( 1, ctrl%n_outer_iter ), '  n_correct=', ctrl%n_correct
         write(*,'(a,es10.3,a,es10.3)') '  alpha_u=', ctrl%alpha_u, '   flux_ref =', flux_ref
      end if
      call case_create( m, ctrl, bcs, fld, s )
      t_now = 0.0_dp
      it = 0
      do while ( it < n_steps )
         nchunk = n_steps - it
         if ( n_out > 0 ) nchunk = min( nchunk, n_out - mod( it, n_out ) )
         if ( kind == RUN_PISO ) then
            if ( it == 0 ) then
               call case_check( ms_piso_run( s, int(nchunk, c_int), ierc ), 'ms_piso_run' )
            else
               call case_check( ms_piso_continue( s, int(nchunk, c_int), ierc ), 'ms_piso_continue' )
            end if
         else
            if ( it == 0 ) then
               call case_check( ms_pimple_run( s, int(nchunk, c_int), ierc ), 'ms_pimple_run' )
            else
               call case_check( ms_pimple_continue( s, int(nchunk, c_int), ierc ), 'ms_pimple_continue' )
            end if
         end if
         if ( .not. allocated( h ) ) allocate( h(8) )
         call case_check( ms_get_history( s, h, 0_c_int, nrows ), 'ms_get_history' )
         deallocate( h )
         allocate( h(8 * max(1, nrows)) )
         call case_check( ms_get_history( s, h, nrows, nrows ), 'ms_get_history' )
         nr = nrows
         do k = 1, nr
            t_now = t_now + ctrl%dt
            write(*,'(a,i6,a,es10.3,a,es9.2,a,es9.2,a,es9.2,a,es9.2,a,i5,a,es9.2)') &
               '  step=', it + k, '  t=', t_now, '  mass-imbal=', h(8*(k-1)+1), &
               '  du/dt=', h(8*(k-1)+2), '  max|u|=', h(8*(k-1)+3), '  CFL~=', h(8*(k-1)+4), &
               '  lin-it=', nint( h(8*(k-1)+5) ), '  lin-res=', h(8*(k-1)+6)
         end do
         it = it + nr
         if ( ierc /= 0 ) then
            call case_pull( s, fld, m%ncells, m%nfaces )
            ier = ierc
            if ( ier == NAN_IER ) then
               call case_nan_guard( m, g, fld, m%ncells, trim(tag), it + 1, ier, t_now + ctrl%dt )
               write(*,'(a,a,a,i0,a)') '[uns-mpi] ', trim(tag), ' aborted at time step ', it + 1, &
                                       ' (non-finite solution)'
               ier = NAN_IER
            end if
            call case_destroy( s )
            return
         end if
         if ( n_out > 0 .and. ( mod( it, n_out ) == 0 .or. it == n_steps ) ) then
            call case_pull( s, fld, m%ncells, m%nfaces )
            call make_step_filename( vtu_prefix, it, '.vtu', vtufile )
            call vtk_write( m, c, g, fld, trim(vtufile), ios, out_is_binary(ctrl), out_native_cells(ctrl) )
            if ( ios /= 0 ) write(*,'(a,i0,a)') 'WARNING: snapshot write failed at step ', it, ' (continuing)'
         end if
      end do
      call case_pull( s, fld, m%ncells, m%nfaces )
      call case_destroy( s )
      write(*,'(a,a,i0,a,es12.4)') trim(tag), ' (MPI) finished: ', n_steps, ' time steps, t_end =', t_now
   end subroutine case_run_transient

   ! simple_run_mpi at one partition (mod_uns_simple_mpi.f90:57-289) on the
   ! device; fields back at the end.  Prints the reference's per-iteration
   ! lines at iteration 1, every 50 and at convergence.
   subroutine case_run_steady( m, g, ctrl, bcs, fld, ier )
      type(mesh_t),   intent(in)    :: m
      type(geom_t),   intent(in)    :: g
      type(ctrl_t),   intent(in)    :: ctrl
      type(bc_t),     intent(in)    :: bcs
      type(fields_t), intent(inout) :: fld
      integer,        intent(out)   :: ier
      type(c_ptr) :: s
      integer(c_int) :: ierc, itd, nrows
      real(c_double), allocatable :: h(:)
      real(dp) :: uscale, flux_ref
      integer :: k
      logical :: conv
      call ref_scales( g, bcs, ctrl%rho, uscale, flux_ref )
      write(*,'(a)') ''
      write(*,'(a)') '--- SIMPLE (MPI) iteration ---'
      write(*,'(a,es10.3,a,es10.3)') '  reference: uscale =', uscale, '   flux_ref =', flux_ref
      call case_create( m, ctrl, bcs, fld, s )
      call case_check( ms_simple_run( s, 0_c_int, itd, ierc ), 'ms_simple_run' )
      allocate( h(8) )
      call case_check( ms_get_history( s, h, 0_c_int, nrows ), 'ms_get_history' )
      deallocate( h )
      allocate( h(8 * max(1, nrows)) )
      call case_check( ms_get_history( s, h, nrows, nrows ), 'ms_get_history' )
      call case_pull( s, fld, m%ncells, m%nfaces )
      call case_destroy( s )
      conv = .false.
      do k = 1, nrows
         conv = h(8*(k-1)+1) < ctrl%outer_tol .and. h(8*(k-1)+2) < ctrl%outer_tol .and. &
                h(8*(k-1)+3) < ctrl%outer_tol
         if ( mod( k, 50 ) == 0 .or. k == 1 .or. conv ) &
            write(*,'(a,i6,a,es10.3,a,es10.3,a,es10.3,a,i5,a,es9.2)') &
               '  it=', k, '  mass-imbal=', h(8*(k-1)+1), '  du_max=', h(8*(k-1)+2), &
               '  dT_max=', h(8*(k-1)+3), '  lin-it=', nint( h(8*(k-1)+5) ), '  lin-res=', h(8*(k-1)+6)
      end do
      ier = ierc
      if ( ier == NAN_IER ) then
         call case_nan_guard( m, g, fld, m%ncells, 'SIMPLE', nrows + 1, ier )
         write(*,'(a,i0,a)') '[uns-mpi] SIMPLE aborted at outer iteration ', nrows + 1, ' (non-finite solution)'
         ier = NAN_IER
         return
      end if
      if ( ier /= 0 ) return
      if ( conv ) then
         write(*,'(a,i0,a)') 'CONVERGED after ', nrows, ' outer iterations'
      else if ( nrows >= ctrl%outer_max ) then
         write(*,'(a)') 'WARNING: SIMPLE reached outer_max without converging'
      end if
   end subroutine case_run_steady

   subroutine case_destroy( s )
      type(c_ptr), intent(inout) :: s
      call case_check( ms_system_destroy( s ), 'ms_system_destroy' )
      s = c_null_ptr
   end subroutine case_destroy

end module mixsolver_case
