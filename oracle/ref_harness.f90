!===============================================================================
! ref_harness.f90 -- fixture writer driving the reference's own FP64 modules
!
! Test infrastructure only: this program is linked against the reference
! sources compiled from /root/reference by oracle/build_ref.sh and is used to
! write golden vectors for tests/golden.  It never ships in the product.
!
! usage: ref_harness <mesh.cas> <unMesh.control> <mode> <out.bin> [in.bin] [n]
!   geom   : geometry, lf, connectivity and CSR pattern
!   simple : seed fields from in.bin (u,p,T), run simple_run(nsteps=n)
!   piso   : seed fields from in.bin (u,p,T), run piso_run (n_time_max steps)
!   pimple : seed fields from in.bin (u,p,T), run pimple_run (n_time_max steps)
!   pisor  : a restored state from in.bin (u,p,T, u_old, u_old_old, T_old,
!            T_old_old, then ts_order as int), T_s and its history = T's, then
!            piso_run (n_time_max steps): the first step of a restarted run
!   funcs  : seed fields from in.bin, then compute_gradients, momentum_assembly
!            (comp 1..3), flux_rhiechow, temperature_assembly, correct_fields
!   funcs2 : funcs with T_s / T_s_old / T_s_old_old seeded too, preceded by the
!            resolved cell types and porous coefficients, with the solid-phase
!            gradient and solid_temperature_assembly added
!   forces : seed fields from in.bin (u,p,T), then compute_forces (written),
!            write_cp_distribution (to cp.dat in the working directory) and
!            compute_separation (printed)
!   vtu    : seed fields from in.bin (u,p,T), T_s = 0.5 T + 1, then vtk_write
!            to a.vtu (inline ASCII) and b.vtu (appended binary) in the
!            working directory
!   ifuncs : funcs2 with coupling-interface data: after the funcs2 inputs,
!            in.bin holds the four interface switches (int p_dir, slip, zhang,
!            T_ltne), then over all faces vel(3), p, T, Ts, q, slip_C,
!            peer_vel(3) (doubles) and T_mode (int); they are applied to the
!            faces of the BC_INTERFACE group through set_interface_* and the
!            set_iface_* switches
!   isimple: u, p, T then the interface block as in ifuncs; simple_run(n)
!   solve  : CSR system from in.bin; residual histories of cg_jacobi, cg_ic0,
!            bicgstab_ilu0 for maxit = 1..n, IC0/ILU0 factors, final solutions
! All arrays are written as raw little-endian stream records in the order the
! Python reader (tests/golden/make_golden.py) expects.
!===============================================================================
program ref_harness
   use mod_precision, only: dp
   use mod_uns_mesh
   use mod_uns_cas_reader, only: read_cas
   use mod_uns_connectivity
   use mod_uns_geometry
   use mod_uns_control
   use mod_uns_bc
   use mod_uns_fields
   use mod_uns_linsolver
   use mod_uns_simple
   use mod_uns_forces, only: compute_forces, write_cp_distribution, compute_separation
   use mod_uns_output, only: vtk_write
   implicit none

   character(len=512) :: casf, ctlf, mode, outf, inf, narg
   type(mesh_t)   :: m
   type(conn_t)   :: c
   type(geom_t)   :: g
   type(ctrl_t)   :: ctrl
   type(bc_t)     :: bcs
   type(fields_t) :: fld
   integer :: ier, nsteps, uo, ui, nargs
   integer(8) :: t0, t1, trate

   nargs = command_argument_count()
   call get_command_argument(1, casf)
   call get_command_argument(2, ctlf)
   call get_command_argument(3, mode)
   call get_command_argument(4, outf)
   inf = ''
   if ( nargs >= 5 ) call get_command_argument(5, inf)
   nsteps = 1
   if ( nargs >= 6 ) then
      call get_command_argument(6, narg)
      read(narg, *) nsteps
   end if

   if ( trim(mode) == 'solve' ) then
      call run_solve()
      stop
   end if

   call read_cas( trim(casf), m, ier );            call chk(ier, 'read_cas')
   call build_connectivity( m, c, ier );           call chk(ier, 'connectivity')
   call compute_geometry( m, g, ier );             call chk(ier, 'geometry')
   call read_control( trim(ctlf), ctrl, ier );     call chk(ier, 'control')
   call resolve_cell_zones( m, ctrl, ier );        call chk(ier, 'cell zones')
   call build_bc( m, ctrl, bcs, ier );             call chk(ier, 'bc')
   call init_fields( m, g, fld )
   call setup_porous_fields( m, ctrl, fld )

   open( newunit=uo, file=trim(outf), access='stream', form='unformatted', &
         status='replace', action='write' )

   select case ( trim(mode) )
   case ( 'geom' )
      call dump_geom()
   case ( 'simple' )
      call seed_fields()
      call system_clock( t0, trate )
      call simple_run( m, c, g, ctrl, bcs, fld, ier, nsteps )
      call system_clock( t1 )
      write(*,'(a,f12.3)') 'HARNESS_DRIVER_SECONDS ', real(t1 - t0, dp) / real(trate, dp)
      write(uo) ier
      call dump_state()
   case ( 'piso' )
      call seed_fields()
      call system_clock( t0, trate )
      call piso_run( m, c, g, ctrl, bcs, fld, '/tmp/ref_harness_piso', ier )
      call system_clock( t1 )
      write(*,'(a,f12.3)') 'HARNESS_DRIVER_SECONDS ', real(t1 - t0, dp) / real(trate, dp)
      write(uo) ier
      call dump_state()
   case ( 'pimple' )
      call seed_fields()
      call pimple_run( m, c, g, ctrl, bcs, fld, '/tmp/ref_harness_pimple', ier )
      write(uo) ier
      call dump_state()
   case ( 'pisor' )
      open( newunit=ui, file=trim(inf), access='stream', form='unformatted', &
            status='old', action='read' )
      read(ui) fld%u, fld%p, fld%T, fld%u_old, fld%u_old_old, fld%T_old, fld%T_old_old
      read(ui) fld%ts_order
      close( ui )
      fld%T_s = fld%T
      fld%T_s_old = fld%T_old
      fld%T_s_old_old = fld%T_old_old
      call piso_run( m, c, g, ctrl, bcs, fld, '/tmp/ref_harness_pisor', ier )
      write(uo) ier
      call dump_state()
   case ( 'forces' )
      call seed_fields()
      call run_forces()
   case ( 'vtu' )
      call seed_fields()
      fld%T_s = 0.5_dp * fld%T + 1.0_dp
      call vtk_write( m, c, g, fld, 'a.vtu', ier )
      call chk( ier, 'vtk_write ascii' )
      call vtk_write( m, c, g, fld, 'b.vtu', ier, binary = .true. )
      call chk( ier, 'vtk_write binary' )
      write(uo) ier
   case ( 'funcs' )
      call run_funcs()
   case ( 'funcs2' )
      call run_funcs2()
   case ( 'ifuncs' )
      call run_funcs2( .true. )
   case ( 'isimple' )
      open( newunit=ui, file=trim(inf), access='stream', form='unformatted', &
            status='old', action='read' )
      read(ui) fld%u, fld%p, fld%T
      call read_iface( ui )
      close( ui )
      call simple_run( m, c, g, ctrl, bcs, fld, ier, nsteps )
      write(uo) ier
      call dump_state()
   case default
      write(*,*) 'unknown mode ', trim(mode)
      stop 2
   end select
   close( uo )

contains

   subroutine chk( ierr, what )
      integer, intent(in) :: ierr
      character(len=*), intent(in) :: what
      if ( ierr /= 0 ) then
         write(*,'(a,a,i0)') 'ref_harness: failed in ', what, ier
         stop 3
      end if
   end subroutine chk

   subroutine dump_geom()
      type(csr_t) :: A
      write(uo) m%ncells, m%nfaces, size(c%cf)
      write(uo) g%xf, g%sf, g%area, g%vol, g%xc, fld%lf
      write(uo) c%cf_ptr, c%cf, c%c2c_ptr, c%c2c
      call build_cell_csr( m, c, A )
      write(uo) size(A%col_idx)
      write(uo) A%row_ptr, A%col_idx
      write(uo) bcs%fgrp
   end subroutine dump_geom

   ! in.bin: u(3,nc), p(nc), T(nc)
   subroutine seed_fields()
      if ( len_trim(inf) == 0 ) return
      open( newunit=ui, file=trim(inf), access='stream', form='unformatted', &
            status='old', action='read' )
      read(ui) fld%u, fld%p, fld%T
      close( ui )
      fld%T_s = fld%T
   end subroutine seed_fields

   subroutine run_forces()
      real(dp) :: Fx, Fy, Fz, CD, CL, u_inf, A_ref, p_inf
      call compute_forces( m, g, ctrl, bcs, fld, Fx, Fy, Fz, CD, CL, u_inf, A_ref, p_inf )
      write(uo) Fx, Fy, Fz, CD, CL, u_inf, A_ref, p_inf
      call write_cp_distribution( m, g, ctrl, bcs, fld, 'cp.dat' )
      call compute_separation( m, g, ctrl, bcs, fld )
   end subroutine run_forces

   subroutine dump_state()
      write(uo) fld%u, fld%p, fld%flux, fld%T, fld%gp, fld%gu, fld%apc, fld%gt
      write(uo) fld%T_s
   end subroutine dump_state

   ! in.bin: u, p, T, flux(nf), u_old, u_old_old, T_old, T_old_old, apc(nc),
   !         pp(nc), ts_order (int)
   subroutine run_funcs()
      type(csr_t) :: A
      real(dp), allocatable :: rhs(:), ap(:), pp(:), phif(:), dpp(:,:)
      integer :: comp, tso
      allocate( rhs(m%ncells), ap(m%ncells), pp(m%ncells), phif(m%nfaces), &
                dpp(3,m%ncells) )
      open( newunit=ui, file=trim(inf), access='stream', form='unformatted', &
            status='old', action='read' )
      read(ui) fld%u, fld%p, fld%T, fld%flux, fld%u_old, fld%u_old_old, &
               fld%T_old, fld%T_old_old, fld%apc, pp, tso
      close( ui )
      fld%ts_order = tso
      fld%T_s = fld%T
      call build_cell_csr( m, c, A )
      call compute_gradients( m, g, bcs, fld )
      write(uo) fld%gp, fld%gu, fld%gt
      do comp = 1, 3
         call momentum_assembly( m, g, ctrl, bcs, fld, comp, A, rhs, ap )
         write(uo) A%val, rhs, ap
      end do
      call temperature_assembly( m, g, ctrl, bcs, fld, A, rhs, ap )
      write(uo) A%val, rhs, ap
      call flux_rhiechow( m, g, ctrl, bcs, fld )
      write(uo) fld%flux
      call outflow_mass_rescale( m, g, bcs, fld )
      write(uo) fld%flux
      call correct_fields( m, g, ctrl, bcs, fld, pp, phif, dpp )
      write(uo) fld%u, fld%p, fld%flux, dpp
   end subroutine run_funcs

   ! in.bin: as run_funcs, then T_s(nc), T_s_old(nc), T_s_old_old(nc)
   ! interface block of in.bin (see the usage notes above)
   subroutine read_iface( un )
      integer, intent(in) :: un
      integer :: sw(4), ig, k, nfi, nq
      real(dp), allocatable :: vel(:,:), pv(:), tv(:), tsv(:), qv(:), cv(:), peer(:,:)
      integer, allocatable :: tmode(:), faces(:), qfaces(:)
      allocate( vel(3,m%nfaces), pv(m%nfaces), tv(m%nfaces), tsv(m%nfaces), qv(m%nfaces), &
                cv(m%nfaces), peer(3,m%nfaces), tmode(m%nfaces) )
      read(un) sw
      read(un) vel, pv, tv, tsv, qv, cv, peer, tmode
      ig = 0
      do k = 1, bcs%nb
         if ( bcs%gb(k)%btype == BC_INTERFACE ) ig = k
      end do
      if ( ig == 0 ) stop 4
      nfi = bcs%gb(ig)%nf
      allocate( faces(nfi) )
      faces = bcs%gb(ig)%faces(1:nfi)
      call set_interface_vel( bcs, faces, vel(:, faces) )
      call set_interface_p( bcs, faces, pv(faces) )
      call set_interface_T_ltne( bcs, faces, tv(faces), tsv(faces) )
      call set_interface_slip( bcs, faces, cv(faces), peer(:, faces) )
      nq = count( tmode(faces) == 1 )
      allocate( qfaces(nq) )
      qfaces = pack( faces, tmode(faces) == 1 )
      if ( nq > 0 ) call set_interface_q( bcs, qfaces, qv(qfaces) )
      call set_iface_p_dirichlet( sw(1) /= 0 )
      call set_iface_slip_mode( sw(2) /= 0 )
      call set_iface_zhang_vel( sw(3) /= 0 )
      call set_iface_t_ltne_mode( sw(4) /= 0 )
   end subroutine read_iface

   subroutine run_funcs2( with_iface )
      logical, intent(in), optional :: with_iface
      type(csr_t) :: A
      real(dp), allocatable :: rhs(:), ap(:), pp(:), phif(:), dpp(:,:)
      integer :: comp, tso
      allocate( rhs(m%ncells), ap(m%ncells), pp(m%ncells), phif(m%nfaces), &
                dpp(3,m%ncells) )
      open( newunit=ui, file=trim(inf), access='stream', form='unformatted', &
            status='old', action='read' )
      read(ui) fld%u, fld%p, fld%T, fld%flux, fld%u_old, fld%u_old_old, &
               fld%T_old, fld%T_old_old, fld%apc, pp, tso
      read(ui) fld%T_s, fld%T_s_old, fld%T_s_old_old
      if ( present(with_iface) ) then
         if ( with_iface ) call read_iface( ui )
      end if
      close( ui )
      fld%ts_order = tso
      write(uo) m%cztype
      write(uo) fld%perm_dir, fld%inertial, fld%porosity, fld%disp_l, fld%disp_t, &
                fld%bj_alpha, fld%k_s, fld%cp_s, fld%rho_s, fld%h_sf, fld%a_sf, fld%q_src
      call build_cell_csr( m, c, A )
      call compute_gradients( m, g, bcs, fld )
      write(uo) fld%gp, fld%gu, fld%gt, fld%gts
      do comp = 1, 3
         call momentum_assembly( m, g, ctrl, bcs, fld, comp, A, rhs, ap )
         write(uo) A%val, rhs, ap
      end do
      call temperature_assembly( m, g, ctrl, bcs, fld, A, rhs, ap )
      write(uo) A%val, rhs, ap
      call solid_temperature_assembly( m, g, ctrl, bcs, fld, A, rhs, ap )
      write(uo) A%val, rhs, ap
      call flux_rhiechow( m, g, ctrl, bcs, fld )
      write(uo) fld%flux
      call outflow_mass_rescale( m, g, bcs, fld )
      write(uo) fld%flux
      call correct_fields( m, g, ctrl, bcs, fld, pp, phif, dpp )
      write(uo) fld%u, fld%p, fld%flux, dpp, fld%u_old
   end subroutine run_funcs2

   ! in.bin: n, nnz, row_ptr(n+1), col_idx(nnz), val(nnz), b(n), x0(n),
   !         tol, kmax ;   nsteps = maxit used for the final solves
   subroutine run_solve()
      type(csr_t) :: A, L
      real(dp), allocatable :: b(:), x0(:), x(:)
      real(dp) :: tol, res
      integer :: n, nnz, kmax, k, it, ierr, isolver
      open( newunit=ui, file=trim(inf), access='stream', form='unformatted', &
            status='old', action='read' )
      read(ui) n, nnz
      A%nrows = n
      allocate( A%row_ptr(n+1), A%col_idx(nnz), A%val(nnz), b(n), x0(n), x(n) )
      read(ui) A%row_ptr, A%col_idx, A%val, b, x0, tol, kmax
      close( ui )
      open( newunit=uo, file=trim(outf), access='stream', form='unformatted', &
            status='replace', action='write' )
      L = A
      call ic0_factor( L, ierr )
      write(uo) ierr, L%val
      L = A
      call ilu0_factor( L, ierr )
      write(uo) ierr, L%val
      do isolver = 1, 3
         do k = 1, kmax
            x = x0
            call run_one( isolver, A, b, x, tol, k, it, res, ierr )
            write(uo) it, ierr, res
         end do
         x = x0
         call run_one( isolver, A, b, x, tol, nsteps, it, res, ierr )
         write(uo) it, ierr, res, x
      end do
      close( uo )
   end subroutine run_solve

   subroutine run_one( isolver, A, b, x, tol, maxit, it, res, ierr )
      integer, intent(in) :: isolver, maxit
      type(csr_t), intent(in) :: A
      real(dp), intent(in) :: b(:), tol
      real(dp), intent(inout) :: x(:)
      integer, intent(out) :: it, ierr
      real(dp), intent(out) :: res
      select case ( isolver )
      case ( 1 )
         call cg_jacobi( A, b, x, tol, maxit, it, res, ierr )
      case ( 2 )
         call cg_ic0( A, b, x, tol, maxit, it, res, ierr )
      case default
         call bicgstab_ilu0( A, b, x, tol, maxit, it, res, ierr )
      end select
   end subroutine run_one

end program ref_harness
