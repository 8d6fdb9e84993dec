!===============================================================================
! mod_uns_simple_mpi.f90 -- loop-level drop-in replacement of the reference
! module of the same name (src/unstructured/mod_uns_simple_mpi.f90, publics
! at :50) for a run on ONE partition.
!
! Compiled in place of the reference file (with mixsolver_c.f90,
! mixsolver_case.f90 and libmixsolver_hip.so), it keeps the public names and
! signatures that uns_run_mpi (mod_uns_driver_mpi.f90:319-329) calls, and runs
! the whole loop on the MI355X: the mesh, matrices and fields are created on
! the device once, stay there for every iteration or time step, and come back
! to `fld` at the snapshot steps (n_out_every, written through the reference's
! vtk_write as write_snapshot_mpi does at one partition) and at the end.
!
!   simple_run_mpi (:57)  -> ms_simple_run
!   piso_run_mpi   (:298) -> ms_piso_run, continued with ms_piso_continue
!   pimple_run_mpi (:580) -> ms_pimple_run, continued with ms_pimple_continue
!   nan_guard_mpi  (:1024) -> nan_scan / nan_report of the reference
!
! Error codes are the reference's (20-23, 30-36, 50-56, 60).  At one
! partition the local mesh is the global one, its halo is empty and every
! collective is an identity, so the halo argument is accepted as any type and
! never read.  A partitioned run (halo cells) stops with a message.
!===============================================================================
module mod_uns_simple_mpi
   use mod_precision, only: dp
   use mod_uns_mesh, only: mesh_t
   use mod_uns_connectivity, only: conn_t
   use mod_uns_geometry, only: geom_t
   use mod_uns_control, only: ctrl_t
   use mod_uns_bc, only: bc_t
   use mod_uns_fields, only: fields_t
   use mod_uns_local_mesh, only: local_mesh_t
   use mixsolver_case, only: case_run_transient, case_run_steady, case_nan_guard, RUN_PISO, RUN_PIMPLE
   implicit none
   private
   public :: simple_run_mpi, piso_run_mpi, pimple_run_mpi, nan_guard_mpi

contains

   subroutine one_partition( lm, who )
      type(local_mesh_t), intent(in) :: lm
      character(len=*),   intent(in) :: who
      if ( lm%nghost /= 0 .or. lm%nowned /= lm%m%ncells ) then
         write(*,'(a,a,i0,a,i0)') who, ': the device library runs one partition only; nowned=', &
            lm%nowned, ' nghost=', lm%nghost
         error stop 1
      end if
   end subroutine one_partition

   subroutine simple_run_mpi( lm, hi, ctrl, bcs, fld, ier )
      type(local_mesh_t), intent(in)    :: lm
      type(*),            intent(in)    :: hi
      type(ctrl_t),       intent(in)    :: ctrl
      type(bc_t),         intent(in)    :: bcs
      type(fields_t),     intent(inout) :: fld
      integer,            intent(out)   :: ier
      call one_partition( lm, 'simple_run_mpi' )
      call case_run_steady( lm%m, lm%g, ctrl, bcs, fld, ier )
   end subroutine simple_run_mpi

   subroutine piso_run_mpi( lm, hi, ctrl, bcs, fld, vtu_prefix, m_g, c_g, g_g, part, fld_g, ier )
      type(local_mesh_t), intent(in)    :: lm
      type(*),            intent(in)    :: hi
      type(ctrl_t),       intent(in)    :: ctrl
      type(bc_t),         intent(in)    :: bcs
      type(fields_t),     intent(inout) :: fld
      character(len=*),   intent(in)    :: vtu_prefix
      type(mesh_t),       intent(in)    :: m_g
      type(conn_t),       intent(in)    :: c_g
      type(geom_t),       intent(in)    :: g_g
      integer,            intent(in)    :: part(:)
      type(fields_t),     intent(inout) :: fld_g
      integer,            intent(out)   :: ier
      call one_partition( lm, 'piso_run_mpi' )
      ! one partition: the global mesh is the local one, cell for cell
      call case_run_transient( RUN_PISO, m_g, c_g, g_g, ctrl, bcs, fld, vtu_prefix, ier )
   end subroutine piso_run_mpi

   subroutine pimple_run_mpi( lm, hi, ctrl, bcs, fld, vtu_prefix, m_g, c_g, g_g, part, fld_g, ier )
      type(local_mesh_t), intent(in)    :: lm
      type(*),            intent(in)    :: hi
      type(ctrl_t),       intent(in)    :: ctrl
      type(bc_t),         intent(in)    :: bcs
      type(fields_t),     intent(inout) :: fld
      character(len=*),   intent(in)    :: vtu_prefix
      type(mesh_t),       intent(in)    :: m_g
      type(conn_t),       intent(in)    :: c_g
      type(geom_t),       intent(in)    :: g_g
      integer,            intent(in)    :: part(:)
      type(fields_t),     intent(inout) :: fld_g
      integer,            intent(out)   :: ier
      call one_partition( lm, 'pimple_run_mpi' )
      call case_run_transient( RUN_PIMPLE, m_g, c_g, g_g, ctrl, bcs, fld, vtu_prefix, ier )
   end subroutine pimple_run_mpi

   subroutine nan_guard_mpi( lm, fld, stage, it, ier, t_val )
      type(local_mesh_t), intent(in)  :: lm
      type(fields_t),     intent(in)  :: fld
      character(len=*),   intent(in)  :: stage
      integer,            intent(in)  :: it
      integer,            intent(out) :: ier
      real(dp), optional, intent(in)  :: t_val
      if ( present(t_val) ) then
         call case_nan_guard( lm%m, lm%g, fld, lm%nowned, stage, it, ier, t_val )
      else
         call case_nan_guard( lm%m, lm%g, fld, lm%nowned, stage, it, ier )
      end if
   end subroutine nan_guard_mpi

end module mod_uns_simple_mpi
