! case_dump -- reads a case with the reference's own set-up routines
! (read_cas, read_control, resolve_cell_zones, build_bc, init_fields,
! setup_porous_fields, as uns_run_mpi does) and writes what
! mixsolver_case.f90 hands to libmixsolver_hip.so: the ms_ctrl_t record, the
! ms_bcgroup_t records, the mesh arrays, the cell types and the per-cell
! porous coefficients.  Test infrastructure: it links the reference's
! modules, so oracle/build_ref.sh builds it into oracle/_ref and it runs in
! this container only; tests/test_oracle_golden.py compares its output with
! the Python control parser and mesh reader.
!
! usage: case_dump <mesh.cas> <unMesh.control> <out.bin>
program case_dump
   use iso_c_binding
   use mod_precision, only: dp
   use mod_uns_mesh
   use mod_uns_cas_reader, only: read_cas
   use mod_uns_connectivity
   use mod_uns_geometry
   use mod_uns_control
   use mod_uns_bc
   use mod_uns_fields
   use mixsolver_c, only: ms_ctrl_t, ms_bcgroup_t
   use mixsolver_case, only: case_convert
   implicit none
   character(len=512) :: casf, ctlf, outf
   type(mesh_t) :: m
   type(conn_t) :: c
   type(geom_t) :: g
   type(ctrl_t) :: ctrl
   type(bc_t) :: bcs
   type(fields_t) :: fld
   type(ms_ctrl_t) :: cc
   type(ms_bcgroup_t), allocatable :: groups(:)
   real(c_double), allocatable :: x(:)
   integer(c_long_long), allocatable :: fptr(:)
   integer(c_int), allocatable :: fnodes(:), c0(:), c1(:), zone(:)
   integer(c_int8_t), allocatable :: raw(:)
   integer :: ier, uo, i
   call get_command_argument(1, casf)
   call get_command_argument(2, ctlf)
   call get_command_argument(3, outf)
   call read_cas( trim(casf), m, ier );            call chk( ier, 'read_cas' )
   call build_connectivity( m, c, ier );           call chk( ier, 'connectivity' )
   call compute_geometry( m, g, ier );             call chk( ier, 'geometry' )
   call read_control( trim(ctlf), ctrl, ier );     call chk( ier, 'control' )
   call resolve_cell_zones( m, ctrl, ier );        call chk( ier, 'cell zones' )
   call build_bc( m, ctrl, bcs, ier );             call chk( ier, 'bc' )
   call init_fields( m, g, fld )
   call setup_porous_fields( m, ctrl, fld )
   call case_convert( m, ctrl, bcs, cc, groups, x, fptr, fnodes, c0, c1, zone )
   open( newunit=uo, file=trim(outf), access='stream', form='unformatted', status='replace' )
   write(uo) int(m%nnodes, c_int), int(m%nfaces, c_int), int(m%ncells, c_int), int(bcs%nb, c_int)
   raw = transfer( cc, raw, int(c_sizeof(cc)) )
   write(uo) raw
   do i = 1, bcs%nb
      raw = transfer( groups(i), raw, int(c_sizeof(groups(i))) )
      write(uo) raw
   end do
   write(uo) x, fptr, fnodes, c0, c1, zone
   if ( allocated( m%cztype ) ) then
      write(uo) int(1, c_int), int(m%cztype(1:m%ncells), c_int)
   else
      write(uo) int(0, c_int)
   end if
   write(uo) fld%porosity, fld%perm_dir, fld%inertial, fld%k_s, fld%q_src
   close( uo )
contains
   subroutine chk( e, what )
      integer, intent(in) :: e
      character(len=*), intent(in) :: what
      if ( e /= 0 ) then
         write(*,'(a,a,i0)') what, ' failed: ', e
         error stop 2
      end if
   end subroutine chk
end program case_dump
