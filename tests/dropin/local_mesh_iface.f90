! local_mesh_iface.f90 -- the components of local_mesh_t
! (mod_uns_local_mesh.f90:38-56) that the loop-level drop-in
! mixsolver_amd/fortran/mod_uns_simple_mpi.f90 reads (m, c, g, nowned,
! nghost), so that drop-in can be compiled and type-checked in a build
! without MPI.  The real module needs MPI and is the one the reference's
! build links; this file is never linked into a program.
module mod_uns_local_mesh
   use mod_uns_mesh, only: mesh_t
   use mod_uns_connectivity, only: conn_t
   use mod_uns_geometry, only: geom_t
   implicit none
   private
   public :: local_mesh_t
   type :: local_mesh_t
      type(mesh_t) :: m
      type(conn_t) :: c
      type(geom_t) :: g
      integer :: nowned = 0
      integer :: nghost = 0
   end type local_mesh_t
end module mod_uns_local_mesh
