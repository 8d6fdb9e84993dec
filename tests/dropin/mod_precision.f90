! Kinds module for the drop-in check program.  The drop-in solver modules
! name their real kind `dp` from a module called mod_precision, as the
! reference's callers do; this one defines the kinds the check needs and
! nothing else.
module mod_precision
   use iso_fortran_env, only: real64, int32
   implicit none
   integer, parameter :: dp = real64
   integer, parameter :: ip = int32
   real(dp), parameter :: pi = 3.14159265358979323846_dp
end module mod_precision
