! Host-only check of the Fortran binding: struct sizes equal the C sizes
! passed on the command line, and a host-only entry point is callable.
program check_binding
   use iso_c_binding
   use mixsolver_c
   implicit none
   type(ms_ctrl_t) :: ctrl
   type(ms_bcgroup_t) :: g(2)
   real(c_double) :: xc(3, 8)
   integer(c_int) :: agg(8), rc, i
   character(len=32) :: a1, a2
   call get_command_argument(1, a1)
   call get_command_argument(2, a2)
   do i = 1, 8
      xc(:, i) = [real(i, c_double), 0.0_c_double, 0.0_c_double]
   end do
   rc = ms_coarse_aggregates(8_c_int, xc, 4_c_int, agg)
   print '(a,i0,a,i0,a,i0)', 'sizes ', c_sizeof(ctrl), ' ', c_sizeof(g(1)), ' rc ', rc
   print '(a,8i2)', 'agg', agg
   if (trim(a1) /= '') then
      if (c_sizeof(ctrl) /= str2int(a1) .or. c_sizeof(g(1)) /= str2int(a2)) stop 2
   end if
   if (rc /= MS_OK .or. any(agg /= [0, 0, 1, 1, 2, 2, 3, 3])) stop 3
contains
   integer function str2int(s)
      character(len=*), intent(in) :: s
      read (s, *) str2int
   end function
end program check_binding
