! Calls the hx3d C ABI from Fortran through hx3d_binding without touching a GPU: descriptor
! validation happens before any device call, so invalid descriptors must come back as
! HX3D_EINVAL with a message, and a null engine must be rejected.
program abi_check
   use iso_c_binding
   use hx3d_binding
   implicit none
   type(hx3d_case_desc) :: d
   type(c_ptr) :: eng
   integer(c_int) :: rc
   integer(c_int), target :: N(3) = [3, 3, 3]
   real(c_double) :: dt, res(5)

   rc = hx3d_engine_create(d, eng)            ! K = 0
   if (rc /= HX3D_EINVAL) error stop "K = 0 not rejected"
   write(*, '(A)') "K=0: " // hx3d_error_message(eng)
   rc = hx3d_engine_destroy(eng)

   d % K = 1
   d % N = c_loc(N)                           ! connectivity and geometry missing
   rc = hx3d_engine_create(d, eng)
   if (rc /= HX3D_EINVAL) error stop "missing arrays not rejected"
   write(*, '(A)') "missing arrays: " // hx3d_error_message(eng)
   rc = hx3d_engine_destroy(eng)

   if (hx3d_max_dt(c_null_ptr, 0.4_c_double, 0.4_c_double, dt) /= HX3D_EINVAL) error stop "null engine"
   if (hx3d_residuals(c_null_ptr, res) /= HX3D_EINVAL) error stop "null engine"
   if (hx3d_num_nodes(c_null_ptr) /= -1) error stop "null engine"
   write(*, '(A)') "abi_check ok"
end program abi_check
