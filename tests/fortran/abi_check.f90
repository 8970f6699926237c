! Calls the hx3d C ABI from Fortran through hx3d_binding without touching a GPU: descriptor
! validation happens before any device call, so invalid descriptors must come back as
! HX3D_EINVAL with a message, and a null engine must be rejected.  Then the host-only .control
! reader on the case.control and box.mesh of the working directory.
program abi_check
   use iso_c_binding
   use hx3d_binding
   implicit none
   type(hx3d_case_desc) :: d
   type(c_ptr) :: eng
   integer(c_int) :: rc
   integer(c_int), target :: N(3) = [3, 3, 3]
   real(c_double) :: dt, res(5)
   type(c_ptr) :: ctl, mesh
   type(hx3d_case_desc) :: dc
   type(hx3d_control_run) :: run
   character(kind=c_char, len=32) :: btype, bcoupled

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

   rc = hx3d_control_read("case.control" // c_null_char, ctl)
   if (rc /= HX3D_OK) error stop "control file not read"
   rc = hx3d_control_case(ctl, c_null_char, mesh, dc, run)
   if (rc /= HX3D_OK) error stop "control case not built"
   if (dc % K /= 8 .or. dc % F /= 24 .or. dc % mach /= 0.1_c_double .or. dc % reynolds /= 1600.0_c_double) &
      error stop "control case settings"
   if (run % steps /= 10 .or. run % output_interval /= 5 .or. run % dt /= 1.0e-3_c_double) error stop "run keys"
   rc = hx3d_control_boundary(ctl, "LEFT" // c_null_char, btype, bcoupled, len(btype, c_int))
   if (rc /= HX3D_OK .or. btype(1:9) /= "periodic" // c_null_char .or. bcoupled(1:6) /= "right" // c_null_char) &
      error stop "boundary block"
   rc = hx3d_mesh_destroy(mesh)
   rc = hx3d_control_destroy(ctl)
   write(*, '(A)') "control ok"
   write(*, '(A)') "abi_check ok"
end program abi_check
