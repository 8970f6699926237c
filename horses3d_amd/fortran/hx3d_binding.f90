!
!  hx3d_binding: ISO_C_BINDING interface to the hx3d engine C ABI (include/hx3d.h).
!
!  Standalone (no HORSES3D modules needed): every entry point of the C ABI with its C
!  signature, the case descriptor as an interoperable derived type, and the selector and
!  status constants.  Pointer members of the descriptor are type(c_ptr) and are set with
!  c_loc() of TARGET arrays that the caller keeps alive for the duration of
!  hx3d_engine_create.
!
module hx3d_binding
   use iso_c_binding
   implicit none

   integer(c_int), parameter :: HX3D_OK = 0, HX3D_EINVAL = 1, HX3D_EUNSUPPORTED = 2, &
                                HX3D_EDEVICE = 3, HX3D_ENOMEM = 4, HX3D_ENAN = 5
   ! "riemann solver"
   integer(c_int), parameter :: HX3D_RIEMANN_ROE = 0, HX3D_RIEMANN_LXF = 1
   ! "inviscid discretization"
   integer(c_int), parameter :: HX3D_INVISCID_STANDARD = 0, HX3D_INVISCID_SPLIT = 1
   ! "averaging"
   integer(c_int), parameter :: HX3D_AVG_STANDARD = 0, HX3D_AVG_DUCROS = 1, &
                                HX3D_AVG_KENNEDYGRUBER = 2, HX3D_AVG_PIROZZOLI = 3, &
                                HX3D_AVG_MORINISHI = 4
   ! "les model"
   integer(c_int), parameter :: HX3D_LES_NONE = 0, HX3D_LES_SMAGORINSKY = 1, HX3D_LES_WALE = 2
   ! "discretization nodes"
   integer(c_int), parameter :: HX3D_NODES_GAUSSLOBATTO = 0, HX3D_NODES_GAUSS = 1

   type, bind(C) :: hx3d_case_desc
      integer(c_int) :: K = 0
      type(c_ptr)    :: N = c_null_ptr          ! int [3K]
      type(c_ptr)    :: faceID = c_null_ptr     ! int [6K], 0-based
      type(c_ptr)    :: faceSide = c_null_ptr   ! int [6K], 1 left / 2 right
      integer(c_int) :: F = 0
      type(c_ptr)    :: fElem = c_null_ptr      ! int [2F], 0-based
      type(c_ptr)    :: fESide = c_null_ptr     ! int [2F], 1..6
      type(c_ptr)    :: fRot = c_null_ptr       ! int [F], 0..7
      type(c_ptr)    :: Ja = c_null_ptr         ! double [9*nodes]
      type(c_ptr)    :: J = c_null_ptr          ! double [nodes]
      type(c_ptr)    :: fNormal = c_null_ptr    ! double [3*F*(N+1)^2] or null
      type(c_ptr)    :: fT1 = c_null_ptr
      type(c_ptr)    :: fT2 = c_null_ptr
      type(c_ptr)    :: fJac = c_null_ptr       ! double [F*(N+1)^2] or null
      real(c_double) :: mach = 0.0_c_double, reynolds = 0.0_c_double, prandtl = 0.72_c_double
      integer(c_int) :: euler = 0
      integer(c_int) :: riemann = HX3D_RIEMANN_ROE, inviscid = HX3D_INVISCID_STANDARD
      integer(c_int) :: averaging = HX3D_AVG_STANDARD
      real(c_double) :: lambda_stab = 1.0_c_double
      integer(c_int) :: les = HX3D_LES_NONE
      real(c_double) :: les_c = 0.0_c_double
      integer(c_int) :: nodes = HX3D_NODES_GAUSSLOBATTO
   end type hx3d_case_desc

   ! volume monitors (hx3d_volume_monitor kinds)
   integer(c_int), parameter :: HX3D_MON_KINETIC_ENERGY = 0, HX3D_MON_KINETIC_ENERGY_RATE = 1, &
                                HX3D_MON_ENSTROPHY = 2, HX3D_MON_MEAN_VELOCITY = 3, &
                                HX3D_MON_INTERNAL_ENERGY = 4, HX3D_MON_ENTROPY = 5, &
                                HX3D_MON_MATH_ENTROPY = 6

   ! time-integration keys of a .control file (hx3d_control_case)
   type, bind(C) :: hx3d_control_run
      real(c_double) :: cfl = 0.0_c_double, dcfl = 0.0_c_double, dt = 0.0_c_double
      real(c_double) :: tolerance = 0.0_c_double
      integer(c_int) :: steps = 0, output_interval = 0, restart = 0
   end type hx3d_control_run

   interface
      integer(c_int) function hx3d_engine_create(desc, eng) bind(C, name="hx3d_engine_create")
         import :: c_int, c_ptr, hx3d_case_desc
         type(hx3d_case_desc), intent(in) :: desc
         type(c_ptr), intent(out)         :: eng
      end function
      integer(c_int) function hx3d_engine_destroy(eng) bind(C, name="hx3d_engine_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
      end function
      type(c_ptr) function hx3d_last_error(eng) bind(C, name="hx3d_last_error")
         import :: c_ptr
         type(c_ptr), value :: eng
      end function
      integer(c_long) function hx3d_num_nodes(eng) bind(C, name="hx3d_num_nodes")
         import :: c_long, c_ptr
         type(c_ptr), value :: eng
      end function
      integer(c_int) function hx3d_set_state(eng, Q) bind(C, name="hx3d_set_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(in) :: Q(*)
      end function
      integer(c_int) function hx3d_get_state(eng, Q) bind(C, name="hx3d_get_state")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double), intent(out) :: Q(*)
      end function
      integer(c_int) function hx3d_set_state_device(eng, Qd) bind(C, name="hx3d_set_state_device")
         import :: c_int, c_ptr
         type(c_ptr), value :: eng, Qd
      end function
      integer(c_int) function hx3d_get_state_device(eng, Qd) bind(C, name="hx3d_get_state_device")
         import :: c_int, c_ptr
         type(c_ptr), value :: eng, Qd
      end function
      integer(c_int) function hx3d_compute_qdot(eng, t, QDot) bind(C, name="hx3d_compute_qdot")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value    :: eng
         real(c_double), value :: t
         real(c_double)        :: QDot(*)
      end function
      integer(c_int) function hx3d_rk3_step(eng, t, dt) bind(C, name="hx3d_rk3_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value    :: eng
         real(c_double), value :: t, dt
      end function
      integer(c_int) function hx3d_set_ctd_after_steps(eng, on) bind(C, name="hx3d_set_ctd_after_steps")
         import :: c_int, c_ptr
         type(c_ptr), value    :: eng
         integer(c_int), value :: on
      end function
      integer(c_int) function hx3d_get_qdot(eng, QDot) bind(C, name="hx3d_get_qdot")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: eng
         real(c_double)     :: QDot(*)
      end function
      integer(c_int) function hx3d_max_dt(eng, cfl, dcfl, dt) bind(C, name="hx3d_max_dt")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: eng
         real(c_double), value       :: cfl, dcfl
         real(c_double), intent(out) :: dt
      end function
      integer(c_int) function hx3d_advance(eng, nsteps, cfl, dcfl, t) bind(C, name="hx3d_advance")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value            :: eng
         integer(c_int), value         :: nsteps
         real(c_double), value         :: cfl, dcfl
         real(c_double), intent(inout) :: t
      end function
      integer(c_int) function hx3d_advance_dt(eng, nsteps, dt, t) bind(C, name="hx3d_advance_dt")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value            :: eng
         integer(c_int), value         :: nsteps
         real(c_double), value         :: dt
         real(c_double), intent(inout) :: t
      end function
      integer(c_int) function hx3d_residuals(eng, res) bind(C, name="hx3d_residuals")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: eng
         real(c_double), intent(out) :: res(5)
      end function
      integer(c_int) function hx3d_kinetic_energy(eng, ke) bind(C, name="hx3d_kinetic_energy")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: eng
         real(c_double), intent(out) :: ke
      end function
      integer(c_int) function hx3d_check_nan(eng, has_nan) bind(C, name="hx3d_check_nan")
         import :: c_int, c_ptr
         type(c_ptr), value          :: eng
         integer(c_int), intent(out) :: has_nan
      end function
      integer(c_int) function hx3d_synchronize(eng) bind(C, name="hx3d_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: eng
      end function
      integer(c_int) function hx3d_query(eng, done) bind(C, name="hx3d_query")
         import :: c_int, c_ptr
         type(c_ptr), value          :: eng
         integer(c_int), intent(out) :: done
      end function
      integer(c_int) function hx3d_get_time(eng, t) bind(C, name="hx3d_get_time")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: eng
         real(c_double), intent(out) :: t
      end function
      integer(c_long) function hx3d_engine_device_bytes(eng) bind(C, name="hx3d_engine_device_bytes")
         import :: c_long, c_ptr
         type(c_ptr), value :: eng
      end function
      integer(c_int) function hx3d_device_memory(free_bytes, total_bytes) bind(C, name="hx3d_device_memory")
         import :: c_int, c_size_t
         integer(c_size_t), intent(out) :: free_bytes, total_bytes
      end function
      integer(c_int) function hx3d_set_step_graph(eng, on) bind(C, name="hx3d_set_step_graph")
         import :: c_int, c_ptr
         type(c_ptr), value    :: eng
         integer(c_int), value :: on
      end function
      integer(c_int) function hx3d_set_timing(eng, on) bind(C, name="hx3d_set_timing")
         import :: c_int, c_ptr
         type(c_ptr), value    :: eng
         integer(c_int), value :: on
      end function
      integer(c_int) function hx3d_kernel_times(eng, out_ms) bind(C, name="hx3d_kernel_times")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: eng
         real(c_double), intent(out) :: out_ms(4)
      end function
      ! mesh front end: path and periodic pairs as NUL-terminated strings
      integer(c_int) function hx3d_mesh_read_specmesh(path, N, orders, periodic, mesh) &
            bind(C, name="hx3d_mesh_read_specmesh")
         import :: c_int, c_ptr, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value              :: N
         type(c_ptr), value                 :: orders
         character(kind=c_char), intent(in) :: periodic(*)
         type(c_ptr), intent(out)           :: mesh
      end function
      integer(c_int) function hx3d_mesh_read_specmesh_nodes(path, N, orders, periodic, nodes, mesh) &
            bind(C, name="hx3d_mesh_read_specmesh_nodes")
         import :: c_int, c_ptr, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value              :: N
         type(c_ptr), value                 :: orders
         character(kind=c_char), intent(in) :: periodic(*)
         integer(c_int), value              :: nodes
         type(c_ptr), intent(out)           :: mesh
      end function
      type(c_ptr) function hx3d_mesh_error(mesh) bind(C, name="hx3d_mesh_error")
         import :: c_ptr
         type(c_ptr), value :: mesh
      end function
      integer(c_long) function hx3d_mesh_num_nodes(mesh) bind(C, name="hx3d_mesh_num_nodes")
         import :: c_long, c_ptr
         type(c_ptr), value :: mesh
      end function
      integer(c_long) function hx3d_mesh_num_face_nodes(mesh) bind(C, name="hx3d_mesh_num_face_nodes")
         import :: c_long, c_ptr
         type(c_ptr), value :: mesh
      end function
      type(c_ptr) function hx3d_mesh_coords(mesh) bind(C, name="hx3d_mesh_coords")
         import :: c_ptr
         type(c_ptr), value :: mesh
      end function
      integer(c_int) function hx3d_mesh_fill_desc(mesh, desc) bind(C, name="hx3d_mesh_fill_desc")
         import :: c_int, c_ptr, hx3d_case_desc
         type(c_ptr), value                  :: mesh
         type(hx3d_case_desc), intent(inout) :: desc
      end function
      integer(c_int) function hx3d_mesh_destroy(mesh) bind(C, name="hx3d_mesh_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: mesh
      end function
      integer(c_int) function hx3d_volume_monitor(eng, kind, value) bind(C, name="hx3d_volume_monitor")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: eng
         integer(c_int), value       :: kind
         real(c_double), intent(out) :: value
      end function
      ! solution files (.hsol); strings NUL-terminated, header_name may be c_null_char alone
      type(c_ptr) function hx3d_hsol_error() bind(C, name="hx3d_hsol_error")
         import :: c_ptr
      end function
      integer(c_int) function hx3d_hsol_reference_values(mach, refs) bind(C, name="hx3d_hsol_reference_values")
         import :: c_int, c_double
         real(c_double), value       :: mach
         real(c_double), intent(out) :: refs(6)
      end function
      integer(c_int) function hx3d_hsol_write(path, header_name, K, N, iter, time, refs, Q, grad) &
            bind(C, name="hx3d_hsol_write")
         import :: c_int, c_ptr, c_char, c_double
         character(kind=c_char), intent(in) :: path(*), header_name(*)
         integer(c_int), value              :: K, iter
         integer(c_int), intent(in)         :: N(3, *)
         real(c_double), value              :: time
         real(c_double), intent(in)         :: refs(6)
         type(c_ptr), value                 :: Q, grad
      end function
      integer(c_int) function hx3d_hsol_read_header(path, type, node_type, K, iter, time, refs) &
            bind(C, name="hx3d_hsol_read_header")
         import :: c_int, c_char, c_double
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), intent(out)        :: type, node_type, K, iter
         real(c_double), intent(out)        :: time, refs(6)
      end function
      integer(c_int) function hx3d_hsol_read(path, K, N, Q, grad) bind(C, name="hx3d_hsol_read")
         import :: c_int, c_ptr, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value              :: K
         integer(c_int), intent(in)         :: N(3, *)
         type(c_ptr), value                 :: Q, grad
      end function
      integer(c_int) function hx3d_save_solution(eng, path, header_name, iter, time) bind(C, name="hx3d_save_solution")
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value                 :: eng
         character(kind=c_char), intent(in) :: path(*), header_name(*)
         integer(c_int), value              :: iter
         real(c_double), value              :: time
      end function
      integer(c_int) function hx3d_load_solution(eng, path, iter, time) bind(C, name="hx3d_load_solution")
         import :: c_int, c_ptr, c_char, c_double
         type(c_ptr), value                 :: eng
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), intent(out)        :: iter
         real(c_double), intent(out)        :: time
      end function
      ! .control files; root may be c_null_char alone (working directory)
      integer(c_int) function hx3d_control_read(path, ctl) bind(C, name="hx3d_control_read")
         import :: c_int, c_ptr, c_char
         character(kind=c_char), intent(in) :: path(*)
         type(c_ptr), intent(out)           :: ctl
      end function
      type(c_ptr) function hx3d_control_error(ctl) bind(C, name="hx3d_control_error")
         import :: c_ptr
         type(c_ptr), value :: ctl
      end function
      type(c_ptr) function hx3d_control_value(ctl, key) bind(C, name="hx3d_control_value")
         import :: c_ptr, c_char
         type(c_ptr), value                 :: ctl
         character(kind=c_char), intent(in) :: key(*)
      end function
      integer(c_int) function hx3d_control_boundary(ctl, zone, type, coupled, len) bind(C, name="hx3d_control_boundary")
         import :: c_int, c_ptr, c_char
         type(c_ptr), value                  :: ctl
         character(kind=c_char), intent(in)  :: zone(*)
         character(kind=c_char), intent(out) :: type(*), coupled(*)
         integer(c_int), value               :: len
      end function
      integer(c_int) function hx3d_control_case(ctl, root, mesh, desc, run) bind(C, name="hx3d_control_case")
         import :: c_int, c_ptr, c_char, hx3d_case_desc, hx3d_control_run
         type(c_ptr), value                  :: ctl
         character(kind=c_char), intent(in)  :: root(*)
         type(c_ptr), intent(out)            :: mesh
         type(hx3d_case_desc), intent(out)   :: desc
         type(hx3d_control_run), intent(out) :: run
      end function
      integer(c_int) function hx3d_control_destroy(ctl) bind(C, name="hx3d_control_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctl
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: s
      end function
   end interface

contains

   ! hx3d_last_error as a Fortran string
   function hx3d_error_message(eng) result(msg)
      type(c_ptr), intent(in)       :: eng
      character(len=:), allocatable :: msg
      type(c_ptr)                   :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: i, n
      p = hx3d_last_error(eng)
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function hx3d_error_message

end module hx3d_binding
