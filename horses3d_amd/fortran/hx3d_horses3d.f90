!
!  hx3d_horses3d: drop-in adapters that run HORSES3D's compressible Navier-Stokes time
!  derivative and its low-storage RK3 step on the hx3d engine.
!
!  Compiled inside a HORSES3D build (its modules SMConstants, HexMeshClass, ParticlesClass,
!  DGSEMClass on the module path, Includes.h on the include path, -DNAVIERSTOKES).
!
!    hx3d_attach                 builds the engine from the HexMesh element/face storage
!                                (HexMesh.f90:71-180, HexElementClass.f90:58-104,
!                                FaceClass.f90:63-101, MappedGeometry.f90)
!    hx3d_ComputeTimeDerivative  conforms to ComputeTimeDerivative_f (DGSEMClass.f90:77-97)
!    hx3d_TakeRK3Step            conforms to TimeStep_FCN (TimeIntegratorDefinitions.f90:11-34),
!                                the interface of TimeIntegrator_t % RKStep
!                                (TimeIntegrator.f90:63, assigned at :215 / :233)
!    hx3d_detach                 destroys the engine
!
!  Errors end the run with ERROR STOP and the engine's message, as the reference does on its
!  own errors (SpatialDiscretization.f90:97-98).
!
#include "Includes.h"
module hx3d_horses3d
   use iso_c_binding
   use SMConstants
   use HexMeshClass
#ifdef FLOW
   use ParticlesClass, only: Particles_t
#endif
   use DGSEMClass, only: ComputeTimeDerivative_f
   use ExplicitMethods, only: CTD_AFTER_STEPS
   use NodalStorageClass, only: CurrentNodes, GAUSS
   use hx3d_binding
   implicit none
   private
   public :: hx3d_attach, hx3d_detach, hx3d_ComputeTimeDerivative, hx3d_TakeRK3Step, hx3d_handle

   type(c_ptr), save                         :: eng = c_null_ptr
   integer, allocatable, save                :: off(:)    ! first packed node of element e
   real(c_double), allocatable, target, save :: Qp(:)     ! packed state Q(5*g + v)

contains

   function hx3d_handle() result(p)
      type(c_ptr) :: p
      p = eng
   end function hx3d_handle

   subroutine check(status)
      integer(c_int), intent(in) :: status
      if (status /= HX3D_OK) then
         write(*, '(A,I0,A,A)') "hx3d error ", status, ": ", hx3d_error_message(eng)
         error stop "hx3d engine call failed"
      end if
   end subroutine check

!
!  Engine from the mesh.  Selectors take the HX3D_* values of the .control keys "riemann
!  solver", "inviscid discretization", "averaging", "les model" (parsed by the caller).
!
   subroutine hx3d_attach(mesh, mach, reynolds, prandtl, riemann, inviscid, averaging, les, les_c)
      type(HexMesh), intent(in)  :: mesh
      real(kind=RP), intent(in)  :: mach, reynolds, prandtl, les_c
      integer, intent(in)        :: riemann, inviscid, averaging, les
      integer(c_int), allocatable, target :: N(:), faceID(:), faceSide(:), fElem(:), fESide(:), fRot(:)
      real(c_double), allocatable, target :: Ja(:), Jac(:), fN(:), fT1(:), fT2(:), fJ(:)
      type(hx3d_case_desc) :: d
      integer :: K, F, e, f_, i, j, k_, g, q, nq

      K = size(mesh % elements)
      F = size(mesh % faces)
      allocate(N(3*K), faceID(6*K), faceSide(6*K), fElem(2*F), fESide(2*F), fRot(F), off(K+1))
      off(1) = 0
      do e = 1, K
         N(3*e-2:3*e)        = mesh % elements(e) % Nxyz
         faceID(6*e-5:6*e)   = mesh % elements(e) % faceIDs - 1
         faceSide(6*e-5:6*e) = mesh % elements(e) % faceSide
         off(e+1) = off(e) + product(mesh % elements(e) % Nxyz + 1)
      end do
      allocate(Ja(9*off(K+1)), Jac(off(K+1)), Qp(5*off(K+1)))
      do e = 1, K
         associate(el => mesh % elements(e))
         g = off(e)
         do k_ = 0, el % Nxyz(3) ; do j = 0, el % Nxyz(2) ; do i = 0, el % Nxyz(1)
            Ja(9*g+1:9*g+3) = el % geom % jGradXi(:, i, j, k_)
            Ja(9*g+4:9*g+6) = el % geom % jGradEta(:, i, j, k_)
            Ja(9*g+7:9*g+9) = el % geom % jGradZeta(:, i, j, k_)
            Jac(g+1)        = el % geom % jacobian(i, j, k_)
            g = g + 1
         end do ; end do ; end do
         end associate
      end do
      ! face frames in the left element's face coordinates, (Nf1+1)(Nf2+1) nodes per face
      nq = 0
      do f_ = 1, F
         nq = nq + product(mesh % faces(f_) % Nf + 1)
      end do
      allocate(fN(3*nq), fT1(3*nq), fT2(3*nq), fJ(nq))
      q = 0
      do f_ = 1, F
         associate(fc => mesh % faces(f_))
         fElem(2*f_-1:2*f_)  = fc % elementIDs - 1
         fESide(2*f_-1:2*f_) = fc % elementSide
         fRot(f_)            = fc % rotation
         do j = 0, fc % Nf(2) ; do i = 0, fc % Nf(1)
            fN(3*q+1:3*q+3)  = fc % geom % normal(:, i, j)
            fT1(3*q+1:3*q+3) = fc % geom % t1(:, i, j)
            fT2(3*q+1:3*q+3) = fc % geom % t2(:, i, j)
            fJ(q+1)          = fc % geom % jacobian(i, j)
            q = q + 1
         end do ; end do
         end associate
      end do

      d % K = K ; d % N = c_loc(N) ; d % faceID = c_loc(faceID) ; d % faceSide = c_loc(faceSide)
      d % F = F ; d % fElem = c_loc(fElem) ; d % fESide = c_loc(fESide) ; d % fRot = c_loc(fRot)
      d % Ja = c_loc(Ja) ; d % J = c_loc(Jac)
      d % fNormal = c_loc(fN) ; d % fT1 = c_loc(fT1) ; d % fT2 = c_loc(fT2) ; d % fJac = c_loc(fJ)
      d % mach = mach ; d % reynolds = reynolds ; d % prandtl = prandtl
      d % riemann = riemann ; d % inviscid = inviscid ; d % averaging = averaging
      d % les = les ; d % les_c = les_c
      ! the node type the reference built its NodalStorage and geometry on
      d % nodes = merge(HX3D_NODES_GAUSS, HX3D_NODES_GAUSSLOBATTO, CurrentNodes == GAUSS)
      call check(hx3d_engine_create(d, eng))
   end subroutine hx3d_attach

   subroutine hx3d_detach()
      integer(c_int) :: status
      if (c_associated(eng)) status = hx3d_engine_destroy(eng)
      eng = c_null_ptr
      if (allocated(off)) deallocate(off)
      if (allocated(Qp)) deallocate(Qp)
   end subroutine hx3d_detach

   ! element storage Q(1:5,0:Nx,0:Ny,0:Nz) is variable-fastest, the packed layout itself
   subroutine pack_state(mesh)
      type(HexMesh), intent(in) :: mesh
      integer :: e
      do e = 1, size(mesh % elements)
         Qp(5*off(e)+1:5*off(e+1)) = reshape(mesh % elements(e) % storage % Q, [5*(off(e+1)-off(e))])
      end do
   end subroutine pack_state

!
!  ComputeTimeDerivative_f: QDot of the elements' Q at `time` (no source terms on this path)
!
   subroutine hx3d_ComputeTimeDerivative(mesh, particles, time, mode, HO_Elements, element_mask, Level)
      type(HexMesh), target           :: mesh
#ifdef FLOW
      type(Particles_t)               :: particles
#else
      logical                         :: particles
#endif
      real(kind=RP)                   :: time
      integer,             intent(in) :: mode
      logical, intent(in), optional   :: HO_Elements
      logical, intent(in), optional   :: element_mask(:)
      integer, intent(in), optional   :: Level
      integer :: e
      call pack_state(mesh)
      call check(hx3d_set_state(eng, Qp))
      call check(hx3d_compute_qdot(eng, real(time, c_double), Qp))
      do e = 1, size(mesh % elements)
         mesh % elements(e) % storage % QDot = reshape(Qp(5*off(e)+1:5*off(e+1)), &
                                                       shape(mesh % elements(e) % storage % QDot))
      end do
   end subroutine hx3d_ComputeTimeDerivative

!
!  TimeStep_FCN: one Williamson RK3 step of the elements' Q (ExplicitMethods.f90:667-788).
!  The engine keeps its own RK register; ComputeTimeDerivative is not called.  On return the
!  elements hold the new Q and, as after the reference's step, the QDot its integrator reads
!  next (ComputeMaxResiduals, the monitors and the steady-state tolerance,
!  TimeIntegrator.f90:852-874): the last stage's, or with CTD_AFTER_STEPS (:784) the derivative
!  of the new state.  A NaN in the new state stops the run (checkForNan, :786).
!
   subroutine hx3d_TakeRK3Step(mesh, particles, t, deltaT, ComputeTimeDerivative, dt_vec, dts, &
                               global_dt, iter, dtAdaptation)
      type(HexMesh)              :: mesh
#ifdef FLOW
      type(Particles_t)          :: particles
#else
      logical                    :: particles
#endif
      real(kind=RP)              :: t, deltaT
      procedure(ComputeTimeDerivative_f) :: ComputeTimeDerivative
      real(kind=RP), allocatable, dimension(:), intent(in), optional :: dt_vec
      logical, intent(in), optional       :: dts
      real(kind=RP), intent(in), optional :: global_dt
      integer, intent(in), optional       :: iter
      logical, intent(in), optional       :: dtAdaptation
      integer :: e
      integer(c_int) :: bad
      call pack_state(mesh)
      call check(hx3d_set_state(eng, Qp))
      call check(hx3d_set_ctd_after_steps(eng, merge(1_c_int, 0_c_int, CTD_AFTER_STEPS)))
      call check(hx3d_rk3_step(eng, real(t, c_double), real(deltaT, c_double)))
      call check(hx3d_get_state(eng, Qp))
      do e = 1, size(mesh % elements)
         mesh % elements(e) % storage % Q = reshape(Qp(5*off(e)+1:5*off(e+1)), &
                                                    shape(mesh % elements(e) % storage % Q))
      end do
      call check(hx3d_get_qdot(eng, Qp))
      do e = 1, size(mesh % elements)
         mesh % elements(e) % storage % QDot = reshape(Qp(5*off(e)+1:5*off(e+1)), &
                                                       shape(mesh % elements(e) % storage % QDot))
      end do
      ! checkForNan (ExplicitMethods.f90:786, :1856-1905)
      call check(hx3d_check_nan(eng, bad))
      if (bad /= 0) then
         write(*, '(A)') "Numerical divergence obtained in solver."
         error stop 99
      end if
   end subroutine hx3d_TakeRK3Step

end module hx3d_horses3d
