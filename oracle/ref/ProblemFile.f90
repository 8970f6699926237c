!
!  User plugin for the reference solver, written for this project against the reference's
!  ProblemFile API (Solver/src/libs/problemfile/ProblemFile.f90:22-237).  It is test
!  infrastructure only: it turns the reference build into an oracle that writes golden
!  vectors.
!
!  * Initial condition: the Taylor-Green vortex
!        rho = 1, u = sin(x+px) cos(y+py) cos(z+pz), v = -cos(x+px) sin(y+py) cos(z+pz), w = 0,
!        p   = 1/(gamma M^2) + (cos 2(x+px) + cos 2(y+py)) (cos 2(z+pz) + 2) / 16
!    (phases px,py,pz from HX3D_TGV_PHASE, default 0), or, when HX3D_IC_FILE names a file,
!    the state read from that file (stream binary, float64, element order, Q(1:5,0:Nx,0:Ny,0:Nz)).
!  * Dumps: when HX3D_DUMP_PREFIX is set, the state, its time derivative, the element and
!    face geometry and the connectivity are written (stream binary) before the time steps
!    listed in HX3D_DUMP_STEPS (comma separated step counts) and at the end; every step's
!    (time, dt) is appended to <prefix>_dt.txt.
!
#include "Includes.h"

module hx3dDumpModule
   use SMConstants
   use HexMeshClass
   implicit none
   integer, save :: stepCounter = 0
contains
   subroutine hx3d_dump(mesh, time, tag)
      class(HexMesh)              :: mesh
      real(kind=RP), intent(in)   :: time
      character(len=*), intent(in):: tag
      character(len=1024)         :: prefix
      integer :: u, eID, fID, stat
      call get_environment_variable("HX3D_DUMP_PREFIX", prefix, status=stat)
      if (stat /= 0) return
      open(newunit=u, file=trim(prefix)//"_"//trim(tag)//".bin", access="stream", &
           form="unformatted", status="replace")
      write(u) int(mesh % no_of_elements, 4), int(size(mesh % faces), 4), time
      do eID = 1, mesh % no_of_elements
         associate(e => mesh % elements(eID))
         write(u) int(e % Nxyz, 4), int(e % faceIDs, 4), int(e % faceSide, 4)
         write(u) e % geom % x, e % geom % jGradXi, e % geom % jGradEta, e % geom % jGradZeta, &
                  e % geom % jacobian, e % storage % Q, e % storage % QDot
         end associate
      end do
      do fID = 1, size(mesh % faces)
         associate(f => mesh % faces(fID))
         write(u) int(f % Nf, 4), int(f % elementIDs, 4), int(f % elementSide, 4), &
                  int(f % rotation, 4), int(f % FaceType, 4), int(f % projectionType, 4), &
                  int(f % NfLeft, 4), int(f % NfRight, 4)
         write(u) f % geom % x, f % geom % normal, f % geom % t1, f % geom % t2, f % geom % jacobian
         end associate
      end do
      close(u)
   end subroutine hx3d_dump

   logical function hx3d_dump_now(step)
      integer, intent(in) :: step
      character(len=1024) :: steps
      integer :: stat, i0, i, v
      hx3d_dump_now = .false.
      call get_environment_variable("HX3D_DUMP_STEPS", steps, status=stat)
      if (stat /= 0) return
      i0 = 1
      do i = 1, len_trim(steps) + 1
         if (i > len_trim(steps) .or. steps(i:i) == ",") then
            if (i > i0) then
               read(steps(i0:i-1), *) v
               if (v == step) hx3d_dump_now = .true.
            end if
            i0 = i + 1
         end if
      end do
   end function hx3d_dump_now
end module hx3dDumpModule

         SUBROUTINE UserDefinedStartup
            IMPLICIT NONE
         END SUBROUTINE UserDefinedStartup

         SUBROUTINE UserDefinedFinalSetup(mesh, thermodynamics_, dimensionless_, refValues_)
            USE HexMeshClass
            use PhysicsStorage
            use FluidData
            IMPLICIT NONE
            CLASS(HexMesh)                      :: mesh
            type(Thermodynamics_t), intent(in)  :: thermodynamics_
            type(Dimensionless_t),  intent(in)  :: dimensionless_
            type(RefValues_t),      intent(in)  :: refValues_
         END SUBROUTINE UserDefinedFinalSetup

         subroutine UserDefinedInitialCondition(mesh, thermodynamics_, dimensionless_, refValues_)
            use smconstants
            use physicsstorage
            use hexmeshclass
            use fluiddata
            implicit none
            class(hexmesh)                      :: mesh
            type(Thermodynamics_t), intent(in)  :: thermodynamics_
            type(Dimensionless_t),  intent(in)  :: dimensionless_
            type(RefValues_t),      intent(in)  :: refValues_
            integer        :: eid, i, j, k, u, stat
            real(kind=RP)  :: x, y, z, rho, uu, vv, ww, p, ph(3)
            character(len=1024) :: fname

            call get_environment_variable("HX3D_IC_FILE", fname, status=stat)
            if (stat == 0) then
               open(newunit=u, file=trim(fname), access="stream", form="unformatted", status="old")
               do eid = 1, mesh % no_of_elements
                  read(u) mesh % elements(eid) % storage % Q
               end do
               close(u)
               return
            end if

            ph = 0.0_RP
            call get_environment_variable("HX3D_TGV_PHASE", fname, status=stat)
            if (stat == 0) read(fname, *) ph

            do eid = 1, mesh % no_of_elements
               associate(e => mesh % elements(eid))
               do k = 0, e % Nxyz(3) ; do j = 0, e % Nxyz(2) ; do i = 0, e % Nxyz(1)
                  x = e % geom % x(1,i,j,k) + ph(1)
                  y = e % geom % x(2,i,j,k) + ph(2)
                  z = e % geom % x(3,i,j,k) + ph(3)
                  rho = 1.0_RP
                  uu  =  sin(x) * cos(y) * cos(z)
                  vv  = -cos(x) * sin(y) * cos(z)
                  ww  = 0.0_RP
                  p   = 1.0_RP / dimensionless_ % gammaM2 &
                        + (cos(2.0_RP*x) + cos(2.0_RP*y)) * (cos(2.0_RP*z) + 2.0_RP) / 16.0_RP
                  e % storage % Q(1,i,j,k) = rho
                  e % storage % Q(2,i,j,k) = rho * uu
                  e % storage % Q(3,i,j,k) = rho * vv
                  e % storage % Q(4,i,j,k) = rho * ww
                  e % storage % Q(5,i,j,k) = p / thermodynamics_ % gammaMinus1 &
                                             + 0.5_RP * rho * (uu*uu + vv*vv + ww*ww)
               end do ; end do ; end do
               end associate
            end do
         end subroutine UserDefinedInitialCondition

         subroutine UserDefinedState1(x, t, nHat, Q, thermodynamics_, dimensionless_, refValues_)
            use SMConstants
            use PhysicsStorage
            use FluidData
            implicit none
            real(kind=RP), intent(in)     :: x(NDIM)
            real(kind=RP), intent(in)     :: t
            real(kind=RP), intent(in)     :: nHat(NDIM)
            real(kind=RP), intent(inout)  :: Q(NCONS)
            type(Thermodynamics_t),    intent(in)  :: thermodynamics_
            type(Dimensionless_t),     intent(in)  :: dimensionless_
            type(RefValues_t),         intent(in)  :: refValues_
         end subroutine UserDefinedState1

         subroutine UserDefinedGradVars1(x, t, nHat, Q, U, GetGradients, thermodynamics_, dimensionless_, refValues_)
            use SMConstants
            use PhysicsStorage
            use FluidData
            use VariableConversion, only: GetGradientValues_f
            implicit none
            real(kind=RP), intent(in)          :: x(NDIM)
            real(kind=RP), intent(in)          :: t
            real(kind=RP), intent(in)          :: nHat(NDIM)
            real(kind=RP), intent(in)          :: Q(NCONS)
            real(kind=RP), intent(inout)       :: U(NGRAD)
            procedure(GetGradientValues_f)     :: GetGradients
            type(Thermodynamics_t), intent(in) :: thermodynamics_
            type(Dimensionless_t),  intent(in) :: dimensionless_
            type(RefValues_t),      intent(in) :: refValues_
         end subroutine UserDefinedGradVars1

         subroutine UserDefinedNeumann1(x, t, nHat, Q, U_x, U_y, U_z, flux, thermodynamics_, dimensionless_, refValues_)
            use SMConstants
            use PhysicsStorage
            use FluidData
            implicit none
            real(kind=RP), intent(in)    :: x(NDIM)
            real(kind=RP), intent(in)    :: t
            real(kind=RP), intent(in)    :: nHat(NDIM)
            real(kind=RP), intent(in)    :: Q(NCONS)
            real(kind=RP), intent(in)    :: U_x(NGRAD)
            real(kind=RP), intent(in)    :: U_y(NGRAD)
            real(kind=RP), intent(in)    :: U_z(NGRAD)
            real(kind=RP), intent(inout) :: flux(NCONS)
            type(Thermodynamics_t), intent(in) :: thermodynamics_
            type(Dimensionless_t),  intent(in) :: dimensionless_
            type(RefValues_t),      intent(in) :: refValues_
         end subroutine UserDefinedNeumann1

         SUBROUTINE UserDefinedPeriodicOperation(mesh, time, dt, Monitors, thermodynamics_, dimensionless_, refValues_)
            use SMConstants
            USE HexMeshClass
            use MonitorsClass
            use fluiddata
            use physicsstorage
            use hx3dDumpModule
            IMPLICIT NONE
            CLASS(HexMesh)               :: mesh
            REAL(KIND=RP)                :: time
            REAL(KIND=RP)                :: dt
            type(Monitor_t), intent(in) :: monitors
            type(Thermodynamics_t), intent(in)    :: thermodynamics_
            type(Dimensionless_t),  intent(in)    :: dimensionless_
            type(RefValues_t),      intent(in)    :: refValues_
            character(len=1024) :: prefix
            character(len=32)   :: tag
            integer :: u, stat
            call get_environment_variable("HX3D_DUMP_PREFIX", prefix, status=stat)
            if (stat == 0) then
               open(newunit=u, file=trim(prefix)//"_dt.txt", position="append", status="unknown")
               write(u, '(I8,2ES26.17E3)') stepCounter, time, dt
               close(u)
               if (hx3d_dump_now(stepCounter)) then
                  write(tag, '(A,I0)') "step", stepCounter
                  call hx3d_dump(mesh, time, tag)
               end if
            end if
            stepCounter = stepCounter + 1
         END SUBROUTINE UserDefinedPeriodicOperation

         subroutine UserDefinedSourceTermNS(x, Q, time, S, thermodynamics_, dimensionless_, refValues_)
            use SMConstants
            USE HexMeshClass
            use PhysicsStorage
            use FluidData
            IMPLICIT NONE
            real(kind=RP),             intent(in)  :: x(NDIM)
            real(kind=RP),             intent(in)  :: Q(NCONS)
            real(kind=RP),             intent(in)  :: time
            real(kind=RP),             intent(inout) :: S(NCONS)
            type(Thermodynamics_t), intent(in)  :: thermodynamics_
            type(Dimensionless_t),  intent(in)  :: dimensionless_
            type(RefValues_t),      intent(in)  :: refValues_
         end subroutine UserDefinedSourceTermNS

         SUBROUTINE UserDefinedFinalize(mesh, time, iter, maxResidual, thermodynamics_, dimensionless_, &
                                        refValues_, monitors, elapsedTime, CPUTime)
            use SMConstants
            USE HexMeshClass
            use PhysicsStorage
            use FluidData
            use MonitorsClass
            use hx3dDumpModule
            IMPLICIT NONE
            CLASS(HexMesh)                        :: mesh
            REAL(KIND=RP)                         :: time
            integer                               :: iter
            real(kind=RP)                         :: maxResidual
            type(Thermodynamics_t), intent(in)    :: thermodynamics_
            type(Dimensionless_t),  intent(in)    :: dimensionless_
            type(RefValues_t),      intent(in)    :: refValues_
            type(Monitor_t),        intent(in)    :: monitors
            real(kind=RP),             intent(in) :: elapsedTime
            real(kind=RP),             intent(in) :: CPUTime
            call hx3d_dump(mesh, time, "final")
         END SUBROUTINE UserDefinedFinalize

      SUBROUTINE UserDefinedTermination
         IMPLICIT NONE
      END SUBROUTINE UserDefinedTermination
