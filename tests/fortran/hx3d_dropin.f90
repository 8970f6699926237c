!
!  Test driver (test infrastructure, not product): the reference's compressible Navier-Stokes
!  solver with its explicit RK3 step taken by the hx3d engine through the drop-in adapter.
!
!  The set-up is the reference's own: control file, physics, nodal storage, DGSEM, initial
!  condition (the ProblemFile plugin of the reference build), spatial discretisation and time
!  integrator, in the order main.f90 builds them (Solver/src/NavierStokesSolver/main.f90:47-162).
!  Once the time integrator is built, its RKStep pointer (TimeIntegrator.f90:63, set at :215-292)
!  is pointed at hx3d_TakeRK3Step and the engine is attached to the mesh.  The reference's own
!  loop (TimeIntegrator % integrate, TimeIntegrator.f90:737-890) then runs: it takes each step
!  through the pointer, and its ComputeMaxResiduals, monitors and output read the state and QDot
!  the adapter leaves in the element storage.  The control file must give every key the
!  reference's CheckInputIntegrity would default (main.f90:238-318; it is part of main.f90's
!  program unit, which this driver replaces).
!
!  usage: hx3d_dropin.ns case.control
!
#include "Includes.h"
program hx3d_dropin
   use SMConstants
   use FTValueDictionaryClass
   use PhysicsStorage
   use SharedBCModule
   use DGSEMClass
   use TimeIntegratorClass
   use MPI_Process_Info
   use SpatialDiscretization
   use pAdaptationClass,      only: GetMeshPolynomialOrders
   use NodalStorageClass
   use FluidData
   use FileReaders,           only: ReadControlFile
   use InterpolationMatrices, only: Initialize_InterpolationMatrices, Finalize_InterpolationMatrices
   use ProblemFileFunctions
   use BoundaryConditions,    only: DestructBoundaryConditions
   use Utilities,             only: toLower
   use hx3d_binding
   use hx3d_horses3d
   implicit none

   procedure(UserDefinedStartup_f)     :: UserDefinedStartup
   procedure(UserDefinedFinalSetup_f)  :: UserDefinedFinalSetup
   procedure(UserDefinedFinalize_f)    :: UserDefinedFinalize
   procedure(UserDefinedTermination_f) :: UserDefinedTermination
   type(FTValueDictionary) :: controlVariables
   type(DGSem)             :: sem
   type(TimeIntegrator_t)  :: timeIntegrator
   logical                 :: success
   integer                 :: initial_iteration
   real(kind=RP)           :: initial_time
   integer, allocatable    :: Nx(:), Ny(:), Nz(:)
   integer                 :: Nmax
   integer                 :: riemann, inviscid, averaging
   character(len=LINE_LENGTH) :: s

   call SetSolver(NAVIERSTOKES_SOLVER)
   call MPI_Process % Init
   call controlVariables % initWithSize(16)
   call UserDefinedStartup
   call ConstructSharedBCModule
   call ReadControlFile(controlVariables)

   call ConstructPhysicsStorage(controlVariables, success)
   if (.not. success) error stop "Physics parameters input error"
   sem % ManufacturedSol = .false.
   call GetMeshPolynomialOrders(controlVariables, Nx, Ny, Nz, Nmax)
   call InitializeNodalStorage(controlVariables, Nmax)
   call Initialize_InterpolationMatrices(Nmax)
   call sem % construct(controlVariables = controlVariables, Nx_ = Nx, Ny_ = Ny, Nz_ = Nz, &
                        success = success, generateMonitor = .true.)
   if (.not. success) error stop "Mesh reading error"
   call UserDefinedFinalSetup(sem % mesh, thermodynamics, dimensionless, refValues)
   call sem % SetInitialCondition(controlVariables, initial_iteration, initial_time)
   call Initialize_SpaceAndTimeMethods(controlVariables, sem)
   sem % particles % active = .false.
   call timeIntegrator % construct(controlVariables, sem, initial_iteration, initial_time)

   ! the drop-in: the engine from the mesh storage, and the integrator's RK step through it
   s = controlVariables % stringValueForKey("riemann solver", LINE_LENGTH)
   call toLower(s)
   select case (trim(s))
   case ("roe");            riemann = HX3D_RIEMANN_ROE
   case ("lax-friedrichs"); riemann = HX3D_RIEMANN_LXF
   case default;            error stop "hx3d_dropin: riemann solver not on this path"
   end select
   s = controlVariables % stringValueForKey("inviscid discretization", LINE_LENGTH)
   call toLower(s)
   inviscid = merge(HX3D_INVISCID_SPLIT, HX3D_INVISCID_STANDARD, trim(s) == "split-form")
   s = controlVariables % stringValueForKey("averaging", LINE_LENGTH)
   call toLower(s)
   select case (trim(s))
   case ("ducros");         averaging = HX3D_AVG_DUCROS
   case ("kennedy-gruber"); averaging = HX3D_AVG_KENNEDYGRUBER
   case ("pirozzoli");      averaging = HX3D_AVG_PIROZZOLI
   case ("morinishi");      averaging = HX3D_AVG_MORINISHI
   case default;            averaging = HX3D_AVG_STANDARD
   end select
   call hx3d_attach(sem % mesh, controlVariables % doublePrecisionValueForKey("mach number"), &
                    controlVariables % doublePrecisionValueForKey("reynolds number"), &
                    controlVariables % doublePrecisionValueForKey("prandtl number"), &
                    riemann, inviscid, averaging, HX3D_LES_NONE, 0.0_RP)
   timeIntegrator % RKStep => hx3d_TakeRK3Step
   write(STD_OUT, '(A)') "hx3d_dropin: RK3 steps on the hx3d engine"

   call timeIntegrator % integrate(sem, controlVariables, sem % monitors, sem % samplings, &
                                   ComputeTimeDerivative, ComputeTimeDerivativeIsolated)

   call UserDefinedFinalize(sem % mesh, timeIntegrator % time, sem % numberOfTimeSteps, &
                            sem % maxResidual, thermodynamics, dimensionless, refValues, &
                            sem % monitors, 0.0_RP, 0.0_RP)
   call hx3d_detach()
   write(STD_OUT, '(A,I0,A)') "hx3d_dropin: ", sem % numberOfTimeSteps, " steps done"
   call timeIntegrator % destruct()
   call sem % destruct()
   call DestructBoundaryConditions
   call Finalize_SpaceAndTimeMethods
   call Finalize_InterpolationMatrices
   call DestructGlobalNodalStorage()
   call destructSharedBCModule
   call UserDefinedTermination
   call MPI_Process % Close
end program hx3d_dropin
