/*
 * hx3d -- MI355X-native DGSEM time-derivative engine for HORSES3D's compressible
 * Navier-Stokes solver.  C ABI (plain pointers and sizes; no C++ or torch types).
 *
 * The engine replaces two reference procedures:
 *   - ComputeTimeDerivative(mesh, particles, time, mode)
 *       Solver/src/NavierStokesSolver/SpatialDiscretization.f90:227-320, interface
 *       ComputeTimeDerivative_f in Solver/src/libs/discretization/DGSEMClass.f90:77-97
 *       -> hx3d_compute_qdot
 *   - TakeRK3Step(mesh, particles, t, deltaT, ComputeTimeDerivative, ...)
 *       Solver/src/libs/timeintegrator/ExplicitMethods.f90:667-788, interface TimeStep_FCN
 *       (TimeIntegratorDefinitions.f90:11-34) of the RKStep pointer (TimeIntegrator.f90:63)
 *       -> hx3d_rk3_step
 * together with the step-loop helpers the time integrator calls around them:
 *   - MaxTimeStep (DGSEMClass.f90:870-1035)               -> hx3d_max_dt
 *   - ComputeMaxResiduals (DGSEMClass.f90:770-860)        -> hx3d_residuals
 *   - checkForNan (ExplicitMethods.f90:1856-1905)         -> hx3d_check_nan
 *   - kinetic-energy volume monitor (VolumeIntegrals.f90:179-195,
 *     VolumeMonitor.f90:300)                              -> hx3d_kinetic_energy
 *   - the IntegrateInTime step loop (TimeIntegrator.f90:737-812) for CFL-bounded
 *     explicit RK3                                        -> hx3d_advance
 *
 * Storage contract (the reference's HexMesh element/face storage, StorageClass.f90):
 *   element e (0-based) has orders N[3e..3e+2]; its node n = i + (Nx+1)*(j + (Ny+1)*k)
 *   is global node g = sum_{e'<e} (Nx'+1)(Ny'+1)(Nz'+1) + n; host state arrays are packed
 *   per node, variable fastest, Q[5*g + v] (the reference's QNS layout,
 *   StorageClass.f90:138-173).  Element face slots 0..5 are EFRONT, EBACK, EBOTTOM,
 *   ERIGHT, ETOP, ELEFT (MeshTypes.f90:17-18).  Faces follow Face (FaceClass.f90:63-101):
 *   left/right element, the element face slot (1..6) each side sits on, and the
 *   rotation 0..7 (leftIndexes2Right, MeshTypes.f90:70-108).
 *
 * Status codes: 0 = OK, > 0 = error (text from hx3d_last_error).  No exceptions or
 * process exits cross the ABI.  One host thread per engine at a time; engines own their
 * device memory and their HIP stream.
 */
#ifndef HX3D_H
#define HX3D_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define HX3D_OK 0
#define HX3D_EINVAL 1
#define HX3D_EUNSUPPORTED 2
#define HX3D_EDEVICE 3
#define HX3D_ENOMEM 4
#define HX3D_ENAN 5

/* selectors (keys of the reference's .control file) */
enum { HX3D_RIEMANN_ROE = 0, HX3D_RIEMANN_LXF = 1 };                    /* "riemann solver" */
enum { HX3D_INVISCID_STANDARD = 0, HX3D_INVISCID_SPLIT = 1 };           /* "inviscid discretization" */
enum { HX3D_AVG_STANDARD = 0, HX3D_AVG_DUCROS = 1, HX3D_AVG_KENNEDYGRUBER = 2,
       HX3D_AVG_PIROZZOLI = 3, HX3D_AVG_MORINISHI = 4 };                 /* "averaging" */
enum { HX3D_LES_NONE = 0, HX3D_LES_SMAGORINSKY = 1, HX3D_LES_WALE = 2 }; /* "les model" */
enum { HX3D_NODES_GAUSSLOBATTO = 0, HX3D_NODES_GAUSS = 1 };             /* "discretization nodes" */

typedef struct {
    /* mesh (ConstructMesh_FromSpecMeshFile_ / HexMesh, HexMesh.f90:71-180) */
    int K;                  /* number of elements */
    const int *N;           /* [3K] polynomial orders (Nx, Ny, Nz) per element */
    const int *faceID;      /* [6K] 0-based face of each element face slot */
    const int *faceSide;    /* [6K] 1 = element is the face's left side, 2 = right */
    int F;                  /* number of (interior, periodic-paired) faces */
    const int *fElem;       /* [2F] 0-based left, right element */
    const int *fESide;      /* [2F] element face slot 1..6 of the left, right element */
    const int *fRot;        /* [F] rotation 0..7 */
    /* geometry (MappedGeometry.f90:64-386), per global node g */
    const double *Ja;       /* [9*nodes] Ja^d_c at 9g + 3d + c (d = xi, eta, zeta) */
    const double *J;        /* [nodes] Jacobian */
    /* face geometry (MappedGeometryFace, MappedGeometry.f90:482-760) per face node
     * q = f*(N+1)^2 + i + (N+1)*j in the left element's face coordinates: unit normal,
     * tangents and surface Jacobian |Ja|.  Optional (all NULL): then the normal and |Ja| are
     * taken from the left element's metrics at the face node (identical for GL nodes) and the
     * tangents from a fixed orthonormal completion of the normal. */
    const double *fNormal;  /* [3*F*(N+1)^2] */
    const double *fT1;      /* [3*F*(N+1)^2] */
    const double *fT2;      /* [3*F*(N+1)^2] */
    const double *fJac;     /* [F*(N+1)^2] */
    /* physics (ConstructPhysicsStorage_NS, PhysicsStorage_NS.f90:145-439) */
    double mach, reynolds, prandtl;
    int euler;              /* 1: "flow equations" = Euler (no viscous terms) */
    /* discretisation (Initialize_SpaceAndTimeMethods, SpatialDiscretization.f90:44-210) */
    int riemann, inviscid, averaging;
    double lambda_stab;     /* "lambda stabilization" (LxF), default 1 */
    int les;                /* "les model" */
    double les_c;           /* "les model intensity" (Smagorinsky 0.2, WALE 0.325) */
    /* "discretization nodes" (NodalStorageClass.f90:80-105): HX3D_NODES_GAUSSLOBATTO (0, so a
     * zero-filled descriptor keeps Gauss-Lobatto) or HX3D_NODES_GAUSS, the reference's default
     * (main.f90:278-281; standard form only, and the face geometry must be given) */
    int nodes;
} hx3d_case_desc;

typedef struct hx3d_engine hx3d_engine;

/* Creates an engine on the current HIP device: uploads geometry, builds the
 * element/face gather tables.  Supported: Gauss-Lobatto or Gauss nodes, orders 1 <= N <= 7
 * (one order everywhere, or per-element (Nx, Ny, Nz) with mortar faces, whose face frames must
 * then be given at the face orders), every face interior or periodic. */
int hx3d_engine_create(const hx3d_case_desc *desc, hx3d_engine **out);
int hx3d_engine_destroy(hx3d_engine *eng);
/* Batching (no single counterpart in the reference: HORSES3D runs one case per process): ncases
 * independent cases of one order N in one engine, stepped together in every launch.  The cases
 * share the discretisation (Riemann solver, volume form, averaging, LES model, lambda) and may
 * differ in mesh size and geometry and in the physics numbers (Mach, Reynolds and Prandtl
 * numbers, LES constant, ConstructPhysicsStorage_NS, PhysicsStorage_NS.f90:145-439).  Each
 * case's dt is MaxTimeStep of its own state (DGSEMClass.f90:870-1035) and each has its own time
 * and divergence flag, so every case evolves exactly as it does alone, bit for bit.  The state
 * arrays are the cases' packed states one after the other.  hx3d_advance and hx3d_advance_dt
 * take t_inout = NULL (they queue the steps); hx3d_max_dt is not available. */
int hx3d_engine_create_multi(const hx3d_case_desc *descs, int ncases, hx3d_engine **out);
/* case c takes part in the following steps (on = 1) or keeps its state and time (on = 0);
 * stream-ordered */
int hx3d_multi_set_active(hx3d_engine *eng, int c, int on);
/* waits for the queued work; then times[c] and diverged[c] (1: a NaN appeared in case c's state)
 * for every case (either may be NULL) */
int hx3d_multi_status(hx3d_engine *eng, double *times, int *diverged);
const char *hx3d_last_error(const hx3d_engine *eng);
long hx3d_num_nodes(const hx3d_engine *eng);

/* State transfer in the packed host layout Q[5g+v] (LoadSolutionForRestart / the QNS copy). */
int hx3d_set_state(hx3d_engine *eng, const double *Q_host);
int hx3d_get_state(hx3d_engine *eng, double *Q_host);
/* count nodes of the resident state from node first on, packed (a case's part of a
 * several-case engine); waits */
int hx3d_get_state_nodes(hx3d_engine *eng, long first, long count, double *Q_host);
/* The same on device memory (packed layout, stream-ordered on the engine stream). */
int hx3d_set_state_device(hx3d_engine *eng, const double *Q_dev);
int hx3d_get_state_device(hx3d_engine *eng, double *Q_dev);

/* Drop-in for ComputeTimeDerivative: QDot(Q, t) of the resident state.  QDot_host may be
 * NULL (the derivative then stays on the device for hx3d_residuals). */
int hx3d_compute_qdot(hx3d_engine *eng, double t, double *QDot_host);
/* Drop-in for TakeRK3Step: one Williamson RK3 step of size dt on the resident state.  Like the
 * reference, it leaves the last stage's time derivative as the current QDot (hx3d_get_qdot,
 * hx3d_residuals), or, with hx3d_set_ctd_after_steps(eng, 1), the derivative of the new state
 * (the "compute time derivative after timestep" key, CTD_AFTER_STEPS, ExplicitMethods.f90:784).
 * Queued on the engine stream; the RK update sets a sticky divergence flag when a NaN appears
 * (checkForNan, ExplicitMethods.f90:786), reported by hx3d_query, hx3d_advance and
 * hx3d_check_nan. */
int hx3d_rk3_step(hx3d_engine *eng, double t, double dt);
/* on = 1: after every step (hx3d_rk3_step, hx3d_advance, hx3d_advance_dt) the time derivative is
 * recomputed at the new state (CTD_AFTER_STEPS, TimeIntegrator.f90 "compute time derivative
 * after timestep"); on = 0 (the reference's default): the last stage's derivative is kept. */
int hx3d_set_ctd_after_steps(hx3d_engine *eng, int on);
/* The current time derivative (storage % QDot after ComputeTimeDerivative or a step) in the
 * packed host layout; computed from the resident state if none is current. */
int hx3d_get_qdot(hx3d_engine *eng, double *QDot_host);
/* MaxTimeStep(cfl, dcfl) of the resident state. */
int hx3d_max_dt(hx3d_engine *eng, double cfl, double dcfl, double *dt);
/* The step loop: nsteps RK3 steps, each with dt = MaxTimeStep(cfl, dcfl) of the state at
 * the start of the step (computed on the device; no host round trip).  t_inout advances.
 * With t_inout != NULL the call waits for the steps and returns HX3D_ENAN when a NaN appeared
 * in the state (the reference stops there, ExplicitMethods.f90:786 and :1856-1905); with
 * t_inout = NULL it only queues them (see hx3d_query). */
int hx3d_advance(hx3d_engine *eng, int nsteps, double cfl, double dcfl, double *t_inout);
/* The step loop with a fixed step: nsteps RK3 steps of size dt (the .control file's "dt" key
 * instead of "cfl", TimeIntegrator.f90:196-198); t_inout advances by dt per step. */
int hx3d_advance_dt(hx3d_engine *eng, int nsteps, double dt, double *t_inout);
/* on = 1: hx3d_advance issues each step as a HIP graph (captured on the first call for a given
 * (cfl, dcfl), replayed once per step; kernel launches only; not while timing is on).  on = 0
 * (the default): the same launches issued one by one, which measured 2-5% faster on MI355X
 * from 4^3 x N=3 to 16^3 x N=4. */
int hx3d_set_step_graph(hx3d_engine *eng, int on);
/* ComputeMaxResiduals (DGSEMClass.f90:770-860, called after each step at
 * TimeIntegrator.f90:852): max |QDot| per variable of the current time derivative (after a step,
 * the last stage's, or the new state's with CTD_AFTER_STEPS). */
int hx3d_residuals(hx3d_engine *eng, double res[5]);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): hx3d_residuals with
 * the result left in device memory.  res_dev[0..4] (caller-owned device memory) receives max |QDot|
 * per variable of the current time derivative, computed first from the resident state if none is
 * current; queued on the engine stream, no wait, no allocation.  The engine's device must be
 * current.  Partial maxima go to the library's one static device array per GPU (2048 blocks x 5
 * doubles, 80 KB, held until the process ends); the kernels take at most 2048 blocks that stride
 * over the nodes, so every node count an engine holds fits it.  Calls that overlap on one GPU
 * (any engines, any streams) race for that array. */
int hx3d_residuals_noalloc(hx3d_engine *eng, double *res_dev);
/* Kinetic-energy monitor of the resident state. */
int hx3d_kinetic_energy(hx3d_engine *eng, double *ke);
/* Volume monitors (VolumeMonitor_Update, VolumeMonitor.f90:278-330, with the integrals of
 * VolumeIntegrals.f90:179-375): the volume average of the selected quantity of the resident
 * state (enstrophy: half the average of |curl u|^2).  The rate and the enstrophy use the time
 * derivative and gradients of the resident state. */
enum { HX3D_MON_KINETIC_ENERGY = 0, HX3D_MON_KINETIC_ENERGY_RATE = 1, HX3D_MON_ENSTROPHY = 2,
       HX3D_MON_MEAN_VELOCITY = 3, HX3D_MON_INTERNAL_ENERGY = 4, HX3D_MON_ENTROPY = 5,
       HX3D_MON_MATH_ENTROPY = 6 };
int hx3d_volume_monitor(hx3d_engine *eng, int kind, double *value);
/* checkForNan: *has_nan = 1 if any state value is NaN. */
int hx3d_check_nan(hx3d_engine *eng, int *has_nan);
/* Blocks until the engine's queued work is done. */
int hx3d_synchronize(hx3d_engine *eng);
/* Without blocking: *done = 1 when the engine's queued work is done, else 0.  With
 * t_inout = NULL, hx3d_advance only queues the steps, so a host can keep several engines
 * (each with its own stream) busy at once and poll them.  Returns HX3D_ENAN (with *done = 1)
 * when the finished steps diverged (a NaN appeared in the state since it was last set). */
int hx3d_query(hx3d_engine *eng, int *done);
/* The engine's time (advanced on the device by hx3d_advance); waits for queued work. */
int hx3d_get_time(hx3d_engine *eng, double *t);
/* Device memory held by the engine, in bytes. */
long hx3d_engine_device_bytes(const hx3d_engine *eng);
/* Free and total memory of the current HIP device (admission by free HBM). */
int hx3d_device_memory(size_t *free_bytes, size_t *total_bytes);
/* No counterpart in the reference (its elements run one at a time).  Mixed orders: the element
 * kernels are persistent, one launch per stage over tiles of elements (one thread per node, tiles
 * of 256 or 512 threads).  The tiling (which elements share a block) is a policy, packing +
 * 4 * sizes.  Packing: 0 first-fit decreasing by size, 1 elements of one order triple together,
 * 2 one element per tile.  Sizes: 0 256-thread tiles when every element has at most 256 nodes,
 * else 512-thread tiles (the default); 1 512-thread tiles; 2 256-thread tiles for the elements
 * that fit and 512-thread tiles for the others.  hx3d_autotune times reps evaluations of the
 * resident state's time derivative under each packing (sizes 0; ms[p] per evaluation when
 * ms != NULL, 3 entries), keeps the fastest and returns it in *choice; results do not depend on
 * the policy.  A uniform-order engine has one tiling: *choice = -1.  hx3d_set_tiling /
 * hx3d_get_tiling set and read the policy (mixed orders). */
int hx3d_autotune(hx3d_engine *eng, int reps, int *choice, double *ms);
int hx3d_set_tiling(hx3d_engine *eng, int policy);
int hx3d_get_tiling(const hx3d_engine *eng, int *policy, int *ntiles);
/* Average duration (ms) per launch of each stage kernel over the last hx3d_advance call,
 * measured with HIP events on the engine stream when timing was enabled with
 * hx3d_set_timing(eng, 1): out[0..2] = gradient+lifting, face flux, volume+surface+RK;
 * out[3] = their sum (one stage). */
int hx3d_set_timing(hx3d_engine *eng, int on);
int hx3d_kernel_times(hx3d_engine *eng, double out_ms[4]);

/* ---- Solution files (.hsol, SolutionFile.f90:61-213): the restart and output files of
 * HexMesh_SaveSolution / HexMesh_LoadSolution (HexMesh.f90:3376-3556, 4582-4912).  Host
 * arrays in the packed layout Q[5g+v]; gradients (type 3 files) as three such arrays U_x, U_y,
 * U_z one after the other.  Errors: text from hx3d_hsol_error (per thread). */
const char *hx3d_hsol_error(void);
/* gamma, R, V, rho, T, Mach of the default reference state (PhysicsStorage_NS.f90:292-306) */
int hx3d_hsol_reference_values(double mach, double refs[6]);
/* header_name: the name stored in the file (NULL: path); grad may be NULL (type 2 file) */
int hx3d_hsol_write(const char *path, const char *header_name, int K, const int *N, int iter, double time,
                    const double refs[6], const double *Q, const double *grad);
/* the same with the node type field of the file set from nodes (HX3D_NODES_*); hx3d_hsol_write
 * writes Gauss-Lobatto files.  A reader takes a file of either type as it is (node_type in
 * the header: 1 Gauss, 2 Gauss-Lobatto), as HexMesh_LoadSolution does without interpolation. */
int hx3d_hsol_write_nodes(const char *path, const char *header_name, int K, const int *N, int iter, double time,
                          const double refs[6], const double *Q, const double *grad, int nodes);
int hx3d_hsol_read_header(const char *path, int *type, int *node_type, int *K, int *iter, double *time,
                          double refs[6]);
/* checks the element count and each element's orders against N; grad may be NULL */
int hx3d_hsol_read(const char *path, int K, const int *N, double *Q, double *grad);
/* the resident state to / from a solution file (SaveSolution, LoadSolutionForRestart) */
int hx3d_save_solution(hx3d_engine *eng, const char *path, const char *header_name, int iter, double time);
int hx3d_load_solution(hx3d_engine *eng, const char *path, int *iter, double *time);

/* ---- Mesh front end (host only, no GPU): a SpecMesh file read into the contract above with
 * the reference's topology and geometry.
 *   ConstructMesh_FromSpecMeshFile_ (Read_SpecMesh.f90:36-348) with ConstructFaces,
 *   ConstructPeriodicFaces, DeletePeriodicMinusFaces (HexMesh.f90:356-944),
 *   HexMesh_ConstructGeometry (HexMesh.f90:2677-3059), ConstructMappedGeometry and
 *   ConstructMappedGeometryFace (MappedGeometry.f90:64-386, 482-760).
 * orders: [3K] per-element (Nx, Ny, Nz) (the "polynomial order file") or NULL for N everywhere.
 * periodic: the .control file's periodic boundaries as "name=coupled;name=coupled;..." (each
 * "#define boundary name / type = periodic / coupled boundary = coupled" block); every other
 * boundary makes the call fail with HX3D_EUNSUPPORTED.  On failure *out still holds the mesh
 * object with the message (hx3d_mesh_error); destroy it. */
typedef struct hx3d_mesh hx3d_mesh;
int hx3d_mesh_read_specmesh(const char *path, int N, const int *orders, const char *periodic, hx3d_mesh **out);
/* The same with the geometry built on the given "discretization nodes" (HX3D_NODES_*): node
 * coordinates, metrics and face frames at Gauss or Gauss-Lobatto points (the NodalStorage the
 * reference's ConstructMappedGeometry uses); hx3d_mesh_read_specmesh is the Gauss-Lobatto form.
 * hx3d_mesh_fill_desc then sets desc->nodes. */
int hx3d_mesh_read_specmesh_nodes(const char *path, int N, const int *orders, const char *periodic, int nodes,
                                  hx3d_mesh **out);
const char *hx3d_mesh_error(const hx3d_mesh *m);
long hx3d_mesh_num_nodes(const hx3d_mesh *m);
long hx3d_mesh_num_face_nodes(const hx3d_mesh *m);
/* node coordinates [3*nodes], x of node g at 3g + c (HexElement geom % x) */
const double *hx3d_mesh_coords(const hx3d_mesh *m);
/* sets the mesh and geometry fields of desc (K, N, faces, Ja, J, face frames) to arrays the
 * mesh owns; the physics and discretisation fields are left to the caller */
int hx3d_mesh_fill_desc(const hx3d_mesh *m, hx3d_case_desc *desc);
int hx3d_mesh_destroy(hx3d_mesh *m);

/* ---- .control files (host only) ----------------------------------------------------------
 * hx3d_control_read: ReadControlFile (ReadInputFile.f90:20-125): '!' starts a comment, a line
 * starting with '/' is blank, "keyword = value" lines (keys lower-cased, quoted values taken
 * between the quotes), "#define ... #end" blocks skipped; the first entry of a key is the one
 * found (FTDictionaryClass.f90:371-535).  Then CheckInputIntegrity's defaults
 * (NavierStokesSolver/main.f90:239-318).  On failure *out holds the object with the message. */
typedef struct hx3d_control hx3d_control;
int hx3d_control_read(const char *path, hx3d_control **out);
const char *hx3d_control_error(const hx3d_control *c);
/* stringValueForKey: the value of a key (any case), NULL when absent; valid until destroy */
const char *hx3d_control_value(const hx3d_control *c, const char *key);
/* boundary condition of one mesh zone from the "#define boundary a__b" blocks: GetZoneType
 * (BoundaryConditions.f90:170-260) and ConstructPeriodicBC (PeriodicBC.f90:73-151); type and,
 * for periodic zones, the coupled zone, lower-cased, each into len characters (NUL-ended) */
int hx3d_control_boundary(hx3d_control *c, const char *zone, char *type, char *coupled, int len);

/* the time-integration keys (constructTimeIntegrator, TimeIntegrator.f90:178-214) */
typedef struct {
    double cfl, dcfl;       /* "cfl", "dcfl" (dt from MaxTimeStep); 0 when "dt" is given */
    double dt;              /* "dt" (fixed step); 0 when "cfl" is given */
    double tolerance;       /* "convergence tolerance" */
    int steps;              /* "number of time steps" */
    int output_interval;    /* "output interval" */
    int restart;            /* "restart" = .true.: start from "restart file name" */
} hx3d_control_run;

/* The whole case of a .control file: the physics and discretisation keys
 * (ConstructPhysicsStorage_NS, SetRiemannSolver, the inviscid / viscous discretisation, the
 * LES model, the node type), the polynomial orders (GetMeshPolynomialOrders,
 * pAdaptationClass.f90:198-240: "polynomial order file", "polynomial order", or
 * "polynomial order i/j/k"), and the mesh ("mesh file name", read by the front end with the
 * boundary blocks' periodic pairs).  Relative file names are taken from root when it is given,
 * else from the working directory (the reference's behaviour).  desc points into *mesh, which
 * the caller destroys after hx3d_engine_create; *mesh is set whenever the mesh was read, also
 * on failure.  Keys the engine does not implement give HX3D_EUNSUPPORTED, unknown names and
 * missing keys HX3D_EINVAL, with the message in hx3d_control_error. */
int hx3d_control_case(hx3d_control *c, const char *root, hx3d_mesh **mesh, hx3d_case_desc *desc,
                      hx3d_control_run *run);
int hx3d_control_destroy(hx3d_control *c);

#ifdef __cplusplus
}
#endif
#endif
