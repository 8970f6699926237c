/*
 * hx3d oracle: a plain-C, FP64 restatement of HORSES3D's compressible Navier-Stokes time
 * derivative (ComputeTimeDerivative, Solver/src/NavierStokesSolver/SpatialDiscretization.f90:227-320)
 * and the low-storage RK3 step that drives it (TakeRK3Step,
 * Solver/src/libs/timeintegrator/ExplicitMethods.f90:667-788).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it.  The product path
 * never links or calls it.
 *
 * Data contract (mirrors the reference's HexMesh element/face storage, StorageClass.f90):
 *  - elements e = 0..K-1, orders N[3e+0..2]; node n = i + (Nx+1)*(j + (Ny+1)*k) of element e
 *    lives at global node g = off[e] + n;
 *  - state arrays are packed per node, variable fastest: Q[5*g + v]  (the reference's QNS
 *    layout, StorageClass.f90:138-173);
 *  - metrics Ja[9*g + 3*d + c] = Ja^d_c (d = xi, eta, zeta) and J[g];
 *  - element face slots 0..5 = EFRONT, EBACK, EBOTTOM, ERIGHT, ETOP, ELEFT (MeshTypes.f90:17-18);
 *    faceID[6e+s] (0-based face), faceSide[6e+s] (1 = left, 2 = right);
 *  - faces f = 0..F-1: fElem[2f+0/1] (0-based left/right element), fESide[2f+0/1] (element
 *    face slot 1..6 of that element), fRot[f] (0..7), fNf[2f+0/1] (face order), face node
 *    arrays at foff[f] + i + (Nf1+1)*j: normal[3*], t1[3*], t2[3*], fjac[].
 */
#ifndef HX3D_ORACLE_H
#define HX3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum { OX_ROE = 0, OX_LXF = 1 };
enum { OX_STANDARD = 0, OX_SPLIT = 1 };
enum { OX_AVG_STANDARD = 0, OX_AVG_DUCROS = 1, OX_AVG_KG = 2, OX_AVG_PIROZZOLI = 3, OX_AVG_MORINISHI = 4 };
enum { OX_LES_NONE = 0, OX_LES_SMAGORINSKY = 1, OX_LES_WALE = 2 };
enum { OX_NODES_GAUSSLOBATTO = 0, OX_NODES_GAUSS = 1 };   /* "discretization nodes" */

typedef struct {
    double mach, reynolds, prandtl;
    int riemann, inviscid, averaging, les;
    double lambda_stab;   /* LxF stabilisation factor (default 1) */
    double les_c;         /* Smagorinsky Cs or WALE Cw */
    int euler;            /* 1: inviscid (flow equations = "Euler") */
    int nodes;            /* OX_NODES_GAUSSLOBATTO or OX_NODES_GAUSS (standard form only) */
} ox_params;

typedef struct ox_case ox_case;

/* Spectral operators of NodalStorage (GL nodes), NodalStorageClass.f90:170-337.
 * Arrays are (N+1) or (N+1)^2 row-major: D[i*(N+1)+j] = D(i,j). */
int ox_nodal_gl(int N, double *x, double *w, double *D, double *hatD, double *sharpD,
                double *vL, double *vR, double *bL, double *bR);
/* the same on Gauss(-Legendre) nodes (no sharpD: split forms run on Gauss-Lobatto nodes) */
int ox_nodal_gauss(int N, double *x, double *w, double *D, double *hatD,
                   double *vL, double *vR, double *bL, double *bR);
int ox_nodal_cgl(int N, double *xcgl, double *DCGL, double *Tcheb2gauss);
/* Tset(N->M): interpolation (N<M) or weighted projection (N>M), InterpolationMatrices.f90:42-108.
 * T[(M+1)x(N+1)] row-major. */
int ox_tset(int N, int M, double *T);
/* the same on the given node type (OX_NODES_*) */
int ox_tset_nodes(int nodes, int N, int M, double *T);

ox_case *ox_case_create(int K, const int *N, const long *off,
                        const int *faceID, const int *faceSide,
                        const double *Ja, const double *J,
                        int F, const int *fElem, const int *fESide, const int *fRot,
                        const int *fNf, const long *foff,
                        const double *normal, const double *t1, const double *t2,
                        const double *fjac, const ox_params *p);
void ox_case_destroy(ox_case *c);
long ox_case_nodes(const ox_case *c);

/* QDot = F(Q, t) (ComputeTimeDerivative).  Optional outputs may be NULL:
 * grad[15*g + 5*d + v] = U_d of variable v after BR1 lifting. */
int ox_compute_qdot(ox_case *c, const double *Q, double t, double *QDot, double *grad);
/* One Williamson RK3 step in place (TakeRK3Step); G is the RK register (zeroed by caller
 * before the first step, as the reference's allocation does). */
int ox_rk3_step(ox_case *c, double *Q, double *G, double t, double dt);
/* MaxTimeStep (DGSEMClass.f90:870-1035). */
double ox_max_dt(ox_case *c, const double *Q, double cfl, double dcfl);
/* Volume-averaged kinetic energy monitor (VolumeIntegrals.f90:179-195, VolumeMonitor.f90:300). */
double ox_kinetic_energy(ox_case *c, const double *Q);

#ifdef __cplusplus
}
#endif
#endif
