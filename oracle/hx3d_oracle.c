/*
 * hx3d oracle -- plain-C FP64 restatement of the HORSES3D compressible NS time derivative.
 * TEST INFRASTRUCTURE ONLY (see hx3d_oracle.h).  Every routine cites the reference routine
 * whose algorithm and operation order it follows; file paths are relative to Solver/src/.
 */
#include "hx3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NC 5
#define MAXN 16
#define POW2(x) ((x) * (x))

/* ------------------------------------------------------------------------------------
 * Spectral operators: libs/spectral/LegendreAlgorithms.f90, InterpolationAndDerivatives.f90,
 * NodalStorageClass.f90
 * ---------------------------------------------------------------------------------- */
static void legendre_poly_and_derivative(int N, double x, double *LN, double *dLN)
{
    /* LegendreAlgorithms.f90 LegendrePolyAndDerivative (three-term recursion) */
    if (N == 0) { *LN = 1.0; *dLN = 0.0; return; }
    if (N == 1) { *LN = x; *dLN = 1.0; return; }
    double LNm2 = 1.0, LNm1 = x, dLNm2 = 0.0, dLNm1 = 1.0, L = 0, dL = 0;
    for (int k = 2; k <= N; ++k) {
        L = ((2 * k - 1) * x * LNm1 - (k - 1) * LNm2) / k;
        dL = dLNm2 + (2 * k - 1) * LNm1;
        LNm2 = LNm1; LNm1 = L;
        dLNm2 = dLNm1; dLNm1 = dL;
    }
    *LN = L; *dLN = dL;
}

static void q_and_L(int N, double x, double *q, double *dq, double *LN)
{
    /* LegendreAlgorithms.f90:216-271 qAndLEvaluation */
    double LkM2 = 1.0, dLkM2 = 0.0, LkM1 = x, dLkM1 = 1.0, Lk, dLk;
    for (int k = 2; k <= N; ++k) {
        Lk = ((2 * k - 1) * x * LkM1 - (k - 1) * LkM2) / k;
        dLk = dLkM2 + (2 * k - 1) * LkM1;
        LkM2 = LkM1; LkM1 = Lk; dLkM2 = dLkM1; dLkM1 = dLk;
    }
    int k = N + 1;
    Lk = ((2 * k - 1) * x * LkM1 - (k - 1) * LkM2) / k;
    dLk = dLkM2 + (2 * k - 1) * LkM1;
    *q = Lk - LkM2;
    *dq = dLk - dLkM2;
    *LN = LkM1;
}

static void gl_nodes(int N, double *x, double *w)
{
    /* LegendreAlgorithms.f90:275-355 LegendreLobattoNodesAndWeights (Newton, 4 eps) */
    const double tol = 4.0 * 2.220446049250313e-16;
    if (N == 1) { x[0] = -1; w[0] = 1; x[1] = 1; w[1] = 1; return; }
    x[0] = -1.0; w[0] = 2.0 / (N * (N + 1)); x[N] = 1.0; w[N] = w[0];
    int nd2 = (N + 1) / 2;
    for (int j = 1; j <= nd2 - 1; ++j) {
        double xj = -cos((j + 0.25) * M_PI / N - 3.0 / (8 * N * M_PI * (j + 0.25)));
        double q, dq, LN;
        for (int k = 0; k <= 10; ++k) {
            q_and_L(N, xj, &q, &dq, &LN);
            double delta = -q / dq;
            xj += delta;
            if (fabs(delta) <= tol * fabs(xj)) break;
        }
        q_and_L(N, xj, &q, &dq, &LN);
        x[j] = xj; w[j] = 2.0 / (N * (N + 1) * LN * LN);
        x[N - j] = -xj; w[N - j] = w[j];
    }
    if (N % 2 == 0) {
        double LN, dLN;
        legendre_poly_and_derivative(N, 0.0, &LN, &dLN);
        x[N / 2] = 0.0; w[N / 2] = 2.0 / (N * (N + 1) * LN * LN);
    }
}

static void gauss_nodes(int N, double *x, double *w)
{
    /* LegendreAlgorithms.f90:137-214 GaussLegendreNodesAndWeights: Newton on L_{N+1} from
     * the Chebyshev-Gauss points, 4 eps relative, half the nodes mirrored */
    const double tol = 4.0 * 2.220446049250313e-16;
    double L, dL;
    if (N == 0) { x[0] = 0.0; w[0] = 2.0; return; }
    if (N == 1) {
        x[0] = -sqrt(1.0 / 3.0); w[0] = 1.0; x[1] = -x[0]; w[1] = w[0];
    } else {
        int nd2 = (N + 1) / 2;
        for (int j = 0; j < nd2; ++j) {
            double xj = -cos((2 * j + 1) * M_PI / (2 * N + 2));
            for (int k = 0; k <= 10; ++k) {
                legendre_poly_and_derivative(N + 1, xj, &L, &dL);
                double delta = -L / dL;
                xj = xj + delta;
                if (fabs(delta) <= tol * fabs(xj)) break;
            }
            legendre_poly_and_derivative(N + 1, xj, &L, &dL);
            x[j] = xj; w[j] = 2.0 / ((1.0 - xj * xj) * (dL * dL));
            x[N - j] = -xj; w[N - j] = w[j];
        }
    }
    if (N % 2 == 0) {
        legendre_poly_and_derivative(N + 1, 0.0, &L, &dL);
        x[N / 2] = 0.0; w[N / 2] = 2.0 / (dL * dL);
    }
}

/* the nodes of the case's node type (NodalStorageClass.f90:208-226) */
static void nodes_of(int type, int N, double *x, double *w)
{
    if (type == OX_NODES_GAUSS) gauss_nodes(N, x, w);
    else gl_nodes(N, x, w);
}

static int almost_equal(double a, double b)
{
    /* libs/foundation/Utilities.f90:110-139 AlmostEqual (tol = 2 eps) */
    const double tol = 2.0 * 2.220446049250313e-16;
    if (a == 0.0 || b == 0.0) return fabs(a - b) <= tol;
    double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return fabs(b - a) <= tol * m;
}

static void bary_weights(int N, const double *x, double *wb)
{
    /* InterpolationAndDerivatives.f90:230-262 BarycentricWeights */
    for (int j = 0; j <= N; ++j) wb[j] = 1.0;
    for (int j = 1; j <= N; ++j)
        for (int k = 0; k <= j - 1; ++k) {
            wb[k] *= (x[k] - x[j]);
            wb[j] *= (x[j] - x[k]);
        }
    for (int j = 0; j <= N; ++j) wb[j] = 1.0 / wb[j];
}

static void deriv_matrix(int N, const double *x, double *D)
{
    /* InterpolationAndDerivatives.f90:799-834 PolynomialDerivativeMatrix */
    double wb[MAXN];
    bary_weights(N, x, wb);
    for (int i = 0; i <= N; ++i) {
        D[i * (N + 1) + i] = 0.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) {
                D[i * (N + 1) + j] = wb[j] / (wb[i] * (x[i] - x[j]));
                D[i * (N + 1) + i] -= D[i * (N + 1) + j];
            }
    }
}

static void interp_vector(double xe, int N, const double *x, const double *wb, double *p)
{
    /* InterpolationAndDerivatives.f90:109-160 InterpolatingPolynomialVector */
    int match = 0;
    for (int j = 0; j <= N; ++j) {
        p[j] = 0.0;
        if (almost_equal(xe, x[j])) { p[j] = 1.0; match = 1; }
    }
    if (match) return;
    double d = 0.0;
    for (int j = 0; j <= N; ++j) { double t = wb[j] / (xe - x[j]); p[j] = t; d += t; }
    for (int j = 0; j <= N; ++j) p[j] /= d;
}

static void interp_matrix(int N, int M, const double *xo, const double *wb, const double *xn,
                          double *T)
{
    /* InterpolationAndDerivatives.f90:413-461 PolynomialInterpolationMatrix; T is (M+1)x(N+1) */
    for (int k = 0; k <= M; ++k) {
        int row = 0;
        for (int j = 0; j <= N; ++j) {
            T[k * (N + 1) + j] = 0.0;
            if (almost_equal(xn[k], xo[j])) { row = 1; T[k * (N + 1) + j] = 1.0; }
        }
        if (!row) {
            double s = 0.0;
            for (int j = 0; j <= N; ++j) {
                double t = wb[j] / (xn[k] - xo[j]);
                T[k * (N + 1) + j] = t; s += t;
            }
            for (int j = 0; j <= N; ++j) T[k * (N + 1) + j] /= s;
        }
    }
}

int ox_nodal_gl(int N, double *x, double *w, double *D, double *hatD, double *sharpD,
                double *vL, double *vR, double *bL, double *bR)
{
    /* NodalStorageClass.f90:170-337 ConstructNodalStorage (GAUSSLOBATTO) */
    if (N < 1 || N >= MAXN) return 1;
    int n1 = N + 1;
    double wb[MAXN];
    gl_nodes(N, x, w);
    deriv_matrix(N, x, D);
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) hatD[i * n1 + j] = D[j * n1 + i] * w[j] / w[i];
    for (int k = 0; k < n1 * n1; ++k) sharpD[k] = 2.0 * D[k];
    sharpD[0] = 2.0 * D[0] + 1.0 / w[0];
    sharpD[N * n1 + N] = 2.0 * D[N * n1 + N] - 1.0 / w[N];
    bary_weights(N, x, wb);
    interp_vector(1.0, N, x, wb, vR);
    interp_vector(-1.0, N, x, wb, vL);
    for (int i = 0; i <= N; ++i) { bL[i] = vL[i] / w[i]; bR[i] = vR[i] / w[i]; }
    return 0;
}

int ox_nodal_gauss(int N, double *x, double *w, double *D, double *hatD, double *vL, double *vR,
                   double *bL, double *bR)
{
    /* NodalStorageClass.f90:170-337 ConstructNodalStorage (GAUSS): no sharpD, the boundary
     * values v(+-1) and b = v / w are dense */
    if (N < 1 || N >= MAXN) return 1;
    int n1 = N + 1;
    double wb[MAXN];
    gauss_nodes(N, x, w);
    deriv_matrix(N, x, D);
    for (int j = 0; j <= N; ++j)
        for (int i = 0; i <= N; ++i) hatD[i * n1 + j] = D[j * n1 + i] * w[j] / w[i];
    bary_weights(N, x, wb);
    interp_vector(1.0, N, x, wb, vR);
    interp_vector(-1.0, N, x, wb, vL);
    for (int i = 0; i <= N; ++i) { bL[i] = vL[i] / w[i]; bR[i] = vR[i] / w[i]; }
    return 0;
}

int ox_nodal_cgl(int N, double *xcgl, double *DCGL, double *T)
{
    /* NodalStorageClass.f90:282-296 (Chebyshev-Gauss-Lobatto framework) */
    if (N < 1 || N >= MAXN) return 1;
    double x[MAXN], w[MAXN], wb[MAXN];
    gl_nodes(N, x, w);
    for (int i = 0; i <= N; ++i) xcgl[i] = -cos(1.0 * i * M_PI / N);
    bary_weights(N, xcgl, wb);
    deriv_matrix(N, xcgl, DCGL);
    interp_matrix(N, N, xcgl, wb, x, T);
    return 0;
}

int ox_tset(int N, int M, double *T) { return ox_tset_nodes(OX_NODES_GAUSSLOBATTO, N, M, T); }

int ox_tset_nodes(int nodes, int N, int M, double *T)
{
    /* libs/spectral/InterpolationMatrices.f90:42-108 InterpolationMatrix_Construct on the
     * case's nodes:
     * N < M: Lagrange interpolation from N-nodes to M-nodes;
     * N > M: L2 projection = transpose of the M->N interpolation weighted by w_N(j)/w_M(i). */
    if (N < 1 || M < 1 || N >= MAXN || M >= MAXN) return 1;
    double xo[MAXN], wo[MAXN], xd[MAXN], wd[MAXN], wb[MAXN];
    nodes_of(nodes, N, xo, wo);
    nodes_of(nodes, M, xd, wd);
    if (N == M) {
        for (int i = 0; i <= M; ++i)
            for (int j = 0; j <= N; ++j) T[i * (N + 1) + j] = (i == j) ? 1.0 : 0.0;
        return 0;
    }
    if (N < M) {
        bary_weights(N, xo, wb);
        interp_matrix(N, M, xo, wb, xd, T);
    } else {
        double Tin[MAXN * MAXN];            /* (N+1) x (M+1): interpolation M -> N */
        bary_weights(M, xd, wb);
        interp_matrix(M, N, xd, wb, xo, Tin);
        for (int i = 0; i <= M; ++i)
            for (int j = 0; j <= N; ++j)
                T[i * (N + 1) + j] = Tin[j * (M + 1) + i] * wo[j] / wd[i];
    }
    return 0;
}

/* ------------------------------------------------------------------------------------
 * Case
 * ---------------------------------------------------------------------------------- */
typedef struct {
    int N;
    double x[MAXN], w[MAXN], D[MAXN * MAXN], hatD[MAXN * MAXN], sharpD[MAXN * MAXN];
    double vL[MAXN], vR[MAXN], bL[MAXN], bR[MAXN];
} nodal_t;

typedef struct {
    /* per side s = 0 (left), 1 (right) */
    int NelL[2], NelR[2], NfL[2], NfR[2], Nf[2], proj[2];
    double *Q[2], *Ux[2], *Uy[2], *Uz[2], *mu[2];   /* face-order arrays (Nf) */
    double *fStar[2];                                /* element-face-order arrays */
    double *unStar[2];                               /* (5,3) per element-face node */
} facestore_t;

struct ox_case {
    int K, F;
    int *N;
    long *off;
    int *faceID, *faceSide;
    double *Ja, *J, *invJ, *vol;
    int *fElem, *fESide, *fRot, *fNf;
    long *foff;
    double *normal, *t1, *t2, *fjac, *fsurf;
    ox_params p;
    /* derived physics constants, PhysicsStorage_NS.f90:145-439 */
    double gamma, gm1, gammaM2, mu, kappa, mu_to_kappa, S_div_T, Trenorm;
    nodal_t nod[MAXN];
    double *Tset[MAXN][MAXN];
    facestore_t *fs;
    long nodes;
    /* element work arrays */
    double *U_x, *U_y, *U_z, *muNS, *muT;
};

/* element face slot (1..6) -> fixed axis, face orientation index axes (axisMap,
 * HexElementConnectivityDefinitions.f90) */
static const int SLOT_AXIS[7] = {0, 1, 1, 2, 0, 2, 0};     /* EFRONT..ELEFT -> xi=0, eta=1, zeta=2 */
static const int SLOT_PLUS[7] = {0, 0, 1, 0, 1, 1, 0};     /* 1 if the face is at +1 */
static const int SLOT_UV[7][2] = {{0, 0}, {0, 2}, {0, 2}, {0, 1}, {1, 2}, {0, 1}, {1, 2}};

static nodal_t *nodal(ox_case *c, int N) { return &c->nod[N]; }

static double *tset(ox_case *c, int N, int M)
{
    if (!c->Tset[N][M]) {
        c->Tset[N][M] = (double *)malloc(sizeof(double) * (M + 1) * (N + 1));
        ox_tset_nodes(c->p.nodes, N, M, c->Tset[N][M]);
    }
    return c->Tset[N][M];
}

static void left_to_right(int i, int j, int Nx, int Ny, int rot, int *ii, int *jj)
{
    /* libs/mesh/MeshTypes.f90:70-108 leftIndexes2Right */
    switch (rot) {
    case 0: *ii = i; *jj = j; break;
    case 1: *ii = Ny - j; *jj = i; break;
    case 2: *ii = Nx - i; *jj = Ny - j; break;
    case 3: *ii = j; *jj = Nx - i; break;
    case 4: *ii = j; *jj = i; break;
    case 5: *ii = Nx - i; *jj = j; break;
    case 6: *ii = Ny - j; *jj = Nx - i; break;
    default: *ii = i; *jj = Ny - j; break;
    }
}

long ox_case_nodes(const ox_case *c) { return c->nodes; }

ox_case *ox_case_create(int K, const int *N, const long *off, const int *faceID,
                        const int *faceSide, const double *Ja, const double *J, int F,
                        const int *fElem, const int *fESide, const int *fRot, const int *fNf,
                        const long *foff, const double *normal, const double *t1,
                        const double *t2, const double *fjac, const ox_params *p)
{
    /* split forms need the Gauss-Lobatto SBP operators (the reference switches a split-form
     * case on Gauss nodes to Gauss-Lobatto, main.f90:305-313) */
    if (p->nodes == OX_NODES_GAUSS && p->inviscid == OX_SPLIT) return NULL;
    ox_case *c = (ox_case *)calloc(1, sizeof(ox_case));
    c->K = K; c->F = F; c->p = *p;
    c->nodes = off[K];
    long nn = c->nodes, fn = foff[F];
#define DUP(dst, src, n, T) do { dst = (T *)malloc(sizeof(T) * (n)); memcpy(dst, src, sizeof(T) * (n)); } while (0)
    DUP(c->N, N, 3 * K, int);
    DUP(c->off, off, K + 1, long);
    DUP(c->faceID, faceID, 6 * K, int);
    DUP(c->faceSide, faceSide, 6 * K, int);
    DUP(c->Ja, Ja, 9 * nn, double);
    DUP(c->J, J, nn, double);
    DUP(c->fElem, fElem, 2 * F, int);
    DUP(c->fESide, fESide, 2 * F, int);
    DUP(c->fRot, fRot, F, int);
    DUP(c->fNf, fNf, 2 * F, int);
    DUP(c->foff, foff, F + 1, long);
    DUP(c->normal, normal, 3 * fn, double);
    DUP(c->t1, t1, 3 * fn, double);
    DUP(c->t2, t2, 3 * fn, double);
    DUP(c->fjac, fjac, fn, double);
#undef DUP
    for (int n = 1; n < MAXN; ++n) {
        nodal_t *nd = &c->nod[n];
        nd->N = n;
        if (p->nodes == OX_NODES_GAUSS)
            ox_nodal_gauss(n, nd->x, nd->w, nd->D, nd->hatD, nd->vL, nd->vR, nd->bL, nd->bR);
        else
            ox_nodal_gl(n, nd->x, nd->w, nd->D, nd->hatD, nd->sharpD, nd->vL, nd->vR, nd->bL, nd->bR);
    }
    /* PhysicsStorage_NS.f90:145-439 with ThermodynamicsAir (:115-132) */
    c->gamma = 1.4; c->gm1 = 1.4 - 1.0;
    double M = p->mach;
    c->gammaM2 = c->gamma * POW2(M);
    if (p->euler) { c->mu = 0; c->kappa = 0; c->mu_to_kappa = 0; }
    else {
        c->mu = 1.0 / p->reynolds;
        c->kappa = 1.0 / (c->gm1 * POW2(M) * p->reynolds * p->prandtl);
        c->mu_to_kappa = 1.0 / (c->gm1 * POW2(M) * p->prandtl);
    }
    double Tref = 520.0 * 5.0 / 9.0, S = 198.6 * 5.0 / 9.0;
    c->S_div_T = S / Tref;
    c->Trenorm = Tref / Tref;
    c->invJ = (double *)malloc(sizeof(double) * nn);
    for (long g = 0; g < nn; ++g) c->invJ[g] = 1.0 / c->J[g];
    /* element volumes (MappedGeometry.f90:130-137) and face surfaces (:750-756) */
    c->vol = (double *)calloc(K, sizeof(double));
    for (int e = 0; e < K; ++e) {
        const int *Ne = &c->N[3 * e];
        nodal_t *a = nodal(c, Ne[0]), *b = nodal(c, Ne[1]), *cc = nodal(c, Ne[2]);
        double v = 0;
        for (int k = 0; k <= Ne[2]; ++k) for (int j = 0; j <= Ne[1]; ++j) for (int i = 0; i <= Ne[0]; ++i)
            v += a->w[i] * b->w[j] * cc->w[k] * c->J[c->off[e] + i + (Ne[0] + 1) * (j + (Ne[1] + 1) * k)];
        c->vol[e] = v;
    }
    c->fsurf = (double *)calloc(F, sizeof(double));
    c->fs = (facestore_t *)calloc(F, sizeof(facestore_t));
    for (int f = 0; f < F; ++f) {
        facestore_t *s = &c->fs[f];
        int nf0 = c->fNf[2 * f], nf1 = c->fNf[2 * f + 1];
        double sf = 0;
        for (int j = 0; j <= nf1; ++j) for (int i = 0; i <= nf0; ++i)
            sf += c->nod[nf0].w[i] * c->nod[nf1].w[j] * c->fjac[c->foff[f] + i + (nf0 + 1) * j];
        c->fsurf[f] = sf;
        /* Face_LinkWithElements, FaceClass.f90:187-278 */
        int rot = c->fRot[f];
        for (int side = 0; side < 2; ++side) {
            int e = c->fElem[2 * f + side], slot = c->fESide[2 * f + side];
            int *Nel = side == 0 ? s->NelL : s->NelR;
            Nel[0] = c->N[3 * e + SLOT_UV[slot][0]];
            Nel[1] = c->N[3 * e + SLOT_UV[slot][1]];
        }
        s->NfL[0] = s->NelL[0]; s->NfL[1] = s->NelL[1];
        if (rot == 0 || rot == 2 || rot == 5 || rot == 7) { s->NfR[0] = s->NelR[0]; s->NfR[1] = s->NelR[1]; }
        else { s->NfR[0] = s->NelR[1]; s->NfR[1] = s->NelR[0]; }
        s->Nf[0] = s->NfL[0] > s->NfR[0] ? s->NfL[0] : s->NfR[0];
        s->Nf[1] = s->NfL[1] > s->NfR[1] ? s->NfL[1] : s->NfR[1];
        for (int side = 0; side < 2; ++side) {
            const int *Nfs = side == 0 ? s->NfL : s->NfR;
            int a = Nfs[0] != s->Nf[0], b = Nfs[1] != s->Nf[1];
            s->proj[side] = (!a && !b) ? 0 : (a && !b) ? 1 : (!a && b) ? 2 : 3;
        }
        int nfn = (s->Nf[0] + 1) * (s->Nf[1] + 1);
        for (int side = 0; side < 2; ++side) {
            const int *Nel = side == 0 ? s->NelL : s->NelR;
            int nen = (Nel[0] + 1) * (Nel[1] + 1);
            s->Q[side] = (double *)calloc(NC * nfn, sizeof(double));
            s->Ux[side] = (double *)calloc(NC * nfn, sizeof(double));
            s->Uy[side] = (double *)calloc(NC * nfn, sizeof(double));
            s->Uz[side] = (double *)calloc(NC * nfn, sizeof(double));
            s->mu[side] = (double *)calloc(2 * nfn, sizeof(double));
            int big = nen > nfn ? nen : nfn;
            s->fStar[side] = (double *)calloc(NC * big, sizeof(double));
            s->unStar[side] = (double *)calloc(NC * 3 * big, sizeof(double));
        }
    }
    c->U_x = (double *)calloc(NC * nn, sizeof(double));
    c->U_y = (double *)calloc(NC * nn, sizeof(double));
    c->U_z = (double *)calloc(NC * nn, sizeof(double));
    c->muNS = (double *)calloc(2 * nn, sizeof(double));
    c->muT = (double *)calloc(nn, sizeof(double));
    return c;
}

void ox_case_destroy(ox_case *c)
{
    if (!c) return;
    free(c->N); free(c->off); free(c->faceID); free(c->faceSide); free(c->Ja); free(c->J);
    free(c->invJ); free(c->vol); free(c->fElem); free(c->fESide); free(c->fRot); free(c->fNf);
    free(c->foff); free(c->normal); free(c->t1); free(c->t2); free(c->fjac); free(c->fsurf);
    for (int f = 0; f < c->F; ++f)
        for (int s = 0; s < 2; ++s) {
            free(c->fs[f].Q[s]); free(c->fs[f].Ux[s]); free(c->fs[f].Uy[s]); free(c->fs[f].Uz[s]);
            free(c->fs[f].mu[s]); free(c->fs[f].fStar[s]); free(c->fs[f].unStar[s]);
        }
    free(c->fs);
    for (int a = 0; a < MAXN; ++a) for (int b = 0; b < MAXN; ++b) free(c->Tset[a][b]);
    free(c->U_x); free(c->U_y); free(c->U_z); free(c->muNS); free(c->muT);
    free(c);
}

/* ------------------------------------------------------------------------------------
 * Physics: libs/physics/navierstokes/{Physics_NS,VariableConversion_NS,RiemannSolvers_NS}.f90
 * ---------------------------------------------------------------------------------- */
static double pressure(const ox_case *c, const double *Q)
{   /* VariableConversion_NS.f90:64 Pressure */
    return c->gm1 * (Q[4] - 0.5 * (POW2(Q[1]) + POW2(Q[2]) + POW2(Q[3])) / Q[0]);
}

static double sutherland(const ox_case *c, double T)
{   /* VariableConversion_NS.f90:167-186 SutherlandsLaw */
    double tT = T * c->Trenorm;
    return (1.0 + c->S_div_T) / (tT + c->S_div_T) * tT * sqrt(tT);
}

static void laminar_mu_kappa(const ox_case *c, const double *Q, double *mu, double *kappa)
{   /* VariableConversion_NS.f90:146-162 get_laminar_mu_kappa */
    double T = c->gammaM2 * pressure(c, Q) / Q[0];
    double s = sutherland(c, T);
    *mu = c->mu * s;
    *kappa = *mu * c->mu_to_kappa;
}

static void euler_flux(const ox_case *c, const double *Q, double F[NC][3])
{   /* Physics_NS.f90:51-97 EulerFlux */
    double u = Q[1] / Q[0], v = Q[2] / Q[0], w = Q[3] / Q[0];
    double p = c->gm1 * (Q[4] - 0.5 * (Q[1] * u + Q[2] * v + Q[3] * w));
    F[0][0] = Q[1]; F[1][0] = Q[1] * u + p; F[2][0] = Q[1] * v; F[3][0] = Q[1] * w; F[4][0] = (Q[4] + p) * u;
    F[0][1] = Q[2]; F[1][1] = F[2][0]; F[2][1] = Q[2] * v + p; F[3][1] = Q[2] * w; F[4][1] = (Q[4] + p) * v;
    F[0][2] = Q[3]; F[1][2] = F[3][0]; F[2][2] = F[3][1]; F[3][2] = Q[3] * w + p; F[4][2] = (Q[4] + p) * w;
}

static void viscous_flux(const ox_case *c, const double *Q, const double *Qx, const double *Qy,
                         const double *Qz, double mu, double beta, double kappa, double F[NC][3])
{   /* Physics_NS.f90:246-304 ViscousFlux_STATE */
    double invRho = 1.0 / Q[0];
    double u = Q[1] * invRho, v = Q[2] * invRho, w = Q[3] * invRho;
    double uDR[3] = {u * invRho, v * invRho, w * invRho};
    double ux[3], uy[3], uz[3];
    for (int d = 0; d < 3; ++d) {
        ux[d] = invRho * Qx[1 + d] - uDR[d] * Qx[0];
        uy[d] = invRho * Qy[1 + d] - uDR[d] * Qy[0];
        uz[d] = invRho * Qz[1 + d] - uDR[d] * Qz[0];
    }
    double g = c->gm1 * c->gammaM2;
    double nTx = g * (invRho * Qx[4] - Q[4] * invRho * invRho * Qx[0] - u * ux[0] - v * ux[1] - w * ux[2]);
    double nTy = g * (invRho * Qy[4] - Q[4] * invRho * invRho * Qy[0] - u * uy[0] - v * uy[1] - w * uy[2]);
    double nTz = g * (invRho * Qz[4] - Q[4] * invRho * invRho * Qz[0] - u * uz[0] - v * uz[1] - w * uz[2]);
    double divV = ux[0] + uy[1] + uz[2];
    F[0][0] = 0.0;
    F[1][0] = mu * (2.0 * ux[0] - 2.0 / 3.0 * divV) + beta * divV;
    F[2][0] = mu * (ux[1] + uy[0]);
    F[3][0] = mu * (ux[2] + uz[0]);
    F[4][0] = F[1][0] * u + F[2][0] * v + F[3][0] * w + kappa * nTx;
    F[0][1] = 0.0;
    F[1][1] = F[2][0];
    F[2][1] = mu * (2.0 * uy[1] - 2.0 / 3.0 * divV) + beta * divV;
    F[3][1] = mu * (uy[2] + uz[1]);
    F[4][1] = F[1][1] * u + F[2][1] * v + F[3][1] * w + kappa * nTy;
    F[0][2] = 0.0;
    F[1][2] = F[3][0];
    F[2][2] = F[3][1];
    F[3][2] = mu * (2.0 * uz[2] - 2.0 / 3.0 * divV) + beta * divV;
    F[4][2] = F[1][2] * u + F[2][2] * v + F[3][2] * w + kappa * nTz;
}

static void velocity_gradients(const double *Q, const double *Qx, const double *Qy,
                               const double *Qz, double ux[3], double uy[3], double uz[3])
{   /* VariableConversion_NS.f90:373-392 getVelocityGradients_State */
    double invRho = 1.0 / Q[0], invRho2 = invRho * invRho;
    double uDR[3] = {Q[1] * invRho2, Q[2] * invRho2, Q[3] * invRho2};
    for (int d = 0; d < 3; ++d) {
        ux[d] = invRho * Qx[1 + d] - uDR[d] * Qx[0];
        uy[d] = invRho * Qy[1 + d] - uDR[d] * Qy[0];
        uz[d] = invRho * Qz[1 + d] - uDR[d] * Qz[0];
    }
}

static double les_viscosity(const ox_case *c, double delta, const double *Q, const double *Qx,
                            const double *Qy, const double *Qz)
{
    double ux[3], uy[3], uz[3];
    velocity_gradients(Q, Qx, Qy, Qz, ux, uy, uz);
    if (c->p.les == OX_LES_SMAGORINSKY) {
        /* LESModels.f90:256-307 Smagorinsky_ComputeViscosity (no wall model) */
        double S[3][3];
        for (int a = 0; a < 3; ++a) { S[a][0] = ux[a]; S[a][1] = uy[a]; S[a][2] = uz[a]; }
        for (int b = 0; b < 3; ++b) { S[0][b] += ux[b]; S[1][b] += uy[b]; S[2][b] += uz[b]; }
        double ss = 0;
        /* sum(S*S) in Fortran column-major order */
        for (int b = 0; b < 3; ++b) for (int a = 0; a < 3; ++a) { S[a][b] *= 0.5; }
        for (int b = 0; b < 3; ++b) for (int a = 0; a < 3; ++a) ss += S[a][b] * S[a][b];
        double normS = sqrt(2.0 * ss);
        double LS = c->p.les_c * delta;
        return Q[0] * POW2(LS) * normS;
    }
    if (c->p.les == OX_LES_WALE) {
        /* LESModels.f90:358-~440 WALE_ComputeViscosity */
        double gV[3][3], S[3][3], gV2[3][3], Sd[3][3];
        for (int b = 0; b < 3; ++b) { gV[0][b] = ux[b]; gV[1][b] = uy[b]; gV[2][b] = uz[b]; }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                S[i][j] = 0.5 * (gV[i][j] + gV[j][i]);
                gV2[i][j] = 0;
                for (int k = 0; k < 3; ++k) gV2[i][j] += gV[i][k] * gV[k][j];
            }
        double divV2 = gV2[0][0] + gV2[1][1] + gV2[2][2];
        double normS = 0;
        for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) normS += S[i][j] * S[i][j];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) Sd[i][j] = 0.5 * (gV2[i][j] + gV2[j][i]);
        for (int i = 0; i < 3; ++i) Sd[i][i] = Sd[i][i] - 1.0 / 3.0 * divV2;
        double normSd = 0;
        for (int j = 0; j < 3; ++j) for (int i = 0; i < 3; ++i) normSd += Sd[i][j] * Sd[i][j];
        double LS = c->p.les_c * delta;
        double mu = Q[0] * POW2(LS) * (pow(normSd, 3.0 / 2.0) / (pow(normS, 5.0 / 2.0) + pow(normSd, 5.0 / 4.0)));
        if (normS < 1.0e-8 && normSd < 1.0e-8) mu = 0.0;
        return mu;
    }
    return 0.0;
}

/* AveragedStates (RiemannSolvers_NS.f90:1784-2077); states are in the rotated frame */
static void averaged_states(const ox_case *c, const double *QL, const double *QR, double pL,
                            double pR, double iL, double iR, double *f)
{
    double uL = iL * QL[1], uR = iR * QR[1], vL = iL * QL[2], vR = iR * QR[2];
    double wL = iL * QL[3], wR = iR * QR[3];
    switch (c->p.averaging) {
    case OX_AVG_DUCROS:
        f[0] = 0.25 * (QL[0] + QR[0]) * (uL + uR);
        f[1] = 0.25 * (QL[1] + QR[1]) * (uL + uR) + 0.5 * (pL + pR);
        f[2] = 0.25 * (QL[2] + QR[2]) * (uL + uR);
        f[3] = 0.25 * (QL[3] + QR[3]) * (uL + uR);
        f[4] = 0.25 * (QL[4] + pL + QR[4] + pR) * (uL + uR);
        break;
    case OX_AVG_KG: {
        double rho = 0.5 * (QL[0] + QR[0]), u = 0.5 * (uL + uR), v = 0.5 * (vL + vR);
        double w = 0.5 * (wL + wR), p = 0.5 * (pL + pR), e = 0.5 * (QL[4] * iL + QR[4] * iR);
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = rho * u * w;
        f[4] = rho * u * e + p * u;
        break; }
    case OX_AVG_PIROZZOLI: {
        double rho = 0.5 * (QL[0] + QR[0]), u = 0.5 * (uL + uR), v = 0.5 * (vL + vR);
        double w = 0.5 * (wL + wR), p = 0.5 * (pL + pR);
        double h = 0.5 * ((QL[4] + pL) * iL + (QR[4] + pR) * iR);
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = rho * u * w;
        f[4] = rho * u * h;
        break; }
    case OX_AVG_MORINISHI: {
        double cp = c->gamma * (1.0 / (1.4 - 1.0));
        double hL = cp * pL, hR = cp * pR;
        f[0] = 0.5 * (QL[1] + QR[1]);
        f[1] = 0.25 * (QL[1] + QR[1]) * (uL + uR) + 0.5 * (pL + pR);
        f[2] = 0.25 * (QL[1] + QR[1]) * (vL + vR);
        f[3] = 0.25 * (QL[1] + QR[1]) * (wL + wR);
        f[4] = 0.5 * (uL * hL + uR * hR) + 0.25 * (QL[1] * uL + QR[1] * uR) * (uL + uR)
             + 0.25 * (QL[1] * vL + QR[1] * vR) * (vL + vR)
             + 0.25 * (QL[1] * wL + QR[1] * wR) * (wL + wR)
             - 0.25 * (QL[1] * POW2(uL) + QR[1] * POW2(uR))
             - 0.25 * (QL[1] * POW2(vL) + QR[1] * POW2(vR))
             - 0.25 * (QL[1] * POW2(wL) + QR[1] * POW2(wR));
        break; }
    default: /* StandardAverage */
        f[0] = 0.5 * (QL[1] + QR[1]);
        f[1] = 0.5 * (QL[1] * uL + QR[1] * uR + pL + pR);
        f[2] = 0.5 * (QL[1] * vL + QR[1] * vR);
        f[3] = 0.5 * (QL[1] * wL + QR[1] * wR);
        f[4] = 0.5 * (uL * (QL[4] + pL) + uR * (QR[4] + pR));
    }
}

static void lxf_flux(const ox_case *c, const double *QL, const double *QR, const double *n,
                     const double *t1, const double *t2, double *flux)
{   /* RiemannSolvers_NS.f90:1251-1334 LxFRiemannSolver */
    double rhoL = QL[0], rhoR = QR[0], iL = 1.0 / rhoL, iR = 1.0 / rhoR;
    double ruL = QL[1] * n[0] + QL[2] * n[1] + QL[3] * n[2];
    double ruR = QR[1] * n[0] + QR[2] * n[1] + QR[3] * n[2];
    double rvL = QL[1] * t1[0] + QL[2] * t1[1] + QL[3] * t1[2];
    double rvR = QR[1] * t1[0] + QR[2] * t1[1] + QR[3] * t1[2];
    double rwL = QL[1] * t2[0] + QL[2] * t2[1] + QL[3] * t2[2];
    double rwR = QR[1] * t2[0] + QR[2] * t2[1] + QR[3] * t2[2];
    double V2L = (POW2(ruL) + POW2(rvL) + POW2(rwL)) * iL;
    double V2R = (POW2(ruR) + POW2(rvR) + POW2(rwR)) * iR;
    double pL = c->gm1 * (QL[4] - 0.5 * V2L), pR = c->gm1 * (QR[4] - 0.5 * V2R);
    double aL = sqrt(c->gamma * pL * iL), aR = sqrt(c->gamma * pR * iR);
    double l1 = fabs(ruL * iL) + aL, l2 = fabs(ruR * iR) + aR;
    double lambda = l1 > l2 ? l1 : l2;
    double QLr[5] = {rhoL, ruL, rvL, rwL, QL[4]}, QRr[5] = {rhoR, ruR, rvR, rwR, QR[4]};
    averaged_states(c, QLr, QRr, pL, pR, iL, iR, flux);
    for (int v = 0; v < 5; ++v) flux[v] = flux[v] - c->p.lambda_stab * (0.5 * lambda * (QRr[v] - QLr[v]));
    double f1 = flux[1], f2 = flux[2], f3 = flux[3];
    for (int d = 0; d < 3; ++d) flux[1 + d] = n[d] * f1 + t1[d] * f2 + t2[d] * f3;
}

static void roe_flux(const ox_case *c, const double *QL, const double *QR, const double *n,
                     double *flux)
{   /* RiemannSolvers_NS.f90:1541-1656 RoeRiemannSolver */
    double g = c->gamma;
    double rL = QL[0], mxL = QL[1], myL = QL[2], mzL = QL[3], EL = QL[4];
    double rR = QR[0], mxR = QR[1], myR = QR[2], mzR = QR[3], ER = QR[4];
    double uL = mxL / rL, vL = myL / rL, wL = mzL / rL;
    double pL = (g - 1.0) * (EL - 0.5 / rL * (POW2(mxL) + POW2(myL) + POW2(mzL)));
    double uR = mxR / rR, vR = myR / rR, wR = mzR / rR;
    double pR = (g - 1.0) * (ER - 0.5 / rR * (POW2(mxR) + POW2(myR) + POW2(mzR)));
    double unL = n[0] * uL + n[1] * vL + n[2] * wL;
    double unR = n[0] * uR + n[1] * vR + n[2] * wR;
    double HL = 0.5 * (uL * uL + vL * vL + wL * wL) + g / (g - 1.0) * pL / rL;
    double HR = 0.5 * (uR * uR + vR * vR + wR * wR) + g / (g - 1.0) * pR / rR;
    double srr = sqrt(rL * rR);
    double wtL = rL / (rL + srr), wtR = 1.0 - wtL;
    double ua = wtL * uL + wtR * uR, va = wtL * vL + wtR * vR, wa = wtL * wL + wtR * wR;
    double Ha = wtL * HL + wtR * HR;
    double c2 = (g - 1.0) * (Ha - 0.5 * (ua * ua + va * va + wa * wa));
    double ca = sqrt(c2);
    double una = ua * n[0] + va * n[1] + wa * n[2];
    if (una >= 0.0) {
        double a1 = 0.5 * ((pR - pL) / c2 - (unR - unL) * srr / ca);
        double lam1 = una - ca;
        double lam1n = lam1 < 0.0 ? lam1 : 0.0;
        double fix1 = ((g + 1.0) / 4.0 * ca / srr) * a1;
        double x1 = -fabs(lam1) - fix1;
        double eps1 = x1 > 0.0 ? x1 : 0.0;
        double s1 = a1 * (lam1n - 0.5 * eps1);
        double mnL = rL * unL;
        flux[0] = (mnL + s1);
        flux[1] = (mnL * uL + pL * n[0] + s1 * (ua - ca * n[0]));
        flux[2] = (mnL * vL + pL * n[1] + s1 * (va - ca * n[1]));
        flux[3] = (mnL * wL + pL * n[2] + s1 * (wa - ca * n[2]));
        flux[4] = (mnL * HL + s1 * (Ha - una * ca));
    } else {
        double a4 = 0.5 * ((pR - pL) / c2 + (unR - unL) * srr / ca);
        double lam4 = una + ca;
        double lam4p = lam4 > 0.0 ? lam4 : 0.0;
        double fix4 = ((g + 1.0) / 4.0 * ca / srr) * a4;
        double x4 = -fabs(lam4) + fix4;
        double eps4 = x4 > 0.0 ? x4 : 0.0;
        double s4 = a4 * (lam4p + 0.5 * eps4);
        double mnR = rR * unR;
        flux[0] = (mnR - s4);
        flux[1] = (mnR * uR + pR * n[0] - s4 * (ua + ca * n[0]));
        flux[2] = (mnR * vR + pR * n[1] - s4 * (va + ca * n[1]));
        flux[3] = (mnR * wR + pR * n[2] - s4 * (wa + ca * n[2]));
        flux[4] = (mnR * HR - s4 * (Ha + una * ca));
    }
}

/* Two-point volume fluxes, RiemannSolvers_NS.f90:2085-2438 */
static void two_point_flux(const ox_case *c, const double *QL, const double *QR,
                           const double *JaL, const double *JaR, double *fS)
{
    double iL = 1.0 / QL[0], iR = 1.0 / QR[0];
    double uL = iL * QL[1], uR = iR * QR[1], vL = iL * QL[2], vR = iR * QR[2];
    double wL = iL * QL[3], wR = iR * QR[3];
    double pL = c->gm1 * (QL[4] - 0.5 * (QL[1] * uL + QL[2] * vL + QL[3] * wL));
    double pR = c->gm1 * (QR[4] - 0.5 * (QR[1] * uR + QR[2] * vR + QR[3] * wR));
    double f[5], g[5], h[5], Ja[3];
    switch (c->p.averaging) {
    case OX_AVG_DUCROS:
        for (int d = 0; d < 3; ++d) Ja[d] = 0.5 * (JaL[d] + JaR[d]);
        f[0] = 0.25 * (QL[0] + QR[0]) * (uL + uR);
        f[1] = 0.25 * (QL[1] + QR[1]) * (uL + uR) + 0.5 * (pL + pR);
        f[2] = 0.25 * (QL[2] + QR[2]) * (uL + uR);
        f[3] = 0.25 * (QL[3] + QR[3]) * (uL + uR);
        f[4] = 0.25 * (QL[4] + pL + QR[4] + pR) * (uL + uR);
        g[0] = 0.25 * (QL[0] + QR[0]) * (vL + vR);
        g[1] = 0.25 * (QL[1] + QR[1]) * (vL + vR);
        g[2] = 0.25 * (QL[2] + QR[2]) * (vL + vR) + 0.5 * (pL + pR);
        g[3] = 0.25 * (QL[3] + QR[3]) * (vL + vR);
        g[4] = 0.25 * (QL[4] + pL + QR[4] + pR) * (vL + vR);
        h[0] = 0.25 * (QL[0] + QR[0]) * (wL + wR);
        h[1] = 0.25 * (QL[1] + QR[1]) * (wL + wR);
        h[2] = 0.25 * (QL[2] + QR[2]) * (wL + wR);
        h[3] = 0.25 * (QL[3] + QR[3]) * (wL + wR) + 0.5 * (pL + pR);
        h[4] = 0.25 * (QL[4] + pL + QR[4] + pR) * (wL + wR);
        for (int v = 0; v < 5; ++v) fS[v] = f[v] * Ja[0] + g[v] * Ja[1] + h[v] * Ja[2];
        return;
    case OX_AVG_KG:
    case OX_AVG_PIROZZOLI: {
        double rho = 0.5 * (QL[0] + QR[0]), u = 0.5 * (uL + uR), v = 0.5 * (vL + vR);
        double w = 0.5 * (wL + wR), p = 0.5 * (pL + pR);
        double e = c->p.averaging == OX_AVG_KG ? 0.5 * (QL[4] * iL + QR[4] * iR)
                                               : 0.5 * ((QL[4] + pL) * iL + (QR[4] + pR) * iR);
        for (int d = 0; d < 3; ++d) Ja[d] = 0.5 * (JaL[d] + JaR[d]);
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = rho * u * w;
        g[0] = rho * v; g[1] = rho * v * u; g[2] = rho * v * v + p; g[3] = rho * v * w;
        h[0] = rho * w; h[1] = rho * w * u; h[2] = rho * w * v; h[3] = rho * w * w + p;
        if (c->p.averaging == OX_AVG_KG) {
            f[4] = rho * u * e + p * u; g[4] = rho * v * e + p * v; h[4] = rho * w * e + p * w;
        } else {
            f[4] = rho * u * e; g[4] = rho * v * e; h[4] = rho * w * e;
        }
        for (int q = 0; q < 5; ++q) fS[q] = f[q] * Ja[0] + g[q] * Ja[1] + h[q] * Ja[2];
        return; }
    case OX_AVG_MORINISHI: {
        /* Morinishi_TwoPointFlux (RiemannSolvers_NS.f90:2147-2229): enthalpy without the
         * kinetic energy, h = cp p */
        double cp = c->gamma * (1.0 / (1.4 - 1.0));
        double hL = cp * pL, hR = cp * pR;
        const double *m[3] = {&QL[1], &QL[2], &QL[3]};
        for (int d = 0; d < 3; ++d) Ja[d] = 0.5 * (JaL[d] + JaR[d]);
        double *out[3] = {f, g, h};
        double vel[3][2] = {{uL, uR}, {vL, vR}, {wL, wR}};
        for (int d = 0; d < 3; ++d) {
            /* direction d: the mass flux rho u_d of each side */
            double mL = *m[d], mR = QR[1 + d];
            double *o = out[d];
            o[0] = 0.5 * (mL + mR);
            o[1] = 0.25 * (mL + mR) * (uL + uR);
            o[2] = 0.25 * (mL + mR) * (vL + vR);
            o[3] = 0.25 * (mL + mR) * (wL + wR);
            o[1 + d] = o[1 + d] + 0.5 * (pL + pR);
            o[4] = 0.5 * (vel[d][0] * hL + vel[d][1] * hR) + 0.25 * (mL * uL + mR * uR) * (uL + uR)
                 + 0.25 * (mL * vL + mR * vR) * (vL + vR)
                 + 0.25 * (mL * wL + mR * wR) * (wL + wR)
                 - 0.25 * (mL * POW2(uL) + mR * POW2(uR))
                 - 0.25 * (mL * POW2(vL) + mR * POW2(vR))
                 - 0.25 * (mL * POW2(wL) + mR * POW2(wR));
        }
        for (int q = 0; q < 5; ++q) fS[q] = f[q] * Ja[0] + g[q] * Ja[1] + h[q] * Ja[2];
        return; }
    default: /* StandardDG_TwoPointFlux */
        for (int d = 0; d < 3; ++d) Ja[d] = (JaL[d] + JaR[d]);
        f[0] = (QL[1] + QR[1]);
        f[1] = (QL[1] * uL + QR[1] * uR + pL + pR);
        f[2] = (QL[1] * vL + QR[1] * vR);
        f[3] = (QL[1] * wL + QR[1] * wR);
        f[4] = (uL * (QL[4] + pL) + uR * (QR[4] + pR));
        g[0] = (QL[2] + QR[2]);
        g[1] = (QL[2] * uL + QR[2] * uR);
        g[2] = (QL[2] * vL + QR[2] * vR + pL + pR);
        g[3] = (QL[2] * wL + QR[2] * wR);
        g[4] = (vL * (QL[4] + pL) + vR * (QR[4] + pR));
        h[0] = (QL[3] + QR[3]);
        h[1] = (QL[3] * uL + QR[3] * uR);
        h[2] = (QL[3] * vL + QR[3] * vR);
        h[3] = (QL[3] * wL + QR[3] * wR + pL + pR);
        h[4] = (wL * (QL[4] + pL) + wR * (QR[4] + pR));
        for (int q = 0; q < 5; ++q) fS[q] = 0.25 * (f[q] * Ja[0] + g[q] * Ja[1] + h[q] * Ja[2]);
        return;
    }
}

/* ------------------------------------------------------------------------------------
 * Element/face plumbing
 * ---------------------------------------------------------------------------------- */
#define NIDX(Ne, i, j, k) ((i) + ((Ne)[0] + 1) * ((j) + ((Ne)[1] + 1) * (k)))

/* Face_AdaptSolutionToFace / Face_AdaptGradientsToFace (FaceClass.f90:281-512): element face
 * trace qe (nv values per node, element face orientation, dims Nel) -> face storage qf (Nf). */
static void adapt_to_face(ox_case *c, int f, int side, int nv, const double *qe, double *qf)
{
    facestore_t *s = &c->fs[f];
    int rot = c->fRot[f];
    const int *Nfs = side == 0 ? s->NfL : s->NfR;
    int n0 = s->Nf[0], n1 = s->Nf[1];
    double rotbuf[MAXN * MAXN * 15];
    const double *qin = qe;
    if (side == 1) {
        const int *Nel = s->NelR;
        for (int j = 0; j <= Nfs[1]; ++j)
            for (int i = 0; i <= Nfs[0]; ++i) {
                int ii, jj;
                left_to_right(i, j, Nfs[0], Nfs[1], rot, &ii, &jj);
                for (int v = 0; v < nv; ++v)
                    rotbuf[(i + (Nfs[0] + 1) * j) * nv + v] = qe[(ii + (Nel[0] + 1) * jj) * nv + v];
            }
        qin = rotbuf;
    }
    int pt = s->proj[side];
    if (pt == 0) {
        memcpy(qf, qin, sizeof(double) * nv * (n0 + 1) * (n1 + 1));
        return;
    }
    memset(qf, 0, sizeof(double) * nv * (n0 + 1) * (n1 + 1));
    double *T0 = tset(c, Nfs[0], n0), *T1 = tset(c, Nfs[1], n1);
    for (int j = 0; j <= n1; ++j)
        for (int i = 0; i <= n0; ++i)
            for (int v = 0; v < nv; ++v) {
                double acc = 0.0;
                if (pt == 1) {
                    for (int l = 0; l <= Nfs[0]; ++l) acc += T0[i * (Nfs[0] + 1) + l] * qin[(l + (Nfs[0] + 1) * j) * nv + v];
                } else if (pt == 2) {
                    for (int l = 0; l <= Nfs[1]; ++l) acc += T1[j * (Nfs[1] + 1) + l] * qin[(i + (Nfs[0] + 1) * l) * nv + v];
                } else {
                    for (int l = 0; l <= Nfs[1]; ++l)
                        for (int m = 0; m <= Nfs[0]; ++m)
                            acc += T0[i * (Nfs[0] + 1) + m] * T1[j * (Nfs[1] + 1) + l] * qin[(m + (Nfs[0] + 1) * l) * nv + v];
                }
                qf[(i + (n0 + 1) * j) * nv + v] = acc;
            }
}

/* Face_ProjectFluxToElements / Face_ProjectGradientFluxToElements (FaceClass.f90:597-700,
 * 865-964): face flux (Nf) -> element face orientation (Nel); right side scaled by fac. */
static void project_to_element(ox_case *c, int f, int side, int nv, const double *flux,
                               double *out, double fac)
{
    facestore_t *s = &c->fs[f];
    int rot = c->fRot[f];
    const int *Nfs = side == 0 ? s->NfL : s->NfR;
    int n0 = s->Nf[0], n1 = s->Nf[1];
    double aux[MAXN * MAXN * 15];
    int pt = s->proj[side];
    if (pt == 0) memcpy(aux, flux, sizeof(double) * nv * (n0 + 1) * (n1 + 1));
    else {
        double *T0 = tset(c, n0, Nfs[0]), *T1 = tset(c, n1, Nfs[1]);
        for (int j = 0; j <= Nfs[1]; ++j)
            for (int i = 0; i <= Nfs[0]; ++i)
                for (int v = 0; v < nv; ++v) {
                    double acc = 0.0;
                    if (pt == 1) {
                        for (int l = 0; l <= n0; ++l) acc += T0[i * (n0 + 1) + l] * flux[(l + (n0 + 1) * j) * nv + v];
                    } else if (pt == 2) {
                        for (int l = 0; l <= n1; ++l) acc += T1[j * (n1 + 1) + l] * flux[(i + (n0 + 1) * l) * nv + v];
                    } else {
                        for (int l = 0; l <= n1; ++l)
                            for (int m = 0; m <= n0; ++m)
                                acc += T0[i * (n0 + 1) + m] * T1[j * (n1 + 1) + l] * flux[(m + (n0 + 1) * l) * nv + v];
                    }
                    aux[(i + (Nfs[0] + 1) * j) * nv + v] = acc;
                }
    }
    if (side == 0) {
        memcpy(out, aux, sizeof(double) * nv * (Nfs[0] + 1) * (Nfs[1] + 1));
        return;
    }
    const int *Nel = s->NelR;
    for (int j = 0; j <= Nfs[1]; ++j)
        for (int i = 0; i <= Nfs[0]; ++i) {
            int ii, jj;
            left_to_right(i, j, Nfs[0], Nfs[1], rot, &ii, &jj);
            for (int v = 0; v < nv; ++v)
                out[(ii + (Nel[0] + 1) * jj) * nv + v] = fac * aux[(i + (Nfs[0] + 1) * j) * nv + v];
        }
}

/* Element trace on slot (HexElement_ProlongSolutionToFaces, HexElementClass.f90:233-302):
 * out[(a + (Na+1) b) * nv + v] = sum_l src[node] * v(l, side) along the slot's axis. */
static void element_trace(ox_case *c, int e, int slot, int nv, const double *src, double *out)
{
    const int *Ne = &c->N[3 * e];
    int ax = SLOT_AXIS[slot], plus = SLOT_PLUS[slot];
    int ua = SLOT_UV[slot][0], va = SLOT_UV[slot][1];
    nodal_t *nd = nodal(c, Ne[ax]);
    const double *vv = plus ? nd->vR : nd->vL;
    int Na = Ne[ua], Nb = Ne[va];
    long o = c->off[e];
    for (int b = 0; b <= Nb; ++b)
        for (int a = 0; a <= Na; ++a)
            for (int v = 0; v < nv; ++v) {
                double acc = 0.0;
                for (int l = 0; l <= Ne[ax]; ++l) {
                    int idx[3];
                    idx[ax] = l; idx[ua] = a; idx[va] = b;
                    acc += src[(o + NIDX(Ne, idx[0], idx[1], idx[2])) * nv + v] * vv[l];
                }
                out[(a + (Na + 1) * b) * nv + v] = acc;
            }
}

/* ------------------------------------------------------------------------------------
 * The stage
 * ---------------------------------------------------------------------------------- */
static void prolong_solution(ox_case *c, const double *Q)
{   /* HexMesh_ProlongSolutionToFaces (HexMesh.f90:948-1036) */
    double tr[MAXN * MAXN * NC];
    for (int e = 0; e < c->K; ++e)
        for (int s = 1; s <= 6; ++s) {
            int f = c->faceID[6 * e + s - 1], side = c->faceSide[6 * e + s - 1] - 1;
            element_trace(c, e, s, NC, Q, tr);
            adapt_to_face(c, f, side, NC, tr, c->fs[f].Q[side]);
        }
}

static void local_gradient(ox_case *c, const double *Q)
{   /* HexElement_ComputeLocalGradient (HexElementClass.f90:427-531), U = Q (state variables) */
    for (int e = 0; e < c->K; ++e) {
        const int *Ne = &c->N[3 * e];
        long o = c->off[e];
        nodal_t *a = nodal(c, Ne[0]), *b = nodal(c, Ne[1]), *cc = nodal(c, Ne[2]);
        for (int k = 0; k <= Ne[2]; ++k)
            for (int j = 0; j <= Ne[1]; ++j)
                for (int i = 0; i <= Ne[0]; ++i) {
                    long g = o + NIDX(Ne, i, j, k);
                    double Uxi[NC] = {0}, Ueta[NC] = {0}, Uzeta[NC] = {0};
                    for (int l = 0; l <= Ne[0]; ++l)
                        for (int v = 0; v < NC; ++v) Uxi[v] += Q[(o + NIDX(Ne, l, j, k)) * NC + v] * a->D[i * (Ne[0] + 1) + l];
                    for (int l = 0; l <= Ne[1]; ++l)
                        for (int v = 0; v < NC; ++v) Ueta[v] += Q[(o + NIDX(Ne, i, l, k)) * NC + v] * b->D[j * (Ne[1] + 1) + l];
                    for (int l = 0; l <= Ne[2]; ++l)
                        for (int v = 0; v < NC; ++v) Uzeta[v] += Q[(o + NIDX(Ne, i, j, l)) * NC + v] * cc->D[k * (Ne[2] + 1) + l];
                    const double *Ja = &c->Ja[9 * g];
                    double iJ = c->invJ[g];
                    for (int v = 0; v < NC; ++v) {
                        c->U_x[g * NC + v] = (Uxi[v] * Ja[0] + Ueta[v] * Ja[3] + Uzeta[v] * Ja[6]) * iJ;
                        c->U_y[g * NC + v] = (Uxi[v] * Ja[1] + Ueta[v] * Ja[4] + Uzeta[v] * Ja[7]) * iJ;
                        c->U_z[g * NC + v] = (Uxi[v] * Ja[2] + Ueta[v] * Ja[5] + Uzeta[v] * Ja[8]) * iJ;
                    }
                }
    }
}

static void br1_interface(ox_case *c)
{   /* BR1_ComputeElementInterfaceAverage (EllipticBR1.f90:571-627) */
    double un[MAXN * MAXN * NC * 3];
    for (int f = 0; f < c->F; ++f) {
        facestore_t *s = &c->fs[f];
        int n0 = s->Nf[0], n1 = s->Nf[1];
        for (int j = 0; j <= n1; ++j)
            for (int i = 0; i <= n0; ++i) {
                int q = i + (n0 + 1) * j;
                long fg = c->foff[f] + q;
                const double *n = &c->normal[3 * fg];
                double jac = c->fjac[fg];
                for (int v = 0; v < NC; ++v) {
                    double uS = 0.5 * (s->Q[1][q * NC + v] - s->Q[0][q * NC + v]) * jac;
                    /* unStar(v, d, i, j) layout: (q*3 + d)*NC + v  -> stored as [q][d][v] */
                    un[(q * 3 + 0) * NC + v] = uS * n[0];
                    un[(q * 3 + 1) * NC + v] = uS * n[1];
                    un[(q * 3 + 2) * NC + v] = uS * n[2];
                }
            }
        project_to_element(c, f, 0, 3 * NC, un, s->unStar[0], 1.0);
        project_to_element(c, f, 1, 3 * NC, un, s->unStar[1], 1.0);
    }
}

static void br1_lift_and_prolong(ox_case *c)
{   /* BR1_GradientFaceLoop (EllipticBR1.f90:531-569) with VectorWeakIntegrals_StdFace
     * (DGIntegrals.f90:365-443), then HexElement_ProlongGradientsToFaces (:304-372) */
    double tr[MAXN * MAXN * NC];
    for (int e = 0; e < c->K; ++e) {
        const int *Ne = &c->N[3 * e];
        long o = c->off[e];
        const double *H[7];
        for (int s = 1; s <= 6; ++s) {
            int f = c->faceID[6 * e + s - 1], side = c->faceSide[6 * e + s - 1] - 1;
            H[s] = c->fs[f].unStar[side];
        }
        nodal_t *A = nodal(c, Ne[0]), *B = nodal(c, Ne[1]), *C = nodal(c, Ne[2]);
        for (int k = 0; k <= Ne[2]; ++k)
            for (int j = 0; j <= Ne[1]; ++j)
                for (int i = 0; i <= Ne[0]; ++i) {
                    long g = o + NIDX(Ne, i, j, k);
                    double fi[3][NC];
                    /* order: L, R (xi), FR, BK (eta), BOT, T (zeta) */
                    const double *hL = &H[6][((j + (Ne[1] + 1) * k) * 3) * NC];
                    const double *hR = &H[4][((j + (Ne[1] + 1) * k) * 3) * NC];
                    const double *hF = &H[1][((i + (Ne[0] + 1) * k) * 3) * NC];
                    const double *hB = &H[2][((i + (Ne[0] + 1) * k) * 3) * NC];
                    const double *hBo = &H[3][((i + (Ne[0] + 1) * j) * 3) * NC];
                    const double *hT = &H[5][((i + (Ne[0] + 1) * j) * 3) * NC];
                    for (int d = 0; d < 3; ++d)
                        for (int v = 0; v < NC; ++v) {
                            double x = hL[d * NC + v] * A->bL[i];
                            x = x + hR[d * NC + v] * A->bR[i];
                            x = x + hF[d * NC + v] * B->bL[j];
                            x = x + hB[d * NC + v] * B->bR[j];
                            x = x + hBo[d * NC + v] * C->bL[k];
                            x = x + hT[d * NC + v] * C->bR[k];
                            fi[d][v] = x;
                        }
                    double iJ = c->invJ[g];
                    for (int v = 0; v < NC; ++v) {
                        c->U_x[g * NC + v] = c->U_x[g * NC + v] + fi[0][v] * iJ;
                        c->U_y[g * NC + v] = c->U_y[g * NC + v] + fi[1][v] * iJ;
                        c->U_z[g * NC + v] = c->U_z[g * NC + v] + fi[2][v] * iJ;
                    }
                }
        for (int s = 1; s <= 6; ++s) {
            int f = c->faceID[6 * e + s - 1], side = c->faceSide[6 * e + s - 1] - 1;
            element_trace(c, e, s, NC, c->U_x, tr);
            adapt_to_face(c, f, side, NC, tr, c->fs[f].Ux[side]);
            element_trace(c, e, s, NC, c->U_y, tr);
            adapt_to_face(c, f, side, NC, tr, c->fs[f].Uy[side]);
            element_trace(c, e, s, NC, c->U_z, tr);
            adapt_to_face(c, f, side, NC, tr, c->fs[f].Uz[side]);
        }
    }
}

static void viscosity(ox_case *c, const double *Q)
{   /* SD.f90:403-441: element and face viscosity (+ LES) */
    if (c->p.euler) return;
    for (long g = 0; g < c->nodes; ++g)
        laminar_mu_kappa(c, &Q[g * NC], &c->muNS[2 * g], &c->muNS[2 * g + 1]);
    if (c->p.les != OX_LES_NONE) {
        for (int e = 0; e < c->K; ++e) {
            const int *Ne = &c->N[3 * e];
            double delta = pow(c->vol[e] / ((Ne[0] + 1) * (Ne[1] + 1) * (Ne[2] + 1)), 1.0 / 3.0);
            for (long g = c->off[e]; g < c->off[e + 1]; ++g) {
                double mt = les_viscosity(c, delta, &Q[g * NC], &c->U_x[g * NC], &c->U_y[g * NC], &c->U_z[g * NC]);
                c->muT[g] = mt;
                c->muNS[2 * g] = c->muNS[2 * g] + mt;
                c->muNS[2 * g + 1] = c->muNS[2 * g + 1] + mt * c->mu_to_kappa;
            }
        }
    }
    for (int f = 0; f < c->F; ++f) {
        facestore_t *s = &c->fs[f];
        int nfn = (s->Nf[0] + 1) * (s->Nf[1] + 1);
        for (int q = 0; q < nfn; ++q)
            for (int side = 0; side < 2; ++side)
                laminar_mu_kappa(c, &s->Q[side][q * NC], &s->mu[side][2 * q], &s->mu[side][2 * q + 1]);
        if (c->p.les != OX_LES_NONE) {
            double delta = sqrt(c->fsurf[f] / ((s->Nf[0] + 1) * (s->Nf[1] + 1)));
            for (int q = 0; q < nfn; ++q)
                for (int side = 0; side < 2; ++side) {
                    double mt = les_viscosity(c, delta, &s->Q[side][q * NC], &s->Ux[side][q * NC],
                                              &s->Uy[side][q * NC], &s->Uz[side][q * NC]);
                    s->mu[side][2 * q] = s->mu[side][2 * q] + mt;
                    s->mu[side][2 * q + 1] = s->mu[side][2 * q + 1] + mt * c->mu_to_kappa;
                }
        }
    }
}

static void volume_integrals(ox_case *c, const double *Q, double *QDot)
{   /* TimeDerivative_VolumetricContribution (SD.f90:1602-1682) */
    int split = c->p.inviscid == OX_SPLIT;
    for (int e = 0; e < c->K; ++e) {
        const int *Ne = &c->N[3 * e];
        int nx = Ne[0] + 1, ny = Ne[1] + 1, nz = Ne[2] + 1, ne = nx * ny * nz;
        long o = c->off[e];
        nodal_t *A = nodal(c, Ne[0]), *B = nodal(c, Ne[1]), *C = nodal(c, Ne[2]);
        double *Finv = (double *)malloc(sizeof(double) * NC * ne * 3);
        double *Fvis = (double *)malloc(sizeof(double) * NC * ne * 3);
        for (int n = 0; n < ne; ++n) {
            long g = o + n;
            double F[NC][3], Fv[NC][3];
            const double *Ja = &c->Ja[9 * g];
            /* BaseClass_ComputeInnerFluxes (HyperbolicDiscretizationClass.f90:83-152) */
            euler_flux(c, &Q[g * NC], F);
            for (int v = 0; v < NC; ++v)
                for (int d = 0; d < 3; ++d)
                    Finv[(d * ne + n) * NC + v] = F[v][0] * Ja[3 * d + 0] + F[v][1] * Ja[3 * d + 1] + F[v][2] * Ja[3 * d + 2];
            /* BR1_ComputeInnerFluxes (EllipticBR1.f90:740-814) */
            if (!c->p.euler) {
                viscous_flux(c, &Q[g * NC], &c->U_x[g * NC], &c->U_y[g * NC], &c->U_z[g * NC],
                             c->muNS[2 * g], 0.0, c->muNS[2 * g + 1], Fv);
                for (int v = 0; v < NC; ++v)
                    for (int d = 0; d < 3; ++d)
                        Fvis[(d * ne + n) * NC + v] = Fv[v][0] * Ja[3 * d + 0] + Fv[v][1] * Ja[3 * d + 1] + Fv[v][2] * Ja[3 * d + 2];
            } else {
                for (int v = 0; v < NC; ++v) for (int d = 0; d < 3; ++d) Fvis[(d * ne + n) * NC + v] = 0.0;
            }
        }
        double *out = &QDot[o * NC];
        for (int n = 0; n < ne * NC; ++n) out[n] = 0.0;
        if (!split) {
            /* contravariantFlux = inviscid - viscous; ScalarWeakIntegrals_StdVolumeGreen (DGIntegrals.f90:56-87) */
            for (int n = 0; n < 3 * ne * NC; ++n) Finv[n] = Finv[n] - Fvis[n];
            for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int l = 0; l < nx; ++l) for (int i = 0; i < nx; ++i)
                for (int v = 0; v < NC; ++v)
                    out[NIDX(Ne, i, j, k) * NC + v] += A->hatD[i * nx + l] * Finv[(0 * ne + NIDX(Ne, l, j, k)) * NC + v];
            for (int k = 0; k < nz; ++k) for (int l = 0; l < ny; ++l) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i)
                for (int v = 0; v < NC; ++v)
                    out[NIDX(Ne, i, j, k) * NC + v] += B->hatD[j * ny + l] * Finv[(1 * ne + NIDX(Ne, i, l, k)) * NC + v];
            for (int l = 0; l < nz; ++l) for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i)
                for (int v = 0; v < NC; ++v)
                    out[NIDX(Ne, i, j, k) * NC + v] += C->hatD[k * nz + l] * Finv[(2 * ne + NIDX(Ne, i, j, l)) * NC + v];
        } else {
            /* SplitDG_ComputeSplitFormFluxes (HyperbolicSplitForm.f90:64-116) consumed on the fly,
             * ScalarWeakIntegrals_SplitVolumeDivergence (DGIntegrals.f90:92-129), QDot = -volInt */
            double fS[NC];
            for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int l = 0; l < nx; ++l) for (int i = 0; i < nx; ++i) {
                int ni = NIDX(Ne, i, j, k), nl = NIDX(Ne, l, j, k);
                if (l == i) for (int v = 0; v < NC; ++v) fS[v] = Finv[(0 * ne + ni) * NC + v];
                else {
                    int lo = i < l ? i : l, hi = i < l ? l : i;
                    int a0 = NIDX(Ne, lo, j, k), a1 = NIDX(Ne, hi, j, k);
                    two_point_flux(c, &Q[(o + a0) * NC], &Q[(o + a1) * NC], &c->Ja[9 * (o + a0) + 0], &c->Ja[9 * (o + a1) + 0], fS);
                }
                for (int v = 0; v < NC; ++v)
                    out[ni * NC + v] = out[ni * NC + v] + A->sharpD[i * nx + l] * fS[v] + A->hatD[i * nx + l] * Fvis[(0 * ne + nl) * NC + v];
            }
            for (int k = 0; k < nz; ++k) for (int l = 0; l < ny; ++l) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
                int ni = NIDX(Ne, i, j, k), nl = NIDX(Ne, i, l, k);
                if (l == j) for (int v = 0; v < NC; ++v) fS[v] = Finv[(1 * ne + ni) * NC + v];
                else {
                    int lo = j < l ? j : l, hi = j < l ? l : j;
                    int a0 = NIDX(Ne, i, lo, k), a1 = NIDX(Ne, i, hi, k);
                    two_point_flux(c, &Q[(o + a0) * NC], &Q[(o + a1) * NC], &c->Ja[9 * (o + a0) + 3], &c->Ja[9 * (o + a1) + 3], fS);
                }
                for (int v = 0; v < NC; ++v)
                    out[ni * NC + v] = out[ni * NC + v] + B->sharpD[j * ny + l] * fS[v] + B->hatD[j * ny + l] * Fvis[(1 * ne + nl) * NC + v];
            }
            for (int l = 0; l < nz; ++l) for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
                int ni = NIDX(Ne, i, j, k), nl = NIDX(Ne, i, j, l);
                if (l == k) for (int v = 0; v < NC; ++v) fS[v] = Finv[(2 * ne + ni) * NC + v];
                else {
                    int lo = k < l ? k : l, hi = k < l ? l : k;
                    int a0 = NIDX(Ne, i, j, lo), a1 = NIDX(Ne, i, j, hi);
                    two_point_flux(c, &Q[(o + a0) * NC], &Q[(o + a1) * NC], &c->Ja[9 * (o + a0) + 6], &c->Ja[9 * (o + a1) + 6], fS);
                }
                for (int v = 0; v < NC; ++v)
                    out[ni * NC + v] = out[ni * NC + v] + C->sharpD[k * nz + l] * fS[v] + C->hatD[k * nz + l] * Fvis[(2 * ne + nl) * NC + v];
            }
            for (int n = 0; n < ne * NC; ++n) out[n] = -out[n];
        }
        free(Finv); free(Fvis);
    }
}

static void interface_fluxes(ox_case *c)
{   /* computeElementInterfaceFlux (SD.f90:1710-1799) with BR1_RiemannSolver (EllipticBR1.f90:816-868) */
    double flux[MAXN * MAXN * NC];
    for (int f = 0; f < c->F; ++f) {
        facestore_t *s = &c->fs[f];
        int n0 = s->Nf[0], n1 = s->Nf[1];
        for (int j = 0; j <= n1; ++j)
            for (int i = 0; i <= n0; ++i) {
                int q = i + (n0 + 1) * j;
                long fg = c->foff[f] + q;
                const double *n = &c->normal[3 * fg];
                double visc[NC] = {0, 0, 0, 0, 0}, inv[NC];
                if (!c->p.euler) {
                    double fL[NC][3], fR[NC][3];
                    viscous_flux(c, &s->Q[0][q * NC], &s->Ux[0][q * NC], &s->Uy[0][q * NC], &s->Uz[0][q * NC],
                                 s->mu[0][2 * q], 0.0, s->mu[0][2 * q + 1], fL);
                    viscous_flux(c, &s->Q[1][q * NC], &s->Ux[1][q * NC], &s->Uy[1][q * NC], &s->Uz[1][q * NC],
                                 s->mu[1][2 * q], 0.0, s->mu[1][2 * q + 1], fR);
                    for (int v = 0; v < NC; ++v) {
                        double a = 0.5 * (fL[v][0] + fR[v][0]), b = 0.5 * (fL[v][1] + fR[v][1]), cz = 0.5 * (fL[v][2] + fR[v][2]);
                        visc[v] = a * n[0] + b * n[1] + cz * n[2];
                    }
                }
                if (c->p.riemann == OX_LXF)
                    lxf_flux(c, &s->Q[0][q * NC], &s->Q[1][q * NC], n, &c->t1[3 * fg], &c->t2[3 * fg], inv);
                else
                    roe_flux(c, &s->Q[0][q * NC], &s->Q[1][q * NC], n, inv);
                for (int v = 0; v < NC; ++v) flux[q * NC + v] = (inv[v] - visc[v]) * c->fjac[fg];
            }
        project_to_element(c, f, 0, NC, flux, s->fStar[0], 1.0);
        project_to_element(c, f, 1, NC, flux, s->fStar[1], -1.0);
    }
}

static void surface_and_scale(ox_case *c, double *QDot)
{   /* TimeDerivative_FacesContribution (SD.f90:1686-1701) with ScalarWeakIntegrals_StdFace
     * (DGIntegrals.f90:214-273), then QDot /= J (SD.f90:483-494) */
    for (int e = 0; e < c->K; ++e) {
        const int *Ne = &c->N[3 * e];
        long o = c->off[e];
        const double *H[7];
        for (int s = 1; s <= 6; ++s) {
            int f = c->faceID[6 * e + s - 1], side = c->faceSide[6 * e + s - 1] - 1;
            H[s] = c->fs[f].fStar[side];
        }
        nodal_t *A = nodal(c, Ne[0]), *B = nodal(c, Ne[1]), *C = nodal(c, Ne[2]);
        for (int k = 0; k <= Ne[2]; ++k)
            for (int j = 0; j <= Ne[1]; ++j)
                for (int i = 0; i <= Ne[0]; ++i) {
                    long g = o + NIDX(Ne, i, j, k);
                    for (int v = 0; v < NC; ++v) {
                        double x = H[6][(j + (Ne[1] + 1) * k) * NC + v] * A->bL[i];
                        x = x + H[4][(j + (Ne[1] + 1) * k) * NC + v] * A->bR[i];
                        x = x + H[1][(i + (Ne[0] + 1) * k) * NC + v] * B->bL[j];
                        x = x + H[2][(i + (Ne[0] + 1) * k) * NC + v] * B->bR[j];
                        x = x + H[3][(i + (Ne[0] + 1) * j) * NC + v] * C->bL[k];
                        x = x + H[5][(i + (Ne[0] + 1) * j) * NC + v] * C->bR[k];
                        QDot[g * NC + v] = (QDot[g * NC + v] - x) / c->J[g];
                    }
                }
    }
}

int ox_compute_qdot(ox_case *c, const double *Q, double t, double *QDot, double *grad)
{
    (void)t;
    prolong_solution(c, Q);
    if (!c->p.euler) {
        local_gradient(c, Q);
        br1_interface(c);
        br1_lift_and_prolong(c);
    }
    viscosity(c, Q);
    volume_integrals(c, Q, QDot);
    interface_fluxes(c);
    surface_and_scale(c, QDot);
    if (grad)
        for (long g = 0; g < c->nodes; ++g)
            for (int v = 0; v < NC; ++v) {
                grad[15 * g + v] = c->U_x[g * NC + v];
                grad[15 * g + 5 + v] = c->U_y[g * NC + v];
                grad[15 * g + 10 + v] = c->U_z[g * NC + v];
            }
    return 0;
}

int ox_rk3_step(ox_case *c, double *Q, double *G, double t, double dt)
{   /* TakeRK3Step (ExplicitMethods.f90:667-788), Williamson low-storage RK3 */
    static const double a[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
    static const double b[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
    static const double cc[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
    long n = c->nodes * NC;
    double *QDot = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < 3; ++k) {
        ox_compute_qdot(c, Q, t + b[k] * dt, QDot, NULL);
        for (long i = 0; i < n; ++i) {
            G[i] = a[k] * G[i] + QDot[i];
            Q[i] = Q[i] + cc[k] * dt * G[i];
        }
    }
    free(QDot);
    return 0;
}

double ox_max_dt(ox_case *c, const double *Q, double cfl, double dcfl)
{   /* MaxTimeStep (DGSEMClass.f90:870-1035) with ComputeEigenvaluesForState (Physics_NS.f90:927-962) */
    double dtC = 1.0e300, dtV = 1.0e300;
    for (int e = 0; e < c->K; ++e) {
        const int *Ne = &c->N[3 * e];
        double dc = 1.0 / fabs(c->nod[Ne[0]].x[1] - c->nod[Ne[0]].x[0]);
        double de = 1.0 / fabs(c->nod[Ne[1]].x[1] - c->nod[Ne[1]].x[0]);
        double dz = 1.0 / fabs(c->nod[Ne[2]].x[1] - c->nod[Ne[2]].x[0]);
        double dc2 = dc * dc, de2 = de * de, dz2 = dz * dz;
        for (long g = c->off[e]; g < c->off[e + 1]; ++g) {
            const double *q = &Q[g * NC];
            double u = fabs(q[1] / q[0]), v = fabs(q[2] / q[0]), w = fabs(q[3] / q[0]);
            double p = pressure(c, q);
            double a = sqrt(c->gamma * p / q[0]);
            double ev[3] = {u + a, v + a, w + a};
            const double *Ja = &c->Ja[9 * g];
            double jac = c->J[g];
            double lx = fabs(Ja[0] * ev[0] + Ja[1] * ev[1] + Ja[2] * ev[2]) * dc;
            double ly = fabs(Ja[3] * ev[0] + Ja[4] * ev[1] + Ja[5] * ev[2]) * de;
            double lz = fabs(Ja[6] * ev[0] + Ja[7] * ev[1] + Ja[8] * ev[2]) * dz;
            double t = cfl * fabs(jac) / (lx + ly + lz);
            if (t < dtC) dtC = t;
            if (!c->p.euler) {
                double T = c->gammaM2 * p / q[0];
                double mu = sutherland(c, T);
                double vx = mu * dc2 * fabs(Ja[0] + Ja[1] + Ja[2]);
                double vy = mu * de2 * fabs(Ja[3] + Ja[4] + Ja[5]);
                double vz = mu * dz2 * fabs(Ja[6] + Ja[7] + Ja[8]);
                double tv = dcfl * fabs(jac) / (vx + vy + vz);
                if (tv < dtV) dtV = tv;
            }
        }
    }
    return dtC < dtV ? dtC : dtV;
}

double ox_kinetic_energy(ox_case *c, const double *Q)
{   /* ScalarVolumeIntegral(KINETIC_ENERGY) / ScalarVolumeIntegral(VOLUME) */
    double ke = 0.0, vol = 0.0;
    for (int e = 0; e < c->K; ++e) {
        const int *Ne = &c->N[3 * e];
        nodal_t *A = nodal(c, Ne[0]), *B = nodal(c, Ne[1]), *C = nodal(c, Ne[2]);
        double val = 0.0, vv = 0.0;
        for (int k = 0; k <= Ne[2]; ++k) for (int j = 0; j <= Ne[1]; ++j) for (int i = 0; i <= Ne[0]; ++i) {
            long g = c->off[e] + NIDX(Ne, i, j, k);
            const double *q = &Q[g * NC];
            double K2 = POW2(q[1]);
            K2 = K2 + POW2(q[2]);
            K2 = K2 + POW2(q[3]);
            K2 = 0.5 * K2 / q[0];
            val = val + A->w[i] * B->w[j] * C->w[k] * c->J[g] * K2;
            vv = vv + A->w[i] * B->w[j] * C->w[k] * c->J[g];
        }
        ke += val; vol += vv;
    }
    return ke / vol;
}
