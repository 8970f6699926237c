// Host-side spectral building blocks shared by the engine and the mesh front end: Gauss-Lobatto
// and Gauss nodes and weights, barycentric interpolation, derivative matrices and the mortar matrices,
// each restated from the reference routine it names (operation order kept: the geometry and the
// operators must come out with the reference's bits).
#pragma once
#include <algorithm>
#include <cmath>

namespace hxs {

inline void gl_nodes(int N, double *x, double *w)
{
    // LegendreLobattoNodesAndWeights (LegendreAlgorithms.f90:275-355)
    const double tol = 4.0 * 2.220446049250313e-16;
    auto qL = [&](double xx, double &q, double &dq, double &LN) {
        double a2 = 1.0, da2 = 0.0, a1 = xx, da1 = 1.0, Lk = 0, dLk = 0;
        for (int k = 2; k <= N; ++k) {
            Lk = ((2 * k - 1) * xx * a1 - (k - 1) * a2) / k;
            dLk = da2 + (2 * k - 1) * a1;
            a2 = a1; a1 = Lk; da2 = da1; da1 = dLk;
        }
        int k = N + 1;
        Lk = ((2 * k - 1) * xx * a1 - (k - 1) * a2) / k;
        dLk = da2 + (2 * k - 1) * a1;
        q = Lk - a2; dq = dLk - da2; LN = a1;
    };
    if (N == 1) { x[0] = -1; x[1] = 1; w[0] = w[1] = 1; return; }
    x[0] = -1.0; x[N] = 1.0; w[0] = w[N] = 2.0 / (N * (N + 1));
    for (int j = 1; j <= (N + 1) / 2 - 1; ++j) {
        double xj = -std::cos((j + 0.25) * M_PI / N - 3.0 / (8 * N * M_PI * (j + 0.25)));
        double q, dq, LN;
        for (int k = 0; k <= 10; ++k) {
            qL(xj, q, dq, LN);
            double d = -q / dq;
            xj += d;
            if (std::fabs(d) <= tol * std::fabs(xj)) break;
        }
        qL(xj, q, dq, LN);
        x[j] = xj; w[j] = 2.0 / (N * (N + 1) * LN * LN);
        x[N - j] = -xj; w[N - j] = w[j];
    }
    if (N % 2 == 0) {
        double a2 = 1.0, a1 = 0.0, L = (N == 0) ? 1.0 : 0.0;
        for (int k = 2; k <= N; ++k) { L = ((2 * k - 1) * 0.0 * a1 - (k - 1) * a2) / k; a2 = a1; a1 = L; }
        x[N / 2] = 0.0; w[N / 2] = 2.0 / (N * (N + 1) * L * L);
    }
}

inline void legendre(int N, double x, double &L, double &dL)
{
    // LegendrePolyAndDerivative (LegendreAlgorithms.f90:88-128), three-term recursion
    if (N == 0) { L = 1.0; dL = 0.0; return; }
    if (N == 1) { L = x; dL = 1.0; return; }
    double Lm2 = 1.0, dLm2 = 0.0, Lm1 = x, dLm1 = 1.0;
    for (int k = 2; k <= N; ++k) {
        L = ((2 * k - 1) * x * Lm1 - (k - 1) * Lm2) / k;
        dL = dLm2 + (2 * k - 1) * Lm1;
        Lm2 = Lm1; Lm1 = L;
        dLm2 = dLm1; dLm1 = dL;
    }
}

inline void gauss_nodes(int N, double *x, double *w)
{
    // GaussLegendreNodesAndWeights (LegendreAlgorithms.f90:137-214): Newton on L_{N+1} from the
    // Chebyshev-Gauss points, 4 eps relative; the other half mirrored, the middle node apart
    const double tol = 4.0 * 2.220446049250313e-16;
    double L = 0.0, dL = 0.0;
    if (N == 0) { x[0] = 0.0; w[0] = 2.0; return; }
    if (N == 1) {
        x[0] = -std::sqrt(1.0 / 3.0); w[0] = 1.0; x[1] = -x[0]; w[1] = w[0];
    } else {
        for (int j = 0; j < (N + 1) / 2; ++j) {
            double xj = -std::cos((2 * j + 1) * M_PI / (2 * N + 2));
            for (int k = 0; k <= 10; ++k) {
                legendre(N + 1, xj, L, dL);
                const double d = -L / dL;
                xj = xj + d;
                if (std::fabs(d) <= tol * std::fabs(xj)) break;
            }
            legendre(N + 1, xj, L, dL);
            x[j] = xj; w[j] = 2.0 / ((1.0 - xj * xj) * (dL * dL));
            x[N - j] = -xj; w[N - j] = w[j];
        }
    }
    if (N % 2 == 0) {
        legendre(N + 1, 0.0, L, dL);
        x[N / 2] = 0.0; w[N / 2] = 2.0 / (dL * dL);
    }
}

// the nodes and weights of "discretization nodes": Gauss (gauss = true) or Gauss-Lobatto
// (NodalStorageClass.f90:208-226)
inline void nodes_of(bool gauss, int N, double *x, double *w)
{
    if (gauss) gauss_nodes(N, x, w);
    else gl_nodes(N, x, w);
}

inline void left_to_right(int i, int j, int Nx, int Ny, int rot, int &ii, int &jj)
{
    // leftIndexes2Right (MeshTypes.f90:70-108)
    switch (rot) {
    case 0: ii = i; jj = j; break;
    case 1: ii = Ny - j; jj = i; break;
    case 2: ii = Nx - i; jj = Ny - j; break;
    case 3: ii = j; jj = Nx - i; break;
    case 4: ii = j; jj = i; break;
    case 5: ii = Nx - i; jj = j; break;
    case 6: ii = Ny - j; jj = Nx - i; break;
    default: ii = i; jj = Ny - j; break;
    }
}

inline bool almost_equal(double a, double b)
{
    // AlmostEqual (Utilities.f90:110-139), tolerance 2 epsilon
    const double tol = 2.0 * 2.220446049250313e-16;
    if (a == 0.0 || b == 0.0) return std::fabs(a - b) <= tol;
    return std::fabs(b - a) <= tol * std::max(std::fabs(a), std::fabs(b));
}

inline void bary(int N, const double *x, double *wb)
{
    // BarycentricWeights (InterpolationAndDerivatives.f90:230-262)
    for (int j = 0; j <= N; ++j) wb[j] = 1.0;
    for (int j = 1; j <= N; ++j)
        for (int k = 0; k < j; ++k) { wb[k] *= (x[k] - x[j]); wb[j] *= (x[j] - x[k]); }
    for (int j = 0; j <= N; ++j) wb[j] = 1.0 / wb[j];
}

// PolynomialInterpolationMatrix (InterpolationAndDerivatives.f90:413-461): T (M+1)x(N+1)
inline void interp_matrix(int N, int M, const double *xo, const double *wb, const double *xn, double *T)
{
    for (int k = 0; k <= M; ++k) {
        bool match = false;
        for (int j = 0; j <= N; ++j) {
            T[k * (N + 1) + j] = 0.0;
            if (almost_equal(xn[k], xo[j])) { match = true; T[k * (N + 1) + j] = 1.0; }
        }
        if (!match) {
            double s = 0.0;
            for (int j = 0; j <= N; ++j) {
                const double t = wb[j] / (xn[k] - xo[j]);
                T[k * (N + 1) + j] = t;
                s = s + t;
            }
            for (int j = 0; j <= N; ++j) T[k * (N + 1) + j] = T[k * (N + 1) + j] / s;
        }
    }
}

// Tset(N -> M) (InterpolationMatrix_Construct, InterpolationMatrices.f90:42-108): Lagrange
// interpolation when N < M, otherwise the transpose of the M -> N interpolation weighted by
// w_N(j) / w_M(i) (the L2 projection)
inline void tset_host(int N, int M, double *T, bool gauss = false)
{
    double xo[8], wo[8], xd[8], wd[8], wb[8];
    nodes_of(gauss, N, xo, wo);
    nodes_of(gauss, M, xd, wd);
    if (N < M) {
        bary(N, xo, wb);
        interp_matrix(N, M, xo, wb, xd, T);
    } else {
        double Tin[64];
        bary(M, xd, wb);
        interp_matrix(M, N, xd, wb, xo, Tin);
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= M; ++i) T[i * (N + 1) + j] = Tin[j * (M + 1) + i] * wo[j] / wd[i];
    }
}


// Chebyshev-Gauss-Lobatto nodes -cos(i pi / N) (NodalStorageClass.f90:249-251)
inline void cgl_nodes(int N, double *x)
{
    for (int i = 0; i <= N; ++i) x[i] = N ? -std::cos(1.0 * i * M_PI / N) : 0.0;
}

// PolynomialDerivativeMatrix (InterpolationAndDerivatives.f90:799-834): D[i*(N+1)+j]
inline void deriv_matrix(int N, const double *x, double *D)
{
    double wb[16];
    bary(N, x, wb);
    for (int i = 0; i <= N; ++i) {
        D[i * (N + 1) + i] = 0.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) {
                D[i * (N + 1) + j] = wb[j] / (wb[i] * (x[i] - x[j]));
                D[i * (N + 1) + i] = D[i * (N + 1) + i] - D[i * (N + 1) + j];
            }
    }
}

// InterpolatingPolynomialVector (InterpolationAndDerivatives.f90:109-159): the Lagrange basis
// of the nodes x (barycentric weights wb) at the point t
inline void interp_vector(double t, int N, const double *x, const double *wb, double *p)
{
    bool match = false;
    for (int j = 0; j <= N; ++j) {
        p[j] = 0.0;
        if (almost_equal(t, x[j])) { p[j] = 1.0; match = true; }
    }
    if (match) return;
    double d = 0.0;
    for (int j = 0; j <= N; ++j) {
        const double q = wb[j] / (t - x[j]);
        p[j] = q;
        d = d + q;
    }
    for (int j = 0; j <= N; ++j) p[j] = p[j] / d;
}

}  // namespace hxs
