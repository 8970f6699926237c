// Stage kernels of the hx3d engine (one explicit RK3 stage = k_grad, k_face, k_volume).
// Included by hx3d_engine.hip.  See that file for the device data layout.
#pragma once
#include <hip/hip_runtime.h>

#include "physics.hpp"

namespace hxk {

using hx::NV;
using hx::Phys;

struct Ops {  // NodalStorage for GL nodes (NodalStorageClass.f90:170-337), row-major (i,l)
    double D[64], hatD[64], sharpD[64], w[8];
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of
// elements so that face neighbours (gathered through L2) mostly share an XCD.  The grid is a
// multiple of 8, which makes this a bijection of [0, grid); logical blocks past the last
// element exit.
__host__ __device__ inline int xcd_block(int b, int grid)
{
    const int per = grid >> 3;
    return (b & 7) * per + (b >> 3);
}

// Averaging selectors of the templated split-form kernel (RiemannSolvers_NS.f90:229-285)
enum { AVG_STANDARD = 0, AVG_DUCROS = 1, AVG_KG = 2, AVG_PIROZZOLI = 3, FORM_STANDARD = -1 };

// ------------------------------------------------------------------------------------
// k_grad: per element, local gradient (HexElement_ComputeLocalGradient,
// HexElementClass.f90:427-531) + BR1 lifting (EllipticBR1.f90:168-397), then the viscous
// state of ViscousFlux_STATE (velocity gradients, grad T) at every node.
// ------------------------------------------------------------------------------------
template <int N, int EPB>
__global__ void __launch_bounds__(((EPB * (N + 1) * (N + 1) * (N + 1) + 63) / 64) * 64)
k_grad(const double *__restrict__ Q, const double *__restrict__ Ja, const double *__restrict__ invJ,
       const int *__restrict__ nbr, const int *__restrict__ efn, const double *__restrict__ fgeo,
       double *__restrict__ vgrad, int K, long S, long FS, Ops ops, Phys ph)
{
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1;
    __shared__ double sQ[EPB][NV][ne];
    __shared__ double sD[n1 * n1];
    const int nb = (K + EPB - 1) / EPB;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int t = threadIdx.x;
    const int el = t / ne, n = t - el * ne;
    const int e = lb * EPB + el;
    const bool active = lb < nb && el < EPB && e < K;
    for (int x = t; x < n1 * n1; x += blockDim.x) sD[x] = ops.D[(x / n1) * 8 + (x % n1)];
    const long g = (long)e * ne + n;
    double q[NV];
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) { q[v] = Q[v * S + g]; sQ[el][v][n] = q[v]; }
    }
    __syncthreads();
    if (!active) return;
    const int i = n % n1, j = (n / n1) % n1, k = n / nf;
    double uxi[NV], ueta[NV], uzeta[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { uxi[v] = 0.0; ueta[v] = 0.0; uzeta[v] = 0.0; }
#pragma unroll
    for (int l = 0; l < n1; ++l) {
        const double di = sD[i * n1 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v) uxi[v] = uxi[v] + sQ[el][v][l + n1 * (j + n1 * k)] * di;
    }
#pragma unroll
    for (int l = 0; l < n1; ++l) {
        const double dj = sD[j * n1 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v) ueta[v] = ueta[v] + sQ[el][v][i + n1 * (l + n1 * k)] * dj;
    }
#pragma unroll
    for (int l = 0; l < n1; ++l) {
        const double dk = sD[k * n1 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v) uzeta[v] = uzeta[v] + sQ[el][v][i + n1 * (j + n1 * l)] * dk;
    }
    double ja[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = Ja[c * S + g];
    const double iJ = invJ[g];
    double gr[3][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        gr[0][v] = (uxi[v] * ja[0] + ueta[v] * ja[3] + uzeta[v] * ja[6]) * iJ;
        gr[1][v] = (uxi[v] * ja[1] + ueta[v] * ja[4] + uzeta[v] * ja[7]) * iJ;
        gr[2][v] = (uxi[v] * ja[2] + ueta[v] * ja[5] + uzeta[v] * ja[8]) * iJ;
    }
    // BR1 lifting (BR1_ComputeElementInterfaceAverage, EllipticBR1.f90:571-627, and
    // BR1_GradientFaceLoop, :531-569): on each face the node lies on,
    // unStar = 1/2 (U_R - U_L) |Ja| n with the face frame of the left element, the same value
    // for both sides; the GL surface integral weights it by b = 1/w_0.
    const bool onface = (i == 0) | (i == N) | (j == 0) | (j == N) | (k == 0) | (k == N);
    if (onface) {
        const double b0 = 1.0 / ops.w[0];
        double fi[3][NV];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) fi[d][v] = 0.0;
        auto lift = [&](int slot, int a, int b) {
            const long x = ((long)e * 6 + slot) * nf + a + n1 * b;
            const int gn = nbr[x];
            const int code = efn[x];
            const long fq = code >> 1;
            const bool right = code & 1;
            const double jac = fgeo[9 * FS + fq];
            const double n0 = fgeo[fq], n1_ = fgeo[FS + fq], n2 = fgeo[2 * FS + fq];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double qn = Q[v * S + gn];
                const double uS = 0.5 * (right ? q[v] - qn : qn - q[v]) * jac;
                fi[0][v] = fi[0][v] + (uS * n0) * b0;
                fi[1][v] = fi[1][v] + (uS * n1_) * b0;
                fi[2][v] = fi[2][v] + (uS * n2) * b0;
            }
        };
        if (i == 0) lift(5, j, k);
        if (i == N) lift(3, j, k);
        if (j == 0) lift(0, i, k);
        if (j == N) lift(1, i, k);
        if (k == 0) lift(2, i, j);
        if (k == N) lift(4, i, j);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) gr[d][v] = gr[d][v] + fi[d][v] * iJ;
    }
    double vg[12];
    hx::viscous_state(ph, q, gr[0], gr[1], gr[2], vg);
#pragma unroll
    for (int c = 0; c < 12; ++c) vgrad[c * S + g] = vg[c];
}

// ------------------------------------------------------------------------------------
// k_face: one thread per face node (computeElementInterfaceFlux, SD.f90:1710-1799)
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(256)
k_face(const double *__restrict__ Q, const double *__restrict__ vgrad, const double *__restrict__ fgeo,
       const int2 *__restrict__ fnode, const double *__restrict__ fdelta, double *__restrict__ fstar,
       long FS, long S, Phys ph)
{
    constexpr int n1 = N + 1, nf = n1 * n1;
    const long fq = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (fq >= FS) return;
    const int2 gg = fnode[fq];
    // face frame: fgeo[c*FS + fq], c = n (0..2), t1 (3..5), t2 (6..8), |Ja| (9)
    double nrm[3], t1[3], t2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nrm[c] = fgeo[c * FS + fq];
    const double jac = fgeo[9 * FS + fq];
    double qL[NV], qR[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { qL[v] = Q[v * S + gg.x]; qR[v] = Q[v * S + gg.y]; }
    double visc[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!ph.euler) {
        // BR1_RiemannSolver (EllipticBR1.f90:816-868) with the face viscosities of
        // compute_viscosity_at_faces (SpatialDiscretization.f90:1345-1398)
        double FL[NV][3], FR[NV][3];
        {
            double vg[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) vg[c] = vgrad[c * S + gg.x];
            double mu, kappa;
            hx::laminar_mu_kappa(ph, qL, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_viscosity(ph, fdelta[fq / nf], qL, vg);
                mu = mu + mt; kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / qL[0];
            hx::viscous_flux(mu, kappa, qL[1] * ir, qL[2] * ir, qL[3] * ir, vg, FL);
        }
        {
            double vg[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) vg[c] = vgrad[c * S + gg.y];
            double mu, kappa;
            hx::laminar_mu_kappa(ph, qR, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_viscosity(ph, fdelta[fq / nf], qR, vg);
                mu = mu + mt; kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / qR[0];
            hx::viscous_flux(mu, kappa, qR[1] * ir, qR[2] * ir, qR[3] * ir, vg, FR);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
            visc[v] = 0.5 * (FL[v][0] + FR[v][0]) * nrm[0] + 0.5 * (FL[v][1] + FR[v][1]) * nrm[1] +
                      0.5 * (FL[v][2] + FR[v][2]) * nrm[2];
    }
    double inv[NV];
    if (ph.riemann == 1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) { t1[c] = fgeo[(3 + c) * FS + fq]; t2[c] = fgeo[(6 + c) * FS + fq]; }
        hx::lxf_flux(ph, qL, qR, nrm, t1, t2, inv);
    } else {
        hx::roe_flux(ph, qL, qR, nrm, inv);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) fstar[v * FS + fq] = (inv[v] - visc[v]) * jac;
}

// ------------------------------------------------------------------------------------
// k_volume: per element, the volume integral (standard or split form), the surface integral
// of the stored face fluxes, 1/J, the Williamson RK3 register update, and (last stage) the
// MaxTimeStep reduction of the updated state for the next step.
// ------------------------------------------------------------------------------------
struct DtArgs {
    const unsigned long long *in;   // [2]: convective, viscous limit bit patterns (or null)
    unsigned long long *out;        // [2]: accumulate the next step's limits (or null)
    unsigned long long *reset;      // [2]: set to +inf by one thread (or null)
    double *time;                   // t += dt by one thread (or null)
    double host_dt;                 // used when in == null
    double cfl, dcfl, dcsi;         // MaxTimeStep parameters (DGSEMClass.f90:870-1035)
};

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

template <int N, int EPB, int FORM>
__global__ void __launch_bounds__(((EPB * (N + 1) * (N + 1) * (N + 1) + 63) / 64) * 64)
k_volume(double *__restrict__ Q, const double *__restrict__ vgrad, const double *__restrict__ Ja,
         const double *__restrict__ Jac, const double *__restrict__ edelta,
         const int *__restrict__ efn, const double *__restrict__ fstar, double *__restrict__ G,
         double *__restrict__ qdot_out, double a_k, double c_k, int update, DtArgs dta, int K,
         long S, long FS, Ops ops, Phys ph)
{
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1;
    constexpr bool SPLIT = FORM >= 0;
    // sF: contravariant flux of the current direction (standard: inviscid - viscous; split:
    // viscous); sP: two-point node data (q, u, v, w, p, 1/rho); sJ: metric row of the direction.
    // sPair: the two-point fluxes of the current direction, one per unordered node pair of
    // each line (computed once: the flux formulas are symmetric in IEEE arithmetic).
    constexpr int NPAIR = N * (N + 1) / 2;
    __shared__ double sF[EPB][NV][ne];
    __shared__ double sP[SPLIT ? EPB : 1][SPLIT ? 10 : 1][SPLIT ? ne : 1];
    __shared__ double sJ[SPLIT ? EPB : 1][SPLIT ? 3 : 1][SPLIT ? ne : 1];
    __shared__ double sPair[SPLIT ? EPB : 1][SPLIT ? nf : 1][SPLIT ? NPAIR : 1][SPLIT ? NV : 1];
    __shared__ double sHat[n1 * n1], sSharp[n1 * n1];
    __shared__ double sMin[2][16];
    const int nb = (K + EPB - 1) / EPB;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int t = threadIdx.x;
    const int el = t / ne, n = t - el * ne;
    const int e = lb * EPB + el;
    const bool active = lb < nb && el < EPB && e < K;
    for (int x = t; x < n1 * n1; x += blockDim.x) {
        sHat[x] = ops.hatD[(x / n1) * 8 + (x % n1)];
        sSharp[x] = ops.sharpD[(x / n1) * 8 + (x % n1)];
    }
    if (dta.reset && blockIdx.x == 0 && t < 2) dta.reset[t] = 0x7FF0000000000000ull;
    const long g = (long)e * ne + n;
    const int i = n % n1, j = (n / n1) % n1, k = n / nf;
    double q[NV], ja[9], Fi[3][NV], Fc[3][NV];
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = Q[v * S + g];
#pragma unroll
        for (int c = 0; c < 9; ++c) ja[c] = Ja[c * S + g];
        hx::euler_contravariant(ph, q, ja, Fi);
        double Fvc[3][NV];
        if (!ph.euler) {
            // BR1_ComputeInnerFluxes (EllipticBR1.f90:740-814) with mu, kappa of SD.f90:403-436
            double vg[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) vg[c] = vgrad[c * S + g];
            double mu, kappa;
            hx::laminar_mu_kappa(ph, q, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_viscosity(ph, edelta[e], q, vg);
                mu = mu + mt;
                kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / q[0];
            double Fv[NV][3];
            hx::viscous_flux(mu, kappa, q[1] * ir, q[2] * ir, q[3] * ir, vg, Fv);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    Fvc[d][v] = Fv[v][0] * ja[3 * d] + Fv[v][1] * ja[3 * d + 1] + Fv[v][2] * ja[3 * d + 2];
        } else {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < NV; ++v) Fvc[d][v] = 0.0;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) Fc[d][v] = SPLIT ? Fvc[d][v] : Fi[d][v] - Fvc[d][v];
        if (SPLIT) {
            // two-point node data (invRho, u, v, w, p as RiemannSolvers_NS.f90:2323-2329)
            const double ir = 1.0 / q[0];
            const double u = ir * q[1], v = ir * q[2], w = ir * q[3];
            const double p = ph.gm1 * (q[4] - 0.5 * (q[1] * u + q[2] * v + q[3] * w));
#pragma unroll
            for (int c = 0; c < NV; ++c) sP[el][c][n] = q[c];
            sP[el][5][n] = u; sP[el][6][n] = v; sP[el][7][n] = w; sP[el][8][n] = p; sP[el][9][n] = ir;
        }
    }
    double qd[NV], pa[5];
#pragma unroll
    for (int v = 0; v < NV; ++v) qd[v] = 0.0;
    // ScalarWeakIntegrals_StdVolumeGreen (DGIntegrals.f90:56-87) or
    // ScalarWeakIntegrals_SplitVolumeDivergence (:92-129): xi, then eta, then zeta
    if (SPLIT && active) {
#pragma unroll
        for (int c = 0; c < 5; ++c) pa[c] = sP[el][5 + c][n];
    }
    // Every thread reaches every barrier below; inactive threads (block padding) only skip
    // the work between them.
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (d) __syncthreads();     // previous direction's reads are done
        if (active) {
#pragma unroll
            for (int v = 0; v < NV; ++v) sF[el][v][n] = Fc[d][v];
            if (SPLIT) {
#pragma unroll
                for (int c = 0; c < 3; ++c) sJ[el][c][n] = ja[3 * d + c];
            }
        }
        __syncthreads();
        const int own = d == 0 ? i : (d == 1 ? j : k);
        const int line = d == 0 ? j + n1 * k : (d == 1 ? i + n1 * k : i + n1 * j);
        auto node_at = [&](int l) {
            return d == 0 ? l + n1 * (j + n1 * k) : (d == 1 ? i + n1 * (l + n1 * k) : i + n1 * (j + n1 * l));
        };
        if (!SPLIT) {
            if (active) {
#pragma unroll
                for (int l = 0; l < n1; ++l) {
                    const double hh = sHat[own * n1 + l];
                    const int m = node_at(l);
#pragma unroll
                    for (int v = 0; v < NV; ++v) qd[v] = qd[v] + hh * sF[el][v][m];
                }
            }
        } else {
            // SplitDG_ComputeSplitFormFluxes (HyperbolicSplitForm.f90:64-116): each unordered
            // pair {a, b} of the line is evaluated once, by node a for b = a + s (mod N+1),
            // s = 1..N/2 (and s = (N+1)/2 for a < (N+1)/2 when N is odd).
            if (active) {
#pragma unroll
                for (int s = 1; s <= (N + 1) / 2; ++s) {
                    if (2 * s == N + 1 && own >= s) continue;
                    const int l = (own + s) % n1;
                    const int m = node_at(l);
                    double qb[NV], pb[5], jb[3], fs[NV];
#pragma unroll
                    for (int c = 0; c < NV; ++c) qb[c] = sP[el][c][m];
#pragma unroll
                    for (int c = 0; c < 5; ++c) pb[c] = sP[el][5 + c][m];
#pragma unroll
                    for (int c = 0; c < 3; ++c) jb[c] = sJ[el][c][m];
                    hx::two_point_flux_t<FORM>(q, pa, &ja[3 * d], qb, pb, jb, fs);
                    const int lo = own < l ? own : l, hi = own < l ? l : own;
                    const int slot = lo * n1 - lo * (lo + 1) / 2 + (hi - lo - 1);
#pragma unroll
                    for (int v = 0; v < NV; ++v) sPair[el][line][slot][v] = fs[v];
                }
            }
            __syncthreads();
            // ScalarWeakIntegrals_SplitVolumeDivergence (DGIntegrals.f90:92-129), in l order;
            // the diagonal is the inviscid contravariant flux
            if (active) {
#pragma unroll
                for (int l = 0; l < n1; ++l) {
                    const double sh = sSharp[own * n1 + l], hh = sHat[own * n1 + l];
                    const int m = node_at(l);
                    if (l == own) {
#pragma unroll
                        for (int v = 0; v < NV; ++v) qd[v] = qd[v] + sh * Fi[d][v] + hh * sF[el][v][m];
                    } else {
                        const int lo = own < l ? own : l, hi = own < l ? l : own;
                        const int slot = lo * n1 - lo * (lo + 1) / 2 + (hi - lo - 1);
#pragma unroll
                        for (int v = 0; v < NV; ++v)
                            qd[v] = qd[v] + sh * sPair[el][line][slot][v] + hh * sF[el][v][m];
                    }
                }
            }
        }
    }
    double tc = INFINITY, tv = INFINITY;   // next step's time-step limits of this node
    if (active) {
        if (SPLIT) {
#pragma unroll
            for (int v = 0; v < NV; ++v) qd[v] = -qd[v];
        }
        // Surface integral (TimeDerivative_FacesContribution, SD.f90:1686-1701): for GL nodes
        // only boundary nodes receive b = 1/w_0 times the stored face flux (negated on the
        // right side).
        const bool onface = (i == 0) | (i == N) | (j == 0) | (j == N) | (k == 0) | (k == N);
        if (onface) {
            const double b0 = 1.0 / ops.w[0];
            double x[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
            auto surf = [&](int slot, int a, int b) {
                const int code = efn[((long)e * 6 + slot) * nf + a + n1 * b];
                const long fq = code >> 1;
                const double sg = (code & 1) ? -b0 : b0;
#pragma unroll
                for (int v = 0; v < NV; ++v) x[v] = x[v] + fstar[v * FS + fq] * sg;
            };
            if (i == 0) surf(5, j, k);
            if (i == N) surf(3, j, k);
            if (j == 0) surf(0, i, k);
            if (j == N) surf(1, i, k);
            if (k == 0) surf(2, i, j);
            if (k == N) surf(4, i, j);
#pragma unroll
            for (int v = 0; v < NV; ++v) qd[v] = qd[v] - x[v];
        }
        // QDot / J (SD.f90:490) as a product with 1/J: the rounding difference is one ulp of the
        // final derivative, taken after the cancellation of volume and surface terms
        const double jg = Jac[g];
        const double ijg = 1.0 / jg;
#pragma unroll
        for (int v = 0; v < NV; ++v) qd[v] = qd[v] * ijg;
        if (qdot_out) {
#pragma unroll
            for (int v = 0; v < NV; ++v) qdot_out[v * S + g] = qd[v];
        }
        if (update) {
            const double dt = dta.in ? fmin(bits_to_double(dta.in[0]), bits_to_double(dta.in[1])) : dta.host_dt;
            if (dta.time && lb == 0 && t == 0) dta.time[0] += dt;
            const double cdt = c_k * dt;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double gv = a_k * G[v * S + g] + qd[v];
                G[v * S + g] = gv;
                q[v] = q[v] + cdt * gv;
                Q[v * S + g] = q[v];
            }
            if (dta.out) {
                // MaxTimeStep of the new state (DGSEMClass.f90:870-1035 with
                // ComputeEigenvaluesForState, Physics_NS.f90:927-962)
                const double u = fabs(q[1] / q[0]), v = fabs(q[2] / q[0]), w = fabs(q[3] / q[0]);
                const double p = hx::pressure(ph, q);
                const double a = sqrt(ph.gamma * p / q[0]);
                const double e0 = u + a, e1 = v + a, e2 = w + a;
                const double lx = fabs(ja[0] * e0 + ja[1] * e1 + ja[2] * e2) * dta.dcsi;
                const double ly = fabs(ja[3] * e0 + ja[4] * e1 + ja[5] * e2) * dta.dcsi;
                const double lz = fabs(ja[6] * e0 + ja[7] * e1 + ja[8] * e2) * dta.dcsi;
                tc = dta.cfl * fabs(jg) / (lx + ly + lz);
                if (!ph.euler) {
                    const double mu = hx::sutherland(ph, ph.gammaM2 * p / q[0]);
                    const double d2 = dta.dcsi * dta.dcsi;
                    const double vx = mu * d2 * fabs(ja[0] + ja[1] + ja[2]);
                    const double vy = mu * d2 * fabs(ja[3] + ja[4] + ja[5]);
                    const double vz = mu * d2 * fabs(ja[6] + ja[7] + ja[8]);
                    tv = dta.dcfl * fabs(jg) / (vx + vy + vz);
                }
            }
        }
    }
    if (update && dta.out) {
        // every lane of the (full, 64-lane) wave takes part; inactive lanes hold +inf.  The
        // minimum over bit patterns is exact and order independent for non-negative doubles.
        for (int o = 32; o > 0; o >>= 1) {
            tc = fmin(tc, __shfl_xor(tc, o, 64));
            tv = fmin(tv, __shfl_xor(tv, o, 64));
        }
        if ((t & 63) == 0) {
            atomicMin(&dta.out[0], (unsigned long long)__double_as_longlong(tc));
            atomicMin(&dta.out[1], (unsigned long long)__double_as_longlong(tv));
        }
    }
}

}  // namespace hxk
