// Uniform-order kernels of the hx3d engine on Gauss nodes (standard form), one element per block.
// On Gauss nodes no node lies on a face: a face value is the Lagrange basis at the face's end of
// the normal direction applied along it (HexElement_ProlongSolutionToFaces /
// _ProlongGradientsToFaces, HexElementClass.f90:233-372), and the BR1 lift and the surface
// integral reach every node with b = v(+-1)/w (VectorWeakIntegrals_StdFace /
// ScalarWeakIntegrals_StdFace, DGIntegrals.f90:214-443).  Here the traces are formed where the
// element's values are in LDS and stored in element-face order, so the face kernel reads them as
// the Gauss-Lobatto path reads boundary nodes:
//   kg_prolong  Q traces of the resident state (after the state is set from outside);
//   kg_grad     local gradient, BR1 lift from the six faces' unStar, viscous state at the nodes,
//               gradient traces;
//   kg_face     Riemann + BR1 fluxes at the face nodes from the traces (both sides);
//   kg_volume   volume integral, surface integral of the six faces with b, 1/J, the RK update,
//               and the updated state's Q traces.
// Included by hx3d_engine.hip after stage_kernels.hpp.
#pragma once
#include <hip/hip_runtime.h>

#include "physics.hpp"
#include "stage_kernels.hpp"

namespace hxg {

using hx::NV;
using hx::Phys;
using hxk::ldg;
using hxk::stg;

// v(-1), v(+1) and b = v / w of the Gauss nodes of order N (NodalStorageClass.f90:257-273)
struct GaussOps {
    double vL[8], vR[8], bL[8], bR[8];
};

template <int N>
constexpr int gauss_tpb() { return (((N + 1) * (N + 1) * (N + 1) + 63) / 64) * 64; }

// node index of the element node at position l along slot s's axis and face coordinates (a, b)
// (slots 0..5 = EFRONT, EBACK, EBOTTOM, ERIGHT, ETOP, ELEFT; their axes y, y, z, x, z, x)
template <int N>
__device__ __forceinline__ int line_node(int s, int a, int b, int l)
{
    constexpr int n1 = N + 1;
    int i, j, k;
    if (s == 0 || s == 1) { i = a; j = l; k = b; }
    else if (s == 2 || s == 4) { i = a; j = b; k = l; }
    else { i = l; j = a; k = b; }
    return i + n1 * (j + n1 * k);
}

// The traces of NVAL node arrays sV[c*ne + n] (LDS) of element e on its six faces:
// out[c*ES + (6e + s) nf + a + (N+1) b] = acc, acc = acc + value(l) v(l) for l = 0..N along the
// slot's axis (the reference's order).  One work item is a line of nodes along an axis at face
// coordinates (a, b) and five of the values: it gives both ends of the line, the slot pair
// 5 / 3 (xi), 0 / 1 (eta) or 2 / 4 (zeta) with v(-1) and v(+1), from one read of each node
// value.  Consecutive threads take consecutive lines of one axis (consecutive face nodes).
template <int N, int NVAL>
__device__ __forceinline__ void write_traces(const double *sV, double *out, long ES, int e, const GaussOps &go,
                                             int t, int nthr)
{
    static_assert(NVAL % NV == 0, "values in groups of five");
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1, NL = 3 * nf, NG = NVAL / NV;
    for (int it = t; it < NL * NG; it += nthr) {
        const int grp = it / NL, L = it - grp * NL;
        const int ax = L / nf, r = L - ax * nf;
        const int sm = ax == 0 ? 5 : (ax == 1 ? 0 : 2), sp = ax == 0 ? 3 : (ax == 1 ? 1 : 4);
        const int n0 = line_node<N>(sm, r % n1, r / n1, 0), st = ax == 0 ? 1 : (ax == 1 ? n1 : nf);
        const long om = (long)(6 * e + sm) * nf + r, op = (long)(6 * e + sp) * nf + r;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int c = grp * NV + v;
            double am = 0.0, ap = 0.0;
#pragma unroll
            for (int l = 0; l < n1; ++l) {
                const double x = sV[c * ne + n0 + l * st];
                am = am + x * go.vL[l];
                ap = ap + x * go.vR[l];
            }
            out[(long)c * ES + om] = am;   // c varies across the wave: 64-bit addresses
            out[(long)c * ES + op] = ap;
        }
    }
}

// ------------------------------------------------------------------------------------
// kg_prolong: Q traces of the resident state (HexMesh_ProlongSolutionToFaces)
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(gauss_tpb<N>())
kg_prolong(const double *__restrict__ Q, double *__restrict__ Qtr, int K, long S, long ES, GaussOps go)
{
    constexpr int ne = (N + 1) * (N + 1) * (N + 1);
    __shared__ double sQ[NV * ne];
    const int e = hxk::xcd_block(blockIdx.x, gridDim.x), t = threadIdx.x;
    if (e >= K) return;   // uniform per block
    if (t < ne) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sQ[v * ne + t] = ldg(Q + v * S, (unsigned)e * ne + t);
    }
    __syncthreads();
    write_traces<N, NV>(sQ, Qtr, ES, e, go, t, blockDim.x);
}

// ------------------------------------------------------------------------------------
// kg_grad: local gradient (HexElement_ComputeLocalGradient), BR1 lift (BR1_ComputeElementInterface-
// Average, EllipticBR1.f90:571-627, and BR1_GradientFaceLoop, :531-569): unStar =
// 1/2 (Q_R - Q_L) |Ja| n at every face node of the element, from its own and its neighbour's
// trace, then every node takes the six faces' unStar with its weights b; the viscous state at
// the node for the volume kernel; the gradient traces for the face kernel
// (HexElement_ProlongGradientsToFaces).
//   nbx[x]  2 * (the neighbour's element face node matching x) + (1 if this element is the face's
//           right side)
//   egeo    normal and |Ja| of the face at x (the left element's face frame on both sides)
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(gauss_tpb<N>())
kg_grad(const double *__restrict__ Q, const double *__restrict__ Qtr, const double *__restrict__ Ja,
        const double *__restrict__ Jac, const int *__restrict__ nbx, const double *__restrict__ egeo,
        double *__restrict__ Gtr, double *__restrict__ vgrad, int K, long S, long ES, hxk::Ops ops, GaussOps go,
        Phys ph, hxk::DtLimits dtl)
{
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1;
    // one LDS space, in turn: the element's state [5][ne]; the faces' unStar [6 nf][15]; the
    // gradient [15][ne]
    constexpr int NW0 = 3 * NV * 6 * nf > 3 * NV * ne ? 3 * NV * 6 * nf : 3 * NV * ne;
    constexpr int NW = NW0 > NV * ne ? NW0 : NV * ne;
    __shared__ double sW[NW];
    __shared__ double sD[n1 * n1];
    double *sQ = sW;
    const int lb = hxk::xcd_block(blockIdx.x, gridDim.x), t = threadIdx.x;
    const int e = lb;
    if (e >= K) {   // a padding block
        if (dtl.part && t == 0) { dtl.part[2 * lb] = INFINITY; dtl.part[2 * lb + 1] = INFINITY; }
        return;
    }
    const bool active = t < ne;
    const int n = active ? t : 0;
    const int i = n % n1, j = (n / n1) % n1, k = n / nf;
    const unsigned g = (unsigned)e * ne + n;
    const double dop = hxk::operator_entry<n1>(ops.D, t);
    double q[NV], ja[9];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = ldg(Q + v * S, g);
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = ldg(Ja + c * S, g);
    const double jg = ldg(Jac, g);
    const double iJ = 1.0 / jg;
    if (t < n1 * n1) sD[t] = dop;
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sQ[v * ne + n] = q[v];
    }
    // unStar at the element's face nodes, the same value on both sides of a face.  The P passes
    // over the face nodes are unrolled so that every pass's index and geometry loads, then its
    // trace loads, are issued before the first is used (here, to arrive during the gradient).
    constexpr int TPB = gauss_tpb<N>(), NFN = 6 * nf, P = (NFN + TPB - 1) / TPB;
    int c2[P];
    double fg[P][4], own[P][NV], nb[P][NV];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int f = t + p * TPB;
        const unsigned x = (unsigned)(6 * e) * nf + (f < NFN ? f : 0);
        c2[p] = ldg(nbx, x);
#pragma unroll
        for (int c = 0; c < 4; ++c) fg[p][c] = ldg(egeo + c * ES, x);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int f = t + p * TPB;
        const unsigned x = (unsigned)(6 * e) * nf + (f < NFN ? f : 0), xn = (unsigned)c2[p] >> 1;
#pragma unroll
        for (int v = 0; v < NV; ++v) { own[p][v] = ldg(Qtr + v * ES, x); nb[p][v] = ldg(Qtr + v * ES, xn); }
    }
    __syncthreads();
    double tc = INFINITY, tv = INFINITY;
    double gr[3][NV];
    if (active) {
        double uxi[NV], ueta[NV], uzeta[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) { uxi[v] = 0.0; ueta[v] = 0.0; uzeta[v] = 0.0; }
#pragma unroll
        for (int l = 0; l < n1; ++l) {
            const double di = sD[i * n1 + l];
#pragma unroll
            for (int v = 0; v < NV; ++v) uxi[v] = uxi[v] + sQ[v * ne + l + n1 * (j + n1 * k)] * di;
        }
#pragma unroll
        for (int l = 0; l < n1; ++l) {
            const double dj = sD[j * n1 + l];
#pragma unroll
            for (int v = 0; v < NV; ++v) ueta[v] = ueta[v] + sQ[v * ne + i + n1 * (l + n1 * k)] * dj;
        }
#pragma unroll
        for (int l = 0; l < n1; ++l) {
            const double dk = sD[k * n1 + l];
#pragma unroll
            for (int v = 0; v < NV; ++v) uzeta[v] = uzeta[v] + sQ[v * ne + i + n1 * (j + n1 * l)] * dk;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gr[0][v] = (uxi[v] * ja[0] + ueta[v] * ja[3] + uzeta[v] * ja[6]) * iJ;
            gr[1][v] = (uxi[v] * ja[1] + ueta[v] * ja[4] + uzeta[v] * ja[7]) * iJ;
            gr[2][v] = (uxi[v] * ja[2] + ueta[v] * ja[5] + uzeta[v] * ja[8]) * iJ;
        }
    }
    __syncthreads();   // the state has been read: the space takes the faces' unStar
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int f = t + p * TPB;
        if (f < NFN) {
            const bool right = c2[p] & 1;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double uS = 0.5 * (right ? own[p][v] - nb[p][v] : nb[p][v] - own[p][v]) * fg[p][3];
#pragma unroll
                for (int d = 0; d < 3; ++d) sW[f * 3 * NV + d * NV + v] = uS * fg[p][d];
            }
        }
    }
    __syncthreads();
    if (active) {
        // the six faces in the order L, R (xi), FR, BK (eta), BOT, T (zeta): slots 5, 3, 0, 1, 2, 4
        const int fl = 5 * nf + j + n1 * k, fr = 3 * nf + j + n1 * k;
        const int ff = 0 * nf + i + n1 * k, fb = 1 * nf + i + n1 * k;
        const int fo = 2 * nf + i + n1 * j, ft = 4 * nf + i + n1 * j;
        const double bl = go.bL[i], br = go.bR[i], bf = go.bL[j], bb = go.bR[j], bo = go.bL[k], bt = go.bR[k];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int c = d * NV + v;
                double fi = sW[fl * 3 * NV + c] * bl;
                fi = fi + sW[fr * 3 * NV + c] * br;
                fi = fi + sW[ff * 3 * NV + c] * bf;
                fi = fi + sW[fb * 3 * NV + c] * bb;
                fi = fi + sW[fo * 3 * NV + c] * bo;
                fi = fi + sW[ft * 3 * NV + c] * bt;
                gr[d][v] = gr[d][v] + fi * iJ;
            }
        double vg[12];
        hx::viscous_state(ph, q, gr[0], gr[1], gr[2], vg);
#pragma unroll
        for (int c = 0; c < 12; ++c) stg(vgrad + c * S, g, vg[c]);
        if (ph.les) stg(vgrad + 12 * S, g, hx::les_factor(ph, q, gr[0], gr[1], gr[2]));
        if (dtl.part) hx::dt_limits(ph, q, ja, jg, dtl.dcsi, dtl.dcsi, dtl.dcsi, dtl.cfl, dtl.dcfl, tc, tv);
    }
    __syncthreads();   // the unStar has been read: the space takes the gradient
    if (active) {
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) sW[(d * NV + v) * ne + n] = gr[d][v];
    }
    __syncthreads();
    write_traces<N, 3 * NV>(sW, Gtr, ES, e, go, t, blockDim.x);
    if (dtl.part) hxk::block_min_pair(tc, tv, dtl.part, t, lb);
}

// ------------------------------------------------------------------------------------
// kg_face: one thread per face node (computeElementInterfaceFlux, SD.f90:1710-1799) from the
// two sides' Q and gradient traces at fnode[q].z / .w (element face nodes); the flux times |Ja|
// for both sides in element-face order, negated for the right side (the surface weights b are
// applied per node by kg_volume).
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(hxk::FACE_TPB)
kg_face(const double *__restrict__ Qtr, const double *__restrict__ Gtr, const double *__restrict__ fgeo,
        const int4 *__restrict__ fnode, const double *__restrict__ fdelta, double *__restrict__ fe, long FS, long ES,
        Phys ph, hxk::DtReduce dtr)
{
    constexpr int nf = (N + 1) * (N + 1);
    const int lb = hxk::xcd_block(blockIdx.x, gridDim.x);
    if (dtr.bits && blockIdx.x == 0) hxk::reduce_dt_partials(dtr.part, dtr.n, dtr.bits, threadIdx.x);
    const unsigned fq = (unsigned)lb * blockDim.x + threadIdx.x;
    if (fq >= FS) return;
    const int4 gi = ldg(fnode, fq);
    const unsigned xL = gi.z, xR = gi.w;
    double nrm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nrm[c] = ldg(fgeo + c * FS, fq);
    const double jac = ldg(fgeo + 9 * FS, fq);
    double qL[NV], qR[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { qL[v] = ldg(Qtr + v * ES, xL); qR[v] = ldg(Qtr + v * ES, xR); }
    double visc[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!ph.euler) {
        // BR1_RiemannSolver (EllipticBR1.f90:816-868): each side's viscous flux from its traces,
        // with the face viscosities of compute_viscosity_at_faces (SD.f90:1345-1398)
        const double fd = ldg(fdelta, fq / nf);
        double F[2][NV][3];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const double *qs = s ? qR : qL;
            const unsigned xs = s ? xR : xL;
            double gg[3 * NV];
#pragma unroll
            for (int c = 0; c < 3 * NV; ++c) gg[c] = ldg(Gtr + c * ES, xs);
            double vg[12];
            hx::viscous_state(ph, qs, gg, gg + NV, gg + 2 * NV, vg);
            double mu, kappa;
            hx::laminar_mu_kappa(ph, qs, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_mu(ph, fd, qs[0], hx::les_factor(ph, qs, gg, gg + NV, gg + 2 * NV));
                mu = mu + mt;
                kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / qs[0];
            hx::viscous_flux(mu, kappa, qs[1] * ir, qs[2] * ir, qs[3] * ir, vg, F[s]);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
            visc[v] = 0.5 * (F[0][v][0] + F[1][v][0]) * nrm[0] + 0.5 * (F[0][v][1] + F[1][v][1]) * nrm[1] +
                      0.5 * (F[0][v][2] + F[1][v][2]) * nrm[2];
    }
    double inv[NV];
    if (ph.riemann == 1) {
        double t1[3], t2[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) { t1[c] = ldg(fgeo + (3 + c) * FS, fq); t2[c] = ldg(fgeo + (6 + c) * FS, fq); }
        hx::lxf_flux(ph, qL, qR, nrm, t1, t2, inv);
    } else {
        hx::roe_flux(ph, qL, qR, nrm, inv);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double fl = (inv[v] - visc[v]) * jac;
        stg(fe + v * ES, xL, fl);
        stg(fe + v * ES, xR, -1.0 * fl);
    }
}

// ------------------------------------------------------------------------------------
// kg_volume: the standard-form volume integral with hatD (TimeDerivative_VolumetricContribution),
// the surface integral of the six faces' fluxes with the node's weights b, 1/J, the Williamson RK3
// update (TakeRK3Step) and, after an update, the new state's Q traces.
// ------------------------------------------------------------------------------------
template <int N>
__global__ void __launch_bounds__(gauss_tpb<N>())
kg_volume(double *__restrict__ Q, double *__restrict__ Qtr, const double *__restrict__ vgrad,
          const double *__restrict__ Ja, const double *__restrict__ Jac, const double *__restrict__ edelta,
          const double *__restrict__ fe, double *__restrict__ G, double *__restrict__ qdot_out, double a_k,
          double c_k, int update, hxk::DtArgs dta, int K, long S, long ES, hxk::Ops ops, GaussOps go, Phys ph)
{
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1, NFN = 6 * nf;
    __shared__ double sF[3 * NV * ne];   // contravariant fluxes [d][v][ne]; then the new state [v][ne]
    __shared__ double sE[NV * NFN];      // the element's face fluxes [v][6 nf] (element-face order)
    __shared__ double sHat[n1 * n1];
    const int e = hxk::xcd_block(blockIdx.x, gridDim.x), t = threadIdx.x;
    if (e >= K) return;   // uniform per block
    const bool active = t < ne;
    const int n = active ? t : 0;
    const int i = n % n1, j = (n / n1) % n1, k = n / nf;
    const unsigned g = (unsigned)e * ne + n;
    const double hop = hxk::operator_entry<n1>(ops.hatD, t);
    // the element's face fluxes, contiguous in element-face order, staged once (each is used by
    // the N+1 nodes of its line)
    for (int it = t; it < NV * NFN; it += blockDim.x) {
        const int v = it / NFN, f = it - v * NFN;
        sE[it] = fe[(long)v * ES + (long)(6 * e) * nf + f];
    }
    double q[NV], ja[9], vg[12];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = ldg(Q + v * S, g);
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = ldg(Ja + c * S, g);
    if (!ph.euler) {
#pragma unroll
        for (int c = 0; c < 12; ++c) vg[c] = ldg(vgrad + c * S, g);
    }
    const double dl = ldg(edelta, (unsigned)e);
    double lx = 0.0;
    if (ph.les) lx = ldg(vgrad + 12 * S, g);
    const double jg = ldg(Jac, g);
    double gk[NV];
    if (update && a_k != 0.0) {
#pragma unroll
        for (int v = 0; v < NV; ++v) gk[v] = ldg(G + v * S, g);
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) gk[v] = 0.0;
    }
    if (active) {
        double Fvc[3][NV];
        if (!ph.euler) {
            // BR1_ComputeInnerFluxes (EllipticBR1.f90:740-814) with mu, kappa of SD.f90:403-436
            double mu, kappa;
            hx::laminar_mu_kappa(ph, q, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_mu(ph, dl, q[0], lx);
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
        double Fi[3][NV];
        hx::euler_contravariant(ph, q, ja, Fi);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) sF[(d * NV + v) * ne + n] = Fi[d][v] - Fvc[d][v];
    }
    if (t < n1 * n1) sHat[t] = hop;
    __syncthreads();
    if (active) {
        // surface integral (ScalarWeakIntegrals_StdFace): L, R, FR, BK, BOT, T with the node's b
        double xs[NV];
        {
            const int xl = 5 * nf + j + n1 * k, xr = 3 * nf + j + n1 * k;
            const int xf = 0 * nf + i + n1 * k, xb = 1 * nf + i + n1 * k;
            const int xo = 2 * nf + i + n1 * j, xt = 4 * nf + i + n1 * j;
            const double bl = go.bL[i], br = go.bR[i], bf = go.bL[j], bb = go.bR[j], bo = go.bL[k], bt = go.bR[k];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double *fv = sE + v * NFN;
                double x = fv[xl] * bl;
                x = x + fv[xr] * br;
                x = x + fv[xf] * bf;
                x = x + fv[xb] * bb;
                x = x + fv[xo] * bo;
                x = x + fv[xt] * bt;
                xs[v] = x;
            }
        }
        double qd[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) qd[v] = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int own = d == 0 ? i : (d == 1 ? j : k);
            const int stride = d == 0 ? 1 : (d == 1 ? n1 : nf);
            const int base = n - own * stride;
#pragma unroll
            for (int l = 0; l < n1; ++l) {
                const double hh = sHat[own * n1 + l];
                const int m = base + l * stride;
#pragma unroll
                for (int v = 0; v < NV; ++v) qd[v] = qd[v] + hh * sF[(d * NV + v) * ne + m];
            }
        }
        const double ijg = 1.0 / jg;
#pragma unroll
        for (int v = 0; v < NV; ++v) qd[v] = (qd[v] - xs[v]) * ijg;
        if (qdot_out) {
#pragma unroll
            for (int v = 0; v < NV; ++v) stg(qdot_out + v * S, g, qd[v]);
        }
        if (update) {
            const double dt = fmin(hxk::bits_to_double(dta.in[0]), hxk::bits_to_double(dta.in[1]));
            if (dta.time && e == 0 && t == 0) dta.time[0] += dt;
            const double cdt = c_k * dt;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double gv = a_k * gk[v] + qd[v];
                stg(G + v * S, g, gv);
                q[v] = q[v] + cdt * gv;
                stg(Q + v * S, g, q[v]);
            }
            if (dta.nan && (isnan(q[0]) || isnan(q[1]) || isnan(q[2]) || isnan(q[3]) || isnan(q[4])))
                atomicOr(dta.nan, 1);
        }
    }
    if (update) {   // uniform: the new state's traces for the next stage
        __syncthreads();   // the fluxes have been read
        if (active) {
#pragma unroll
            for (int v = 0; v < NV; ++v) sF[v * ne + n] = q[v];
        }
        __syncthreads();
        write_traces<N, NV>(sF, Qtr, ES, e, go, t, blockDim.x);
    }
}

}  // namespace hxg
