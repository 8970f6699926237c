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

// Keeps a value opaque to the compiler until this point: the consumers of a value loaded early
// (to overlap its latency with other work) cannot be hoisted above it, so the wait for the
// load stays here.  No instruction is emitted.
__device__ __forceinline__ void hold(double &x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(x));
#else
    (void)x;
#endif
}

// Element i of an array through its (wave-uniform) base and a 32-bit offset: the access takes
// the base from scalar registers and one 32-bit vector offset, with no 64-bit address
// arithmetic per access.  Every array plane the kernels index holds fewer than 2^29 elements
// (checked at engine creation), so the byte offset fits 32 bits.
// The base is made opaque in scalar registers (global address space kept), so that the
// compiler does not fold the plane offset into a per-lane 64-bit address.
#if defined(__HIP_DEVICE_COMPILE__)
#define HX_GLOBAL __attribute__((address_space(1)))
#else
#define HX_GLOBAL
#endif
template <class T>
__device__ __forceinline__ HX_GLOBAL T *sbase(T *x)
{
    auto g = (HX_GLOBAL T *)x;
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(g));
#endif
    return g;
}
template <class T>
__device__ __forceinline__ T ldg(const T *x, unsigned i)
{
    return *(const HX_GLOBAL T *)((const HX_GLOBAL char *)sbase(x) + i * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stg(T *x, unsigned i, T v)
{
    *(HX_GLOBAL T *)((HX_GLOBAL char *)sbase(x) + i * (unsigned)sizeof(T)) = v;
}

// Entry t of an (N+1) x (N+1) operator (row stride 8 in Ops); threads t >= (N+1)^2 get 0.
// Loaded before the node data and stored to LDS after it has been issued, so the wait for it
// does not hold back the other loads.
template <int n1>
__device__ __forceinline__ double operator_entry(const double *src, int t)
{
    const int x = t < n1 * n1 ? t : 0;   // an unconditional load (see k_volume)
    return src[(x / n1) * 8 + (x % n1)];
}

__device__ __forceinline__ void block_min_pair(double a, double b, double *part, int t, int slot);

// Several independent cases of one order in one engine (batching): each case has its own
// physics constants, dt, time, divergence flag and active mask; element blocks never straddle
// cases, so a block's case and constants are uniform.  All pointers null: one case.
struct CaseConst {
    double gammaM2, mu, mu_to_kappa, les_c;   // the constants that differ between cases
};
struct Multi {
    const int4 *blocks;     // per logical block: first element, element count, case, first block of its case
    const CaseConst *cc;    // [cases]
    const int *active;      // [cases]: 0 = finished, its blocks and faces do no work
    const int *fcase;       // [faces]: the case of each face
    const int *pfirst;      // [cases + 1]: each case's range of logical blocks (its dt partials)
    int cases;
};

__device__ __forceinline__ void case_physics(Phys &ph, const CaseConst &c)
{
    ph.gammaM2 = c.gammaM2;
    ph.mu = c.mu;
    ph.mu_to_kappa = c.mu_to_kappa;
    ph.les_c = c.les_c;
}

// The elements of logical block lb: first element e0, count cnt (0: nothing to do), case cs,
// and whether the block is its case's first.  One case: EPB elements per block.
template <int EPB>
__device__ __forceinline__ void block_elements(const Multi &mc, int lb, int K, int &e0, int &cnt, int &cs, bool &first,
                                               Phys &ph)
{
    if (mc.blocks) {
        const int4 b = mc.blocks[lb];
        e0 = b.x;
        cs = b.z;
        cnt = mc.active[cs] ? b.y : 0;
        first = b.w != 0;
        case_physics(ph, mc.cc[cs]);
    } else {
        e0 = lb * EPB;
        cnt = e0 < K ? (K - e0 < EPB ? K - e0 : EPB) : 0;
        cs = 0;
        first = lb == 0;
    }
}

// MaxTimeStep of the state a stage starts from, accumulated by the stage's gradient kernel
// (part == null: not this stage): block minima of the convective and viscous limits to
// part[2*block + 0/1], reduced by the face kernel that follows (DtReduce).
struct DtLimits {
    double *part;
    double cfl, dcfl, dcsi;
};

// Averaging selectors of the templated split-form kernel (RiemannSolvers_NS.f90:229-285)
enum { AVG_STANDARD = 0, AVG_DUCROS = 1, AVG_KG = 2, AVG_PIROZZOLI = 3, AVG_MORINISHI = 4, FORM_STANDARD = -1 };

// Element face node x = (6e + slot) N1^2 + a + N1 b (a, b the slot's face coordinates); the
// face quantities the element kernels read are stored in this order, so an element reads its
// own contiguous range without an index lookup.
template <int N>
__device__ __forceinline__ unsigned efx(int e, int slot, int a, int b)
{
    return ((unsigned)e * 6 + slot) * ((N + 1) * (N + 1)) + a + (N + 1) * b;
}

// ------------------------------------------------------------------------------------
// k_grad: per element, local gradient (HexElement_ComputeLocalGradient,
// HexElementClass.f90:427-531) + BR1 lifting (EllipticBR1.f90:168-397), then the viscous
// state of ViscousFlux_STATE (velocity gradients, grad T) at every node.
//   nbr[x]        2 * (neighbour's node) + (1 if this element is the face's right side)
//   egeo[c*ES+x]  face normal (c = 0..2) and surface Jacobian |Ja| (c = 3) of the face
// ------------------------------------------------------------------------------------
#ifndef HX3D_GRAD_WAVES
#define HX3D_GRAD_WAVES 1
#endif
template <int N, int EPB>
__global__ void __launch_bounds__(((EPB * (N + 1) * (N + 1) * (N + 1) + 63) / 64) * 64)
__attribute__((amdgpu_waves_per_eu(HX3D_GRAD_WAVES)))
k_grad(const double *__restrict__ Q, const double *__restrict__ Ja, const double *__restrict__ Jac,
       const int *__restrict__ nbr, const double *__restrict__ egeo, double *__restrict__ vgrad,
       int K, long S, long ES, Ops ops, Phys ph, DtLimits dtl, Multi mc)
{
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1;
    __shared__ double sQ[EPB][NV][ne];
    __shared__ double sD[n1 * n1];
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int t = threadIdx.x;
    int e0, cnt, cs;
    bool first;
    block_elements<EPB>(mc, lb, K, e0, cnt, cs, first, ph);
    if (cnt == 0) {   // uniform: a padding block or a finished case
        if (dtl.part && t == 0) { dtl.part[2 * lb] = INFINITY; dtl.part[2 * lb + 1] = INFINITY; }
        return;
    }
    const int el = t / ne, n = t - el * ne;
    const int e = e0 + el;
    const bool active = el < cnt;
    const unsigned g = (unsigned)e * ne + n;
    const int i = n % n1, j = (n / n1) % n1, k = n / nf;
    // Every load is unconditional (padding threads read a clamped element and discard it; see
    // k_volume).  The node's (at most three) faces: x-face slot 5 / 3, y-face 0 / 1, z-face
    // 2 / 4; faces the node is not on get zero traces and zero geometry, so their lifting term
    // is an exact zero and the accumulation below is branch-free.
    const int ec = active ? e : e0 + cnt - 1;
    const unsigned gc = (unsigned)ec * ne + n;
    const bool onx = i == 0 || i == N, ony = j == 0 || j == N, onz = k == 0 || k == N;
    const unsigned fx = efx<N>(ec, i == 0 ? 5 : 3, j, k), fy = efx<N>(ec, j == 0 ? 0 : 1, i, k),
                   fz = efx<N>(ec, k == 0 ? 2 : 4, i, j);
    const double dop = operator_entry<n1>(ops.D, t);
    const int cx = ldg(nbr, fx), cy = ldg(nbr, fy), cz = ldg(nbr, fz);
    double q[NV], ja[9], gx[4], gy[4], gz[4];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = ldg(Q + v * S, gc);
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = ldg(Ja + c * S, gc);
    const double jg = ldg(Jac, gc);
    const double iJ = 1.0 / jg;   // the same bits as a stored 1/J
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double a = ldg(egeo + c * ES, fx), b = ldg(egeo + c * ES, fy), d = ldg(egeo + c * ES, fz);
        gx[c] = onx ? a : 0.0;
        gy[c] = ony ? b : 0.0;
        gz[c] = onz ? d : 0.0;
    }
    // neighbour traces, issued before the barrier so that they arrive during the derivative
    double qx[NV], qy[NV], qz[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double a = ldg(Q + v * S, (unsigned)cx >> 1), b = ldg(Q + v * S, (unsigned)cy >> 1),
                     d = ldg(Q + v * S, (unsigned)cz >> 1);
        qx[v] = onx ? a : 0.0;
        qy[v] = ony ? b : 0.0;
        qz[v] = onz ? d : 0.0;
    }
    if (t < n1 * n1) sD[t] = dop;
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sQ[el][v][n] = q[v];
    }
    __syncthreads();
    double tc = INFINITY, tv = INFINITY;   // MaxTimeStep limits of this node (stage 1)
    if (active) {
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
    // for both sides; the GL surface integral weights it by b = 1/w_0.  Faces in slot order
    // 5, 3, 0, 1, 2, 4 (ELEFT/ERIGHT, EFRONT/EBACK, EBOTTOM/ETOP).
    {
        const double b0 = 1.0 / ops.w[0];
        double fi[3][NV];
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int v = 0; v < NV; ++v) fi[d][v] = 0.0;
        auto lift = [&](const double *fg, int code, const double *qn) {
            const bool right = code & 1;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double uS = 0.5 * (right ? q[v] - qn[v] : qn[v] - q[v]) * fg[3];
                fi[0][v] = fi[0][v] + (uS * fg[0]) * b0;
                fi[1][v] = fi[1][v] + (uS * fg[1]) * b0;
                fi[2][v] = fi[2][v] + (uS * fg[2]) * b0;
            }
        };
        lift(gx, cx, qx);
        lift(gy, cy, qy);
        lift(gz, cz, qz);
        if (onx || ony || onz) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < NV; ++v) gr[d][v] = gr[d][v] + fi[d][v] * iJ;
        }
    }
    double vg[12];
    hx::viscous_state(ph, q, gr[0], gr[1], gr[2], vg);
#pragma unroll
    for (int c = 0; c < 12; ++c) stg(vgrad + c * S, g, vg[c]);
    // LES: the filter-width-free factor of the eddy viscosity, from the conservative gradients
    // (the face and volume kernels apply their own filter widths)
    if (ph.les) stg(vgrad + 12 * S, g, hx::les_factor(ph, q, gr[0], gr[1], gr[2]));
    if (dtl.part) hx::dt_limits(ph, q, ja, jg, dtl.dcsi, dtl.dcsi, dtl.dcsi, dtl.cfl, dtl.dcfl, tc, tv);
    }
    if (dtl.part) block_min_pair(tc, tv, dtl.part, t, lb);
}

// ------------------------------------------------------------------------------------
// k_face: one thread per face node (computeElementInterfaceFlux, SD.f90:1710-1799).
// fnode[q] = (left node, right node, left element face node, right element face node).
// The flux times the surface weight b = 1/w_0 is stored for both sides in element face order,
// negated for the right side (TimeDerivative_FacesContribution, SD.f90:1686-1701).
// ------------------------------------------------------------------------------------
__device__ __forceinline__ void reduce_dt_partials(const double *part, int n, unsigned long long *bits, int t);

// dtred: (partials, count, slot) of the step's dt limits, reduced by block 0 first (or null);
// with several cases, logical block c < cases reduces case c's partials into slot c
struct DtReduce {
    const double *part;
    int n;
    unsigned long long *bits;
};

#ifndef HX3D_FACE_WAVES
#define HX3D_FACE_WAVES 1
#endif
// threads (face nodes) per block of k_face (128 and 512 measured 12% and 18% slower at 16^3 x N=4)
constexpr int FACE_TPB = 256;
// T2X: the stored second tangent is n x t1 (vCross, MappedGeometry.f90:745, :807-817) bit for
// bit at every face node (checked on the host), so it is formed here instead of read.  LXF: the
// engine's Riemann solver is Lax-Friedrichs, the one that reads the tangents (Roe does not).
template <int N, bool T2X, bool LXF>
__global__ void __launch_bounds__(FACE_TPB)
__attribute__((amdgpu_waves_per_eu(HX3D_FACE_WAVES)))
k_face(const double *__restrict__ Q, const double *__restrict__ vgrad, const double *__restrict__ fgeo,
       const int4 *__restrict__ fnode, const double *__restrict__ fdelta, double *__restrict__ fe,
       long FS, long S, long ES, double b0, Phys ph, DtReduce dtr, Multi mc)
{
    constexpr int n1 = N + 1, nf = n1 * n1;
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    if (dtr.bits) {
        if (!mc.blocks) {
            if (blockIdx.x == 0) reduce_dt_partials(dtr.part, dtr.n, dtr.bits, threadIdx.x);
        } else if (lb < mc.cases) {   // case lb's dt (its blocks' partials)
            const int p0 = mc.pfirst[lb];
            reduce_dt_partials(dtr.part + 2 * p0, mc.pfirst[lb + 1] - p0, dtr.bits + 2 * lb, threadIdx.x);
        }
    }
    const unsigned fq = (unsigned)lb * blockDim.x + threadIdx.x;
    if (fq >= FS) return;
    if (mc.blocks) {
        const int cs = mc.fcase[fq / nf];
        if (!mc.active[cs]) return;
        case_physics(ph, mc.cc[cs]);
    }
    const int4 gi = ldg(fnode, fq);
    const unsigned gx = gi.x, gy = gi.y, gz = gi.z, gw = gi.w;
    // face frame: fgeo[c*FS + fq], c = n (0..2), t1 (3..5), t2 (6..8), |Ja| (9)
    double nrm[3], t1[3], t2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nrm[c] = ldg(fgeo + c * FS, fq);
    const double jac = ldg(fgeo + 9 * FS, fq);
    // all loads unconditional (see k_volume)
    double qL[NV], qR[NV], vgL[12], vgR[12];
#pragma unroll
    for (int v = 0; v < NV; ++v) { qL[v] = ldg(Q + v * S, gx); qR[v] = ldg(Q + v * S, gy); }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t1[c] = LXF ? ldg(fgeo + (3 + c) * FS, fq) : 0.0;
        t2[c] = (LXF && !T2X) ? ldg(fgeo + (6 + c) * FS, fq) : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 12; ++c) { vgL[c] = ldg(vgrad + c * S, gx); vgR[c] = ldg(vgrad + c * S, gy); }
    const double fd = ldg(fdelta, fq / nf);
    double xL = 0.0, xR = 0.0;   // LES factors of the two sides (the gradient kernel's plane 12)
    if (ph.les) { xL = ldg(vgrad + 12 * S, gx); xR = ldg(vgrad + 12 * S, gy); }
    double visc[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!ph.euler) {
        // BR1_RiemannSolver (EllipticBR1.f90:816-868) with the face viscosities of
        // compute_viscosity_at_faces (SpatialDiscretization.f90:1345-1398)
        double FL[NV][3], FR[NV][3];
        {
            double mu, kappa;
            hx::laminar_mu_kappa(ph, qL, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_mu(ph, fd, qL[0], xL);
                mu = mu + mt; kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / qL[0];
            hx::viscous_flux(mu, kappa, qL[1] * ir, qL[2] * ir, qL[3] * ir, vgL, FL);
        }
        {
            double mu, kappa;
            hx::laminar_mu_kappa(ph, qR, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_mu(ph, fd, qR[0], xR);
                mu = mu + mt; kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / qR[0];
            hx::viscous_flux(mu, kappa, qR[1] * ir, qR[2] * ir, qR[3] * ir, vgR, FR);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v)
            visc[v] = 0.5 * (FL[v][0] + FR[v][0]) * nrm[0] + 0.5 * (FL[v][1] + FR[v][1]) * nrm[1] +
                      0.5 * (FL[v][2] + FR[v][2]) * nrm[2];
    }
    double inv[NV];
    if (LXF) {
        if (T2X) {
            t2[0] = nrm[1] * t1[2] - t1[1] * nrm[2];
            t2[1] = nrm[2] * t1[0] - t1[2] * nrm[0];
            t2[2] = nrm[0] * t1[1] - t1[0] * nrm[1];
        }
        hx::lxf_flux(ph, qL, qR, nrm, t1, t2, inv);
    } else {
        hx::roe_flux(ph, qL, qR, nrm, inv);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        const double fb = ((inv[v] - visc[v]) * jac) * b0;
        stg(fe + v * ES, gz, fb);
        stg(fe + v * ES, gw, -fb);
    }
}

// ------------------------------------------------------------------------------------
// k_volume: per element, the volume integral (standard or split form), the surface integral
// of the stored face fluxes, 1/J, the Williamson RK3 register update, and (last stage) the
// MaxTimeStep reduction of the updated state for the next step.
// ------------------------------------------------------------------------------------
struct DtArgs {
    const unsigned long long *in;   // [2]: convective, viscous limit bit patterns of this step's dt
    double *time;                   // t += dt by one thread (or null)
    int *nan;                       // set to 1 when an updated state value is NaN (or null)
    DtLimits lim;                   // consumed by the stage's gradient kernel
    DtReduce red;                   // consumed by the stage's face kernel
};

__device__ __forceinline__ double bits_to_double(unsigned long long b) { return __longlong_as_double((long long)b); }

// Block minimum of (a, b) written to part[2*blockIdx.x + 0/1]; every thread of the block calls
// it (inactive threads with +inf).  The minimum of doubles is exact and independent of order,
// so the two-level reduction (here, then k_dt_reduce) gives the same bits as a serial one,
// without contended atomics.
__device__ __forceinline__ void block_min_pair(double a, double b, double *part, int t, int slot)
{
    __shared__ double sm[2][16];
    for (int o = 32; o > 0; o >>= 1) {
        a = fmin(a, __shfl_xor(a, o, 64));
        b = fmin(b, __shfl_xor(b, o, 64));
    }
    if ((t & 63) == 0) { sm[0][t >> 6] = a; sm[1][t >> 6] = b; }
    __syncthreads();
    if (t == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w) { a = fmin(a, sm[0][w]); b = fmin(b, sm[1][w]); }
        part[2 * slot] = a;
        part[2 * slot + 1] = b;
    }
}

// The minima of n (convective, viscous) pairs of part as the bit patterns of a dt slot; every
// thread of the calling block takes part.
__device__ __forceinline__ void reduce_dt_partials(const double *part, int n, unsigned long long *bits, int t)
{
    __shared__ double sm[2][16];
    double a = INFINITY, b = INFINITY;
    for (int x = t; x < n; x += blockDim.x) {
        a = fmin(a, part[2 * x]);
        b = fmin(b, part[2 * x + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        a = fmin(a, __shfl_xor(a, o, 64));
        b = fmin(b, __shfl_xor(b, o, 64));
    }
    if ((t & 63) == 0) { sm[0][t >> 6] = a; sm[1][t >> 6] = b; }
    __syncthreads();
    if (t == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; ++w) { a = fmin(a, sm[0][w]); b = fmin(b, sm[1][w]); }
        bits[0] = (unsigned long long)__double_as_longlong(a);
        bits[1] = (unsigned long long)__double_as_longlong(b);
    }
}

// Offset of the node pair (lo, hi), lo < hi, of an (N+1)-node line in the pair table (pairs
// lo-major).
__host__ __device__ constexpr int pair_slot(int n1, int lo, int hi)
{
    return lo * n1 - (lo * (lo + 1)) / 2 + (hi - lo - 1);
}

// One thread per node, EPB elements per block.  The contravariant fluxes of all three
// directions (split form: the viscous ones, whose mass entry is zero; standard form: inviscid
// minus viscous) go to LDS once.  Per direction each node accumulates, in l order,
// qd = (qd + sharpD F#) + hatD Fv (DGIntegrals.f90:92-129) or qd = qd + hatD F (:56-87).
// Split form, per direction: the node data and the metric rows (times metric_scale) go to LDS;
// each unordered node pair of a line is evaluated once, lower node index first as the reference
// evaluates both (i, l) and (l, i) (HyperbolicSplitForm.f90:64-116), by node lo of the pair
// (lo, lo + s mod N+1), s = 1..(N+1)/2 (for even N+1, s = (N+1)/2 only for lo < (N+1)/2); the
// pair fluxes then replace the node data in LDS and every node reads its N off-diagonal
// entries from there (at N = 5 each node evaluates its own row of N pair fluxes instead: ROW).
// The diagonal entry is the node's inviscid contravariant flux.
// Minimum waves per SIMD asked of the register allocator: 3 for the standard form, and for
// Kennedy-Gruber and Pirozzoli up to N = 5 (no spills in 168 VGPRs up to N = 4; at N = 5 the
// few spilled registers cost less than a third wave gains), else no constraint (at N = 6 the
// forced third wave measured 14% slower than two waves without spills).
#ifndef HX3D_VOL_WAVES
#define HX3D_VOL_WAVES 3
#endif
// Split form at N = 5 without the pair table (each node evaluates its own row); built with
// HX3D_ROW5=0, the pair table as at the other orders (for comparison).
#ifndef HX3D_ROW5
#define HX3D_ROW5 1
#endif
constexpr int volume_waves(int n, int form)
{
    return form == -1 ? 3 : (form == 2 || form == 3) && n <= 5 ? HX3D_VOL_WAVES : 1;
}

template <int N, int EPB, int FORM>
__global__ void __launch_bounds__(((EPB * (N + 1) * (N + 1) * (N + 1) + 63) / 64) * 64)
__attribute__((amdgpu_waves_per_eu(volume_waves(N, FORM))))
k_volume(double *__restrict__ Q, const double *__restrict__ vgrad, const double *__restrict__ Ja,
         const double *__restrict__ Jac, const double *__restrict__ edelta, const double *__restrict__ fe,
         double *__restrict__ G, double *__restrict__ qdot_out, double a_k, double c_k, int update,
         DtArgs dta, int K, long S, long ES, Ops ops, Phys ph, Multi mc)
{
    constexpr int n1 = N + 1, nf = n1 * n1, ne = nf * n1;
    constexpr bool SPLIT = FORM >= 0;
    constexpr int AVG = SPLIT ? FORM : 0;
    constexpr int ND = hx::node_data_size(AVG);
    constexpr bool SYM = hx::two_point_symmetric(AVG);
    constexpr int NF = SPLIT ? 4 : 5, V0 = SPLIT ? 1 : 0;
    // split form: node data [ND][ne] and metric rows [3][ne], then the pair fluxes
    // [pair][NV][line] of the current direction, in one space per element
    constexpr int NPAIR = n1 * (n1 - 1) / 2, NP = n1 / 2;
    // split form at N = 5: each node evaluates its own row of pair fluxes (ROW), without the
    // pair table
    constexpr bool ROW = SPLIT && N == 5 && HX3D_ROW5;
    constexpr int RN = (ND + 3) * ne, RP = ROW ? 0 : NPAIR * NV * nf;
    constexpr int RSZ = SPLIT ? (RN > RP ? RN : RP) : 1;
    __shared__ double sF[EPB][3][NF][ne];
    __shared__ double sR[EPB][RSZ];
    __shared__ double sHat[n1 * n1], sSharp[n1 * n1];
    __shared__ int sSlot[n1 * n1];   // pair table offset of (own, l); 0 on the diagonal
    const int lb = xcd_block(blockIdx.x, gridDim.x);
    const int t = threadIdx.x;
    int e0, cnt, cs;
    bool first;
    block_elements<EPB>(mc, lb, K, e0, cnt, cs, first, ph);
    if (cnt == 0) return;   // uniform: a padding block or a finished case
    const int el = t / ne, n = t - el * ne;
    const int e = e0 + el;
    const bool active = el < cnt;
    const int i = n % n1, j = (n / n1) % n1, k = n / nf;
    const bool onx = i == 0 || i == N, ony = j == 0 || j == N, onz = k == 0 || k == N;
    // ---- Loads are unconditional (padding threads read a clamped element and discard it): a
    // load inside a branch makes the compiler wait for the whole memory queue at the first use
    // of any earlier load.  The face values are issued first and summed once the node's own
    // data has arrived (loads complete in order); G and dt are issued at the last direction
    // and consumed after it.
    const int ec = active ? e : e0 + cnt - 1;
    const unsigned gc = (unsigned)ec * ne + n, gw = (unsigned)e * ne + n;
    const double hop = operator_entry<n1>(ops.hatD, t), sop = operator_entry<n1>(ops.sharpD, t);
    double sx[NV], sy[NV], sz[NV];
    {
        const unsigned xx = efx<N>(ec, i == 0 ? 5 : 3, j, k), xy = efx<N>(ec, j == 0 ? 0 : 1, i, k),
                       xz = efx<N>(ec, k == 0 ? 2 : 4, i, j);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double fx = ldg(fe + v * ES, xx), fy = ldg(fe + v * ES, xy), fz = ldg(fe + v * ES, xz);
            sx[v] = onx ? fx : 0.0;   // faces the node is not on add an exact zero
            sy[v] = ony ? fy : 0.0;
            sz[v] = onz ? fz : 0.0;
        }
    }
    double q[NV], ja[9], vg[12];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = ldg(Q + v * S, gc);
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = ldg(Ja + c * S, gc);
#pragma unroll
    for (int c = 0; c < 12; ++c) vg[c] = ldg(vgrad + c * S, gc);
    const double dl = ldg(edelta, (unsigned)ec);
    double lx = 0.0;   // LES factor (the gradient kernel's plane 12)
    if (ph.les) lx = ldg(vgrad + 12 * S, gc);
    const double jg = ldg(Jac, gc);
    double own[ND] = {};   // this node's two-point data (split form)
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
        if (SPLIT) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 1; v < NV; ++v) sF[el][d][v - V0][n] = Fvc[d][v];
            hx::node_data<AVG>(ph, q, own);
        } else {
            double Fi[3][NV];
            hx::euler_contravariant(ph, q, ja, Fi);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < NV; ++v) sF[el][d][v - V0][n] = Fi[d][v] - Fvc[d][v];
        }
    }
    // Surface integral (TimeDerivative_FacesContribution, SD.f90:1686-1701): the stored face
    // values already carry b = 1/w_0 and the side's sign; faces in slot order 5, 3, 0, 1, 2, 4
    double xs[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) xs[v] = ((0.0 + sx[v]) + sy[v]) + sz[v];
    if (t < n1 * n1) {
        sHat[t] = hop;
        sSharp[t] = sop;
        const int a = t / n1, b = t - a * n1;
        sSlot[t] = a == b ? 0 : pair_slot(n1, a < b ? a : b, a < b ? b : a) * NV * nf;
    }
    double qd[NV], gk[NV], dt0 = 0.0, dt1 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) qd[v] = 0.0;
    double pu, pv, pw, pp;   // the diagonal's primitives, formed once for the three directions
    hx::euler_primitives(ph, q, pu, pv, pw, pp);
    // xi, then eta, then zeta.  Every thread reaches every barrier; inactive threads (block
    // padding) only skip the work between them.
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int own_i = d == 0 ? i : (d == 1 ? j : k);
        const int stride = d == 0 ? 1 : (d == 1 ? n1 : nf);
        const int base = n - own_i * stride;
        const int line = d == 0 ? j + n1 * k : (d == 1 ? i + n1 * k : i + n1 * j);
        double jh[3];   // this node's metric row times metric_scale
#pragma unroll
        for (int c = 0; c < 3; ++c) jh[c] = hx::metric_scale(AVG) * ja[3 * d + c];
        if (SPLIT) {
            if (d) __syncthreads();     // the previous direction's pair fluxes are read
            if (active) {
#pragma unroll
                for (int c = 0; c < ND; ++c) sR[el][c * ne + n] = own[c];
#pragma unroll
                for (int c = 0; c < 3; ++c) sR[el][(ND + c) * ne + n] = jh[c];
            }
        }
        if (SPLIT || d == 0) __syncthreads();
        if (d == 2) {
            // The first RK stage (a_k = 0) does not read the register: the reference's
            // 0 G + QDot is QDot up to the sign of an exact zero, which no state value sees
            // (Q + c dt (+-0) = Q); every later stage reads it.
            if (a_k != 0.0) {
#pragma unroll
                for (int v = 0; v < NV; ++v) gk[v] = ldg(G + v * S, gc);
            } else {
#pragma unroll
                for (int v = 0; v < NV; ++v) gk[v] = 0.0;
            }
            dt0 = bits_to_double(dta.in[2 * cs]);
            dt1 = bits_to_double(dta.in[2 * cs + 1]);
        }
        if (!SPLIT) {
            if (active) {
#pragma unroll
                for (int l = 0; l < n1; ++l) {
                    const double hh = sHat[own_i * n1 + l];
                    const int m = base + l * stride;
#pragma unroll
                    for (int v = 0; v < NV; ++v) qd[v] = qd[v] + hh * sF[el][d][v][m];
                }
            }
            continue;
        }
        if (ROW) {
            // each node evaluates its whole row of two-point fluxes from the line's node data (a
            // pair twice, once by each of its nodes, with the same arithmetic) and accumulates it
            // in l order; the diagonal is the node's inviscid contravariant flux
            if (active) {
                double fi[NV];
                hx::euler_contravariant_row(q, pu, pv, pw, pp, &ja[3 * d], fi);
#pragma unroll 1
                for (int l = 0; l < n1; ++l) {
                    const double sh = sSharp[own_i * n1 + l], hh = sHat[own_i * n1 + l];
                    const int m = base + l * stride;
                    double b[ND], jb[3], tp[NV];
#pragma unroll
                    for (int c = 0; c < ND; ++c) b[c] = sR[el][c * ne + m];
#pragma unroll
                    for (int c = 0; c < 3; ++c) jb[c] = sR[el][(ND + c) * ne + m];
                    if (SYM) {
                        hx::two_point<AVG>(own, jh, b, jb, tp);
                    } else {
                        const bool lo = own_i < l;
                        double x1[ND], x2[ND], j1[3], j2[3];
#pragma unroll
                        for (int c = 0; c < ND; ++c) { x1[c] = lo ? own[c] : b[c]; x2[c] = lo ? b[c] : own[c]; }
#pragma unroll
                        for (int c = 0; c < 3; ++c) { j1[c] = lo ? jh[c] : jb[c]; j2[c] = lo ? jb[c] : jh[c]; }
                        hx::two_point<AVG>(x1, j1, x2, j2, tp);
                    }
                    qd[0] = qd[0] + sh * (l == own_i ? fi[0] : tp[0]);
#pragma unroll
                    for (int v = 1; v < NV; ++v)
                        qd[v] = qd[v] + sh * (l == own_i ? fi[v] : tp[v]) + hh * sF[el][d][v - V0][m];
                }
            }
            continue;
        }
        // this node's pairs (own, own + s mod N+1)
        double pf[NP][NV];
        if (active) {
#pragma unroll
            for (int s = 1; s <= NP; ++s) {
                const int l = own_i + s < n1 ? own_i + s : own_i + s - n1;
                const int m = base + l * stride;
                double b[ND], jb[3];
#pragma unroll
                for (int c = 0; c < ND; ++c) b[c] = sR[el][c * ne + m];
#pragma unroll
                for (int c = 0; c < 3; ++c) jb[c] = sR[el][(ND + c) * ne + m];
                if (SYM) {
                    hx::two_point<AVG>(own, jh, b, jb, pf[s - 1]);
                } else {
                    const bool lo = own_i < l;
                    double x1[ND], x2[ND], j1[3], j2[3];
#pragma unroll
                    for (int c = 0; c < ND; ++c) { x1[c] = lo ? own[c] : b[c]; x2[c] = lo ? b[c] : own[c]; }
#pragma unroll
                    for (int c = 0; c < 3; ++c) { j1[c] = lo ? jh[c] : jb[c]; j2[c] = lo ? jb[c] : jh[c]; }
                    hx::two_point<AVG>(x1, j1, x2, j2, pf[s - 1]);
                }
            }
        }
        __syncthreads();   // node data read: the space takes the pair fluxes
        if (active) {
#pragma unroll
            for (int s = 1; s <= NP; ++s) {
                if (2 * s == n1 && own_i >= s) continue;
                const int l = own_i + s < n1 ? own_i + s : own_i + s - n1;
                const int po = sSlot[own_i * n1 + l] + line;
#pragma unroll
                for (int v = 0; v < NV; ++v) sR[el][po + v * nf] = pf[s - 1][v];
            }
        }
        __syncthreads();
        if (active) {
            double fi[NV];
            hx::euler_contravariant_row(q, pu, pv, pw, pp, &ja[3 * d], fi);
            // unrolled by two at N = 4 (0.8% faster there, same registers); other orders
            // would spill
            constexpr int VU = N == 4 ? 2 : 1;
#pragma unroll VU
            for (int l = 0; l < n1; ++l) {
                const double sh = sSharp[own_i * n1 + l], hh = sHat[own_i * n1 + l];
                const int m = base + l * stride;
                const int po = sSlot[own_i * n1 + l] + line;
                qd[0] = qd[0] + sh * (l == own_i ? fi[0] : sR[el][po]);
#pragma unroll
                for (int v = 1; v < NV; ++v)
                    qd[v] = qd[v] + sh * (l == own_i ? fi[v] : sR[el][po + v * nf]) + hh * sF[el][d][v - V0][m];
            }
        }
    }
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (SPLIT) qd[v] = -qd[v];
            hold(xs[v]);
            qd[v] = qd[v] - xs[v];
        }
        // QDot / J (SD.f90:490) as a product with 1/J: the rounding difference is one ulp of the
        // final derivative, taken after the cancellation of volume and surface terms
        const double ijg = 1.0 / jg;
#pragma unroll
        for (int v = 0; v < NV; ++v) qd[v] = qd[v] * ijg;
        if (qdot_out) {
#pragma unroll
            for (int v = 0; v < NV; ++v) stg(qdot_out + v * S, gw, qd[v]);
        }
        if (update) {
            hold(dt0);
            hold(dt1);
            const double dt = fmin(dt0, dt1);
            if (dta.time && first && t == 0) dta.time[cs] += dt;
            const double cdt = c_k * dt;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                hold(gk[v]);
                const double gv = a_k * gk[v] + qd[v];
                stg(G + v * S, gw, gv);
                q[v] = q[v] + cdt * gv;
                stg(Q + v * S, gw, q[v]);
            }
            // checkForNan (ExplicitMethods.f90:1856-1905) of the updated state: a sticky flag,
            // written only where a NaN appears
            if (dta.nan && (isnan(q[0]) || isnan(q[1]) || isnan(q[2]) || isnan(q[3]) || isnan(q[4])))
                atomicOr(dta.nan + cs, 1);
        }
    }
}

}  // namespace hxk
