// Mixed-order stage kernels of the hx3d engine: elements with their own orders (Nx, Ny, Nz)
// and conforming faces between mismatched orders through the reference's mortar operators
// (Face_AdaptSolutionToFace / Face_AdaptGradientsToFace, FaceClass.f90:281-512, and
// Face_ProjectFluxToElements / Face_ProjectGradientFluxToElements, FaceClass.f90:597-700,
// 865-964).  Included by hx3d_engine.hip after stage_kernels.hpp.
//
// One stage = km_face_grad -> km_grad -> km_face -> km_volume:
//   km_face_grad  one wave per face (or two small faces of one shape, see face_lane): both Q
//                 traces to the face order, BR1 unStar there,
//                 projected back onto each side's element face (lift[15][ES]);
//   km_grad       element tiles: local gradient + lifting -> grad Q (15 values);
//   km_face       as km_face_grad: Q and grad Q traces to the face order, Riemann + BR1
//                 fluxes there, projected back (right side negated) and scaled by the
//                 receiving element's surface weight -> fe[5][ES];
//   km_volume     element tiles: volume integral, surface integral, 1/J, RK update and
//                 the next step's dt limits.
// The element kernels are persistent: one launch per stage sweeps every element, each block
// taking tiles (elements packed into at most 512 nodes, one thread per node) in turn, each
// thread reading what it works on from the tiles' node map.
// The same kernels run Gauss-node cases (GAUSS template flag, one order or per-element
// orders): the traces are the Lagrange basis at the face's end applied along the normal
// direction, and the BR1 lift and the surface integral reach every node with b = v(+-1)/w.
#pragma once
#include <hip/hip_runtime.h>

#include "physics.hpp"
#include "stage_kernels.hpp"

namespace hxm {

using hx::NV;
using hx::Phys;
using hxk::DtArgs;

constexpr int MAXN1 = 8;   // orders 1..7

// Operators of every order N (row i, column j at [N][i*8 + j]) and the mortar matrices
// Tset(N -> M) ((M+1) x (N+1), row-major with row length N+1), InterpolationMatrices.f90:42-108.
struct OpsAll {
    double D[MAXN1][64], hatD[MAXN1][64], sharpD[MAXN1][64];
    double b0[MAXN1];     // 1/w_0 = 1/w_N, the GL surface weight
    double dcsi[MAXN1];   // 1/|x_1 - x_0|, MaxTimeStep (DGSEMClass.f90:948-959)
    double T[MAXN1][MAXN1][64];
    // the Lagrange basis at -1 and +1 and b = v / w (NodalStorageClass.f90:257-273): the face
    // values and surface weights of Gauss nodes, on which no node lies on a face
    double vL[MAXN1][8], vR[MAXN1][8], bL[MAXN1][8], bR[MAXN1][8];
};

// Face record: elements, slots, orders, rotation and projection types (Face_LinkWithElements,
// FaceClass.f90:187-278).  NfL/NfR are each side's orders in the face orientation, Nf their max.
struct FaceRec {
    int e[2], slot[2];          // left, right element and element face slot (0..5)
    int Nel[2][2];              // element face orders (element face orientation)
    int Nfs[2][2];              // the same in the face orientation
    int Nf[2], rot, proj[2];    // proj: 0 none, 1 first index, 2 second, 3 both
    int foff;                   // first face node (face order) of this face
};

// A tile of elements for the persistent element kernels (host side): up to TILE_MAXE elements
// whose nodes, one thread each, fill at most TILE threads; thread t works on node t - off[s] of
// elem[s], off[s] <= t < off[s + 1] (threads from off[count] on are idle).  Which elements share
// a tile (the packing) is chosen on the host (hx3d_autotune); every packing gives the same
// results.  The kernels read the tiles through their node map (NodeMap).
// Tiles of two sizes, each swept by a launch of its own: 256 threads (two independent blocks
// per CU, whose loads and barriers do not wait for each other) for elements of at most 256
// nodes, 512 threads for larger elements (up to 8 x 8 x 8 nodes).  The host picks the sizes.
constexpr int TILE = 512, TILE_SMALL = 256, TILE_MAXE = 30;
struct Tile {
    int count;
    int elem[TILE_MAXE];
    int off[TILE_MAXE + 1];
};

// What thread t of a tile works on, built on the host with the tiles (one entry per thread of
// every tile, at [tile * T + t] for tiles of T threads): node g of element e, and code = i | j << 3 | k << 6 |
// Nx << 9 | Ny << 12 | Nz << 15 | (first thread of the element in the tile) << 18 | active << 27.
// One load gives a thread everything the tile table, the orders and the node offsets would
// give through three dependent ones; the element kernels load the next tile's entry while
// they work on the current one.  Idle threads (active = 0) carry node 0 of the tile's last
// element, so that their unconditional loads stay in bounds.
struct NodeMap {
    int g, e, code;
};
template <int T>
__device__ __forceinline__ NodeMap load_map(const NodeMap *__restrict__ map, int tile, int t)
{
    return map[(long)tile * T + t];
}

// element face slot 0..5 = EFRONT, EBACK, EBOTTOM, ERIGHT, ETOP, ELEFT: the fixed axis, its end,
// and the axes of the slot's face coordinates (a, b)
__device__ __constant__ int M_SLOT_AXIS[6] = {1, 1, 2, 0, 2, 0};
__device__ __constant__ int M_SLOT_PLUS[6] = {0, 1, 0, 1, 1, 0};
__device__ __constant__ int M_SLOT_U[6] = {0, 0, 0, 1, 0, 1};
__device__ __constant__ int M_SLOT_V[6] = {2, 2, 1, 2, 1, 2};

__device__ __forceinline__ void left_to_right_d(int i, int j, int Nx, int Ny, int rot, int &ii, int &jj)
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

// node of element e at element face coordinates (a, b) of slot s
__device__ __forceinline__ long slot_node(const int *N3, const long *off, int e, int s, int a, int b)
{
    const int *Ne = &N3[3 * e];
    int idx[3];
    idx[M_SLOT_AXIS[s]] = M_SLOT_PLUS[s] ? Ne[M_SLOT_AXIS[s]] : 0;
    idx[M_SLOT_U[s]] = a;
    idx[M_SLOT_V[s]] = b;
    return off[e] + idx[0] + (Ne[0] + 1) * (idx[1] + (Ne[1] + 1) * idx[2]);
}

// Face_AdaptSolutionToFace (side trace -> face order Nf) and Face_ProjectFluxToElements (face ->
// the side's element face) at face node (i, j).  Projection type pt: 0 the orders agree (a
// copy), 1 Tset along the first face index, 2 along the second, 3 along both, accumulated as
// acc += T0[i][m] * T1[j][l] * value in the loop order l (outer), m (inner) of the reference.

// The face's mortar matrices in LDS.  Along each face index d at most one side interpolates
// (the face order is the larger of the two sides' orders, so the other side's order is the face
// order): m[d] = Tset(N_s -> n_d) (side to face) and m[2 + d] = Tset(n_d -> N_s) (face to side)
// of that side serve both sides, one entry per lane.  The interpolation loops then read them at
// LDS latency instead of through the cache from the operator table.  (Reading them from the
// operator table instead measured 19% slower in km_face.)
struct Mort {
    const double *a0, *a1, *p0, *p1;   // side to face along the two face indices, face to side
};
__device__ __forceinline__ Mort mort_lds(const double (*m)[64]) { return Mort{m[0], m[1], m[2], m[3]}; }

__device__ __forceinline__ void stage_mortar(const OpsAll &ops, const FaceRec &r, double (*m)[64], int lane)
{
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        const int n = r.Nf[d], Ns = r.Nfs[0][d] < n ? r.Nfs[0][d] : r.Nfs[1][d];
        if (Ns < n) {
            m[d][lane] = ops.T[Ns][n][lane];
            m[2 + d][lane] = ops.T[n][Ns][lane];
        }
    }
}

// Face_AdaptSolutionToFace for K values at once (values k at tr[k*64 + q], the side's trace in
// the face orientation, q = m + (Ns0+1) l), with the matrices staged by stage_mortar: each
// source node's weight is formed once, and every value accumulates in the reference's order
// with the reference's products.
template <int K>
__device__ __forceinline__ void adapt_k(const Mort &m, const FaceRec &r, int side, const double *tr,
                                        int i, int j, double *out)
{
    const int N0 = r.Nfs[side][0], N1 = r.Nfs[side][1], n0 = r.Nf[0];
    const int pt = r.proj[side];
    if (pt == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = tr[k * 64 + i + (n0 + 1) * j];
        return;
    }
    const double *T0 = m.a0, *T1 = m.a1;
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
    if (pt == 1) {
        for (int l = 0; l <= N0; ++l) {
            const double w = T0[i * (N0 + 1) + l];
            const int q = l + (N0 + 1) * j;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += w * tr[k * 64 + q];
        }
    } else if (pt == 2) {
        for (int l = 0; l <= N1; ++l) {
            const double w = T1[j * (N1 + 1) + l];
            const int q = i + (N0 + 1) * l;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += w * tr[k * 64 + q];
        }
    } else {
        for (int l = 0; l <= N1; ++l)
            for (int mm = 0; mm <= N0; ++mm) {
                const double w = T0[i * (N0 + 1) + mm] * T1[j * (N1 + 1) + l];
                const int q = mm + (N0 + 1) * l;
#pragma unroll
                for (int k = 0; k < K; ++k) out[k] += w * tr[k * 64 + q];
            }
    }
}

// Face_ProjectFluxToElements for K values at once (values k at fl[k*64 + q], q = m + (Nf0+1) l)
template <int K>
__device__ __forceinline__ void project_k(const Mort &m, const FaceRec &r, int side, const double *fl,
                                          int i, int j, double *out)
{
    const int n0 = r.Nf[0], n1 = r.Nf[1];
    const int pt = r.proj[side];
    if (pt == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = fl[k * 64 + i + (n0 + 1) * j];
        return;
    }
    const double *T0 = m.p0, *T1 = m.p1;
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
    if (pt == 1) {
        for (int l = 0; l <= n0; ++l) {
            const double w = T0[i * (n0 + 1) + l];
            const int q = l + (n0 + 1) * j;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += w * fl[k * 64 + q];
        }
    } else if (pt == 2) {
        for (int l = 0; l <= n1; ++l) {
            const double w = T1[j * (n1 + 1) + l];
            const int q = i + (n0 + 1) * l;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += w * fl[k * 64 + q];
        }
    } else {
        for (int l = 0; l <= n1; ++l)
            for (int mm = 0; mm <= n0; ++mm) {
                const double w = T0[i * (n0 + 1) + mm] * T1[j * (n1 + 1) + l];
                const int q = mm + (n0 + 1) * l;
#pragma unroll
                for (int k = 0; k < K; ++k) out[k] += w * fl[k * 64 + q];
            }
    }
}

// The two-direction case (projection type 3) of adapt_k / project_k, factorised: along the
// first index into tmp (lanes (a, l) of the half-interpolated grid), then along the second.
// Each pass is the reference's interpolation along one index; the pair does (N0 + N1 + 2) K
// multiply-adds per node instead of (N0 + 1)(N1 + 1) K, in a different rounding order (within
// the 1e-12 parity bar).  Every lane of the wave calls it (it passes two block barriers);
// lanes without a node get 0.  src: K values at src[k*64 + m + (S0+1) l] of a grid of
// (S0+1) x (S1+1) nodes; out: the node (i, j) of the (D0+1) x (D1+1) grid; T0 (D0+1) x (S0+1),
// T1 (D1+1) x (S1+1); tmp: K x 64 doubles of LDS.
template <int K>
__device__ __forceinline__ void interp2_k(const double *T0, const double *T1, int S0, int S1, int D0, int D1,
                                          const double *src, double *tmp, int lane, int i, int j, bool on,
                                          double *out)
{
    if (lane < (D0 + 1) * (S1 + 1)) {
        const int a = lane % (D0 + 1), l = lane / (D0 + 1);
        double acc[K];
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0.0;
        for (int m = 0; m <= S0; ++m) {
            const double w = T0[a * (S0 + 1) + m];
            const int q = m + (S0 + 1) * l;
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] += w * src[k * 64 + q];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) tmp[k * 64 + lane] = acc[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = 0.0;
    if (on) {
        for (int l = 0; l <= S1; ++l) {
            const double w = T1[j * (S1 + 1) + l];
            const int q = i + (D0 + 1) * l;
#pragma unroll
            for (int k = 0; k < K; ++k) out[k] += w * tmp[k * 64 + q];
        }
    }
    __syncthreads();   // tmp read: the next call may overwrite it
}

// Face_AdaptSolutionToFace / Face_ProjectFluxToElements at node (i, j) of the destination grid,
// lanes without a node (on = false) giving 0.  FACT: faces interpolated along both indices take
// the factorised pair of passes.  That is for the values that only reach the viscous terms (the
// BR1 traces and lifting of the gradient, the gradients at the face): there the different
// rounding order stays far inside the parity bar.  The state traces of the Riemann solver and
// the projected face fluxes keep the reference's order: the surface integral cancels against the
// volume integral, and factorising those measured 1.7e-12 on the 20^3 p-anisotropic case.
// Uniform per face: every lane of the wave calls it.
template <int K, bool FACT>
__device__ __forceinline__ void adapt_face(const Mort &m, const FaceRec &r, int side, const double *tr,
                                           double *tmp, int lane, int i, int j, bool on, double *out)
{
    if (FACT && r.proj[side] == 3) {
        interp2_k<K>(m.a0, m.a1, r.Nfs[side][0], r.Nfs[side][1], r.Nf[0], r.Nf[1], tr, tmp, lane, i, j, on, out);
    } else if (on) {
        adapt_k<K>(m, r, side, tr, i, j, out);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = 0.0;
    }
}

template <int K, bool FACT>
__device__ __forceinline__ void project_face(const Mort &m, const FaceRec &r, int side, const double *fl,
                                             double *tmp, int lane, int i, int j, bool on, double *out)
{
    if (FACT && r.proj[side] == 3) {
        interp2_k<K>(m.p0, m.p1, r.Nf[0], r.Nf[1], r.Nfs[side][0], r.Nfs[side][1], fl, tmp, lane, i, j, on, out);
    } else if (on) {
        project_k<K>(m, r, side, fl, i, j, out);
    } else {
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = 0.0;
    }
}

// Stage the side's element trace of K node arrays src[c*S + g] into tr[c*64 + q] in the face
// orientation (the right side is rotated, Face_AdaptSolutionToFace).  Lane q of the face.  On
// Gauss-Lobatto nodes the trace is the face's row of nodes; on Gauss nodes (GAUSS) it is the
// Lagrange basis at the face's end of the normal direction applied along it,
// acc = acc + value * v(l) for l = 0..N (HexElement_ProlongSolutionToFaces,
// HexElementClass.f90:233-302, and _ProlongGradientsToFaces, :304-372).
template <int K, bool GAUSS = false>
__device__ __forceinline__ void gather_trace_k(const FaceRec &r, int side, const int *N3, const long *off,
                                               const double *src, long S, double *tr, int lane,
                                               const OpsAll *ops = nullptr)
{
    const int N0 = r.Nfs[side][0], N1 = r.Nfs[side][1];
    const int nq = (N0 + 1) * (N1 + 1);
    if (lane < nq) {
        const int i = lane % (N0 + 1), j = lane / (N0 + 1);
        int a = i, b = j;
        if (side == 1) left_to_right_d(i, j, N0, N1, r.rot, a, b);
        const long g = slot_node(N3, off, r.e[side], r.slot[side], a, b);
        if (!GAUSS) {
#pragma unroll
            for (int c = 0; c < K; ++c) tr[c * 64 + lane] = src[c * S + g];
        } else {
            // g is the node at the face's end of the normal axis (index 0 or N); walk the axis
            const int s = r.slot[side], ax = M_SLOT_AXIS[s];
            const int *Ne = &N3[3 * r.e[side]];
            const int Na = Ne[ax];
            const long stride = ax == 0 ? 1 : (ax == 1 ? (Ne[0] + 1) : (long)(Ne[0] + 1) * (Ne[1] + 1));
            const long g0 = M_SLOT_PLUS[s] ? g - (long)Na * stride : g;
            const double *vv = M_SLOT_PLUS[s] ? ops->vR[Na] : ops->vL[Na];
            double w[MAXN1];
#pragma unroll
            for (int l = 0; l < MAXN1; ++l) w[l] = l <= Na ? vv[l] : 0.0;
            // the line's loads all issued before the sum (a loop to Na would wait for each load in
            // turn); lanes past the line's end load nothing and add nothing
#pragma unroll
            for (int c = 0; c < K; ++c) {
                double x[MAXN1];
#pragma unroll
                for (int l = 0; l < MAXN1; ++l) x[l] = l <= Na ? src[c * S + g0 + l * stride] : 0.0;
                double acc = 0.0;
#pragma unroll
                for (int l = 0; l < MAXN1; ++l)
                    if (l <= Na) acc = acc + x[l] * w[l];
                tr[c * 64 + lane] = acc;
            }
        }
    }
}

// The faces of one wave of the face kernels: sched[w] = (face, -1) puts one face on all 64
// lanes; (face a, face b) puts two faces of one shape (the same orders on both sides, so the same
// projection types and mortar matrices, at most 32 face nodes) on the two half-waves, so that
// small faces do not leave most of a wave idle.  Each half works in its own 32 columns of the
// per-face LDS rows (every grid of such a face, and every half-interpolated grid, has at most
// 32 nodes); the mortar matrices are shared.  Both halves take the same branches at every
// barrier.
struct FaceLane {
    int f;     // the lane's face
    int q;     // the lane's index within its face
    int col;   // first LDS column of its face
};
__device__ __forceinline__ FaceLane face_lane(const int2 *sched, int w, int lane)
{
    const int2 s = sched[w];
    const bool two = s.y >= 0;
    const int h = two ? lane >> 5 : 0;
    return FaceLane{h ? s.y : s.x, two ? (lane & 31) : lane, 32 * h};
}

// element face node of the side's face-orientation node (i, j)
__device__ __forceinline__ long elem_face_node(const FaceRec &r, int side, const int *efoff, int i, int j)
{
    int a = i, b = j;
    if (side == 1) left_to_right_d(i, j, r.Nfs[1][0], r.Nfs[1][1], r.rot, a, b);
    return (long)efoff[6 * r.e[side] + r.slot[side]] + a + (r.Nel[side][0] + 1) * b;
}

// ------------------------------------------------------------------------------------
// km_face_grad: BR1 unStar = 1/2 (U_R - U_L) |Ja| n at the face nodes
// (BR1_ComputeElementInterfaceAverage, EllipticBR1.f90:571-627), projected onto both sides:
// lift[(3v + d)... laid out as lift[(d*5 + v)*ES + x].
// ------------------------------------------------------------------------------------
template <bool GAUSS>
__global__ void __launch_bounds__(64)
km_face_grad(const double *__restrict__ Q, const int *__restrict__ N3, const long *__restrict__ off,
             const int *__restrict__ efoff, const FaceRec *__restrict__ faces, const int2 *__restrict__ sched,
             const double *__restrict__ fgeo, const OpsAll *__restrict__ opsp, double *__restrict__ lift, int W,
             long S, long ES, long FSm)
{
    // LDS: both sides' Q traces, then one direction's unStar in their place; the mortar
    // matrices; the half-interpolated grid of a factorised interpolation (9.5 KB)
    __shared__ double sT[2][NV * 64];
    __shared__ double sM[4][64];
    __shared__ double sTmp[NV * 64];
    // waves in XCD-contiguous ranges (the grid is a multiple of 8): faces that share elements
    // mostly run on one XCD and find the elements' traces in its L2
    const int w = hxk::xcd_block(blockIdx.x, gridDim.x), lane = threadIdx.x;
    if (w >= W) return;   // uniform per block
    const FaceLane fl = face_lane(sched, w, lane);
    const int ql = fl.q;
    double *const tL = sT[0] + fl.col, *const tR = sT[1] + fl.col, *const tmp = sTmp + fl.col;
    double *const sU = tL;   // [NV][64], once both traces are at the face order
    const FaceRec r = faces[fl.f];
    stage_mortar(*opsp, r, sM, lane);
    const Mort M = mort_lds(sM);
    gather_trace_k<NV, GAUSS>(r, 0, N3, off, Q, S, tL, ql, opsp);
    gather_trace_k<NV, GAUSS>(r, 1, N3, off, Q, S, tR, ql, opsp);
    const int n0 = r.Nf[0], n1 = r.Nf[1], nq = (n0 + 1) * (n1 + 1);
    const bool on = ql < nq;
    const int i = on ? ql % (n0 + 1) : 0, j = on ? ql / (n0 + 1) : 0;
    const long fq = r.foff + (on ? ql : 0);
    double nrm[3];
    nrm[0] = fgeo[fq];
    nrm[1] = fgeo[FSm + fq];
    nrm[2] = fgeo[2 * FSm + fq];
    const double jac = fgeo[9 * FSm + fq];
    __syncthreads();
    double uS[NV], qL[NV], qR[NV];
    adapt_face<NV, true>(M, r, 0, tL, tmp, ql, i, j, on, qL);
    adapt_face<NV, true>(M, r, 1, tR, tmp, ql, i, j, on, qR);
#pragma unroll
    for (int v = 0; v < NV; ++v) uS[v] = 0.5 * (qR[v] - qL[v]) * jac;
    // unStar(v) n_d, one direction at a time, projected onto both sides: lift[(5d + v)*ES + x]
    for (int d = 0; d < 3; ++d) {
        __syncthreads();   // the traces (d = 0), the previous direction's unStar (d > 0) have been read
        if (on) {
#pragma unroll
            for (int v = 0; v < NV; ++v) sU[v * 64 + ql] = uS[v] * nrm[d];
        }
        __syncthreads();
        for (int side = 0; side < 2; ++side) {
            const int N0 = r.Nfs[side][0], N1 = r.Nfs[side][1];
            const bool son = ql < (N0 + 1) * (N1 + 1);
            const int si = son ? ql % (N0 + 1) : 0, sj = son ? ql / (N0 + 1) : 0;
            double lv[NV];
            project_face<NV, true>(M, r, side, sU, tmp, ql, si, sj, son, lv);
            if (son) {
                const long x = elem_face_node(r, side, efoff, si, sj);
#pragma unroll
                for (int v = 0; v < NV; ++v) lift[(d * NV + v) * ES + x] = lv[v];
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// km_grad: local gradient (HexElement_ComputeLocalGradient, HexElementClass.f90:427-531) and
// BR1 lifting (BR1_GradientFaceLoop, EllipticBR1.f90:531-569, with the weights b of
// VectorWeakIntegrals_StdFace): grad Q at every node, U[(5d + v)*S + g].
// ------------------------------------------------------------------------------------
template <bool GAUSS, int T>
__global__ void __launch_bounds__(T)
km_grad(const NodeMap *__restrict__ tmap, int ntiles, const double *__restrict__ Q,
        const int *__restrict__ efoff, const double *__restrict__ Ja,
        const double *__restrict__ Jac, const double *__restrict__ lift, const OpsAll *__restrict__ opsp,
        double *__restrict__ U, long S, long ES, Phys ph, hxk::DtLimits dtl, int pbase)
{
    HIP_DYNAMIC_SHARED(double, smem)
    const OpsAll &ops = *opsp;
    const int t = threadIdx.x;
    double *sD = smem;                   // D of every order, [MAXN1][64] (visible after the first barrier)
    double *sQall = smem + MAXN1 * 64;   // per element [NV][ne], at NV * off[slot]
    for (int x = t; x < MAXN1 * 64; x += blockDim.x) sD[x] = ops.D[x / 64][x % 64];
    double tc = INFINITY, tv = INFINITY;   // MaxTimeStep limits of this thread's nodes (stage 1)
    NodeMap nm = load_map<T>(tmap, blockIdx.x < ntiles ? blockIdx.x : 0, t);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int e = nm.e, code = nm.code;
    const long g = nm.g;   // loads are unconditional (see k_volume)
    const int i = code & 7, j = (code >> 3) & 7, k = (code >> 6) & 7;
    const int Nx = (code >> 9) & 7, Ny = (code >> 12) & 7, Nz = (code >> 15) & 7;
    const int nx = Nx + 1, ny = Ny + 1, nz = Nz + 1, ne = nx * ny * nz;
    const int first = (code >> 18) & 511;
    const bool active = (code >> 27) & 1;
    const int n = i + nx * (j + ny * k);
    double q[NV], ja[9];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = Q[v * S + g];
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = Ja[c * S + g];
    const double jg = Jac[g];
    if (tile + (int)gridDim.x < ntiles) nm = load_map<T>(tmap, tile + gridDim.x, t);   // the next tile's
    const double iJ = 1.0 / jg;   // the same bits as a stored 1/J
    const bool onx = i == 0 || i == Nx, ony = j == 0 || j == Ny, onz = k == 0 || k == Nz;
    const long xx = efoff[6 * e + (i == 0 ? 5 : 3)] + j + ny * k;
    const long xy = efoff[6 * e + (j == 0 ? 0 : 1)] + i + nx * k;
    const long xz = efoff[6 * e + (k == 0 ? 2 : 4)] + i + nx * j;
    double *sQ = sQall + NV * first;   // [NV][ne]
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) sQ[v * ne + n] = q[v];
    }
    __syncthreads();
    if (active) {
    const double *Dx = sD + 64 * Nx, *Dy = sD + 64 * Ny, *Dz = sD + 64 * Nz;
    double uxi[NV], ueta[NV], uzeta[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { uxi[v] = 0.0; ueta[v] = 0.0; uzeta[v] = 0.0; }
    for (int l = 0; l < nx; ++l) {
        const double di = Dx[i * 8 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v) uxi[v] = uxi[v] + sQ[v * ne + l + nx * (j + ny * k)] * di;
    }
    for (int l = 0; l < ny; ++l) {
        const double dj = Dy[j * 8 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v) ueta[v] = ueta[v] + sQ[v * ne + i + nx * (l + ny * k)] * dj;
    }
    for (int l = 0; l < nz; ++l) {
        const double dk = Dz[k * 8 + l];
#pragma unroll
        for (int v = 0; v < NV; ++v) uzeta[v] = uzeta[v] + sQ[v * ne + i + nx * (j + ny * l)] * dk;
    }
    if (GAUSS) {
        // every node takes the lifting values of all six faces with its surface weights b,
        // in the order L, R (xi), FR, BK (eta), BOT, T (zeta) of VectorWeakIntegrals_StdFace
        // (DGIntegrals.f90:365-443); slots 5, 3, 0, 1, 2, 4
        const int *ef = &efoff[6 * e];
        const long xl = ef[5] + j + ny * k, xr = ef[3] + j + ny * k;
        const long xf = ef[0] + i + nx * k, xb = ef[1] + i + nx * k;
        const long xo = ef[2] + i + nx * j, xt = ef[4] + i + nx * j;
        const double bl = ops.bL[Nx][i], br = ops.bR[Nx][i], bf = ops.bL[Ny][j], bb = ops.bR[Ny][j];
        const double bo = ops.bL[Nz][k], bt = ops.bR[Nz][k];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double gr[3];
            gr[0] = (uxi[v] * ja[0] + ueta[v] * ja[3] + uzeta[v] * ja[6]) * iJ;
            gr[1] = (uxi[v] * ja[1] + ueta[v] * ja[4] + uzeta[v] * ja[7]) * iJ;
            gr[2] = (uxi[v] * ja[2] + ueta[v] * ja[5] + uzeta[v] * ja[8]) * iJ;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double *lc = lift + (long)(d * NV + v) * ES;
                double fi = lc[xl] * bl;
                fi = fi + lc[xr] * br;
                fi = fi + lc[xf] * bf;
                fi = fi + lc[xb] * bb;
                fi = fi + lc[xo] * bo;
                fi = fi + lc[xt] * bt;
                U[(d * NV + v) * S + g] = gr[d] + fi * iJ;
            }
        }
    } else {
        const double bx = ops.b0[Nx], by = ops.b0[Ny], bz = ops.b0[Nz];
        // lifting values of the node's faces (slot order 5/3, 0/1, 2/4); absent faces give 0.
        // Loaded after the derivatives, so that they do not hold registers through them.
        double hx_[3 * NV], hy_[3 * NV], hz_[3 * NV];
#pragma unroll
        for (int c = 0; c < 3 * NV; ++c) {
            const double a = lift[c * ES + xx], b = lift[c * ES + xy], d = lift[c * ES + xz];
            hx_[c] = onx ? a : 0.0;
            hy_[c] = ony ? b : 0.0;
            hz_[c] = onz ? d : 0.0;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double gr[3];
            gr[0] = (uxi[v] * ja[0] + ueta[v] * ja[3] + uzeta[v] * ja[6]) * iJ;
            gr[1] = (uxi[v] * ja[1] + ueta[v] * ja[4] + uzeta[v] * ja[7]) * iJ;
            gr[2] = (uxi[v] * ja[2] + ueta[v] * ja[5] + uzeta[v] * ja[8]) * iJ;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double ax = onx ? hx_[d * NV + v] * bx : 0.0;
                const double ay = ony ? hy_[d * NV + v] * by : 0.0;
                const double az = onz ? hz_[d * NV + v] * bz : 0.0;
                const double fi = ((0.0 + ax) + ay) + az;
                U[(d * NV + v) * S + g] = (onx || ony || onz) ? gr[d] + fi * iJ : gr[d];
            }
        }
    }
    if (dtl.part) {
        double c0, c1;
        hx::dt_limits(ph, q, ja, jg, ops.dcsi[Nx], ops.dcsi[Ny], ops.dcsi[Nz], dtl.cfl, dtl.dcfl, c0, c1);
        tc = fmin(tc, c0);
        tv = fmin(tv, c1);
    }
    }
    __syncthreads();   // the tile's sQ has been read: the next tile overwrites it
    }
    if (dtl.part) hxk::block_min_pair(tc, tv, dtl.part, t, pbase + blockIdx.x);
}

// ------------------------------------------------------------------------------------
// km_face: computeElementInterfaceFlux (SD.f90:1710-1799) at the face order with the face
// viscosities of compute_viscosity_at_faces (SD.f90:1345-1398), projected back
// (FaceClass.f90:597-700) and scaled by the receiving element's surface weight:
// fe[v*ES + x] = (+/-) projected flux * b.
// ------------------------------------------------------------------------------------
template <bool GAUSS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
km_face(const double *__restrict__ Q, const double *__restrict__ U, const int *__restrict__ N3,
        const long *__restrict__ off, const int *__restrict__ efoff, const FaceRec *__restrict__ faces,
        const int2 *__restrict__ sched, const double *__restrict__ fgeo, const double *__restrict__ fdelta,
        const OpsAll *__restrict__ opsp, double *__restrict__ fe, int W, long S, long ES, long FSm, Phys ph,
        hxk::DtReduce dtr)
{
    // LDS: both sides' Q traces (rows 0-4, 5-9), then one side's grad Q trace at a time, then
    // the face flux, in one space; the mortar matrices; the half-interpolated grid of a
    // factorised interpolation (12.25 KB: three waves per SIMD)
    __shared__ double sR[3 * NV * 64];
    __shared__ double sM[4][64];
    __shared__ double sTmp[NV * 64];
    // waves in XCD-contiguous ranges (the grid is a multiple of 8): faces that share elements
    // mostly run on one XCD and find the elements' traces in its L2
    const int w = hxk::xcd_block(blockIdx.x, gridDim.x), lane = threadIdx.x;
    if (dtr.bits && w == 0) hxk::reduce_dt_partials(dtr.part, dtr.n, dtr.bits, lane);
    if (w >= W) return;   // uniform per block
    const FaceLane fl = face_lane(sched, w, lane);
    const int f = fl.f, ql = fl.q;
    double *const qLt = sR + fl.col, *const qRt = sR + NV * 64 + fl.col, *const gT = sR + fl.col,
                  *const tmp = sTmp + fl.col;
    const FaceRec r = faces[f];
    stage_mortar(*opsp, r, sM, lane);
    const Mort M = mort_lds(sM);
    gather_trace_k<NV, GAUSS>(r, 0, N3, off, Q, S, qLt, ql, opsp);
    gather_trace_k<NV, GAUSS>(r, 1, N3, off, Q, S, qRt, ql, opsp);
    __syncthreads();
    const int n0 = r.Nf[0], n1 = r.Nf[1], nq = (n0 + 1) * (n1 + 1);
    // lanes past the face's nodes work on node (0, 0) and discard the result (every index stays
    // inside the face's arrays)
    const bool on = ql < nq;
    const int i = on ? ql % (n0 + 1) : 0, j = on ? ql / (n0 + 1) : 0;
    const long fq = r.foff + (on ? ql : 0);
    double nrm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) nrm[c] = fgeo[c * FSm + fq];
    const double jac = fgeo[9 * FSm + fq];
    double qs[2][NV];
    adapt_face<NV, false>(M, r, 0, qLt, tmp, ql, i, j, on, qs[0]);
    adapt_face<NV, false>(M, r, 1, qRt, tmp, ql, i, j, on, qs[1]);
    __syncthreads();   // the Q traces are read: the space takes a grad Q trace or the face flux
    double visc[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (!ph.euler) {
        // BR1_RiemannSolver (EllipticBR1.f90:816-868): each side's viscous flux from its
        // gradients at the face order, the sides one after the other through the trace space; the
        // gradients (15 values) are interpolated five at a time (the size of sTmp)
        auto side_grad = [&](int s, double *g) {
#pragma unroll
            for (int c = 0; c < 3; ++c) adapt_face<NV, true>(M, r, s, gT + c * NV * 64, tmp, ql, i, j, on, g + c * NV);
        };
        auto side_flux = [&](int s, const double *g, double (*Fs)[3]) {
            double vg[12];
            hx::viscous_state(ph, qs[s], g, g + NV, g + 2 * NV, vg);
            double mu, kappa;
            hx::laminar_mu_kappa(ph, qs[s], mu, kappa);
            if (ph.les) {
                const double mt = hx::les_mu(ph, fdelta[f], qs[s][0], hx::les_factor(ph, qs[s], g, g + NV, g + 2 * NV));
                mu = mu + mt;
                kappa = kappa + mt * ph.mu_to_kappa;
            }
            const double ir = 1.0 / qs[s][0];
            hx::viscous_flux(mu, kappa, qs[s][1] * ir, qs[s][2] * ir, qs[s][3] * ir, vg, Fs);
        };
        double Fs[2][NV][3], g[3 * NV];
        gather_trace_k<3 * NV, GAUSS>(r, 0, N3, off, U, S, gT, ql, opsp);
        __syncthreads();
        side_grad(0, g);
        if (on) side_flux(0, g, Fs[0]);
        __syncthreads();   // side 0's gradients consumed
        gather_trace_k<3 * NV, GAUSS>(r, 1, N3, off, U, S, gT, ql, opsp);
        __syncthreads();
        side_grad(1, g);
        if (on) {
            side_flux(1, g, Fs[1]);
#pragma unroll
            for (int v = 0; v < NV; ++v)
                visc[v] = 0.5 * (Fs[0][v][0] + Fs[1][v][0]) * nrm[0] + 0.5 * (Fs[0][v][1] + Fs[1][v][1]) * nrm[1] +
                          0.5 * (Fs[0][v][2] + Fs[1][v][2]) * nrm[2];
        }
        __syncthreads();   // the space takes the face flux next
    }
    double *const sF = gT;   // face flux [NV][64]
    if (on) {
        double inv[NV];
        if (ph.riemann == 1) {
            double t1[3], t2[3];   // the tangents: the Lax-Friedrichs flux alone reads them
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                t1[c] = fgeo[(3 + c) * FSm + fq];
                t2[c] = fgeo[(6 + c) * FSm + fq];
            }
            hx::lxf_flux(ph, qs[0], qs[1], nrm, t1, t2, inv);
        } else {
            hx::roe_flux(ph, qs[0], qs[1], nrm, inv);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) sF[v * 64 + ql] = (inv[v] - visc[v]) * jac;
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        const int N0 = r.Nfs[side][0], N1 = r.Nfs[side][1];
        const bool son = ql < (N0 + 1) * (N1 + 1);
        const int si = son ? ql % (N0 + 1) : 0, sj = son ? ql / (N0 + 1) : 0;
        double a[NV];
        project_face<NV, false>(M, r, side, sF, tmp, ql, si, sj, son, a);
        if (son) {
            const long x = elem_face_node(r, side, efoff, si, sj);
            if (GAUSS) {
                // the surface weights b(i) of every node are applied by the volume kernel
#pragma unroll
                for (int v = 0; v < NV; ++v) fe[v * ES + x] = side == 0 ? a[v] : -1.0 * a[v];
            } else {
                const int e = r.e[side], s = r.slot[side];
                const double b = opsp->b0[N3[3 * e + M_SLOT_AXIS[s]]];
#pragma unroll
                for (int v = 0; v < NV; ++v) fe[v * ES + x] = (side == 0 ? a[v] : -1.0 * a[v]) * b;
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// km_volume: as hxk::k_volume with the element's own orders per direction, the viscous
// state formed from grad Q (ViscousFlux_STATE), and MaxTimeStep's 1/|x_1 - x_0| per direction.
// LDS (dynamic): hatD and sharpD of every order [2][MAXN1][64], then per element of the tile,
// at per_node * off[slot]: sF[3][NF][ne], sN[ND][ne], sJ[3][ne].
// ------------------------------------------------------------------------------------
template <int FORM>
constexpr int mvolume_per_node()
{
    return 3 * (FORM >= 0 ? 4 : 5) + (FORM >= 0 ? hx::node_data_size(FORM) + 3 : 0);
}

template <int FORM, bool GAUSS, int T>
__global__ void __launch_bounds__(T)
km_volume(const NodeMap *__restrict__ tmap, int ntiles, double *__restrict__ Q, const double *__restrict__ U,
          const int *__restrict__ efoff,
          const double *__restrict__ Ja, const double *__restrict__ Jac, const double *__restrict__ edelta,
          const double *__restrict__ fe, const OpsAll *__restrict__ opsp, double *__restrict__ G,
          double *__restrict__ qdot_out, double a_k, double c_k, int update, DtArgs dta, long S, long ES,
          Phys ph)
{
    constexpr bool SPLIT = FORM >= 0;
    constexpr int AVG = SPLIT ? FORM : 0;
    constexpr int ND = hx::node_data_size(AVG);
    constexpr bool SYM = hx::two_point_symmetric(AVG);
    constexpr int NF = SPLIT ? 4 : 5, V0 = SPLIT ? 1 : 0;
    constexpr int PER_NODE = mvolume_per_node<FORM>();
    HIP_DYNAMIC_SHARED(double, smem)
    const OpsAll &ops = *opsp;
    const int t = threadIdx.x;
    // hatD and sharpD of every order (read at LDS latency in the contractions below; visible
    // after the first barrier of the direction loop)
    double *sOp = smem;
    for (int x = t; x < 2 * MAXN1 * 64; x += blockDim.x)
        sOp[x] = x < MAXN1 * 64 ? ops.hatD[x / 64][x % 64] : ops.sharpD[(x / 64) - MAXN1][x % 64];
    double *chunk0 = smem + 2 * MAXN1 * 64;
    NodeMap nm = load_map<T>(tmap, blockIdx.x < ntiles ? blockIdx.x : 0, t);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int e = nm.e, code = nm.code;
    const long g = nm.g;
    const int i = code & 7, j = (code >> 3) & 7, k = (code >> 6) & 7;
    const int Nx = (code >> 9) & 7, Ny = (code >> 12) & 7, Nz = (code >> 15) & 7;
    const int nx = Nx + 1, ny = Ny + 1, nz = Nz + 1, ne = nx * ny * nz;
    const int first = (code >> 18) & 511;
    const bool active = (code >> 27) & 1;
    const int n = i + nx * (j + ny * k);
    const bool onx = i == 0 || i == Nx, ony = j == 0 || j == Ny, onz = k == 0 || k == Nz;
    double *sF = chunk0 + PER_NODE * first;   // [3][NF][ne]
    double *sN = sF + 3 * NF * ne;                  // [ND][ne]
    double *sJ = sN + (SPLIT ? ND * ne : 0);        // [3][ne]
    // loads, unconditional (see hxk::k_volume)
    double sx[NV], sy[NV], sz[NV];
    if (GAUSS) {
        // surface integral on every node: the face fluxes of all six faces with the node's
        // surface weights, L, R, FR, BK, BOT, T (ScalarWeakIntegrals_StdFace, DGIntegrals.f90:
        // 214-273), accumulated in sx (sy, sz unused)
        const int *ef = &efoff[6 * e];
        const long xl = ef[5] + j + ny * k, xr = ef[3] + j + ny * k;
        const long xf = ef[0] + i + nx * k, xb = ef[1] + i + nx * k;
        const long xo = ef[2] + i + nx * j, xt = ef[4] + i + nx * j;
        const double bl = ops.bL[Nx][i], br = ops.bR[Nx][i], bf = ops.bL[Ny][j], bb = ops.bR[Ny][j];
        const double bo = ops.bL[Nz][k], bt = ops.bR[Nz][k];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double *fv = fe + (long)v * ES;
            double x = fv[xl] * bl;
            x = x + fv[xr] * br;
            x = x + fv[xf] * bf;
            x = x + fv[xb] * bb;
            x = x + fv[xo] * bo;
            x = x + fv[xt] * bt;
            sx[v] = x;
            sy[v] = 0.0;
            sz[v] = 0.0;
        }
    } else {
        const long xx = efoff[6 * e + (i == 0 ? 5 : 3)] + j + ny * k;
        const long xy = efoff[6 * e + (j == 0 ? 0 : 1)] + i + nx * k;
        const long xz = efoff[6 * e + (k == 0 ? 2 : 4)] + i + nx * j;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double a = fe[v * ES + xx], b = fe[v * ES + xy], c = fe[v * ES + xz];
            sx[v] = onx ? a : 0.0;
            sy[v] = ony ? b : 0.0;
            sz[v] = onz ? c : 0.0;
        }
    }
    double q[NV], ja[9], gq[3][NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) q[v] = Q[v * S + g];
#pragma unroll
    for (int c = 0; c < 9; ++c) ja[c] = Ja[c * S + g];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int v = 0; v < NV; ++v) gq[d][v] = U[(d * NV + v) * S + g];
    const double dl = edelta[e];
    const double jg = Jac[g];
    if (tile + (int)gridDim.x < ntiles) nm = load_map<T>(tmap, tile + gridDim.x, t);   // the next tile's
    if (active) {
        double Fvc[3][NV];
        if (!ph.euler) {
            double vg[12];
            hx::viscous_state(ph, q, gq[0], gq[1], gq[2], vg);
            double mu, kappa;
            hx::laminar_mu_kappa(ph, q, mu, kappa);
            if (ph.les) {
                const double mt = hx::les_mu(ph, dl, q[0], hx::les_factor(ph, q, gq[0], gq[1], gq[2]));
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
                for (int v = 1; v < NV; ++v) sF[(d * NF + v - V0) * ne + n] = Fvc[d][v];
            double a[ND];
            hx::node_data<AVG>(ph, q, a);
#pragma unroll
            for (int c = 0; c < ND; ++c) sN[c * ne + n] = a[c];
        } else {
            double Fi[3][NV];
            hx::euler_contravariant(ph, q, ja, Fi);
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int v = 0; v < NV; ++v) sF[(d * NF + v - V0) * ne + n] = Fi[d][v] - Fvc[d][v];
        }
    }
    double xs[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) xs[v] = GAUSS ? sx[v] : ((0.0 + sx[v]) + sy[v]) + sz[v];
    double qd[NV], gk[NV], dt0 = 0.0, dt1 = 0.0;
#pragma unroll
    for (int v = 0; v < NV; ++v) qd[v] = 0.0;
    for (int d = 0; d < 3; ++d) {
        double jh[3];   // this node's metric row times metric_scale
#pragma unroll
        for (int c = 0; c < 3; ++c) jh[c] = hx::metric_scale(AVG) * ja[3 * d + c];
        if (SPLIT) {
            if (d) __syncthreads();
            if (active) {
#pragma unroll
                for (int c = 0; c < 3; ++c) sJ[c * ne + n] = jh[c];
            }
        }
        if (SPLIT || d == 0) __syncthreads();
        if (d == 2) {
#pragma unroll
            for (int v = 0; v < NV; ++v) gk[v] = G[v * S + g];
            dt0 = hxk::bits_to_double(dta.in[0]);
            dt1 = hxk::bits_to_double(dta.in[1]);
        }
        if (!active) continue;
        const int own = d == 0 ? i : (d == 1 ? j : k);
        const int Nd = d == 0 ? Nx : (d == 1 ? Ny : Nz);
        const int stride = d == 0 ? 1 : (d == 1 ? nx : nx * ny);
        const int base = n - own * stride;
        const double *hatD = sOp + 64 * Nd, *sharpD = sOp + MAXN1 * 64 + 64 * Nd;
        if (!SPLIT) {
            for (int l = 0; l <= Nd; ++l) {
                const double hh = hatD[own * 8 + l];
                const int m = base + l * stride;
#pragma unroll
                for (int v = 0; v < NV; ++v) qd[v] = qd[v] + hh * sF[(d * NF + v) * ne + m];
            }
        } else {
            double fi[NV], pu, pv, pw, pp;
            hx::euler_primitives(ph, q, pu, pv, pw, pp);
            hx::euler_contravariant_row(q, pu, pv, pw, pp, &ja[3 * d], fi);
            for (int l = 0; l <= Nd; ++l) {
                const double sh = sharpD[own * 8 + l], hh = hatD[own * 8 + l];
                const int m = base + l * stride;
                double a[ND], b[ND], jb[3], fs[NV];
#pragma unroll
                for (int c = 0; c < ND; ++c) { a[c] = sN[c * ne + n]; b[c] = sN[c * ne + m]; }
#pragma unroll
                for (int c = 0; c < 3; ++c) jb[c] = sJ[c * ne + m];
                if (SYM) {
                    hx::two_point<AVG>(a, jh, b, jb, fs);
                } else {
                    const bool lo = own < l;
                    double x1[ND], x2[ND], j1[3], j2[3];
#pragma unroll
                    for (int c = 0; c < ND; ++c) { x1[c] = lo ? a[c] : b[c]; x2[c] = lo ? b[c] : a[c]; }
#pragma unroll
                    for (int c = 0; c < 3; ++c) { j1[c] = lo ? jh[c] : jb[c]; j2[c] = lo ? jb[c] : jh[c]; }
                    hx::two_point<AVG>(x1, j1, x2, j2, fs);
                }
                qd[0] = qd[0] + sh * (l == own ? fi[0] : fs[0]);
#pragma unroll
                for (int v = 1; v < NV; ++v)
                    qd[v] = qd[v] + sh * (l == own ? fi[v] : fs[v]) + hh * sF[(d * NF + v - V0) * ne + m];
            }
        }
    }
    if (active) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            if (SPLIT) qd[v] = -qd[v];
            hxk::hold(xs[v]);
            qd[v] = qd[v] - xs[v];
        }
        const double ijg = 1.0 / jg;
#pragma unroll
        for (int v = 0; v < NV; ++v) qd[v] = qd[v] * ijg;
        if (qdot_out) {
#pragma unroll
            for (int v = 0; v < NV; ++v) qdot_out[v * S + g] = qd[v];
        }
        if (update) {
            hxk::hold(dt0);
            hxk::hold(dt1);
            const double dt = fmin(dt0, dt1);
            if (dta.time && tile == 0 && t == 0) dta.time[0] += dt;   // the host passes the time to one launch per stage
            const double cdt = c_k * dt;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                hxk::hold(gk[v]);
                const double gv = a_k * gk[v] + qd[v];
                G[v * S + g] = gv;
                q[v] = q[v] + cdt * gv;
                Q[v * S + g] = q[v];
            }
            if (dta.nan && (isnan(q[0]) || isnan(q[1]) || isnan(q[2]) || isnan(q[3]) || isnan(q[4])))
                atomicOr(dta.nan, 1);
        }
    }
    __syncthreads();   // the tile's LDS has been read: the next tile overwrites it
    }
}

// MaxTimeStep of the resident state for mixed orders: one thread per node, the element of each
// node from eofn (DGSEMClass.f90:870-1035); block minima to part (see k_dt_reduce).
__global__ void __launch_bounds__(256)
km_maxdt(const double *__restrict__ Q, const double *__restrict__ Ja, const double *__restrict__ Jac,
         const int *__restrict__ eofn, const int *__restrict__ N3, const OpsAll *__restrict__ opsp,
         long nodes, long S, double cfl, double dcfl, Phys ph, double *__restrict__ part)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double tc = INFINITY, tv = INFINITY;
    if (g < nodes) {
        const int e = eofn[g];
        const double dc = opsp->dcsi[N3[3 * e]], de = opsp->dcsi[N3[3 * e + 1]], dz = opsp->dcsi[N3[3 * e + 2]];
        double q[NV], ja[9];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = Q[v * S + g];
#pragma unroll
        for (int c = 0; c < 9; ++c) ja[c] = Ja[c * S + g];
        hx::dt_limits(ph, q, ja, Jac[g], dc, de, dz, cfl, dcfl, tc, tv);
    }
    hxk::block_min_pair(tc, tv, part, threadIdx.x, blockIdx.x);
}

}  // namespace hxm
