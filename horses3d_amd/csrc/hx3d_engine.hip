// hx3d engine: MI355X (gfx950) HIP kernels for one explicit RK3 step of HORSES3D's
// compressible Navier-Stokes DGSEM on Gauss-Lobatto nodes, and the C ABI of include/hx3d.h.
//
// One stage (ComputeTimeDerivative, SpatialDiscretization.f90:227-320) is three launches:
//   k_grad    per element: local gradient (HexElement_ComputeLocalGradient,
//             HexElementClass.f90:427-531) + BR1 lifting (BR1_LiftGradients,
//             EllipticBR1.f90:168-397).  With GL nodes the face trace is the boundary node
//             itself (ProlongSolutionToFaces is a copy) and the lifting only touches
//             boundary nodes, so the neighbour's boundary state is gathered directly.
//   k_face    per face node: Riemann flux (Roe/LxF) + BR1 viscous average
//             (computeElementInterfaceFlux, SD.f90:1710-1799), written once per face.
//   k_volume  per element: viscosity (Sutherland + LES), inviscid/viscous contravariant
//             fluxes, the standard or split-form volume integral
//             (TimeDerivative_VolumetricContribution, SD.f90:1602-1682), the surface
//             integral of the face fluxes, 1/J (SD.f90:483-494), and the Williamson RK3
//             register update (TakeRK3Step, ExplicitMethods.f90:748-775), fused.
// Device layout: structure of arrays, one contiguous array per variable:
//   per node g = e*(N+1)^3 + i + (N+1)(j + (N+1)k), S the padded node count:
//     Q[v*S + g], G[v*S + g], vgrad[c*S + g] (velocity gradients, grad T), Ja[(3d + c)*S + g],
//     J[g], invJ[g]
//   per element face node x = (6e + slot)(N+1)^2 + a + (N+1)b, ES the padded count:
//     fe[v*ES + x] (face flux times 1/w_0, negated on the right side), egeo[c*ES + x]
//     (normal, |Ja|), nbr[x] (2 * neighbour node + side)
//   per face node q = f(N+1)^2 + i + (N+1)j (left element's face coordinates), FS the count:
//     fgeo[c*FS + q] (n, t1, t2, |Ja|), fnode[q] (left node, right node, left x, right x)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/hx3d.h"
#include "physics.hpp"
#include "stage_kernels.hpp"
#include "mixed_kernels.hpp"
#include "gauss_kernels.hpp"
#include "spectral.hpp"

using hx::NV;
using hx::Phys;
using hxk::DtArgs;
using hxk::Ops;
using hxs::almost_equal;
using hxs::bary;
using hxs::gl_nodes;
using hxs::interp_matrix;
using hxs::left_to_right;
using hxs::tset_host;

namespace {

constexpr int MAXN = 7;

// element face slot 0..5 = EFRONT, EBACK, EBOTTOM, ERIGHT, ETOP, ELEFT
constexpr int SLOT_AXIS[6] = {1, 1, 2, 0, 2, 0};
constexpr int SLOT_PLUS[6] = {0, 1, 0, 1, 1, 0};
constexpr int SLOT_U[6] = {0, 0, 0, 1, 0, 1};
constexpr int SLOT_V[6] = {2, 2, 1, 2, 1, 2};

inline int grid_for(int nb) { return ((nb + 7) / 8) * 8; }

// ------------------------------------------------------------------------------------
// reductions and small kernels
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_min(double x)
{
    for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ double wave_max(double x)
{
    for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
    return x;
}
__device__ __forceinline__ double wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// MaxTimeStep (DGSEMClass.f90:870-1035): per-node convective and viscous limits, block
// minima to part, reduced by k_dt_reduce (the minimum is exact and independent of order).
__global__ void __launch_bounds__(256)
k_maxdt(const double *__restrict__ Q, const double *__restrict__ Ja, const double *__restrict__ Jac,
        long nodes, long S, double dcsi, double cfl, double dcfl, Phys ph, double *__restrict__ part)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double tc = INFINITY, tv = INFINITY;
    if (g < nodes) {
        double q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = Q[v * S + g];
        double ja[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) ja[c] = Ja[c * S + g];
        hx::dt_limits(ph, q, ja, Jac[g], dcsi, dcsi, dcsi, cfl, dcfl, tc, tv);
    }
    hxk::block_min_pair(tc, tv, part, threadIdx.x, blockIdx.x);
}

// The minima of n (convective, viscous) pairs as the bit patterns of a dt slot.
__global__ void __launch_bounds__(1024)
k_dt_reduce(const double *__restrict__ part, int n, unsigned long long *__restrict__ bits)
{
    __shared__ double sa[16], sb[16];
    double a = INFINITY, b = INFINITY;
    for (int x = threadIdx.x; x < n; x += blockDim.x) {
        a = fmin(a, part[2 * x]);
        b = fmin(b, part[2 * x + 1]);
    }
    a = wave_min(a);
    b = wave_min(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) { a = fmin(a, sa[x]); b = fmin(b, sb[x]); }
        bits[0] = (unsigned long long)__double_as_longlong(a);
        bits[1] = (unsigned long long)__double_as_longlong(b);
    }
}

// a host-given dt as both limits of the n / 2 pairs of a dt slot
__global__ void k_dt_set(unsigned long long *bits, int n, double dt)
{
    for (int x = threadIdx.x; x < n; x += blockDim.x) bits[x] = (unsigned long long)__double_as_longlong(dt);
}

// dt = min(convective, viscous) (DGSEMClass.f90:1024-1030); with advance != 0 also
// time += dt (device scalars: [0] dt, [1] t)
__global__ void k_dt_finish(const unsigned long long *bits, double *sc, int advance)
{
    if (threadIdx.x == 0) {
        const double a = __longlong_as_double((long long)bits[0]);
        const double b = __longlong_as_double((long long)bits[1]);
        const double dt = a < b ? a : b;
        sc[0] = dt;
        if (advance) sc[1] += dt;
    }
}

// Per-block partial results in a fixed order, summed by k_finish_* in a fixed order, so
// every reduction is deterministic.
__global__ void __launch_bounds__(256)
k_absmax(const double *__restrict__ X, long nodes, long S, double *__restrict__ part)
{
    __shared__ double sm[4][NV];
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double m[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) m[v] = g < nodes ? fabs(X[v * S + g]) : 0.0;
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        m[v] = wave_max(m[v]);
        if (ln == 0) sm[w][v] = m[v];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double r = sm[0][threadIdx.x];
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) r = fmax(r, sm[x][threadIdx.x]);
        part[(long)blockIdx.x * NV + threadIdx.x] = r;
    }
}

__global__ void k_finish_max(const double *part, int nparts, double *out)
{
    if (threadIdx.x < NV) {
        double r = 0.0;
        for (int b = 0; b < nparts; ++b) r = fmax(r, part[(long)b * NV + threadIdx.x]);
        out[threadIdx.x] = r;
    }
}

// The allocation-free residuals (hx3d_residuals_noalloc): max |X| per variable over all nodes
// by at most NOALLOC_BLOCKS blocks, each striding over the nodes, partial maxima to part[b*NV + v];
// then one block reduces the nparts partials into out[0..NV-1].  The maximum of doubles is exact
// and independent of order, so the result has the bits of k_absmax + k_finish_max.
constexpr int NOALLOC_BLOCKS = 2048;

__global__ void __launch_bounds__(256)
k_absmax_strided(const double *__restrict__ X, long nodes, long S, double *__restrict__ part)
{
    __shared__ double sm[4][NV];
    double m[NV] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < nodes; g += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int v = 0; v < NV; ++v) m[v] = fmax(m[v], fabs(X[v * S + g]));
    }
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        m[v] = wave_max(m[v]);
        if (ln == 0) sm[w][v] = m[v];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double r = sm[0][threadIdx.x];
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) r = fmax(r, sm[x][threadIdx.x]);
        part[(long)blockIdx.x * NV + threadIdx.x] = r;
    }
}

__global__ void __launch_bounds__(256) k_finish_max_block(const double *__restrict__ part, int nparts, double *out)
{
    __shared__ double sm[4][NV];
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        double r = 0.0;
        for (int b = threadIdx.x; b < nparts; b += blockDim.x) r = fmax(r, part[(long)b * NV + v]);
        r = wave_max(r);
        if (ln == 0) sm[w][v] = r;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double r = sm[0][threadIdx.x];
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) r = fmax(r, sm[x][threadIdx.x]);
        out[threadIdx.x] = r;
    }
}

// Kinetic energy: sum_e sum_n w_i w_j w_k J |rho u|^2 / (2 rho), and the volume
__global__ void __launch_bounds__(256)
k_kinetic(const double *__restrict__ Q, const double *__restrict__ Jac, const double *__restrict__ wq,
          long nodes, long S, double *__restrict__ part)
{
    __shared__ double sk[4], sv[4];
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double ke = 0.0, vol = 0.0;
    if (g < nodes) {
        const double J = Jac[g];
        const double ww = wq[g];
        const double r = Q[g], a = Q[S + g], b = Q[2 * S + g], c = Q[3 * S + g];
        ke = ww * J * (0.5 * (a * a + b * b + c * c) / r);   // VolumeIntegrals.f90:185-193
        vol = ww * J;
    }
    ke = wave_sum(ke);
    vol = wave_sum(vol);
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) { sk[w] = ke; sv[w] = vol; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) { ke += sk[x]; vol += sv[x]; }
        part[2 * blockIdx.x] = ke;
        part[2 * blockIdx.x + 1] = vol;
    }
}

// Volume integrals of the volume monitors (VolumeIntegrals.f90:179-375): sum_e sum_n w_i w_j w_k
// J f(node) for the monitor's f, and the volume, as (value, volume) block partials.
//   kind 1 kinetic energy rate  u . d(rho u)/dt - |u|^2/2 d(rho)/dt   (needs QDot)
//        2 enstrophy            |curl u|^2                          (needs the gradients)
//        3 mean velocity        |rho u| / rho
//        4 internal energy      rho e
//        5 entropy              log p - gamma log rho
//        6 math entropy         -rho (log p - gamma log rho) / (gamma - 1)
// The velocity gradients are the gradient kernel's viscous state (uniform orders) or formed from
// the lifted state gradients U_x, U_y, U_z as getVelocityGradients_State does (mixed orders).
__global__ void __launch_bounds__(256)
k_monitor(int kind, const double *__restrict__ Q, const double *__restrict__ QD, const double *__restrict__ vgrad,
          const double *__restrict__ U, const double *__restrict__ Jac, const double *__restrict__ wq, long nodes,
          long S, Phys ph, double *__restrict__ part)
{
    __shared__ double sk[4], sv[4];
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double val = 0.0, vol = 0.0;
    if (g < nodes) {
        const double J = Jac[g], ww = wq[g];
        double q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) q[v] = Q[v * S + g];
        double f = 0.0;
        if (kind == 1) {
            double u = q[1] / q[0];
            f = u * QD[S + g] - 0.5 * (u * u) * QD[g];
            u = q[2] / q[0];
            f = f + u * QD[2 * S + g] - 0.5 * (u * u) * QD[g];
            u = q[3] / q[0];
            f = f + u * QD[3 * S + g] - 0.5 * (u * u) * QD[g];
        } else if (kind == 2) {
            double ux[3], uy[3], uz[3];   // d(u, v, w)/dx, /dy, /dz
            if (vgrad) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    ux[c] = vgrad[c * S + g];
                    uy[c] = vgrad[(3 + c) * S + g];
                    uz[c] = vgrad[(6 + c) * S + g];
                }
            } else {
                const double ir = 1.0 / q[0], ir2 = ir * ir;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double ud = q[1 + c] * ir2;
                    ux[c] = ir * U[(0 * NV + 1 + c) * S + g] - ud * U[(0 * NV) * S + g];
                    uy[c] = ir * U[(1 * NV + 1 + c) * S + g] - ud * U[(1 * NV) * S + g];
                    uz[c] = ir * U[(2 * NV + 1 + c) * S + g] - ud * U[(2 * NV) * S + g];
                }
            }
            const double a = uy[2] - uz[1], b = uz[0] - ux[2], c = ux[1] - uy[0];
            f = a * a + b * b + c * c;
        } else if (kind == 3) {
            f = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        } else if (kind == 4) {
            f = q[4];
        } else {
            const double p = hx::pressure(ph, q);
            const double sv_ = log(p) - ph.gamma * log(q[0]);
            f = kind == 5 ? sv_ : -q[0] * sv_ / ph.gm1;
        }
        // the reference's products: w w w J f, and for the mean velocity w w w |rho u| / rho J
        val = kind == 3 ? ww * f / q[0] * J : ww * J * f;
        vol = ww * J;
    }
    val = wave_sum(val);
    vol = wave_sum(vol);
    const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) { sk[w] = val; sv[w] = vol; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int x = 1; x < (int)(blockDim.x >> 6); ++x) { val += sk[x]; vol += sv[x]; }
        part[2 * blockIdx.x] = val;
        part[2 * blockIdx.x + 1] = vol;
    }
}

__global__ void k_finish_sum2(const double *part, int nparts, double *out)
{
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int x = 0; x < nparts; ++x) { a += part[2 * x]; b += part[2 * x + 1]; }
        out[0] = a;
        out[1] = b;
    }
}

__global__ void __launch_bounds__(256)
k_nan(const double *__restrict__ Q, long nodes, long S, int *__restrict__ flag)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nodes) return;
    bool bad = false;
#pragma unroll
    for (int v = 0; v < NV; ++v) bad |= isnan(Q[v * S + g]);
    if (bad) atomicOr(flag, 1);
}

// packed (nodes, 5) <-> SoA [5][S]
__global__ void k_pack_in(const double *__restrict__ in, double *__restrict__ Q, long nodes, long S)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nodes) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) Q[v * S + g] = in[NV * g + v];
}
__global__ void k_pack_out(const double *__restrict__ Q, double *__restrict__ out, long nodes, long S)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nodes) return;
#pragma unroll
    for (int v = 0; v < NV; ++v) out[NV * g + v] = Q[v * S + g];
}
__global__ void k_fill(double *__restrict__ X, long n, double val)
{
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) X[g] = val;
}
__global__ void k_fill_int(int *__restrict__ X, int n, int val)
{
    for (int x = threadIdx.x; x < n; x += blockDim.x) X[x] = val;
}

// ------------------------------------------------------------------------------------
// host side: spectral operators (restated from NodalStorageClass.f90 for the engine)
// ------------------------------------------------------------------------------------

Ops make_ops(int N, bool gauss = false)
{
    Ops o;
    std::memset(&o, 0, sizeof(o));
    double x[8], w[8], wb[8];
    hxs::nodes_of(gauss, N, x, w);
    for (int j = 0; j <= N; ++j) wb[j] = 1.0;
    for (int j = 1; j <= N; ++j)
        for (int k = 0; k < j; ++k) { wb[k] *= (x[k] - x[j]); wb[j] *= (x[j] - x[k]); }
    for (int j = 0; j <= N; ++j) wb[j] = 1.0 / wb[j];
    double D[8][8];
    for (int i = 0; i <= N; ++i) {
        D[i][i] = 0.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) { D[i][j] = wb[j] / (wb[i] * (x[i] - x[j])); D[i][i] -= D[i][j]; }
    }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) {
            o.D[i * 8 + j] = D[i][j];
            o.hatD[i * 8 + j] = D[j][i] * w[j] / w[i];
            o.sharpD[i * 8 + j] = 2.0 * D[i][j];
        }
    o.sharpD[0] = 2.0 * D[0][0] + 1.0 / w[0];
    o.sharpD[N * 8 + N] = 2.0 * D[N][N] - 1.0 / w[N];
    if (gauss) std::memset(o.sharpD, 0, sizeof(o.sharpD));   // split forms run on Gauss-Lobatto nodes
    for (int i = 0; i <= N; ++i) o.w[i] = w[i];
    return o;
}


#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return eng->fail(HX3D_EDEVICE, std::string(#x ": ") + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

// Scratch of the allocation-free entry point (hx3d_residuals_noalloc): NOALLOC_BLOCKS x NV
// partial maxima, 80 KB, one copy per GPU the library runs on, held until the process ends.
// On the host the name `scratch` is only a stand-in, not a device address: host code takes the
// address from hipGetSymbolAddress, once per device with that device current (noalloc_scratch),
// and uses that pointer in kernel arguments.  Every call on a GPU shares that GPU's copy, so
// calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[NOALLOC_BLOCKS * NV];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
double *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// scratch's device address on device dev (the current device), looked up on the device's first
// call; nullptr with err set when the lookup fails
double *noalloc_scratch(int dev, hipError_t *err)
{
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (!noalloc_ptr[dev]) {
        void *p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess || !p) return nullptr;
        noalloc_ptr[dev] = static_cast<double *>(p);
    }
    *err = hipSuccess;
    return noalloc_ptr[dev];
}
}  // namespace

struct hx3d_engine {
    int K = 0, N = 0, n1 = 0, nf = 0, ne = 0, F = 0;
    long nodes = 0, S = 0, FS = 0, ES = 0;
    Phys ph{};
    Ops ops{};
    int inviscid = 0;
    double dcsi = 0;
    hipStream_t stream = nullptr;
    double *dQ = nullptr, *dG = nullptr, *dVgrad = nullptr, *dJa = nullptr;
    double *dFe = nullptr, *dQDot = nullptr, *dEdelta = nullptr, *dFdelta = nullptr, *dWq = nullptr;
    // Gauss nodes, one order: Q and gradient traces in element-face order, the neighbour table
    double *dQtr = nullptr, *dGtr = nullptr;
    int *dNbx = nullptr;
    hxg::GaussOps gops{};
    double *dJ = nullptr, *dFgeo = nullptr, *dEgeo = nullptr;
    double *dScal = nullptr, *dPart = nullptr, *dStage = nullptr, *dDtPart = nullptr;
    unsigned long long *dBits = nullptr;
    int *dNbr = nullptr, *dFlag = nullptr;
    int4 *dFnode = nullptr;
    // mixed orders (p-anisotropic elements, mortar faces)
    bool mixed = false;
    bool t2x = false;     // uniform orders: the face kernel forms t2 = n x t1 (see tangent2_is_cross)
    bool gauss = false;   // Gauss nodes (one order: hxg kernels; per-element orders: the mixed path's)
    long FSm = 0;
    int *dN3 = nullptr, *dEfoff = nullptr, *dEofn = nullptr;
    long *dOff = nullptr;
    hxm::FaceRec *dFaces = nullptr;
    int2 *dSched = nullptr;   // the face kernels' waves (hxm::face_lane): one face, or two of one shape
    int nwaves = 0;
    hxm::OpsAll *dOpsAll = nullptr;
    double *dU = nullptr, *dLift = nullptr;
    // mixed orders: the element tiles of the persistent element kernels (packing policy
    // tile_policy, see hx3d_autotune) and each kernel's grid (blocks resident at once)
    hxm::NodeMap *dTmap = nullptr;   // per tile thread: node, element, node indices and orders
    // tiles of TILE_SMALL threads (elements of at most that many nodes), then of TILE threads;
    // map_big: the first map entry of the larger tiles; grids per kernel and tile size
    int ntiles = 0, tile_policy = 0, ntiles_s = 0, ntiles_b = 0;
    long map_big = 0;
    int grid_mgrad_s = 0, grid_mgrad_b = 0, grid_mvol_s = 0, grid_mvol_b = 0;
    std::vector<int> N3h_mixed;   // [3K] element orders (host copy for re-tiling)
    // side streams for the element-bucket launches (the buckets touch disjoint elements, so
    // they run side by side), forked from and joined back to the engine stream with events
    bool capturing = false;   // a step graph is being captured
    std::vector<int> N3h;   // element orders (Nx, Ny, Nz), host copy for solution files
    double mach = 0;
    bool qdot_valid = false, timing = false;
    // CTD_AFTER_STEPS (ExplicitMethods.f90:784): the time derivative is recomputed at the state a
    // step leaves, instead of keeping the last stage's
    bool ctd_after = false;
    // dFlag[0]: hx3d_check_nan's scan; dFlag[1]: sticky divergence flag set by the RK update
    // kernels when a NaN appears.  hFlag: pinned host copy of dFlag[1], refreshed in stream
    // order after each queued step loop (read by hx3d_query once the stream is idle).
    int *hFlag = nullptr;
    int device = 0;   // the HIP device the engine was created on
    // several independent cases in one engine (hx3d_engine_create_multi): per-case tables; with
    // one case ncases = 1, mc is all null and dTime = dScal + 1
    int ncases = 1, nblocks_multi = 0;
    hxk::Multi mc{};
    int4 *dBlocks = nullptr;
    hxk::CaseConst *dCC = nullptr;
    int *dActive = nullptr, *dFcase = nullptr, *dPfirst = nullptr;
    double *dTimeM = nullptr;
    double *dTime = nullptr;   // [ncases] the time of each case
    long dev_bytes = 0;   // device memory the engine holds
    // hx3d_advance's step as a HIP graph, captured for one (cfl, dcfl) and replayed per step
    // (off by default: replays measured 2-5% slower than the same launches issued directly)
    bool use_graph = false;
    hipGraphExec_t step_graph = nullptr;
    double graph_cfl = 0, graph_dcfl = 0;
    double ktime[4] = {0, 0, 0, 0};
    std::vector<hipEvent_t> evpool;   // timing: 4 events per stage of the last hx3d_advance
    std::string err;

    int fail(int code, const std::string &m) { err = m; return code; }
};

namespace {

template <typename T>
int dalloc(hx3d_engine *eng, T **p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) return eng->fail(HX3D_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    eng->dev_bytes += (long)(n * sizeof(T));
    return 0;
}

// elements per block for the element kernels
// The stage kernels address every array plane through a 32-bit byte offset (hxk::ldg), and the
// neighbour table stores 2 * node + side in an int: each plane must hold fewer than 2^29
// elements.
inline bool offsets_fit(long S, long ES, long FS)
{
    const long lim = 1L << 29;
    return S < lim && ES < lim && FS < lim;
}

#ifndef HX3D_EPB4
#define HX3D_EPB4 2
#endif
constexpr int epb_for(int N) { return N == 1 ? 16 : N == 2 ? 8 : N == 3 ? 4 : N == 4 ? HX3D_EPB4 : 1; }

// blocks of the uniform-order element kernels (a multiple of 8, see xcd_block): EPB elements per
// block, or the block table of several cases
int elem_grid(const hx3d_engine *eng, int epb)
{
    if (eng->mc.blocks) return grid_for(eng->nblocks_multi);
    return grid_for((eng->K + epb - 1) / epb);
}
int grad_grid(const hx3d_engine *eng) { return elem_grid(eng, eng->gauss ? 1 : epb_for(eng->N)); }

// Gauss nodes, one order: the Q traces of the resident state, after it is set from outside
template <int N>
void launch_prolong_n(hx3d_engine *eng)
{
    hipLaunchKernelGGL((hxg::kg_prolong<N>), dim3(elem_grid(eng, 1)), dim3(hxg::gauss_tpb<N>()), 0, eng->stream,
                       eng->dQ, eng->dQtr, eng->K, eng->S, eng->ES, eng->gops);
}
void refresh_traces(hx3d_engine *eng)
{
    if (!eng->gauss || eng->mixed) return;
    switch (eng->N) {
    case 1: launch_prolong_n<1>(eng); break;
    case 2: launch_prolong_n<2>(eng); break;
    case 3: launch_prolong_n<3>(eng); break;
    case 4: launch_prolong_n<4>(eng); break;
    case 5: launch_prolong_n<5>(eng); break;
    case 6: launch_prolong_n<6>(eng); break;
    default: launch_prolong_n<7>(eng); break;
    }
}

template <int N>
void launch_grad(hx3d_engine *eng, const hxk::DtLimits &dtl)
{
    constexpr int EPB = epb_for(N), ne = (N + 1) * (N + 1) * (N + 1);
    constexpr int TPB = ((EPB * ne + 63) / 64) * 64;
    hipLaunchKernelGGL((hxk::k_grad<N, EPB>), dim3(elem_grid(eng, EPB)), dim3(TPB), 0, eng->stream, eng->dQ, eng->dJa,
                       eng->dJ, eng->dNbr, eng->dEgeo, eng->dVgrad, eng->K, eng->S, eng->ES,
                       eng->ops, eng->ph, dtl, eng->mc);
}

template <int N>
void launch_face(hx3d_engine *eng, const hxk::DtReduce &dtr)
{
    const int nb = grid_for((int)((eng->FS + hxk::FACE_TPB - 1) / hxk::FACE_TPB));
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(nb), dim3(hxk::FACE_TPB), 0, eng->stream, eng->dQ, eng->dVgrad, eng->dFgeo,
                           eng->dFnode, eng->dFdelta, eng->dFe, eng->FS, eng->S, eng->ES, 1.0 / eng->ops.w[0],
                           eng->ph, dtr, eng->mc);
    };
    if (eng->ph.riemann != 1) go(hxk::k_face<N, false, false>);   // Roe: no tangents
    else if (eng->t2x) go(hxk::k_face<N, true, true>);
    else go(hxk::k_face<N, false, true>);
}

template <int N, int FORM, int EPB>
void launch_volume_e(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot)
{
    constexpr int ne = (N + 1) * (N + 1) * (N + 1);
    constexpr int TPB = ((EPB * ne + 63) / 64) * 64;
    hipLaunchKernelGGL((hxk::k_volume<N, EPB, FORM>), dim3(elem_grid(eng, EPB)), dim3(TPB), 0, eng->stream, eng->dQ,
                       eng->dVgrad, eng->dJa, eng->dJ, eng->dEdelta, eng->dFe, eng->dG,
                       qdot, a_k, c_k, update, dta, eng->K, eng->S, eng->ES, eng->ops, eng->ph, eng->mc);
}

template <int N, int FORM>
void launch_volume_f(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot)
{
    launch_volume_e<N, FORM, epb_for(N)>(eng, a_k, c_k, update, dta, qdot);
}

template <int N>
void launch_volume(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot)
{
    if (eng->inviscid != HX3D_INVISCID_SPLIT) return launch_volume_f<N, hxk::FORM_STANDARD>(eng, a_k, c_k, update, dta, qdot);
    switch (eng->ph.averaging) {
    case HX3D_AVG_DUCROS: return launch_volume_f<N, hxk::AVG_DUCROS>(eng, a_k, c_k, update, dta, qdot);
    case HX3D_AVG_KENNEDYGRUBER: return launch_volume_f<N, hxk::AVG_KG>(eng, a_k, c_k, update, dta, qdot);
    case HX3D_AVG_PIROZZOLI: return launch_volume_f<N, hxk::AVG_PIROZZOLI>(eng, a_k, c_k, update, dta, qdot);
    case HX3D_AVG_MORINISHI: return launch_volume_f<N, hxk::AVG_MORINISHI>(eng, a_k, c_k, update, dta, qdot);
    default: return launch_volume_f<N, hxk::AVG_STANDARD>(eng, a_k, c_k, update, dta, qdot);
    }
}

// One stage on the engine stream.  ev: NULL, or 4 events recorded around the three launches
// (timing only); like a failed launch, a failed record surfaces through hipGetLastError, which
// the calling entry point checks after its launches.
template <int N>
void stage_n(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot,
             hipEvent_t *ev)
{
    if (eng->gauss) {
        // Gauss nodes: traces in element-face order (hxg)
        constexpr int TPB = hxg::gauss_tpb<N>();
        const int nb = elem_grid(eng, 1);
        if (ev) (void)hipEventRecord(ev[0], eng->stream);
        if (!eng->ph.euler)
            hipLaunchKernelGGL((hxg::kg_grad<N>), dim3(nb), dim3(TPB), 0, eng->stream, eng->dQ, eng->dQtr, eng->dJa,
                               eng->dJ, eng->dNbx, eng->dEgeo, eng->dGtr, eng->dVgrad, eng->K, eng->S, eng->ES,
                               eng->ops, eng->gops, eng->ph, dta.lim);
        if (ev) (void)hipEventRecord(ev[1], eng->stream);
        const int nbf = grid_for((int)((eng->FS + hxk::FACE_TPB - 1) / hxk::FACE_TPB));
        hipLaunchKernelGGL((hxg::kg_face<N>), dim3(nbf), dim3(hxk::FACE_TPB), 0, eng->stream, eng->dQtr, eng->dGtr,
                           eng->dFgeo, eng->dFnode, eng->dFdelta, eng->dFe, eng->FS, eng->ES, eng->ph, dta.red);
        if (ev) (void)hipEventRecord(ev[2], eng->stream);
        hipLaunchKernelGGL((hxg::kg_volume<N>), dim3(nb), dim3(TPB), 0, eng->stream, eng->dQ, eng->dQtr, eng->dVgrad,
                           eng->dJa, eng->dJ, eng->dEdelta, eng->dFe, eng->dG, qdot, a_k, c_k, update, dta, eng->K,
                           eng->S, eng->ES, eng->ops, eng->gops, eng->ph);
        if (ev) (void)hipEventRecord(ev[3], eng->stream);
        return;
    }
    if (ev) (void)hipEventRecord(ev[0], eng->stream);
    if (!eng->ph.euler) launch_grad<N>(eng, dta.lim);
    if (ev) (void)hipEventRecord(ev[1], eng->stream);
    launch_face<N>(eng, dta.red);
    if (ev) (void)hipEventRecord(ev[2], eng->stream);
    launch_volume<N>(eng, a_k, c_k, update, dta, qdot);
    if (ev) (void)hipEventRecord(ev[3], eng->stream);
}

// ev: NULL, or 4 events recorded around the three launches of the stage
void stage_mixed(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot,
                 hipEvent_t *ev);

void stage(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot,
           hipEvent_t *ev = nullptr)
{
    if (eng->mixed) return stage_mixed(eng, a_k, c_k, update, dta, qdot, ev);
    switch (eng->N) {
    case 1: stage_n<1>(eng, a_k, c_k, update, dta, qdot, ev); break;
    case 2: stage_n<2>(eng, a_k, c_k, update, dta, qdot, ev); break;
    case 3: stage_n<3>(eng, a_k, c_k, update, dta, qdot, ev); break;
    case 4: stage_n<4>(eng, a_k, c_k, update, dta, qdot, ev); break;
    case 5: stage_n<5>(eng, a_k, c_k, update, dta, qdot, ev); break;
    case 6: stage_n<6>(eng, a_k, c_k, update, dta, qdot, ev); break;
    default: stage_n<7>(eng, a_k, c_k, update, dta, qdot, ev); break;
    }
}

// dt slots (pairs of bit patterns) in dBits: 0-1 alternate through hx3d_advance, 3 is
// hx3d_max_dt's, 4 holds a host-given dt
constexpr int DT_SLOTS = 5;

// dt slot s of every case: ncases (convective, viscous) pairs
unsigned long long *dt_slot(const hx3d_engine *eng, int s) { return eng->dBits + 2L * s * eng->ncases; }

DtArgs host_dt(hx3d_engine *eng, double dt)
{
    unsigned long long *b = dt_slot(eng, 4);
    hipLaunchKernelGGL(k_dt_set, dim3(1), dim3(256), 0, eng->stream, b, 2 * eng->ncases, dt);
    DtArgs a{};
    a.in = b;
    return a;
}

// Standalone MaxTimeStep of the resident state into dt slot `slot` (bits reset first);
// k_dt_finish then writes dt to dScal[0].
void launch_maxdt_bits(hx3d_engine *eng, double cfl, double dcfl, unsigned long long *bits);

void launch_maxdt(hx3d_engine *eng, double cfl, double dcfl, int slot)
{
    unsigned long long *b = dt_slot(eng, slot);
    launch_maxdt_bits(eng, cfl, dcfl, b);
    hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(64), 0, eng->stream, b, eng->dScal, 0);
}

// ------------------------------------------------------------------------------------
// mixed orders
// ------------------------------------------------------------------------------------
// The persistent element kernels: one launch sweeps every element, grid_m* blocks (as many as
// are resident at once) taking the tiles in turn.
template <int FORM>
size_t mvolume_lds(int T) { return sizeof(double) * (2 * hxm::MAXN1 * 64 + (size_t)hxm::mvolume_per_node<FORM>() * T); }
size_t mgrad_lds(int T) { return sizeof(double) * (hxm::MAXN1 * 64 + (size_t)NV * T); }

template <int FORM, bool GAUSS = false>
void launch_mvolume(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot)
{
    if (eng->ntiles_s)
        hipLaunchKernelGGL((hxm::km_volume<FORM, GAUSS, hxm::TILE_SMALL>), dim3(eng->grid_mvol_s), dim3(hxm::TILE_SMALL),
                           mvolume_lds<FORM>(hxm::TILE_SMALL), eng->stream, eng->dTmap, eng->ntiles_s, eng->dQ, eng->dU,
                           eng->dEfoff, eng->dJa, eng->dJ, eng->dEdelta, eng->dFe, eng->dOpsAll, eng->dG, qdot, a_k, c_k,
                           update, dta, eng->S, eng->ES, eng->ph);
    if (eng->ntiles_b) {
        DtArgs db = dta;   // the time advances once per stage: in the first launch
        if (eng->ntiles_s) db.time = nullptr;
        hipLaunchKernelGGL((hxm::km_volume<FORM, GAUSS, hxm::TILE>), dim3(eng->grid_mvol_b), dim3(hxm::TILE),
                           mvolume_lds<FORM>(hxm::TILE), eng->stream, eng->dTmap + eng->map_big, eng->ntiles_b, eng->dQ,
                           eng->dU, eng->dEfoff, eng->dJa, eng->dJ, eng->dEdelta, eng->dFe, eng->dOpsAll, eng->dG, qdot,
                           a_k, c_k, update, db, eng->S, eng->ES, eng->ph);
    }
}

// One stage of the mixed path; GAUSS: Gauss nodes (standard form only)
template <bool GAUSS>
void stage_mixed_nodes(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot,
                       hipEvent_t *ev)
{
    if (ev) (void)hipEventRecord(ev[0], eng->stream);
    hxk::DtLimits lim = dta.lim;
    if (!eng->ph.euler) {
        hipLaunchKernelGGL(hxm::km_face_grad<GAUSS>, dim3(grid_for(eng->nwaves)), dim3(64), 0, eng->stream, eng->dQ,
                           eng->dN3, eng->dOff, eng->dEfoff, eng->dFaces, eng->dSched, eng->dFgeo, eng->dOpsAll,
                           eng->dLift, eng->nwaves, eng->S, eng->ES, eng->FSm);
        // dt partials: the small tiles' blocks first, then the large tiles'
        if (eng->ntiles_s)
            hipLaunchKernelGGL((hxm::km_grad<GAUSS, hxm::TILE_SMALL>), dim3(eng->grid_mgrad_s), dim3(hxm::TILE_SMALL),
                               mgrad_lds(hxm::TILE_SMALL), eng->stream, eng->dTmap, eng->ntiles_s, eng->dQ, eng->dEfoff,
                               eng->dJa, eng->dJ, eng->dLift, eng->dOpsAll, eng->dU, eng->S, eng->ES, eng->ph, lim, 0);
        if (eng->ntiles_b)
            hipLaunchKernelGGL((hxm::km_grad<GAUSS, hxm::TILE>), dim3(eng->grid_mgrad_b), dim3(hxm::TILE),
                               mgrad_lds(hxm::TILE), eng->stream, eng->dTmap + eng->map_big, eng->ntiles_b, eng->dQ,
                               eng->dEfoff, eng->dJa, eng->dJ, eng->dLift, eng->dOpsAll, eng->dU, eng->S, eng->ES,
                               eng->ph, lim, eng->grid_mgrad_s);
    }
    if (ev) (void)hipEventRecord(ev[1], eng->stream);
    hipLaunchKernelGGL(hxm::km_face<GAUSS>, dim3(grid_for(eng->nwaves)), dim3(64), 0, eng->stream, eng->dQ, eng->dU,
                       eng->dN3, eng->dOff, eng->dEfoff, eng->dFaces, eng->dSched, eng->dFgeo, eng->dFdelta,
                       eng->dOpsAll, eng->dFe, eng->nwaves, eng->S, eng->ES, eng->FSm, eng->ph, dta.red);
    if (ev) (void)hipEventRecord(ev[2], eng->stream);
    if (GAUSS) launch_mvolume<hxk::FORM_STANDARD, true>(eng, a_k, c_k, update, dta, qdot);
    else if (eng->inviscid != HX3D_INVISCID_SPLIT) launch_mvolume<hxk::FORM_STANDARD>(eng, a_k, c_k, update, dta, qdot);
    else switch (eng->ph.averaging) {
        case HX3D_AVG_DUCROS: launch_mvolume<hxk::AVG_DUCROS>(eng, a_k, c_k, update, dta, qdot); break;
        case HX3D_AVG_KENNEDYGRUBER: launch_mvolume<hxk::AVG_KG>(eng, a_k, c_k, update, dta, qdot); break;
        case HX3D_AVG_PIROZZOLI: launch_mvolume<hxk::AVG_PIROZZOLI>(eng, a_k, c_k, update, dta, qdot); break;
        case HX3D_AVG_MORINISHI: launch_mvolume<hxk::AVG_MORINISHI>(eng, a_k, c_k, update, dta, qdot); break;
        default: launch_mvolume<hxk::AVG_STANDARD>(eng, a_k, c_k, update, dta, qdot); break;
        }
    if (ev) (void)hipEventRecord(ev[3], eng->stream);
}

void stage_mixed(hx3d_engine *eng, double a_k, double c_k, int update, const DtArgs &dta, double *qdot,
                 hipEvent_t *ev)
{
    if (eng->gauss) stage_mixed_nodes<true>(eng, a_k, c_k, update, dta, qdot, ev);
    else stage_mixed_nodes<false>(eng, a_k, c_k, update, dta, qdot, ev);
}

// MaxTimeStep limits of the resident state into the dt slot bits[0..1]
void launch_maxdt_bits(hx3d_engine *eng, double cfl, double dcfl, unsigned long long *bits)
{
    const long nb = (eng->nodes + 255) / 256;
    if (eng->mixed)
        hipLaunchKernelGGL(hxm::km_maxdt, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, eng->dJa, eng->dJ, eng->dEofn,
                           eng->dN3, eng->dOpsAll, eng->nodes, eng->S, cfl, dcfl, eng->ph, eng->dDtPart);
    else
        hipLaunchKernelGGL(k_maxdt, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, eng->dJa, eng->dJ, eng->nodes,
                           eng->S, eng->dcsi, cfl, dcfl, eng->ph, eng->dDtPart);
    hipLaunchKernelGGL(k_dt_reduce, dim3(1), dim3(1024), 0, eng->stream, eng->dDtPart, (int)nb, bits);
}

// the divergence flag's pinned host copy, and both device flags cleared
int init_flags(hx3d_engine *eng)
{
    if (!eng->dTime) eng->dTime = eng->dScal + 1;
    if (hipHostMalloc((void **)&eng->hFlag, sizeof(int) * eng->ncases, hipHostMallocDefault) != hipSuccess)
        return eng->fail(HX3D_ENOMEM, "hipHostMalloc failed");
    for (int c = 0; c < eng->ncases; ++c) eng->hFlag[c] = 0;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, eng->stream, eng->dFlag, 1 + eng->ncases, 0);
    return HX3D_OK;
}

// hFlag <- dFlag[1] in stream order (read by the host once the stream is idle)
void queue_flag_readback(hx3d_engine *eng)
{
    (void)hipMemcpyAsync(eng->hFlag, eng->dFlag + 1, sizeof(int) * eng->ncases, hipMemcpyDeviceToHost, eng->stream);
}

// After a queued step loop: the divergence flag's read-back, and the time derivative the
// reference's storage holds after TakeRK3Step -- the last stage's (already written to dQDot by
// the caller's last stage), or with CTD_AFTER_STEPS the one of the new state.
void after_steps(hx3d_engine *eng);

// physics constants (PhysicsStorage_NS.f90:145-439, ThermodynamicsAir :115-132)
void set_physics(hx3d_engine *eng, const hx3d_case_desc *d)
{
    Phys &ph = eng->ph;
    ph.gamma = 1.4; ph.gm1 = 1.4 - 1.0;
    // gamma * POW2(Mach) and 1 / (gm1 * POW2(Mach) * Pr) with the square formed first, as the
    // reference does (PhysicsStorage_NS.f90:250, 285): (gamma * M) * M rounds differently
    ph.gammaM2 = ph.gamma * (d->mach * d->mach);
    ph.euler = d->euler ? 1 : 0;
    ph.mu = d->euler ? 0.0 : 1.0 / d->reynolds;
    ph.mu_to_kappa = d->euler ? 0.0 : 1.0 / (ph.gm1 * (d->mach * d->mach) * d->prandtl);
    const double Tref = 520.0 * 5.0 / 9.0, Ssu = 198.6 * 5.0 / 9.0;
    ph.S_div_T = Ssu / Tref;
    ph.Trenorm = Tref / Tref;
    ph.lambda_stab = d->lambda_stab;
    ph.les_c = d->les_c;
    ph.riemann = d->riemann;
    ph.cp = ph.gamma * (1.0 / (1.4 - 1.0));
    ph.averaging = d->averaging;
    ph.les = d->euler ? 0 : d->les;
    eng->inviscid = d->inviscid;
}

// Mixed orders: element and face tables (Face_LinkWithElements, FaceClass.f90:187-278), the
// operators of every order, element buckets, device arrays.  The face frames at the face
// orders must come with the descriptor (MappedGeometryFace of the mortar faces).
// Element tiles of the persistent element kernels under packing policy `policy`:
//   0  largest elements first, each into the first tile with room (first-fit decreasing):
//      the fewest tiles and idle threads;
//   1  elements of one order triple (Nx, Ny, Nz) together, in turn: tiles of like elements, so
//      the contractions' loop bounds agree across a wave;
//   2  one element per tile.
// Tile sizes (policy / 4): by default tiles hold TILE_SMALL threads when every element has at
// most that many nodes, else TILE; the kernels and launches also take both sizes at once (small
// elements in small tiles, the others in large ones, each size swept by its own launch).  Every policy
// gives the same results (each element is computed from its own data alone; the dt partials are
// minima).  Also sets each element kernel's grids: the blocks resident at once.
int set_tiles(hx3d_engine *eng, int policy)
{
    const std::vector<int> &N3 = eng->N3h_mixed;
    const int K = eng->K;
    // policy = packing + 4 * sizes; sizes 0: small tiles when every element fits, else large
    // ones; 1: large tiles for every element; 2: small elements in small tiles, the others in
    // large ones (a launch of each size)
    const int packing = policy % 4, sizes = policy / 4;
    if (policy < 0 || packing > 2 || sizes > 2) return eng->fail(HX3D_EINVAL, "unknown tile packing policy");
    std::vector<int> ne(K);
    for (int e = 0; e < K; ++e) ne[e] = (N3[3 * e] + 1) * (N3[3 * e + 1] + 1) * (N3[3 * e + 2] + 1);
    // the elements of each tile size, packed under the policy into tiles of cap threads
    auto pack = [&](std::vector<int> order, int cap) {
        std::vector<hxm::Tile> tiles;
        auto open_tile = [&]() {
            hxm::Tile t;
            std::memset(&t, 0, sizeof(t));
            tiles.push_back(t);
            return (int)tiles.size() - 1;
        };
        auto add = [&](int ti, int e) {
            hxm::Tile &t = tiles[ti];
            t.elem[t.count] = e;
            t.off[t.count + 1] = t.off[t.count] + ne[e];
            ++t.count;
        };
        auto fits = [&](int ti, int e) {
            const hxm::Tile &t = tiles[ti];
            return t.count < hxm::TILE_MAXE && t.off[t.count] + ne[e] <= cap;
        };
        const int n = (int)order.size();
        if (packing == 0) {
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ne[a] > ne[b]; });
            const int min_ne = n ? ne[order[n - 1]] : 0;
            int first_open = 0;   // tiles before it hold no element of any size left
            for (int e : order) {
                int ti = -1;
                for (int x = first_open; x < (int)tiles.size(); ++x)
                    if (fits(x, e)) { ti = x; break; }
                if (ti < 0) ti = open_tile();
                add(ti, e);
                while (first_open < (int)tiles.size() &&
                       (tiles[first_open].count == hxm::TILE_MAXE || tiles[first_open].off[tiles[first_open].count] + min_ne > cap))
                    ++first_open;
            }
        } else if (packing == 1) {
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
                if (ne[a] != ne[b]) return ne[a] > ne[b];
                for (int c = 0; c < 3; ++c)
                    if (N3[3 * a + c] != N3[3 * b + c]) return N3[3 * a + c] < N3[3 * b + c];
                return false;
            });
            int ti = -1;
            for (int x = 0; x < n; ++x) {
                const int e = order[x];
                const bool same = ti >= 0 && std::equal(&N3[3 * e], &N3[3 * e + 3], &N3[3 * tiles[ti].elem[0]]);
                if (ti < 0 || !same || !fits(ti, e)) ti = open_tile();
                add(ti, e);
            }
        } else {
            for (int e : order) add(open_tile(), e);
        }
        return tiles;
    };
    // one tile size per mesh: small tiles when every element fits (on a 24^3 box with orders on
    // {2..5}, 5% faster per step than 512-thread tiles), else 512-thread tiles for all (with
    // orders up to 7, a launch of each size measured 0.6% slower than the large tiles alone)
    std::vector<int> small, big;
    for (int e = 0; e < K; ++e) (ne[e] <= hxm::TILE_SMALL ? small : big).push_back(e);
    if (sizes == 1 || (sizes == 0 && !big.empty())) {
        big.insert(big.end(), small.begin(), small.end());
        std::sort(big.begin(), big.end());
        small.clear();
    }
    const std::vector<hxm::Tile> ts = pack(small, hxm::TILE_SMALL), tb = pack(big, hxm::TILE);
    for (const hxm::Tile &t : tb)
        if (t.count < 1 || t.off[t.count] > hxm::TILE) return eng->fail(HX3D_EUNSUPPORTED, "an element has more nodes than a tile");
    // the threads' node map (hxm::NodeMap): node offsets as the device's element offsets; the
    // small tiles' entries, then the large tiles'
    std::vector<long> eoff(K + 1, 0);
    for (int e = 0; e < K; ++e) eoff[e + 1] = eoff[e] + ne[e];
    if (eoff[K] > INT_MAX) return eng->fail(HX3D_EUNSUPPORTED, "mixed orders: more than 2^31 - 1 nodes");
    const size_t map_big = ts.size() * (size_t)hxm::TILE_SMALL;
    std::vector<hxm::NodeMap> map(map_big + tb.size() * (size_t)hxm::TILE);
    auto fill = [&](const std::vector<hxm::Tile> &tiles, int cap, size_t base) {
        for (size_t ti = 0; ti < tiles.size(); ++ti) {
            const hxm::Tile &T = tiles[ti];
            int slot = 0;
            for (int t = 0; t < cap; ++t) {
                const bool active = t < T.off[T.count];
                while (active && t >= T.off[slot + 1]) ++slot;
                const int s = active ? slot : T.count - 1;
                const int e = T.elem[s], n = active ? t - T.off[s] : 0;
                const int nx = N3[3 * e] + 1, ny = N3[3 * e + 1] + 1;
                const int i = n % nx, j = (n / nx) % ny, k = n / (nx * ny);
                hxm::NodeMap &m = map[base + ti * cap + t];
                m.g = (int)(eoff[e] + n);
                m.e = e;
                m.code = i | j << 3 | k << 6 | N3[3 * e] << 9 | N3[3 * e + 1] << 12 | N3[3 * e + 2] << 15 |
                         (active ? T.off[s] : 0) << 18 | (active ? 1 : 0) << 27;
            }
        }
    };
    fill(ts, hxm::TILE_SMALL, 0);
    fill(tb, hxm::TILE, map_big);
    if (eng->dTmap) {
        if (hipStreamSynchronize(eng->stream) != hipSuccess) return eng->fail(HX3D_EDEVICE, "synchronize failed");
        (void)hipFree(eng->dTmap);
        eng->dTmap = nullptr;
    }
    int rc;
    if ((rc = dalloc(eng, &eng->dTmap, map.size()))) return rc;
    if (hipMemcpy(eng->dTmap, map.data(), map.size() * sizeof(hxm::NodeMap), hipMemcpyHostToDevice) != hipSuccess)
        return eng->fail(HX3D_EDEVICE, "tile upload failed");
    eng->ntiles_s = (int)ts.size();
    eng->ntiles_b = (int)tb.size();
    eng->ntiles = eng->ntiles_s + eng->ntiles_b;
    eng->map_big = (long)map_big;
    eng->tile_policy = policy;
    // grids: the blocks resident at once on the device (at least one), no more than the tiles
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return eng->fail(HX3D_EDEVICE, "no HIP device");
    hipError_t oe = hipSuccess;
    auto occ = [&](const void *k, int T, size_t lds) {
        int b = 0;
        if (oe == hipSuccess) oe = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, T, lds);
        return std::max(1, b) * cus;
    };
    const int gs = eng->gauss ? occ(reinterpret_cast<const void *>(&hxm::km_grad<true, hxm::TILE_SMALL>), hxm::TILE_SMALL, mgrad_lds(hxm::TILE_SMALL))
                              : occ(reinterpret_cast<const void *>(&hxm::km_grad<false, hxm::TILE_SMALL>), hxm::TILE_SMALL, mgrad_lds(hxm::TILE_SMALL));
    const int gb = eng->gauss ? occ(reinterpret_cast<const void *>(&hxm::km_grad<true, hxm::TILE>), hxm::TILE, mgrad_lds(hxm::TILE))
                              : occ(reinterpret_cast<const void *>(&hxm::km_grad<false, hxm::TILE>), hxm::TILE, mgrad_lds(hxm::TILE));
    const bool split = eng->inviscid == HX3D_INVISCID_SPLIT;
    const int avg = eng->ph.averaging;
    auto vol = [&](auto form_tag, auto gauss_tag) {
        constexpr int FORM = decltype(form_tag)::value;
        constexpr bool GAUSS = decltype(gauss_tag)::value;
        const int a = occ(reinterpret_cast<const void *>(&hxm::km_volume<FORM, GAUSS, hxm::TILE_SMALL>), hxm::TILE_SMALL,
                          mvolume_lds<FORM>(hxm::TILE_SMALL));
        const int b = occ(reinterpret_cast<const void *>(&hxm::km_volume<FORM, GAUSS, hxm::TILE>), hxm::TILE,
                          mvolume_lds<FORM>(hxm::TILE));
        return std::make_pair(a, b);
    };
    using F = std::false_type;
    std::pair<int, int> gv;
    if (eng->gauss) gv = vol(std::integral_constant<int, hxk::FORM_STANDARD>{}, std::true_type{});
    else if (!split) gv = vol(std::integral_constant<int, hxk::FORM_STANDARD>{}, F{});
    else if (avg == HX3D_AVG_DUCROS) gv = vol(std::integral_constant<int, hxk::AVG_DUCROS>{}, F{});
    else if (avg == HX3D_AVG_KENNEDYGRUBER) gv = vol(std::integral_constant<int, hxk::AVG_KG>{}, F{});
    else if (avg == HX3D_AVG_PIROZZOLI) gv = vol(std::integral_constant<int, hxk::AVG_PIROZZOLI>{}, F{});
    else if (avg == HX3D_AVG_MORINISHI) gv = vol(std::integral_constant<int, hxk::AVG_MORINISHI>{}, F{});
    else gv = vol(std::integral_constant<int, hxk::AVG_STANDARD>{}, F{});
    if (oe != hipSuccess) return eng->fail(HX3D_EDEVICE, std::string("occupancy query: ") + hipGetErrorString(oe));
    eng->grid_mgrad_s = eng->ntiles_s ? std::min(eng->ntiles_s, gs) : 0;
    eng->grid_mgrad_b = eng->ntiles_b ? std::min(eng->ntiles_b, gb) : 0;
    eng->grid_mvol_s = eng->ntiles_s ? std::min(eng->ntiles_s, gv.first) : 0;
    eng->grid_mvol_b = eng->ntiles_b ? std::min(eng->ntiles_b, gv.second) : 0;
    return HX3D_OK;
}

// Whether the stored second face tangent is n x t1 with vCross's arithmetic
// (MappedGeometry.f90:745, :807-817) at every face node, bit for bit (fgeo[c*FS + q]: n at c = 0..2,
// t1 at 3..5, t2 at 6..8).  The reference's geometry and the mesh front end's are; a descriptor
// with other tangents keeps them, read by the face kernel.
bool tangent2_is_cross(const std::vector<double> &fgeo, long FS)
{
    for (long q = 0; q < FS; ++q) {
        const double *n0 = &fgeo[q], *t = &fgeo[3 * FS + q], *t2 = &fgeo[6 * FS + q];
        const double n[3] = {n0[0], n0[FS], n0[2 * FS]}, u[3] = {t[0], t[FS], t[2 * FS]};
        const double x[3] = {n[1] * u[2] - u[1] * n[2], n[2] * u[0] - u[2] * n[0], n[0] * u[1] - u[0] * n[1]};
        for (int c = 0; c < 3; ++c)
            if (std::memcmp(&x[c], &t2[c * FS], sizeof(double)) != 0) return false;
    }
    return true;
}

// The waves of the mortar face kernels (hxm::face_lane): a face of more than 32 nodes alone, a
// smaller one with the next face of the same shape (face orders and both sides' orders in the
// face orientation), in face order at the second face of each pair; faces left without a
// partner alone at the end.  Every face is in exactly one wave, and each face's results depend
// only on its own data, so any grouping gives the same bits.  Built with HX3D_FACE_PAIRS=0,
// every face takes a wave of its own (for comparison).
#ifndef HX3D_FACE_PAIRS
#define HX3D_FACE_PAIRS 1
#endif
std::vector<int2> face_waves(const std::vector<hxm::FaceRec> &faces)
{
    const int F = (int)faces.size();
    std::vector<int2> w;
    w.reserve(F);
    std::vector<int> pending(1 << 18, -1);   // by shape: six orders of 3 bits
    for (int f = 0; f < F; ++f) {
        const hxm::FaceRec &r = faces[f];
        if (!HX3D_FACE_PAIRS || (r.Nf[0] + 1) * (r.Nf[1] + 1) > 32) {
            w.push_back(make_int2(f, -1));
            continue;
        }
        const int key = r.Nf[0] | r.Nf[1] << 3 | r.Nfs[0][0] << 6 | r.Nfs[0][1] << 9 | r.Nfs[1][0] << 12 |
                        r.Nfs[1][1] << 15;
        if (pending[key] < 0) {
            pending[key] = f;
        } else {
            w.push_back(make_int2(pending[key], f));
            pending[key] = -1;
        }
    }
    std::vector<int> rest;
    for (int p : pending)
        if (p >= 0) rest.push_back(p);
    std::sort(rest.begin(), rest.end());
    for (int f : rest) w.push_back(make_int2(f, -1));
    return w;
}

int create_mixed(hx3d_engine *eng, const hx3d_case_desc *d)
{
    eng->mixed = true;
    const bool gauss = d->nodes == HX3D_NODES_GAUSS;
    eng->gauss = gauss;
    const int K = d->K, F = d->F;
    eng->K = K; eng->F = F;
    if (!(d->fNormal && d->fT1 && d->fT2 && d->fJac))
        return eng->fail(HX3D_EINVAL, "mixed orders: the descriptor must give the face geometry (fNormal, fT1, fT2, fJac) at the face orders");
    std::vector<int> N3(d->N, d->N + 3 * K);
    std::vector<long> off(K + 1, 0);
    for (int e = 0; e < K; ++e) off[e + 1] = off[e] + (long)(N3[3 * e] + 1) * (N3[3 * e + 1] + 1) * (N3[3 * e + 2] + 1);
    const long nodes = off[K];
    eng->nodes = nodes;
    eng->S = ((nodes + 63) / 64) * 64;
    const long S = eng->S;
    std::vector<int> efoff(6 * K + 1, 0);
    for (int e = 0; e < K; ++e)
        for (int s = 0; s < 6; ++s)
            efoff[6 * e + s + 1] = efoff[6 * e + s] + (N3[3 * e + SLOT_U[s]] + 1) * (N3[3 * e + SLOT_V[s]] + 1);
    eng->ES = (((long)efoff[6 * K] + 63) / 64) * 64;
    const long ES = eng->ES;
    if (!offsets_fit(eng->S, ES, 0))
        return eng->fail(HX3D_EUNSUPPORTED, "mesh too large: more than 2^29 - 1 nodes (or face nodes) per array");
    // faces
    std::vector<hxm::FaceRec> faces(F);
    long fsm = 0;
    for (int f = 0; f < F; ++f) {
        hxm::FaceRec &r = faces[f];
        const int rot = d->fRot[f];
        for (int sd = 0; sd < 2; ++sd) {
            const int e = d->fElem[2 * f + sd], s = d->fESide[2 * f + sd] - 1;
            if (e < 0 || e >= K || s < 0 || s > 5 || rot < 0 || rot > 7)
                return eng->fail(HX3D_EINVAL, "face " + std::to_string(f) + ": bad element/side/rotation");
            if (d->faceID[6 * e + s] != f || d->faceSide[6 * e + s] != sd + 1)
                return eng->fail(HX3D_EINVAL, "face " + std::to_string(f) + ": inconsistent with element faceID/faceSide");
            r.e[sd] = e; r.slot[sd] = s;
            r.Nel[sd][0] = N3[3 * e + SLOT_U[s]];
            r.Nel[sd][1] = N3[3 * e + SLOT_V[s]];
        }
        r.rot = rot;
        r.Nfs[0][0] = r.Nel[0][0]; r.Nfs[0][1] = r.Nel[0][1];
        const bool keep = rot == 0 || rot == 2 || rot == 5 || rot == 7;
        r.Nfs[1][0] = keep ? r.Nel[1][0] : r.Nel[1][1];
        r.Nfs[1][1] = keep ? r.Nel[1][1] : r.Nel[1][0];
        r.Nf[0] = std::max(r.Nfs[0][0], r.Nfs[1][0]);
        r.Nf[1] = std::max(r.Nfs[0][1], r.Nfs[1][1]);
        for (int sd = 0; sd < 2; ++sd) {
            const bool a = r.Nfs[sd][0] != r.Nf[0], b = r.Nfs[sd][1] != r.Nf[1];
            r.proj[sd] = (!a && !b) ? 0 : (a && !b) ? 1 : (!a && b) ? 2 : 3;
        }
        r.foff = (int)fsm;
        fsm += (long)(r.Nf[0] + 1) * (r.Nf[1] + 1);
    }
    eng->FSm = fsm;
    const std::vector<int2> sched = face_waves(faces);
    eng->nwaves = (int)sched.size();
    // operators of every order
    hxm::OpsAll oa;
    std::memset(&oa, 0, sizeof(oa));
    std::vector<double> wq((size_t)S, 0.0);
    double xn[8][8], wn[8][8];
    for (int n = 1; n <= MAXN; ++n) {
        const Ops o = make_ops(n, gauss);
        std::memcpy(oa.D[n], o.D, sizeof(o.D));
        std::memcpy(oa.hatD[n], o.hatD, sizeof(o.hatD));
        std::memcpy(oa.sharpD[n], o.sharpD, sizeof(o.sharpD));
        hxs::nodes_of(gauss, n, xn[n], wn[n]);
        oa.b0[n] = 1.0 / wn[n][0];
        oa.dcsi[n] = 1.0 / std::fabs(xn[n][1] - xn[n][0]);
        // boundary values v(-1), v(+1) of the Lagrange basis and b = v / w (NodalStorageClass.f90:
        // 257-273): one-hot on Gauss-Lobatto nodes, dense on Gauss nodes
        double wb[8];
        hxs::bary(n, xn[n], wb);
        hxs::interp_vector(-1.0, n, xn[n], wb, oa.vL[n]);
        hxs::interp_vector(1.0, n, xn[n], wb, oa.vR[n]);
        for (int i = 0; i <= n; ++i) {
            oa.bL[n][i] = oa.vL[n][i] / wn[n][i];
            oa.bR[n][i] = oa.vR[n][i] / wn[n][i];
        }
        for (int m = 1; m <= MAXN; ++m)
            if (m != n) tset_host(n, m, oa.T[n][m], gauss);
    }
    // geometry, weights, LES filter widths (SD.f90:420, :1378)
    std::vector<double> ja((size_t)9 * S, 0.0), jj((size_t)S, 1.0), edelta(K), fdelta(F),
        fgeo((size_t)10 * fsm, 0.0);
    std::vector<int> eofn((size_t)nodes);
    for (int e = 0; e < K; ++e) {
        const int nx = N3[3 * e] + 1, ny = N3[3 * e + 1] + 1, nz = N3[3 * e + 2] + 1;
        double vol = 0.0;
        for (int kk = 0; kk < nz; ++kk)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const long g = off[e] + i + nx * (j + ny * kk);
                    for (int c = 0; c < 9; ++c) ja[(size_t)c * S + g] = d->Ja[9 * g + c];
                                jj[g] = d->J[g];
                    wq[g] = wn[nx - 1][i] * wn[ny - 1][j] * wn[nz - 1][kk];
                    vol += wq[g] * d->J[g];
                    eofn[g] = e;
                }
        edelta[e] = std::pow(vol / (nx * ny * nz), 1.0 / 3.0);
    }
    for (int f = 0; f < F; ++f) {
        const hxm::FaceRec &r = faces[f];
        double surf = 0.0;
        for (int j = 0; j <= r.Nf[1]; ++j)
            for (int i = 0; i <= r.Nf[0]; ++i) {
                const long q = r.foff + i + (r.Nf[0] + 1) * j;
                for (int c = 0; c < 3; ++c) {
                    fgeo[(size_t)c * fsm + q] = d->fNormal[3 * q + c];
                    fgeo[(size_t)(3 + c) * fsm + q] = d->fT1[3 * q + c];
                    fgeo[(size_t)(6 + c) * fsm + q] = d->fT2[3 * q + c];
                }
                fgeo[(size_t)9 * fsm + q] = d->fJac[q];
                surf += wn[r.Nf[0]][i] * wn[r.Nf[1]][j] * d->fJac[q];
            }
        fdelta[f] = std::sqrt(surf / ((r.Nf[0] + 1) * (r.Nf[1] + 1)));
    }
    {
        // the volume kernel keeps three directions of fluxes (+ two-point node data and a metric
        // row for the split form) of a tile's elements in LDS
        const bool split = d->inviscid == HX3D_INVISCID_SPLIT;
        const long per_node = 3 * (split ? 4 : 5) + (split ? hx::node_data_size(d->averaging) + 3 : 0);
        int dev = 0, maxlds = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return eng->fail(HX3D_EDEVICE, "no HIP device");
        if ((long)sizeof(double) * (per_node * hxm::TILE + 2 * hxm::MAXN1 * 64) > maxlds)
            return eng->fail(HX3D_EUNSUPPORTED, "a tile of elements needs more LDS than a block may use");
    }
    eng->N3h_mixed = N3;
    int rc;
    if ((rc = dalloc(eng, &eng->dQ, (size_t)NV * S)) || (rc = dalloc(eng, &eng->dG, (size_t)NV * S)) ||
        (rc = dalloc(eng, &eng->dU, (size_t)15 * S)) || (rc = dalloc(eng, &eng->dJa, (size_t)9 * S)) ||
        (rc = dalloc(eng, &eng->dWq, (size_t)S)) ||
        (rc = dalloc(eng, &eng->dJ, (size_t)S)) || (rc = dalloc(eng, &eng->dFgeo, (size_t)10 * fsm)) ||
        (rc = dalloc(eng, &eng->dFe, (size_t)NV * ES)) || (rc = dalloc(eng, &eng->dLift, (size_t)15 * ES)) ||
        (rc = dalloc(eng, &eng->dQDot, (size_t)NV * S)) || (rc = dalloc(eng, &eng->dEdelta, (size_t)K)) ||
        (rc = dalloc(eng, &eng->dFdelta, (size_t)F)) || (rc = dalloc(eng, &eng->dScal, 4)) ||
        (rc = dalloc(eng, &eng->dPart, (size_t)NV * ((S + 255) / 256 + 1))) ||
        (rc = dalloc(eng, &eng->dDtPart, (size_t)2 * (std::max((long)K, (S + 255) / 256) + 64))) ||
        (rc = dalloc(eng, &eng->dStage, (size_t)NV * S)) || (rc = dalloc(eng, &eng->dBits, 2 * DT_SLOTS)) ||
        (rc = dalloc(eng, &eng->dFlag, 2)) || (rc = dalloc(eng, &eng->dN3, N3.size())) ||
        (rc = dalloc(eng, &eng->dOff, off.size())) || (rc = dalloc(eng, &eng->dEfoff, efoff.size())) ||
        (rc = dalloc(eng, &eng->dEofn, eofn.size())) ||
        (rc = dalloc(eng, &eng->dFaces, faces.size())) || (rc = dalloc(eng, &eng->dSched, sched.size())) ||
        (rc = dalloc(eng, &eng->dOpsAll, 1)))
        return rc;
    if (hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking) != hipSuccess)
        return eng->fail(HX3D_EDEVICE, "hipStreamCreate failed");
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(eng->dJa, ja.data(), ja.size() * 8) ||
        !up(eng->dJ, jj.data(), jj.size() * 8) || !up(eng->dFgeo, fgeo.data(), fgeo.size() * 8) ||
        !up(eng->dWq, wq.data(), wq.size() * 8) || !up(eng->dEdelta, edelta.data(), edelta.size() * 8) ||
        !up(eng->dFdelta, fdelta.data(), fdelta.size() * 8) || !up(eng->dN3, N3.data(), N3.size() * 4) ||
        !up(eng->dOff, off.data(), off.size() * 8) || !up(eng->dEfoff, efoff.data(), efoff.size() * 4) ||
        !up(eng->dEofn, eofn.data(), eofn.size() * 4) ||
        !up(eng->dFaces, faces.data(), faces.size() * sizeof(hxm::FaceRec)) ||
        !up(eng->dSched, sched.data(), sched.size() * sizeof(int2)) || !up(eng->dOpsAll, &oa, sizeof(oa)))
        return eng->fail(HX3D_EDEVICE, "upload failed");
    const long nb = (NV * S + 255) / 256;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, (long)NV * S, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, eng->stream, eng->dG, (long)NV * S, 0.0);
    hipLaunchKernelGGL(k_fill, dim3((15 * S + 255) / 256), dim3(256), 0, eng->stream, eng->dU, 15L * S, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, eng->stream, eng->dScal, 4L, 0.0);
    if (int rc2 = init_flags(eng)) return rc2;
    if (int rc3 = set_tiles(eng, 0)) return rc3;
    if (hipStreamSynchronize(eng->stream) != hipSuccess) return eng->fail(HX3D_EDEVICE, "initialisation failed");
    return HX3D_OK;
}

// Williamson RK3 coefficients (ExplicitMethods.f90:697-699)
const double RK_A[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
const double RK_C[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

}  // namespace

extern "C" int hx3d_compute_qdot(hx3d_engine *eng, double t, double *QDot_host);

namespace {
void after_steps(hx3d_engine *eng)
{
    eng->qdot_valid = true;
    if (eng->ctd_after) (void)hx3d_compute_qdot(eng, 0.0, nullptr);
    queue_flag_readback(eng);
}
}  // namespace

// ------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------
extern "C" {

const char *hx3d_last_error(const hx3d_engine *eng) { return eng ? eng->err.c_str() : "null engine"; }
long hx3d_num_nodes(const hx3d_engine *eng) { return eng ? eng->nodes : -1; }

int hx3d_engine_create(const hx3d_case_desc *d, hx3d_engine **out)
{
    if (!d || !out) return HX3D_EINVAL;
    *out = nullptr;
    hx3d_engine *eng = new hx3d_engine();
    auto bail = [&](int code) { *out = eng; return code; };
    if (d->K <= 0 || !d->N || !d->faceID || !d->faceSide || !d->fElem || !d->fESide || !d->fRot ||
        !d->Ja || !d->J)
        return bail(eng->fail(HX3D_EINVAL, "descriptor: missing arrays or K <= 0"));
    const int N = d->N[0];
    bool uniform = true;
    for (int e = 0; e < d->K; ++e)
        for (int c = 0; c < 3; ++c) {
            if (d->N[3 * e + c] < 1 || d->N[3 * e + c] > MAXN)
                return bail(eng->fail(HX3D_EUNSUPPORTED, "orders must be 1..7"));
            if (d->N[3 * e + c] != N) uniform = false;
        }
    if (2L * d->F != 6L * d->K) return bail(eng->fail(HX3D_EUNSUPPORTED, "every element face must be an interior (periodic) face: need F = 3K"));
    if (d->riemann < 0 || d->riemann > 1 || d->inviscid < 0 || d->inviscid > 1 || d->averaging < 0 ||
        d->averaging > 4 || d->les < 0 || d->les > 2)
        return bail(eng->fail(HX3D_EINVAL, "descriptor: selector out of range"));
    if (d->mach <= 0 || (!d->euler && d->reynolds <= 0))
        return bail(eng->fail(HX3D_EINVAL, "descriptor: mach and reynolds must be positive"));
    if (d->nodes != HX3D_NODES_GAUSSLOBATTO && d->nodes != HX3D_NODES_GAUSS)
        return bail(eng->fail(HX3D_EINVAL, "descriptor: unknown node type"));
    if (d->nodes == HX3D_NODES_GAUSS && d->inviscid == HX3D_INVISCID_SPLIT)
        return bail(eng->fail(HX3D_EINVAL, "split-form discretizations run on Gauss-Lobatto nodes (the reference "
                                           "switches a split-form case to them)"));
    if (d->nodes == HX3D_NODES_GAUSS && !(d->fNormal && d->fT1 && d->fT2 && d->fJac))
        return bail(eng->fail(HX3D_EINVAL, "Gauss nodes: the descriptor must give the face geometry (fNormal, fT1, "
                                           "fT2, fJac): no element node lies on a face"));
    set_physics(eng, d);
    eng->N3h.assign(d->N, d->N + 3L * d->K);
    eng->mach = d->mach;
    // descriptor checks come first (they need no device); the device is recorded from here on
    if (hipGetDevice(&eng->device) != hipSuccess) return bail(eng->fail(HX3D_EDEVICE, "no HIP device"));
    // per-element orders take the mortar (face-trace) kernels; Gauss nodes of one order the
    // uniform layout with the hxg kernels
    if (!uniform) return bail(create_mixed(eng, d));
    eng->gauss = d->nodes == HX3D_NODES_GAUSS;
    eng->K = d->K; eng->N = N; eng->n1 = N + 1; eng->nf = eng->n1 * eng->n1; eng->ne = eng->nf * eng->n1;
    eng->F = d->F;
    eng->nodes = (long)eng->K * eng->ne;
    eng->S = ((eng->nodes + 63) / 64) * 64;
    eng->FS = (long)eng->F * eng->nf;
    eng->ES = (((long)eng->K * 6 * eng->nf + 63) / 64) * 64;
    eng->inviscid = d->inviscid;
    eng->ops = make_ops(N, eng->gauss);
    {
        double x[8], w[8], wb[8];
        hxs::nodes_of(eng->gauss, N, x, w);
        eng->dcsi = 1.0 / std::fabs(x[1] - x[0]);
        // v(-1), v(+1), b = v / w (NodalStorageClass.f90:257-273)
        hxs::bary(N, x, wb);
        hxs::interp_vector(-1.0, N, x, wb, eng->gops.vL);
        hxs::interp_vector(1.0, N, x, wb, eng->gops.vR);
        for (int i = 0; i <= N; ++i) {
            eng->gops.bL[i] = eng->gops.vL[i] / w[i];
            eng->gops.bR[i] = eng->gops.vR[i] / w[i];
        }
    }

    const int K = eng->K, n1 = eng->n1, nf = eng->nf, ne = eng->ne, F = eng->F;
    const long S = eng->S, nodes = eng->nodes, FS = eng->FS, ES = eng->ES;
    // ---- host tables --------------------------------------------------------------
    auto node_of = [&](int e, int slot, int a, int b) -> long {
        int idx[3];
        idx[SLOT_AXIS[slot]] = SLOT_PLUS[slot] ? N : 0;
        idx[SLOT_U[slot]] = a;
        idx[SLOT_V[slot]] = b;
        return (long)e * ne + idx[0] + n1 * (idx[1] + n1 * idx[2]);
    };
    // per element face node x = (6e + slot) nf + a + (N+1) b: nbr[x] = 2 * (neighbour node) +
    // (1 on the face's right side); per face node: (left node, right node, left x, right x)
    std::vector<int> nbr((size_t)K * 6 * nf, -1);
    std::vector<int4> fnode((size_t)FS);
    std::vector<unsigned char> fslot(F);
    for (int f = 0; f < F; ++f) {
        const int eL = d->fElem[2 * f], eR = d->fElem[2 * f + 1];
        const int sL = d->fESide[2 * f] - 1, sR = d->fESide[2 * f + 1] - 1, rot = d->fRot[f];
        if (eL < 0 || eL >= K || eR < 0 || eR >= K || sL < 0 || sL > 5 || sR < 0 || sR > 5 || rot < 0 || rot > 7)
            return bail(eng->fail(HX3D_EINVAL, "face " + std::to_string(f) + ": bad element/side/rotation"));
        if (d->faceID[6 * eL + sL] != f || d->faceSide[6 * eL + sL] != 1 || d->faceID[6 * eR + sR] != f ||
            d->faceSide[6 * eR + sR] != 2)
            return bail(eng->fail(HX3D_EINVAL, "face " + std::to_string(f) + ": inconsistent with element faceID/faceSide"));
        fslot[f] = (unsigned char)sL;
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                int ii, jj;
                left_to_right(i, j, N, N, rot, ii, jj);
                const long gL = node_of(eL, sL, i, j), gR = node_of(eR, sR, ii, jj);
                const long q = (long)f * nf + i + n1 * j;
                const long xL = ((long)eL * 6 + sL) * nf + i + n1 * j;
                const long xR = ((long)eR * 6 + sR) * nf + ii + n1 * jj;
                fnode[q] = make_int4((int)gL, (int)gR, (int)xL, (int)xR);
                nbr[xL] = (int)(2 * gR);
                nbr[xR] = (int)(2 * gL + 1);
            }
    }
    for (size_t x = 0; x < nbr.size(); ++x)
        if (nbr[x] < 0) return bail(eng->fail(HX3D_EINVAL, "an element face node has no neighbour"));
    // Gauss nodes: 2 * (the neighbour's matching element face node) + (1 on the face's right side)
    std::vector<int> nbx;
    if (eng->gauss) {
        nbx.assign(nbr.size(), 0);
        for (long q = 0; q < FS; ++q) {
            nbx[fnode[q].z] = 2 * fnode[q].w;
            nbx[fnode[q].w] = 2 * fnode[q].z + 1;
        }
    }
    if (!offsets_fit(S, ES, FS))
        return bail(eng->fail(HX3D_EUNSUPPORTED, "mesh too large: more than 2^29 - 1 nodes (or face nodes) per array"));
    // geometry SoA, J, 1/J, quadrature weights, LES filter widths, face frames
    std::vector<double> ja((size_t)9 * S, 0.0), jj((size_t)S, 1.0), wq((size_t)S, 0.0),
        edelta(K), fdelta(F), fgeo((size_t)10 * FS, 0.0), egeo((size_t)4 * ES, 0.0);
    double xg[8], wg[8];
    hxs::nodes_of(eng->gauss, N, xg, wg);
    for (long g = 0; g < nodes; ++g) {
        for (int c = 0; c < 9; ++c) ja[(size_t)c * S + g] = d->Ja[9 * g + c];
        jj[g] = d->J[g];
        const int n = (int)(g % ne);
        wq[g] = wg[n % n1] * wg[(n / n1) % n1] * wg[n / nf];
    }
    for (int e = 0; e < K; ++e) {
        double vol = 0.0;
        for (int n = 0; n < ne; ++n) vol += wq[(long)e * ne + n] * d->J[(long)e * ne + n];
        edelta[e] = std::pow(vol / ne, 1.0 / 3.0);   // SD.f90:420
    }
    // face frames: from the descriptor (the reference's MappedGeometryFace), or, with GL nodes,
    // dS = +-Ja^d of the left element at the face node, |Ja| = |dS|, n = dS/|Ja|
    // (ConstructMappedGeometryFace, MappedGeometry.f90:482-760) and a fixed orthonormal
    // completion t1, t2 of n (the Riemann fluxes do not depend on the tangent choice).
    const bool have_fg = d->fNormal && d->fT1 && d->fT2 && d->fJac;
    if (!have_fg && (d->fNormal || d->fT1 || d->fT2 || d->fJac))
        return bail(eng->fail(HX3D_EINVAL, "descriptor: give all four face geometry arrays or none"));
    for (int f = 0; f < F; ++f) {
        const int sL = fslot[f];
        const int dax = SLOT_AXIS[sL];
        const double sg = SLOT_PLUS[sL] ? 1.0 : -1.0;
        double surf = 0.0;
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                const long q = (long)f * nf + i + n1 * j;
                double nr[3], t1[3], t2[3], jac;
                if (have_fg) {
                    for (int c = 0; c < 3; ++c) { nr[c] = d->fNormal[3 * q + c]; t1[c] = d->fT1[3 * q + c]; t2[c] = d->fT2[3 * q + c]; }
                    jac = d->fJac[q];
                } else {
                    const long g = fnode[q].x;
                    for (int c = 0; c < 3; ++c) nr[c] = sg * d->Ja[9 * g + 3 * dax + c];
                    jac = std::sqrt(nr[0] * nr[0] + nr[1] * nr[1] + nr[2] * nr[2]);
                    for (int c = 0; c < 3; ++c) nr[c] /= jac;
                    // t1 = e_m - (e_m . n) n for the axis e_m least aligned with n, normalised
                    int m = 0;
                    for (int c = 1; c < 3; ++c) if (std::fabs(nr[c]) < std::fabs(nr[m])) m = c;
                    for (int c = 0; c < 3; ++c) t1[c] = (c == m ? 1.0 : 0.0) - nr[m] * nr[c];
                    const double tn = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
                    for (int c = 0; c < 3; ++c) t1[c] /= tn;
                    t2[0] = nr[1] * t1[2] - nr[2] * t1[1];
                    t2[1] = nr[2] * t1[0] - nr[0] * t1[2];
                    t2[2] = nr[0] * t1[1] - nr[1] * t1[0];
                }
                for (int c = 0; c < 3; ++c) {
                    fgeo[(size_t)c * FS + q] = nr[c];
                    fgeo[(size_t)(3 + c) * FS + q] = t1[c];
                    fgeo[(size_t)(6 + c) * FS + q] = t2[c];
                }
                fgeo[(size_t)9 * FS + q] = jac;
                for (int c = 0; c < 3; ++c) {
                    egeo[(size_t)c * ES + fnode[q].z] = nr[c];
                    egeo[(size_t)c * ES + fnode[q].w] = nr[c];
                }
                egeo[(size_t)3 * ES + fnode[q].z] = jac;
                egeo[(size_t)3 * ES + fnode[q].w] = jac;
                surf += wg[i] * wg[j] * jac;
            }
        fdelta[f] = std::sqrt(surf / nf);   // SD.f90:1378
    }
    eng->t2x = tangent2_is_cross(fgeo, FS);
    // ---- device -------------------------------------------------------------------
    int rc;
    if ((rc = dalloc(eng, &eng->dQ, (size_t)NV * S)) || (rc = dalloc(eng, &eng->dG, (size_t)NV * S)) ||
        (rc = dalloc(eng, &eng->dVgrad, (size_t)(eng->ph.les ? 13 : 12) * S)) || (rc = dalloc(eng, &eng->dJa, (size_t)9 * S)) ||
        (rc = dalloc(eng, &eng->dWq, (size_t)S)) ||
        (rc = dalloc(eng, &eng->dJ, (size_t)S)) || (rc = dalloc(eng, &eng->dFgeo, (size_t)10 * FS)) ||
        (rc = dalloc(eng, &eng->dFe, (size_t)NV * ES)) || (rc = dalloc(eng, &eng->dEgeo, (size_t)4 * ES)) || (rc = dalloc(eng, &eng->dQDot, (size_t)NV * S)) ||
        (rc = dalloc(eng, &eng->dEdelta, (size_t)K)) || (rc = dalloc(eng, &eng->dFdelta, (size_t)F)) ||
        (rc = dalloc(eng, &eng->dScal, 4)) || (rc = dalloc(eng, &eng->dPart, (size_t)NV * ((S + 255) / 256 + 1))) ||
        (rc = dalloc(eng, &eng->dDtPart, (size_t)2 * (std::max((long)K, (S + 255) / 256) + 64))) ||
        (rc = dalloc(eng, &eng->dStage, (size_t)NV * S)) || (rc = dalloc(eng, &eng->dBits, 2 * DT_SLOTS)) ||
        (rc = dalloc(eng, &eng->dNbr, nbr.size())) ||
        (rc = dalloc(eng, &eng->dFlag, 2)) || (rc = dalloc(eng, &eng->dFnode, fnode.size())))
        return bail(rc);
    if (eng->gauss && ((rc = dalloc(eng, &eng->dQtr, (size_t)NV * ES)) || (rc = dalloc(eng, &eng->dGtr, (size_t)15 * ES)) ||
                       (rc = dalloc(eng, &eng->dNbx, nbx.size()))))
        return bail(rc);
    if (hipStreamCreateWithFlags(&eng->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(eng->fail(HX3D_EDEVICE, "hipStreamCreate failed"));
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(eng->dJa, ja.data(), ja.size() * 8) ||
        !up(eng->dJ, jj.data(), jj.size() * 8) || !up(eng->dFgeo, fgeo.data(), fgeo.size() * 8) ||
        !up(eng->dWq, wq.data(), wq.size() * 8) || !up(eng->dEdelta, edelta.data(), edelta.size() * 8) ||
        !up(eng->dFdelta, fdelta.data(), fdelta.size() * 8) || !up(eng->dNbr, nbr.data(), nbr.size() * 4) ||
        !up(eng->dEgeo, egeo.data(), egeo.size() * 8) || !up(eng->dFnode, fnode.data(), fnode.size() * 16) ||
        (eng->gauss && !up(eng->dNbx, nbx.data(), nbx.size() * 4)))
        return bail(eng->fail(HX3D_EDEVICE, "upload failed"));
    // zero state, RK register and time
    const long nb = (NV * S + 255) / 256;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, (long)NV * S, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(256), 0, eng->stream, eng->dG, (long)NV * S, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, eng->stream, eng->dScal, 4L, 0.0);
    if (eng->gauss) {
        hipLaunchKernelGGL(k_fill, dim3((15 * ES + 255) / 256), dim3(256), 0, eng->stream, eng->dGtr, 15L * ES, 0.0);
        refresh_traces(eng);
    }
    if (int rc2 = init_flags(eng)) return bail(rc2);
    if (hipStreamSynchronize(eng->stream) != hipSuccess)
        return bail(eng->fail(HX3D_EDEVICE, "initialisation failed"));
    *out = eng;
    return HX3D_OK;
}

// ---- several independent cases of one order in one engine ------------------------------
// The union of the cases' meshes (element, face and node numbers offset case by case; no face
// joins two cases) is built as one case; then the per-case tables: element blocks that never
// straddle cases (so a block's case, constants and dt are uniform), each face's case, the
// per-case constants, dt slots, times, divergence flags and active mask.
int hx3d_engine_create_multi(const hx3d_case_desc *descs, int ncases, hx3d_engine **out)
{
    if (!descs || ncases < 1 || !out) return HX3D_EINVAL;
    *out = nullptr;
    const hx3d_case_desc &d0 = descs[0];
    auto reject = [&](int code, const std::string &m) {
        hx3d_engine *e = new hx3d_engine();
        e->fail(code, m);
        *out = e;
        return code;
    };
    long K = 0, F = 0, nodes = 0, fnodes = 0;
    const bool fg = d0.fNormal && d0.fT1 && d0.fT2 && d0.fJac;
    std::vector<long> Kc(ncases), Fc(ncases), nc(ncases);
    for (int c = 0; c < ncases; ++c) {
        const hx3d_case_desc &d = descs[c];
        if (d.K <= 0 || !d.N || !d.faceID || !d.faceSide || !d.fElem || !d.fESide || !d.fRot || !d.Ja || !d.J)
            return reject(HX3D_EINVAL, "case " + std::to_string(c) + ": missing arrays or K <= 0");
        for (long e = 0; e < 3L * d.K; ++e)
            if (d.N[e] != d0.N[0]) return reject(HX3D_EUNSUPPORTED, "several cases in one engine need one order N everywhere");
        if (d.riemann != d0.riemann || d.inviscid != d0.inviscid || d.averaging != d0.averaging || d.les != d0.les ||
            d.euler != d0.euler || d.lambda_stab != d0.lambda_stab)
            return reject(HX3D_EUNSUPPORTED, "several cases in one engine need the same discretisation (Riemann solver, "
                                             "volume form, averaging, LES model, lambda)");
        if (d.euler) return reject(HX3D_EUNSUPPORTED, "several cases in one engine: Navier-Stokes cases only");
        if (d.nodes != HX3D_NODES_GAUSSLOBATTO)
            return reject(HX3D_EUNSUPPORTED, "several cases in one engine: Gauss-Lobatto nodes only");
        if (fg != (d.fNormal && d.fT1 && d.fT2 && d.fJac))
            return reject(HX3D_EINVAL, "give the face geometry for every case or for none");
        if (d.mach <= 0 || d.reynolds <= 0) return reject(HX3D_EINVAL, "case " + std::to_string(c) + ": mach and reynolds must be positive");
        const int n1 = d0.N[0] + 1;
        Kc[c] = d.K; Fc[c] = d.F; nc[c] = (long)d.K * n1 * n1 * n1;
        K += d.K; F += d.F; nodes += nc[c]; fnodes += (long)d.F * n1 * n1;
    }
    if (K >= (1L << 31) / 8) return reject(HX3D_EUNSUPPORTED, "too many elements");
    const int n1 = d0.N[0] + 1, nf = n1 * n1;
    std::vector<int> N(3 * K), faceID(6 * K), faceSide(6 * K), fElem(2 * F), fESide(2 * F), fRot(F), fcase(F);
    std::vector<double> Ja(9 * nodes), J(nodes), fN, fT1, fT2, fJ;
    if (fg) { fN.resize(3 * fnodes); fT1.resize(3 * fnodes); fT2.resize(3 * fnodes); fJ.resize(fnodes); }
    long k0 = 0, f0 = 0, g0 = 0, q0 = 0;
    for (int c = 0; c < ncases; ++c) {
        const hx3d_case_desc &d = descs[c];
        for (long x = 0; x < 3 * Kc[c]; ++x) N[3 * k0 + x] = d.N[x];
        for (long x = 0; x < 6 * Kc[c]; ++x) { faceID[6 * k0 + x] = d.faceID[x] + (int)f0; faceSide[6 * k0 + x] = d.faceSide[x]; }
        for (long x = 0; x < 2 * Fc[c]; ++x) { fElem[2 * f0 + x] = d.fElem[x] + (int)k0; fESide[2 * f0 + x] = d.fESide[x]; }
        for (long x = 0; x < Fc[c]; ++x) { fRot[f0 + x] = d.fRot[x]; fcase[f0 + x] = c; }
        std::memcpy(&Ja[9 * g0], d.Ja, sizeof(double) * 9 * nc[c]);
        std::memcpy(&J[g0], d.J, sizeof(double) * nc[c]);
        if (fg) {
            const long qn = Fc[c] * nf;
            std::memcpy(&fN[3 * q0], d.fNormal, sizeof(double) * 3 * qn);
            std::memcpy(&fT1[3 * q0], d.fT1, sizeof(double) * 3 * qn);
            std::memcpy(&fT2[3 * q0], d.fT2, sizeof(double) * 3 * qn);
            std::memcpy(&fJ[q0], d.fJac, sizeof(double) * qn);
            q0 += qn;
        }
        k0 += Kc[c]; f0 += Fc[c]; g0 += nc[c];
    }
    hx3d_case_desc u = d0;
    u.K = (int)K; u.F = (int)F; u.N = N.data(); u.faceID = faceID.data(); u.faceSide = faceSide.data();
    u.fElem = fElem.data(); u.fESide = fESide.data(); u.fRot = fRot.data(); u.Ja = Ja.data(); u.J = J.data();
    u.fNormal = fg ? fN.data() : nullptr; u.fT1 = fg ? fT1.data() : nullptr;
    u.fT2 = fg ? fT2.data() : nullptr; u.fJac = fg ? fJ.data() : nullptr;
    hx3d_engine *eng = nullptr;
    int rc = hx3d_engine_create(&u, &eng);
    *out = eng;
    if (rc) return rc;
    if (eng->mixed) return eng->fail(HX3D_EUNSUPPORTED, "several cases in one engine need one order N everywhere");
    // block table: EPB elements per block within each case
    const int epb = epb_for(eng->N);
    std::vector<int4> blocks;
    std::vector<int> pfirst(ncases + 1, 0);
    k0 = 0;
    for (int c = 0; c < ncases; ++c) {
        pfirst[c] = (int)blocks.size();
        for (long e = 0; e < Kc[c]; e += epb)
            blocks.push_back(make_int4((int)(k0 + e), (int)std::min<long>(epb, Kc[c] - e), c, e == 0 ? 1 : 0));
        k0 += Kc[c];
    }
    pfirst[ncases] = (int)blocks.size();
    eng->nblocks_multi = (int)blocks.size();
    while ((long)blocks.size() < grid_for(eng->nblocks_multi)) blocks.push_back(make_int4(0, 0, 0, 0));
    std::vector<hxk::CaseConst> cc(ncases);
    for (int c = 0; c < ncases; ++c) {
        const hx3d_case_desc &d = descs[c];
        const double M2 = d.mach * d.mach;   // as set_physics forms them
        cc[c].gammaM2 = eng->ph.gamma * M2;
        cc[c].mu = 1.0 / d.reynolds;
        cc[c].mu_to_kappa = 1.0 / (eng->ph.gm1 * M2 * d.prandtl);
        cc[c].les_c = d.les_c;
    }
    if ((long)eng->F * nf / hxk::FACE_TPB + 1 < ncases)
        return eng->fail(HX3D_EUNSUPPORTED, "more cases than face blocks to reduce their dt");
    // per-case replacements of the one-case dt slots, flags and time
    (void)hipStreamSynchronize(eng->stream);
    (void)hipFree(eng->dBits);
    (void)hipFree(eng->dFlag);
    (void)hipHostFree(eng->hFlag);
    eng->dBits = nullptr; eng->dFlag = nullptr; eng->hFlag = nullptr;
    eng->ncases = ncases;
    const std::vector<int> ones(ncases, 1);
    if ((rc = dalloc(eng, &eng->dBits, (size_t)2 * DT_SLOTS * ncases)) || (rc = dalloc(eng, &eng->dFlag, (size_t)1 + ncases)) ||
        (rc = dalloc(eng, &eng->dBlocks, blocks.size())) || (rc = dalloc(eng, &eng->dCC, cc.size())) ||
        (rc = dalloc(eng, &eng->dActive, (size_t)ncases)) || (rc = dalloc(eng, &eng->dFcase, fcase.size())) ||
        (rc = dalloc(eng, &eng->dPfirst, pfirst.size())) || (rc = dalloc(eng, &eng->dTimeM, (size_t)ncases)))
        return rc;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up(eng->dBlocks, blocks.data(), blocks.size() * sizeof(int4)) || !up(eng->dCC, cc.data(), cc.size() * sizeof(hxk::CaseConst)) ||
        !up(eng->dActive, ones.data(), ones.size() * 4) || !up(eng->dFcase, fcase.data(), fcase.size() * 4) ||
        !up(eng->dPfirst, pfirst.data(), pfirst.size() * 4))
        return eng->fail(HX3D_EDEVICE, "upload failed");
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, eng->stream, eng->dTimeM, (long)ncases, 0.0);
    eng->dTime = eng->dTimeM;
    eng->mc = hxk::Multi{eng->dBlocks, eng->dCC, eng->dActive, eng->dFcase, eng->dPfirst, ncases};
    if ((rc = init_flags(eng))) return rc;
    if (hipStreamSynchronize(eng->stream) != hipSuccess) return eng->fail(HX3D_EDEVICE, "initialisation failed");
    return HX3D_OK;
}

int hx3d_multi_set_active(hx3d_engine *eng, int c, int on)
{
    if (!eng || !eng->mc.blocks || c < 0 || c >= eng->ncases) return HX3D_EINVAL;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, eng->stream, eng->dActive + c, 1, on ? 1 : 0);
    HIPCHK(hipGetLastError());
    return HX3D_OK;
}

int hx3d_multi_status(hx3d_engine *eng, double *times, int *diverged)
{
    if (!eng || !eng->mc.blocks) return HX3D_EINVAL;
    if (times) HIPCHK(hipMemcpyAsync(times, eng->dTime, sizeof(double) * eng->ncases, hipMemcpyDeviceToHost, eng->stream));
    if (diverged) HIPCHK(hipMemcpyAsync(diverged, eng->dFlag + 1, sizeof(int) * eng->ncases, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_engine_destroy(hx3d_engine *eng)
{
    if (!eng) return HX3D_OK;
    // teardown: nothing is left to report an error to, so the calls' statuses are not checked
    if (eng->stream) (void)hipStreamSynchronize(eng->stream);
    void *ptrs[] = {eng->dQ, eng->dG, eng->dVgrad, eng->dJa, eng->dJ, eng->dFgeo, eng->dEgeo, eng->dFe,
                    eng->dQDot, eng->dEdelta, eng->dFdelta, eng->dWq, eng->dScal, eng->dPart, eng->dStage,
                    eng->dBits, eng->dNbr, eng->dFlag, eng->dFnode, eng->dN3, eng->dEfoff,
                    eng->dEofn, eng->dOff, eng->dFaces, eng->dSched, eng->dOpsAll, eng->dU, eng->dLift, eng->dDtPart, eng->dTmap,
                    eng->dBlocks, eng->dCC, eng->dActive, eng->dFcase, eng->dPfirst, eng->dTimeM,
                    eng->dQtr, eng->dGtr, eng->dNbx};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : eng->evpool) (void)hipEventDestroy(e);
    if (eng->step_graph) (void)hipGraphExecDestroy(eng->step_graph);
    if (eng->hFlag) (void)hipHostFree(eng->hFlag);
    if (eng->stream) (void)hipStreamDestroy(eng->stream);
    delete eng;
    return HX3D_OK;
}

int hx3d_set_state_device(hx3d_engine *eng, const double *Qd)
{
    if (!eng || !Qd) return HX3D_EINVAL;
    const long nb = (eng->nodes + 255) / 256;
    hipLaunchKernelGGL(k_pack_in, dim3(nb), dim3(256), 0, eng->stream, Qd, eng->dQ, eng->nodes, eng->S);
    refresh_traces(eng);   // Gauss nodes: the new state's face traces
    // a new state restarts the RK register (the reference allocates G_NS zeroed)
    const long nb2 = (NV * eng->S + 255) / 256;
    hipLaunchKernelGGL(k_fill, dim3(nb2), dim3(256), 0, eng->stream, eng->dG, (long)NV * eng->S, 0.0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, eng->stream, eng->dFlag + 1, eng->ncases, 0);
    queue_flag_readback(eng);
    eng->qdot_valid = false;
    HIPCHK(hipGetLastError());
    return HX3D_OK;
}

int hx3d_get_state_device(hx3d_engine *eng, double *Qd)
{
    if (!eng || !Qd) return HX3D_EINVAL;
    const long nb = (eng->nodes + 255) / 256;
    hipLaunchKernelGGL(k_pack_out, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, Qd, eng->nodes, eng->S);
    HIPCHK(hipGetLastError());
    return HX3D_OK;
}

int hx3d_get_state_nodes(hx3d_engine *eng, long first, long count, double *Qh)
{
    if (!eng || !Qh || first < 0 || count < 0 || first + count > eng->nodes) return HX3D_EINVAL;
    if (count == 0) return HX3D_OK;
    const long nb = (count + 255) / 256;
    // pack the range: node first + g of the SoA state to Qh-ordered staging
    hipLaunchKernelGGL(k_pack_out, dim3(nb), dim3(256), 0, eng->stream, eng->dQ + first, eng->dStage, count, eng->S);
    HIPCHK(hipMemcpyAsync(Qh, eng->dStage, sizeof(double) * NV * count, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_set_state(hx3d_engine *eng, const double *Qh)
{
    if (!eng || !Qh) return HX3D_EINVAL;
    HIPCHK(hipMemcpyAsync(eng->dStage, Qh, sizeof(double) * NV * eng->nodes, hipMemcpyHostToDevice, eng->stream));
    int rc = hx3d_set_state_device(eng, eng->dStage);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_get_state(hx3d_engine *eng, double *Qh)
{
    if (!eng || !Qh) return HX3D_EINVAL;
    int rc = hx3d_get_state_device(eng, eng->dStage);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(Qh, eng->dStage, sizeof(double) * NV * eng->nodes, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

// HexMesh_SaveSolution (HexMesh.f90:3376-3556) of the resident state: a solution file (type 2)
// with the default reference values of the engine's Mach number
int hx3d_save_solution(hx3d_engine *eng, const char *path, const char *header_name, int iter, double time)
{
    if (!eng || !path) return HX3D_EINVAL;
    std::vector<double> Q((size_t)NV * eng->nodes);
    int rc = hx3d_get_state(eng, Q.data());
    if (rc) return rc;
    double refs[6];
    hx3d_hsol_reference_values(eng->mach, refs);
    rc = hx3d_hsol_write_nodes(path, header_name, eng->K, eng->N3h.data(), iter, time, refs, Q.data(), nullptr,
                               eng->gauss ? HX3D_NODES_GAUSS : HX3D_NODES_GAUSSLOBATTO);
    if (rc) return eng->fail(rc, hx3d_hsol_error());
    return HX3D_OK;
}

// LoadSolutionForRestart: the state of a solution file, checked against the engine's mesh
int hx3d_load_solution(hx3d_engine *eng, const char *path, int *iter, double *time)
{
    if (!eng || !path) return HX3D_EINVAL;
    std::vector<double> Q((size_t)NV * eng->nodes);
    int rc = hx3d_hsol_read(path, eng->K, eng->N3h.data(), Q.data(), nullptr);
    if (!rc) rc = hx3d_hsol_read_header(path, nullptr, nullptr, nullptr, iter, time, nullptr);
    if (rc) return eng->fail(rc, hx3d_hsol_error());
    return hx3d_set_state(eng, Q.data());
}

int hx3d_compute_qdot(hx3d_engine *eng, double t, double *QDot_host)
{
    (void)t;   // no source terms on this path (UserDefinedSourceTermNS is zero)
    if (!eng) return HX3D_EINVAL;
    stage(eng, 0.0, 0.0, 0, host_dt(eng, 0.0), eng->dQDot);
    HIPCHK(hipGetLastError());
    eng->qdot_valid = true;
    if (QDot_host) {
        const long nb = (eng->nodes + 255) / 256;
        hipLaunchKernelGGL(k_pack_out, dim3(nb), dim3(256), 0, eng->stream, eng->dQDot, eng->dStage, eng->nodes, eng->S);
        HIPCHK(hipMemcpyAsync(QDot_host, eng->dStage, sizeof(double) * NV * eng->nodes, hipMemcpyDeviceToHost, eng->stream));
        HIPCHK(hipStreamSynchronize(eng->stream));
    }
    return HX3D_OK;
}

int hx3d_rk3_step(hx3d_engine *eng, double t, double dt)
{
    (void)t;
    if (!eng) return HX3D_EINVAL;
    DtArgs a = host_dt(eng, dt);
    a.nan = eng->dFlag + 1;
    for (int k = 0; k < 3; ++k) stage(eng, RK_A[k], RK_C[k], 1, a, k == 2 ? eng->dQDot : nullptr);
    after_steps(eng);
    HIPCHK(hipGetLastError());
    return HX3D_OK;
}

int hx3d_set_ctd_after_steps(hx3d_engine *eng, int on)
{
    if (!eng) return HX3D_EINVAL;
    eng->ctd_after = on != 0;
    return HX3D_OK;
}

int hx3d_get_qdot(hx3d_engine *eng, double *QDot_host)
{
    if (!eng || !QDot_host) return HX3D_EINVAL;
    if (!eng->qdot_valid) {
        int rc = hx3d_compute_qdot(eng, 0.0, nullptr);
        if (rc) return rc;
    }
    const long nb = (eng->nodes + 255) / 256;
    hipLaunchKernelGGL(k_pack_out, dim3(nb), dim3(256), 0, eng->stream, eng->dQDot, eng->dStage, eng->nodes, eng->S);
    HIPCHK(hipMemcpyAsync(QDot_host, eng->dStage, sizeof(double) * NV * eng->nodes, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

// The end of a step loop: with t_inout, wait, return the time, and report a divergence
// (checkForNan, ExplicitMethods.f90:786: the reference stops there) as HX3D_ENAN.
static int finish_steps(hx3d_engine *eng, double *t_inout)
{
    HIPCHK(hipGetLastError());
    if (!t_inout) return HX3D_OK;
    HIPCHK(hipMemcpyAsync(t_inout, eng->dScal + 1, sizeof(double), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    if (*eng->hFlag) return eng->fail(HX3D_ENAN, "numerical divergence: a NaN appeared in the state");
    return HX3D_OK;
}

int hx3d_advance_dt(hx3d_engine *eng, int nsteps, double dt, double *t_inout)
{
    if (!eng || nsteps < 0 || !(dt > 0.0)) return HX3D_EINVAL;
    if (eng->mc.blocks && t_inout) return eng->fail(HX3D_EINVAL, "several cases: pass no time, read them with hx3d_multi_status");
    if (t_inout) HIPCHK(hipMemcpyAsync(eng->dScal + 1, t_inout, sizeof(double), hipMemcpyHostToDevice, eng->stream));
    DtArgs a = host_dt(eng, dt);
    a.nan = eng->dFlag + 1;
    for (int s = 0; s < nsteps; ++s)
        for (int k = 0; k < 3; ++k) {
            DtArgs b = a;
            if (k == 2) b.time = eng->dTime;   // t += dt once per step, on the device
            stage(eng, RK_A[k], RK_C[k], 1, b, k == 2 && s == nsteps - 1 ? eng->dQDot : nullptr);
        }
    HIPCHK(hipGetLastError());
    if (nsteps > 0) after_steps(eng);
    return finish_steps(eng, t_inout);
}

int hx3d_max_dt(hx3d_engine *eng, double cfl, double dcfl, double *dt)
{
    if (!eng || !dt) return HX3D_EINVAL;
    if (eng->mc.blocks) return eng->fail(HX3D_EUNSUPPORTED, "several cases: each case's dt is formed on the device by hx3d_advance");
    launch_maxdt(eng, cfl, dcfl, 3);
    HIPCHK(hipMemcpyAsync(dt, eng->dScal, sizeof(double), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_advance(hx3d_engine *eng, int nsteps, double cfl, double dcfl, double *t_inout)
{
    if (!eng || nsteps < 0) return HX3D_EINVAL;
    if (eng->mc.blocks && t_inout) return eng->fail(HX3D_EINVAL, "several cases: pass no time, read them with hx3d_multi_status");
    if (t_inout) HIPCHK(hipMemcpyAsync(eng->dScal + 1, t_inout, sizeof(double), hipMemcpyHostToDevice, eng->stream));
    const size_t nev = eng->timing ? (size_t)12 * nsteps : 0;
    // timing events need no system-scope fence: they are read only through
    // hipEventElapsedTime, and the fence adds its own time to every measured launch
    while (eng->evpool.size() < nev) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        eng->evpool.push_back(e);
    }
    // dt of step s (slot 0): MaxTimeStep of the state the step starts from.  The first stage's
    // gradient kernel leaves block minima of the limits (an Euler case has no gradient kernel:
    // a standalone pass writes them), which block 0 of the first stage's face kernel reduces into
    // the slot before any volume kernel of the step reads dt.  Every step issues the same
    // launches with the same arguments: ev, when given, holds 12 events for the step.
    auto issue_step = [eng, cfl, dcfl](hipEvent_t *ev, double *qdot) {
        unsigned long long *slot = dt_slot(eng, 0);
        int nparts;
        if (eng->ph.euler) {
            const long nb = (eng->nodes + 255) / 256;
            if (eng->mixed)
                hipLaunchKernelGGL(hxm::km_maxdt, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, eng->dJa, eng->dJ,
                                   eng->dEofn, eng->dN3, eng->dOpsAll, eng->nodes, eng->S, cfl, dcfl, eng->ph,
                                   eng->dDtPart);
            else
                hipLaunchKernelGGL(k_maxdt, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, eng->dJa, eng->dJ,
                                   eng->nodes, eng->S, eng->dcsi, cfl, dcfl, eng->ph, eng->dDtPart);
            nparts = (int)nb;
        } else {
            nparts = eng->mixed ? eng->grid_mgrad_s + eng->grid_mgrad_b : grad_grid(eng);
        }
        for (int k = 0; k < 3; ++k) {
            DtArgs a{};
            a.in = slot;
            if (k == 0) {
                if (!eng->ph.euler) a.lim = hxk::DtLimits{eng->dDtPart, cfl, dcfl, eng->dcsi};
                a.red = hxk::DtReduce{eng->dDtPart, nparts, slot};
            }
            if (k == 2) a.time = eng->dTime;
            a.nan = eng->dFlag + 1;
            stage(eng, RK_A[k], RK_C[k], 1, a, k == 2 ? qdot : nullptr, ev ? ev + 4 * k : nullptr);
        }
    };
    if (eng->timing || !eng->use_graph) {
        for (int s = 0; s < nsteps; ++s)
            issue_step(eng->timing ? &eng->evpool[(size_t)12 * s] : nullptr, s == nsteps - 1 ? eng->dQDot : nullptr);
    } else if (nsteps > 0) {
        // capture the step once per (cfl, dcfl); the graph holds kernel launches only
        if (!eng->step_graph || eng->graph_cfl != cfl || eng->graph_dcfl != dcfl) {
            if (eng->step_graph) {
                HIPCHK(hipGraphExecDestroy(eng->step_graph));
                eng->step_graph = nullptr;
            }
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(eng->stream, hipStreamCaptureModeThreadLocal));
            eng->capturing = true;
            issue_step(nullptr, nullptr);
            eng->capturing = false;
            const hipError_t ce = hipStreamEndCapture(eng->stream, &g);
            if (ce != hipSuccess || !g) return eng->fail(HX3D_EDEVICE, std::string("step capture: ") + hipGetErrorString(ce));
            const hipError_t ie = hipGraphInstantiate(&eng->step_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);   // the instantiated graph does not need it
            if (ie != hipSuccess) {
                eng->step_graph = nullptr;
                return eng->fail(HX3D_EDEVICE, std::string("step graph: ") + hipGetErrorString(ie));
            }
            eng->graph_cfl = cfl;
            eng->graph_dcfl = dcfl;
        }
        // the last step is issued directly: it also leaves its last stage's time derivative
        for (int s = 0; s + 1 < nsteps; ++s) HIPCHK(hipGraphLaunch(eng->step_graph, eng->stream));
        issue_step(nullptr, eng->dQDot);
    }
    HIPCHK(hipGetLastError());
    if (eng->timing && nsteps > 0) {
        HIPCHK(hipStreamSynchronize(eng->stream));
        double acc[3] = {0, 0, 0};
        for (size_t x = 0; x < (size_t)3 * nsteps; ++x) {
            for (int kk = 0; kk < 3; ++kk) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, eng->evpool[4 * x + kk], eng->evpool[4 * x + kk + 1]));
                acc[kk] += ms;
            }
        }
        for (int x = 0; x < 3; ++x) eng->ktime[x] = acc[x] / (3.0 * nsteps);
        eng->ktime[3] = (eng->ktime[0] + eng->ktime[1] + eng->ktime[2]);
    }
    if (nsteps > 0) after_steps(eng);
    return finish_steps(eng, t_inout);
}

int hx3d_residuals(hx3d_engine *eng, double res[5])
{
    if (!eng || !res) return HX3D_EINVAL;
    if (!eng->qdot_valid) {
        int rc = hx3d_compute_qdot(eng, 0.0, nullptr);
        if (rc) return rc;
    }
    const int nb = (int)((eng->nodes + 255) / 256);
    hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(256), 0, eng->stream, eng->dQDot, eng->nodes, eng->S, eng->dPart);
    hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(64), 0, eng->stream, eng->dPart, nb, eng->dStage);
    HIPCHK(hipMemcpyAsync(res, eng->dStage, sizeof(double) * NV, hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_residuals_noalloc(hx3d_engine *eng, double *res_dev)
{
    if (!eng || !res_dev) return HX3D_EINVAL;
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != eng->device)
        return eng->fail(HX3D_EINVAL, "hx3d_residuals_noalloc: make the engine's device current first");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return eng->fail(HX3D_EUNSUPPORTED, "hx3d_residuals_noalloc: device index beyond the scratch table");
    hipError_t e = hipSuccess;
    double *part = noalloc_scratch(dev, &e);
    if (!part) return eng->fail(HX3D_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    if (!eng->qdot_valid) {   // the stage kernels run on the engine's own buffers (no allocation)
        int rc = hx3d_compute_qdot(eng, 0.0, nullptr);
        if (rc) return rc;
    }
    // at most NOALLOC_BLOCKS blocks stride over the nodes, so the scratch fits every node count
    const int nb = (int)std::min<long>((eng->nodes + 255) / 256, NOALLOC_BLOCKS);
    if (nb < 1 || nb > NOALLOC_BLOCKS)
        return eng->fail(HX3D_EINVAL, "hx3d_residuals_noalloc: partials beyond the scratch");
    hipLaunchKernelGGL(k_absmax_strided, dim3(nb), dim3(256), 0, eng->stream, eng->dQDot, eng->nodes, eng->S, part);
    hipLaunchKernelGGL(k_finish_max_block, dim3(1), dim3(256), 0, eng->stream, part, nb, res_dev);
    HIPCHK(hipGetLastError());
    return HX3D_OK;
}

int hx3d_kinetic_energy(hx3d_engine *eng, double *ke)
{
    if (!eng || !ke) return HX3D_EINVAL;
    const int nb = (int)((eng->nodes + 255) / 256);
    hipLaunchKernelGGL(k_kinetic, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, eng->dJ, eng->dWq, eng->nodes, eng->S, eng->dPart);
    hipLaunchKernelGGL(k_finish_sum2, dim3(1), dim3(64), 0, eng->stream, eng->dPart, nb, eng->dStage);
    double r[2];
    HIPCHK(hipMemcpyAsync(r, eng->dStage, sizeof(r), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    *ke = r[0] / r[1];
    return HX3D_OK;
}

// VolumeMonitor_Update (VolumeMonitor.f90:278-330): the integral over the volume; the enstrophy
// with its factor 1/2.  Rate and enstrophy use the derivative and gradients of the resident
// state (computed here if the last derivative is stale).
int hx3d_volume_monitor(hx3d_engine *eng, int kind, double *value)
{
    if (!eng || !value) return HX3D_EINVAL;
    if (kind == HX3D_MON_KINETIC_ENERGY) return hx3d_kinetic_energy(eng, value);
    if (kind < HX3D_MON_KINETIC_ENERGY_RATE || kind > HX3D_MON_MATH_ENTROPY)
        return eng->fail(HX3D_EINVAL, "unknown volume monitor");
    if (kind == HX3D_MON_ENSTROPHY && eng->ph.euler)
        return eng->fail(HX3D_EUNSUPPORTED, "enstrophy needs the gradients of a Navier-Stokes case");
    if ((kind == HX3D_MON_KINETIC_ENERGY_RATE || kind == HX3D_MON_ENSTROPHY) && !eng->qdot_valid) {
        int rc = hx3d_compute_qdot(eng, 0.0, nullptr);
        if (rc) return rc;
    }
    const int nb = (int)((eng->nodes + 255) / 256);
    hipLaunchKernelGGL(k_monitor, dim3(nb), dim3(256), 0, eng->stream, kind, eng->dQ, eng->dQDot,
                       eng->mixed ? nullptr : eng->dVgrad, eng->dU, eng->dJ, eng->dWq, eng->nodes, eng->S, eng->ph,
                       eng->dPart);
    hipLaunchKernelGGL(k_finish_sum2, dim3(1), dim3(64), 0, eng->stream, eng->dPart, nb, eng->dStage);
    double r[2];
    HIPCHK(hipMemcpyAsync(r, eng->dStage, sizeof(r), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    *value = (kind == HX3D_MON_ENSTROPHY ? 0.5 * r[0] : r[0]) / r[1];
    return HX3D_OK;
}

int hx3d_check_nan(hx3d_engine *eng, int *has_nan)
{
    if (!eng || !has_nan) return HX3D_EINVAL;
    int zero = 0;
    HIPCHK(hipMemcpyAsync(eng->dFlag, &zero, sizeof(int), hipMemcpyHostToDevice, eng->stream));
    const long nb = (eng->nodes + 255) / 256;
    hipLaunchKernelGGL(k_nan, dim3(nb), dim3(256), 0, eng->stream, eng->dQ, eng->nodes, eng->S, eng->dFlag);
    HIPCHK(hipMemcpyAsync(has_nan, eng->dFlag, sizeof(int), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_synchronize(hx3d_engine *eng)
{
    if (!eng) return HX3D_EINVAL;
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

int hx3d_query(hx3d_engine *eng, int *done)
{
    if (!eng || !done) return HX3D_EINVAL;
    const hipError_t e = hipStreamQuery(eng->stream);
    if (e == hipSuccess) *done = 1;
    else if (e == hipErrorNotReady) *done = 0;
    else return eng->fail(HX3D_EDEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(e));
    if (*done && !eng->mc.blocks && *eng->hFlag)
        return eng->fail(HX3D_ENAN, "numerical divergence: a NaN appeared in the state");
    return HX3D_OK;
}

int hx3d_get_time(hx3d_engine *eng, double *t)
{
    if (!eng || !t) return HX3D_EINVAL;
    HIPCHK(hipMemcpyAsync(t, eng->dScal + 1, sizeof(double), hipMemcpyDeviceToHost, eng->stream));
    HIPCHK(hipStreamSynchronize(eng->stream));
    return HX3D_OK;
}

long hx3d_engine_device_bytes(const hx3d_engine *eng) { return eng ? eng->dev_bytes : -1; }

int hx3d_device_memory(size_t *free_bytes, size_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return HX3D_EINVAL;
    return hipMemGetInfo(free_bytes, total_bytes) == hipSuccess ? HX3D_OK : HX3D_EDEVICE;
}

int hx3d_set_step_graph(hx3d_engine *eng, int on)
{
    if (!eng) return HX3D_EINVAL;
    eng->use_graph = on != 0;
    return HX3D_OK;
}

int hx3d_autotune(hx3d_engine *eng, int reps, int *choice, double *ms)
{
    if (!eng || reps < 1) return HX3D_EINVAL;
    if (!eng->mixed) {   // one tiling: nothing to choose
        if (choice) *choice = -1;
        return HX3D_OK;
    }
    constexpr int NPOL = 3;
    hipEvent_t ev[2];
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableSystemFence));
    if (hipEventCreateWithFlags(&ev[1], hipEventDisableSystemFence) != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return eng->fail(HX3D_EDEVICE, "hipEventCreate failed");
    }
    int best = -1, rc = HX3D_OK;
    float best_ms = 0.f;
    for (int pol = 0; pol < NPOL && rc == HX3D_OK; ++pol) {
        if ((rc = set_tiles(eng, pol))) break;
        // one untimed evaluation, then reps timed ones of the resident state's time derivative
        for (int r = 0; r <= reps; ++r) {
            if (r == 1) (void)hipEventRecord(ev[0], eng->stream);
            stage(eng, 0.0, 0.0, 0, host_dt(eng, 0.0), eng->dQDot);
        }
        (void)hipEventRecord(ev[1], eng->stream);
        float t = 0.f;
        if (hipEventSynchronize(ev[1]) != hipSuccess || hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess) {
            rc = eng->fail(HX3D_EDEVICE, "autotune timing failed");
            break;
        }
        if (ms) ms[pol] = t / reps;
        if (best < 0 || t < best_ms) { best = pol; best_ms = t; }
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
    if ((rc = set_tiles(eng, best))) return rc;
    if (eng->step_graph) {   // a captured step holds the old tiles
        HIPCHK(hipGraphExecDestroy(eng->step_graph));
        eng->step_graph = nullptr;
    }
    eng->qdot_valid = true;   // dQDot holds the resident state's derivative
    if (choice) *choice = best;
    return HX3D_OK;
}

int hx3d_set_tiling(hx3d_engine *eng, int policy)
{
    if (!eng || !eng->mixed) return HX3D_EINVAL;
    int rc = set_tiles(eng, policy);
    if (rc) return rc;
    if (eng->step_graph) {
        HIPCHK(hipGraphExecDestroy(eng->step_graph));
        eng->step_graph = nullptr;
    }
    return HX3D_OK;
}

int hx3d_get_tiling(const hx3d_engine *eng, int *policy, int *ntiles)
{
    if (!eng || !eng->mixed) return HX3D_EINVAL;
    if (policy) *policy = eng->tile_policy;
    if (ntiles) *ntiles = eng->ntiles;
    return HX3D_OK;
}

int hx3d_set_timing(hx3d_engine *eng, int on)
{
    if (!eng) return HX3D_EINVAL;
    eng->timing = on != 0;
    return HX3D_OK;
}

int hx3d_kernel_times(hx3d_engine *eng, double out_ms[4])
{
    if (!eng || !out_ms) return HX3D_EINVAL;
    for (int x = 0; x < 4; ++x) out_ms[x] = eng->ktime[x];
    return HX3D_OK;
}

}  // extern "C"
