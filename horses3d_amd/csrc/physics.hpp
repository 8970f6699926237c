// Point-wise compressible Navier-Stokes physics for the hx3d kernels (FP64, device + host).
// Each function restates the reference routine cited beside it (paths relative to
// Solver/src/libs/physics/) with the reference's operation order: the Taylor-Green cases run
// at Mach 0.1, where the background pressure makes the volume and surface terms cancel by
// three orders of magnitude, so the rounding sequence of the large terms is kept identical
// (the library is compiled without floating-point contraction).
#pragma once
#include <hip/hip_runtime.h>

#define HX_HD __host__ __device__ __forceinline__

namespace hx {

constexpr int NV = 5;

struct Phys {
    double gamma, gm1, gammaM2, mu, mu_to_kappa, S_div_T, Trenorm, lambda_stab, les_c, cp;
    int riemann, averaging, les, euler;
};

// navierstokes/VariableConversion_NS.f90:64 Pressure, :99-112 Temperature,
// :146-186 get_laminar_mu_kappa + SutherlandsLaw
HX_HD double pressure(const Phys &ph, const double q[NV])
{
    return ph.gm1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]);
}

HX_HD double sutherland(const Phys &ph, double T)
{
    const double tT = T * ph.Trenorm;
    return (1.0 + ph.S_div_T) / (tT + ph.S_div_T) * tT * sqrt(tT);
}

HX_HD void laminar_mu_kappa(const Phys &ph, const double q[NV], double &mu, double &kappa)
{
    const double T = ph.gammaM2 * pressure(ph, q) / q[0];
    mu = ph.mu * sutherland(ph, T);
    kappa = mu * ph.mu_to_kappa;
}

// MaxTimeStep limits of one node (DGSEMClass.f90:870-1035 with ComputeEigenvaluesForState,
// Physics_NS.f90:927-962): convective cfl |J| / sum_d |Ja_d . lambda| dxi_d^-1 and viscous
// dcfl |J| / sum_d mu(T) dxi_d^-2 |sum Ja_d| -- mu = SutherlandsLaw(T) without the 1/Re factor,
// as the reference has it (DGSEMClass.f90:990, 999-1002).
HX_HD void dt_limits(const Phys &ph, const double q[NV], const double ja[9], double jg, double dc, double de,
                     double dz, double cfl, double dcfl, double &tc, double &tv)
{
    const double u = fabs(q[1] / q[0]), v = fabs(q[2] / q[0]), w = fabs(q[3] / q[0]);
    const double p = pressure(ph, q);
    const double a = sqrt(ph.gamma * p / q[0]);
    const double e0 = u + a, e1 = v + a, e2 = w + a;
    const double lx = fabs(ja[0] * e0 + ja[1] * e1 + ja[2] * e2) * dc;
    const double ly = fabs(ja[3] * e0 + ja[4] * e1 + ja[5] * e2) * de;
    const double lz = fabs(ja[6] * e0 + ja[7] * e1 + ja[8] * e2) * dz;
    tc = cfl * fabs(jg) / (lx + ly + lz);
    tv = INFINITY;
    if (!ph.euler) {
        const double mu = sutherland(ph, ph.gammaM2 * p / q[0]);
        const double vx = mu * (dc * dc) * fabs(ja[0] + ja[1] + ja[2]);
        const double vy = mu * (de * de) * fabs(ja[3] + ja[4] + ja[5]);
        const double vz = mu * (dz * dz) * fabs(ja[6] + ja[7] + ja[8]);
        tv = dcfl * fabs(jg) / (vx + vy + vz);
    }
}

// navierstokes/Physics_NS.f90:51-97 EulerFlux, then the contravariant projection of
// BaseClass_ComputeInnerFluxes (discretization/HyperbolicDiscretizationClass.f90:83-152):
// out[d][v] = F_x[v] Ja^d_x + F_y[v] Ja^d_y + F_z[v] Ja^d_z
HX_HD void euler_contravariant(const Phys &ph, const double q[NV], const double ja[9], double out[3][NV])
{
    const double u = q[1] / q[0], v = q[2] / q[0], w = q[3] / q[0];
    const double p = ph.gm1 * (q[4] - 0.5 * (q[1] * u + q[2] * v + q[3] * w));
    double F[NV][3];
    F[0][0] = q[1]; F[1][0] = q[1] * u + p; F[2][0] = q[1] * v; F[3][0] = q[1] * w; F[4][0] = (q[4] + p) * u;
    F[0][1] = q[2]; F[1][1] = F[2][0]; F[2][1] = q[2] * v + p; F[3][1] = q[2] * w; F[4][1] = (q[4] + p) * v;
    F[0][2] = q[3]; F[1][2] = F[3][0]; F[2][2] = F[3][1]; F[3][2] = q[3] * w + p; F[4][2] = (q[4] + p) * w;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < NV; ++c)
            out[d][c] = F[c][0] * ja[3 * d] + F[c][1] * ja[3 * d + 1] + F[c][2] * ja[3 * d + 2];
}

// navierstokes/Physics_NS.f90:246-304 ViscousFlux_STATE (beta = 0), split in two halves:
//   viscous_state   the velocity gradients u_x, u_y, u_z and grad T it forms from Q and the
//                   lifted state gradients: vg[0..2] = u_x, vg[3..5] = u_y, vg[6..8] = u_z,
//                   vg[9..11] = grad T;
//   viscous_flux    the stress / heat-flux tail F[v][c] from mu, kappa, the velocity and vg.
// Together they perform the reference's operations in its order.
HX_HD void viscous_state(const Phys &ph, const double q[NV], const double gx[NV], const double gy[NV],
                         const double gz[NV], double vg[12])
{
    const double ir = 1.0 / q[0];
    const double u = q[1] * ir, v = q[2] * ir, w = q[3] * ir;
    const double ud[3] = {u * ir, v * ir, w * ir};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        vg[d] = ir * gx[1 + d] - ud[d] * gx[0];
        vg[3 + d] = ir * gy[1 + d] - ud[d] * gy[0];
        vg[6 + d] = ir * gz[1 + d] - ud[d] * gz[0];
    }
    const double g = ph.gm1 * ph.gammaM2;
    vg[9] = g * (ir * gx[4] - q[4] * ir * ir * gx[0] - u * vg[0] - v * vg[1] - w * vg[2]);
    vg[10] = g * (ir * gy[4] - q[4] * ir * ir * gy[0] - u * vg[3] - v * vg[4] - w * vg[5]);
    vg[11] = g * (ir * gz[4] - q[4] * ir * ir * gz[0] - u * vg[6] - v * vg[7] - w * vg[8]);
}

HX_HD void viscous_flux(double mu, double kappa, double u, double v, double w, const double vg[12],
                        double F[NV][3])
{
    const double *ux = vg, *uy = vg + 3, *uz = vg + 6;
    const double divV = ux[0] + uy[1] + uz[2];
    F[0][0] = 0.0;
    F[1][0] = mu * (2.0 * ux[0] - 2.0 / 3.0 * divV);
    F[2][0] = mu * (ux[1] + uy[0]);
    F[3][0] = mu * (ux[2] + uz[0]);
    F[4][0] = F[1][0] * u + F[2][0] * v + F[3][0] * w + kappa * vg[9];
    F[0][1] = 0.0;
    F[1][1] = F[2][0];
    F[2][1] = mu * (2.0 * uy[1] - 2.0 / 3.0 * divV);
    F[3][1] = mu * (uy[2] + uz[1]);
    F[4][1] = F[1][1] * u + F[2][1] * v + F[3][1] * w + kappa * vg[10];
    F[0][2] = 0.0;
    F[1][2] = F[3][0];
    F[2][2] = F[3][1];
    F[3][2] = mu * (2.0 * uz[2] - 2.0 / 3.0 * divV);
    F[4][2] = F[1][2] * u + F[2][2] * v + F[3][2] * w + kappa * vg[11];
}

// common/LESModels.f90:256-307 Smagorinsky and :358-~440 WALE eddy viscosity (no wall model),
// mu_t = rho (C delta)^2 X, split in two: les_factor gives X, the part that does not depend on
// the filter width delta (the element's or the face's), from the state and its conservative
// gradients with the velocity gradients formed as getVelocityGradients_State forms them
// (VariableConversion_NS.f90:373-392: uDivRho = Q(IRHOU:IRHOW) invRho^2); les_mu applies
// rho (C delta)^2 in the reference's order.
HX_HD double les_factor(const Phys &ph, const double q[NV], const double gx[NV], const double gy[NV],
                        const double gz[NV])
{
    const double invRho = 1.0 / q[0], invRho2 = invRho * invRho;
    const double uDR[3] = {q[1] * invRho2, q[2] * invRho2, q[3] * invRho2};
    double ux[3], uy[3], uz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ux[d] = invRho * gx[1 + d] - uDR[d] * gx[0];
        uy[d] = invRho * gy[1 + d] - uDR[d] * gy[0];
        uz[d] = invRho * gz[1 + d] - uDR[d] * gz[0];
    }
    if (ph.les == 1) {
        // S(:,1) = U_x, S(:,2) = U_y, S(:,3) = U_z; S(1,:) += U_x ...; S = S/2; |S| = sqrt(2 sum S*S)
        double S[3][3];
#pragma unroll
        for (int a = 0; a < 3; ++a) { S[a][0] = ux[a]; S[a][1] = uy[a]; S[a][2] = uz[a]; }
#pragma unroll
        for (int b = 0; b < 3; ++b) { S[0][b] = S[0][b] + ux[b]; S[1][b] = S[1][b] + uy[b]; S[2][b] = S[2][b] + uz[b]; }
        double ss = 0.0;
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int a = 0; a < 3; ++a) { const double s = 0.5 * S[a][b]; ss += s * s; }
        return sqrt(2.0 * ss);
    }
    if (ph.les == 2) {
        double G[3][3], S[3][3], G2[3][3], Sd[3][3];
#pragma unroll
        for (int b = 0; b < 3; ++b) { G[0][b] = ux[b]; G[1][b] = uy[b]; G[2][b] = uz[b]; }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                S[i][j] = 0.5 * (G[i][j] + G[j][i]);
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc = acc + G[i][k] * G[k][j];
                G2[i][j] = acc;
            }
        const double div2 = G2[0][0] + G2[1][1] + G2[2][2];
        double normS = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) normS += S[i][j] * S[i][j];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) Sd[i][j] = 0.5 * (G2[i][j] + G2[j][i]);
#pragma unroll
        for (int i = 0; i < 3; ++i) Sd[i][i] = Sd[i][i] - 1.0 / 3.0 * div2;
        double normSd = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) normSd += Sd[i][j] * Sd[i][j];
        if (normS < 1.0e-8 && normSd < 1.0e-8) return 0.0;   // mu = 0
        return pow(normSd, 1.5) / (pow(normS, 2.5) + pow(normSd, 1.25));
    }
    return 0.0;
}

HX_HD double les_mu(const Phys &ph, double delta, double rho, double X)
{
    const double LS = ph.les_c * delta;
    return rho * (LS * LS) * X;
}

// navierstokes/RiemannSolvers_NS.f90:1541-1656 RoeRiemannSolver: the one-sided Roe flux with
// the entropy fix of the one wave that can change sign, the reference's operation order.
HX_HD void roe_flux(const Phys &ph, const double L[NV], const double R[NV], const double n[3], double f[NV])
{
    const double g = ph.gamma;
    const double rL = L[0], mxL = L[1], myL = L[2], mzL = L[3], EL = L[4];
    const double rR = R[0], mxR = R[1], myR = R[2], mzR = R[3], ER = R[4];
    const double uL = mxL / rL, vL = myL / rL, wL = mzL / rL;
    const double pL = (g - 1.0) * (EL - 0.5 / rL * (mxL * mxL + myL * myL + mzL * mzL));
    const double uR = mxR / rR, vR = myR / rR, wR = mzR / rR;
    const double pR = (g - 1.0) * (ER - 0.5 / rR * (mxR * mxR + myR * myR + mzR * mzR));
    const double unL = n[0] * uL + n[1] * vL + n[2] * wL;
    const double unR = n[0] * uR + n[1] * vR + n[2] * wR;
    const double HL = 0.5 * (uL * uL + vL * vL + wL * wL) + g / (g - 1.0) * pL / rL;
    const double HR = 0.5 * (uR * uR + vR * vR + wR * wR) + g / (g - 1.0) * pR / rR;
    // Roe averages: weights sqrt(rho) of each side
    const double srr = sqrt(rL * rR);
    const double wtL = rL / (rL + srr), wtR = 1.0 - wtL;
    const double ua = wtL * uL + wtR * uR, va = wtL * vL + wtR * vR, wa = wtL * wL + wtR * wR;
    const double Ha = wtL * HL + wtR * HR;
    const double c2 = (g - 1.0) * (Ha - 0.5 * (ua * ua + va * va + wa * wa));
    const double ca = sqrt(c2);
    const double una = ua * n[0] + va * n[1] + wa * n[2];
    if (una >= 0.0) {
        // left state plus the u - c wave (amplitude a1, speed lam1, entropy fix fix1)
        const double a1 = 0.5 * ((pR - pL) / c2 - (unR - unL) * srr / ca);
        const double lam1 = una - ca;
        const double lam1n = lam1 < 0.0 ? lam1 : 0.0;
        const double fix1 = ((g + 1.0) / 4.0 * ca / srr) * a1;
        const double x1 = -fabs(lam1) - fix1;
        const double eps1 = x1 > 0.0 ? x1 : 0.0;
        const double s1 = a1 * (lam1n - 0.5 * eps1);
        const double mn = rL * unL;
        f[0] = mn + s1;
        f[1] = mn * uL + pL * n[0] + s1 * (ua - ca * n[0]);
        f[2] = mn * vL + pL * n[1] + s1 * (va - ca * n[1]);
        f[3] = mn * wL + pL * n[2] + s1 * (wa - ca * n[2]);
        f[4] = mn * HL + s1 * (Ha - una * ca);
    } else {
        // right state minus the u + c wave
        const double a4 = 0.5 * ((pR - pL) / c2 + (unR - unL) * srr / ca);
        const double lam4 = una + ca;
        const double lam4p = lam4 > 0.0 ? lam4 : 0.0;
        const double fix4 = ((g + 1.0) / 4.0 * ca / srr) * a4;
        const double x4 = -fabs(lam4) + fix4;
        const double eps4 = x4 > 0.0 ? x4 : 0.0;
        const double s4 = a4 * (lam4p + 0.5 * eps4);
        const double mn = rR * unR;
        f[0] = mn - s4;
        f[1] = mn * uR + pR * n[0] - s4 * (ua + ca * n[0]);
        f[2] = mn * vR + pR * n[1] - s4 * (va + ca * n[1]);
        f[3] = mn * wR + pR * n[2] - s4 * (wa + ca * n[2]);
        f[4] = mn * HR - s4 * (Ha + una * ca);
    }
}

// AveragedStates (RiemannSolvers_NS.f90:1784-2077) of rotated states (rho, rho u_n, rho u_t1,
// rho u_t2, rho e)
HX_HD void averaged_states(const Phys &ph, const double QL[NV], const double QR[NV], double pL,
                           double pR, double iL, double iR, double f[NV])
{
    const double uL = iL * QL[1], uR = iR * QR[1], vL = iL * QL[2], vR = iR * QR[2];
    const double wL = iL * QL[3], wR = iR * QR[3];
    switch (ph.averaging) {
    case 1:
        f[0] = 0.25 * (QL[0] + QR[0]) * (uL + uR);
        f[1] = 0.25 * (QL[1] + QR[1]) * (uL + uR) + 0.5 * (pL + pR);
        f[2] = 0.25 * (QL[2] + QR[2]) * (uL + uR);
        f[3] = 0.25 * (QL[3] + QR[3]) * (uL + uR);
        f[4] = 0.25 * (QL[4] + pL + QR[4] + pR) * (uL + uR);
        break;
    case 2: {
        const double rho = 0.5 * (QL[0] + QR[0]), u = 0.5 * (uL + uR), v = 0.5 * (vL + vR);
        const double w = 0.5 * (wL + wR), p = 0.5 * (pL + pR), e = 0.5 * (QL[4] * iL + QR[4] * iR);
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = rho * u * w;
        f[4] = rho * u * e + p * u;
        break; }
    case 3: {
        const double rho = 0.5 * (QL[0] + QR[0]), u = 0.5 * (uL + uR), v = 0.5 * (vL + vR);
        const double w = 0.5 * (wL + wR), p = 0.5 * (pL + pR);
        const double h = 0.5 * ((QL[4] + pL) * iL + (QR[4] + pR) * iR);
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = rho * u * w;
        f[4] = rho * u * h;
        break; }
    case 4: {
        const double hL = ph.cp * pL, hR = ph.cp * pR;
        f[0] = 0.5 * (QL[1] + QR[1]);
        f[1] = 0.25 * (QL[1] + QR[1]) * (uL + uR) + 0.5 * (pL + pR);
        f[2] = 0.25 * (QL[1] + QR[1]) * (vL + vR);
        f[3] = 0.25 * (QL[1] + QR[1]) * (wL + wR);
        f[4] = 0.5 * (uL * hL + uR * hR) + 0.25 * (QL[1] * uL + QR[1] * uR) * (uL + uR)
             + 0.25 * (QL[1] * vL + QR[1] * vR) * (vL + vR)
             + 0.25 * (QL[1] * wL + QR[1] * wR) * (wL + wR)
             - 0.25 * (QL[1] * (uL * uL) + QR[1] * (uR * uR))
             - 0.25 * (QL[1] * (vL * vL) + QR[1] * (vR * vR))
             - 0.25 * (QL[1] * (wL * wL) + QR[1] * (wR * wR));
        break; }
    default:
        f[0] = 0.5 * (QL[1] + QR[1]);
        f[1] = 0.5 * (QL[1] * uL + QR[1] * uR + pL + pR);
        f[2] = 0.5 * (QL[1] * vL + QR[1] * vR);
        f[3] = 0.5 * (QL[1] * wL + QR[1] * wR);
        f[4] = 0.5 * (uL * (QL[4] + pL) + uR * (QR[4] + pR));
    }
}

// navierstokes/RiemannSolvers_NS.f90:1251-1334 LxFRiemannSolver (rotated face frame)
HX_HD void lxf_flux(const Phys &ph, const double QL[NV], const double QR[NV], const double n[3],
                    const double t1[3], const double t2[3], double f[NV])
{
    const double rhoL = QL[0], rhoR = QR[0];
    const double iL = 1.0 / rhoL, iR = 1.0 / rhoR;
    const double ruL = QL[1] * n[0] + QL[2] * n[1] + QL[3] * n[2];
    const double ruR = QR[1] * n[0] + QR[2] * n[1] + QR[3] * n[2];
    const double rvL = QL[1] * t1[0] + QL[2] * t1[1] + QL[3] * t1[2];
    const double rvR = QR[1] * t1[0] + QR[2] * t1[1] + QR[3] * t1[2];
    const double rwL = QL[1] * t2[0] + QL[2] * t2[1] + QL[3] * t2[2];
    const double rwR = QR[1] * t2[0] + QR[2] * t2[1] + QR[3] * t2[2];
    const double V2L = (ruL * ruL + rvL * rvL + rwL * rwL) * iL;
    const double V2R = (ruR * ruR + rvR * rvR + rwR * rwR) * iR;
    const double pL = ph.gm1 * (QL[4] - 0.5 * V2L), pR = ph.gm1 * (QR[4] - 0.5 * V2R);
    const double aL = sqrt(ph.gamma * pL * iL), aR = sqrt(ph.gamma * pR * iR);
    const double l1 = fabs(ruL * iL) + aL, l2 = fabs(ruR * iR) + aR;
    const double lam = l1 > l2 ? l1 : l2;
    const double Lr[NV] = {rhoL, ruL, rvL, rwL, QL[4]}, Rr[NV] = {rhoR, ruR, rvR, rwR, QR[4]};
    averaged_states(ph, Lr, Rr, pL, pR, iL, iR, f);
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v] = f[v] - ph.lambda_stab * (0.5 * lam * (Rr[v] - Lr[v]));
    const double f1 = f[1], f2 = f[2], f3 = f[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) f[1 + d] = n[d] * f1 + t1[d] * f2 + t2[d] * f3;
}

// Two-point volume fluxes, RiemannSolvers_NS.f90:2085-2438 (Standard, Ducros, Kennedy-Gruber,
// Pirozzoli): fSharp = f Ja_x + g Ja_y + h Ja_z of the averaged states and metrics, with the
// reference's operation order.  node_data() forms, once per node, what the flux of that
// averaging reads of a node (u, v, w, p as RiemannSolvers_NS.f90:2323-2329; the energy factors
// e = rhoE/rho and h = (rhoE + p)/rho of Kennedy-Gruber and Pirozzoli are the products the
// reference forms inside the flux, taken out of the pair loop).  The reference's constant
// factors (1/2 of each average, 1/4 of Ducros' products and of Standard's sum) are powers of two
// and are applied to the node data and to the metric rows (metric_scale) instead: scaling by a
// power of two commutes with rounding, so 0.5 (a + b) and a/2 + b/2 round to the same double
// and every flux keeps the reference's bits, with 9-18 fewer multiplications per pair.
constexpr int node_data_size(int avg) { return avg == 0 || avg == 4 ? 8 : avg == 1 ? 9 : 6; }
constexpr double metric_scale(int avg) { return avg == 0 ? 0.25 : 0.5; }

template <int AVG>
HX_HD void node_data(const Phys &ph, const double q[NV], double *a)
{
    const double ir = 1.0 / q[0];
    const double u = ir * q[1], v = ir * q[2], w = ir * q[3];
    const double p = ph.gm1 * (q[4] - 0.5 * (q[1] * u + q[2] * v + q[3] * w));
    if (AVG == 0) {          // rho u, rho v, rho w, rho e, u, v, w, p
        a[0] = q[1]; a[1] = q[2]; a[2] = q[3]; a[3] = q[4];
        a[4] = u; a[5] = v; a[6] = w; a[7] = p;
    } else if (AVG == 4) {   // rho u, rho v, rho w, u, v, w, p, h = cp p (no kinetic energy)
        a[0] = q[1]; a[1] = q[2]; a[2] = q[3];
        a[3] = u; a[4] = v; a[5] = w; a[6] = p; a[7] = ph.cp * p;
    } else if (AVG == 1) {   // q/4, u, v, w, p/4
        a[0] = 0.25 * q[0]; a[1] = 0.25 * q[1]; a[2] = 0.25 * q[2]; a[3] = 0.25 * q[3];
        a[4] = 0.25 * q[4]; a[5] = u; a[6] = v; a[7] = w; a[8] = 0.25 * p;
    } else {                 // (rho, u, v, w, p, e (KG) or h (Pirozzoli)) / 2
        a[0] = 0.5 * q[0]; a[1] = 0.5 * u; a[2] = 0.5 * v; a[3] = 0.5 * w; a[4] = 0.5 * p;
        a[5] = 0.5 * (AVG == 2 ? q[4] * ir : (q[4] + p) * ir);
    }
}

// Kennedy-Gruber, Pirozzoli and Morinishi are symmetric in (a, b) in IEEE arithmetic (every
// term is a sum or product of one a-term and the matching b-term); Standard and Ducros are not,
// and the reference evaluates every pair with the lower node index as a.
constexpr bool two_point_symmetric(int avg) { return avg >= 2; }

// a, b: node_data of the two nodes; ja, jb: their metric rows times metric_scale
template <int AVG>
HX_HD void two_point(const double *a, const double *ja, const double *b, const double *jb, double fs[NV])
{
    double f[NV], g[NV], h[NV], J[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) J[d] = ja[d] + jb[d];
    if (AVG == 4) {
        // Morinishi_TwoPointFlux (RiemannSolvers_NS.f90:2147-2229): direction d's flux from
        // each side's rho u_d (m), with the reference's factors 1/2 and 1/4 in place
        const double su = a[3] + b[3], sv = a[4] + b[4], sw = a[5] + b[5];
        const double uL = a[3], vL = a[4], wL = a[5], uR = b[3], vR = b[4], wR = b[5];
        const double pa = 0.5 * (a[6] + b[6]);
        double *o3[3] = {f, g, h};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double mL = a[d], mR = b[d];
            double *o = o3[d];
            const double s = 0.25 * (mL + mR);
            o[0] = 0.5 * (mL + mR);
            o[1] = s * su;
            o[2] = s * sv;
            o[3] = s * sw;
            o[1 + d] = o[1 + d] + pa;
            o[4] = 0.5 * (a[3 + d] * a[7] + b[3 + d] * b[7]) + 0.25 * (mL * uL + mR * uR) * su
                 + 0.25 * (mL * vL + mR * vR) * sv
                 + 0.25 * (mL * wL + mR * wR) * sw
                 - 0.25 * (mL * (uL * uL) + mR * (uR * uR))
                 - 0.25 * (mL * (vL * vL) + mR * (vR * vR))
                 - 0.25 * (mL * (wL * wL) + mR * (wR * wR));
        }
    } else if (AVG == 0) {
        const double uL = a[4], vL = a[5], wL = a[6], pL = a[7];
        const double uR = b[4], vR = b[5], wR = b[6], pR = b[7];
        f[0] = (a[0] + b[0]);
        f[1] = (a[0] * uL + b[0] * uR + pL + pR);
        f[2] = (a[0] * vL + b[0] * vR);
        f[3] = (a[0] * wL + b[0] * wR);
        f[4] = (uL * (a[3] + pL) + uR * (b[3] + pR));
        g[0] = (a[1] + b[1]);
        g[1] = (a[1] * uL + b[1] * uR);
        g[2] = (a[1] * vL + b[1] * vR + pL + pR);
        g[3] = (a[1] * wL + b[1] * wR);
        g[4] = (vL * (a[3] + pL) + vR * (b[3] + pR));
        h[0] = (a[2] + b[2]);
        h[1] = (a[2] * uL + b[2] * uR);
        h[2] = (a[2] * vL + b[2] * vR);
        h[3] = (a[2] * wL + b[2] * wR + pL + pR);
        h[4] = (wL * (a[3] + pL) + wR * (b[3] + pR));
    } else if (AVG == 1) {
        const double su = a[5] + b[5], sv = a[6] + b[6], sw = a[7] + b[7];
        const double p2 = 2.0 * (a[8] + b[8]);   // 0.5 (pL + pR)
        const double e4 = a[4] + a[8] + b[4] + b[8];
        f[0] = (a[0] + b[0]) * su;
        f[1] = (a[1] + b[1]) * su + p2;
        f[2] = (a[2] + b[2]) * su;
        f[3] = (a[3] + b[3]) * su;
        f[4] = e4 * su;
        g[0] = (a[0] + b[0]) * sv;
        g[1] = (a[1] + b[1]) * sv;
        g[2] = (a[2] + b[2]) * sv + p2;
        g[3] = (a[3] + b[3]) * sv;
        g[4] = e4 * sv;
        h[0] = (a[0] + b[0]) * sw;
        h[1] = (a[1] + b[1]) * sw;
        h[2] = (a[2] + b[2]) * sw;
        h[3] = (a[3] + b[3]) * sw + p2;
        h[4] = e4 * sw;
    } else {
        const double rho = a[0] + b[0];
        const double u = a[1] + b[1], v = a[2] + b[2], w = a[3] + b[3];
        const double p = a[4] + b[4];
        const double e = a[5] + b[5];
        f[0] = rho * u; f[1] = rho * u * u + p; f[2] = rho * u * v; f[3] = rho * u * w;
        g[0] = rho * v; g[1] = rho * v * u; g[2] = rho * v * v + p; g[3] = rho * v * w;
        h[0] = rho * w; h[1] = rho * w * u; h[2] = rho * w * v; h[3] = rho * w * w + p;
        if (AVG == 2) {
            f[4] = rho * u * e + p * u; g[4] = rho * v * e + p * v; h[4] = rho * w * e + p * w;
        } else {
            f[4] = rho * u * e; g[4] = rho * v * e; h[4] = rho * w * e;
        }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) fs[q] = f[q] * J[0] + g[q] * J[1] + h[q] * J[2];
}

// The velocity and pressure euler_contravariant forms (the same operations).
HX_HD void euler_primitives(const Phys &ph, const double q[NV], double &u, double &v, double &w, double &p)
{
    u = q[1] / q[0];
    v = q[2] / q[0];
    w = q[3] / q[0];
    p = ph.gm1 * (q[4] - 0.5 * (q[1] * u + q[2] * v + q[3] * w));
}

// One direction of euler_contravariant from its primitives (the same operations for row d).
HX_HD void euler_contravariant_row(const double q[NV], double u, double v, double w, double p, const double jd[3],
                                   double out[NV])
{
    const double F[NV][3] = {{q[1], q[2], q[3]},
                             {q[1] * u + p, q[1] * v, q[1] * w},
                             {q[1] * v, q[2] * v + p, q[2] * w},
                             {q[1] * w, q[2] * w, q[3] * w + p},
                             {(q[4] + p) * u, (q[4] + p) * v, (q[4] + p) * w}};
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = F[c][0] * jd[0] + F[c][1] * jd[1] + F[c][2] * jd[2];
}

}  // namespace hx
