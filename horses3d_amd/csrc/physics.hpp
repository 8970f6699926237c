// Point-wise compressible Navier-Stokes physics for the hx3d kernels (FP64, device + host).
// Each function restates the reference routine cited beside it (paths relative to
// Solver/src/libs/physics/).  Where the reference works in a rotated face frame, the
// kernels use the equivalent rotation-invariant Cartesian form (noted per routine).
#pragma once
#include <hip/hip_runtime.h>

#define HX_HD __host__ __device__ __forceinline__

namespace hx {

constexpr int NV = 5;

struct Phys {
    double gamma, gm1, gammaM2, mu, mu_to_kappa, S_div_T, Trenorm, lambda_stab, les_c;
    int riemann, averaging, les, euler;
};

// navierstokes/VariableConversion_NS.f90:146-186 get_laminar_mu_kappa + SutherlandsLaw
HX_HD void laminar_mu_kappa(const Phys &ph, const double q[NV], double &mu, double &kappa)
{
    double p = ph.gm1 * (q[4] - 0.5 * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) / q[0]);
    double T = ph.gammaM2 * p / q[0];
    double tT = T * ph.Trenorm;
    double s = (1.0 + ph.S_div_T) / (tT + ph.S_div_T) * tT * sqrt(tT);
    mu = ph.mu * s;
    kappa = mu * ph.mu_to_kappa;
}

// navierstokes/Physics_NS.f90:51-97 EulerFlux, contracted with the three metric rows:
// out[d][v] = sum_c F_c[v] * Ja^d_c
HX_HD void euler_contravariant(const Phys &ph, const double q[NV], const double ja[9], double out[3][NV])
{
    double ir = 1.0 / q[0];
    double u = q[1] * ir, v = q[2] * ir, w = q[3] * ir;
    double p = ph.gm1 * (q[4] - 0.5 * (q[1] * u + q[2] * v + q[3] * w));
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

// navierstokes/Physics_NS.f90:246-304 ViscousFlux_STATE (beta = 0).  F[v][c].
HX_HD void viscous_flux(const Phys &ph, const double q[NV], const double gx[NV], const double gy[NV],
                        const double gz[NV], double mu, double kappa, double F[NV][3])
{
    double ir = 1.0 / q[0];
    double u = q[1] * ir, v = q[2] * ir, w = q[3] * ir;
    double ud[3] = {u * ir, v * ir, w * ir};
    double ux[3], uy[3], uz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ux[d] = ir * gx[1 + d] - ud[d] * gx[0];
        uy[d] = ir * gy[1 + d] - ud[d] * gy[0];
        uz[d] = ir * gz[1 + d] - ud[d] * gz[0];
    }
    double g = ph.gm1 * ph.gammaM2;
    double eir2 = q[4] * ir * ir;
    double nTx = g * (ir * gx[4] - eir2 * gx[0] - u * ux[0] - v * ux[1] - w * ux[2]);
    double nTy = g * (ir * gy[4] - eir2 * gy[0] - u * uy[0] - v * uy[1] - w * uy[2]);
    double nTz = g * (ir * gz[4] - eir2 * gz[0] - u * uz[0] - v * uz[1] - w * uz[2]);
    double divV = ux[0] + uy[1] + uz[2];
    double t11 = mu * (2.0 * ux[0] - 2.0 / 3.0 * divV);
    double t22 = mu * (2.0 * uy[1] - 2.0 / 3.0 * divV);
    double t33 = mu * (2.0 * uz[2] - 2.0 / 3.0 * divV);
    double t12 = mu * (ux[1] + uy[0]);
    double t13 = mu * (ux[2] + uz[0]);
    double t23 = mu * (uy[2] + uz[1]);
    F[0][0] = 0.0; F[1][0] = t11; F[2][0] = t12; F[3][0] = t13;
    F[4][0] = t11 * u + t12 * v + t13 * w + kappa * nTx;
    F[0][1] = 0.0; F[1][1] = t12; F[2][1] = t22; F[3][1] = t23;
    F[4][1] = t12 * u + t22 * v + t23 * w + kappa * nTy;
    F[0][2] = 0.0; F[1][2] = t13; F[2][2] = t23; F[3][2] = t33;
    F[4][2] = t13 * u + t23 * v + t33 * w + kappa * nTz;
}

// common/LESModels.f90:256-307 Smagorinsky and :358-~440 WALE eddy viscosity (no wall
// model), with the velocity gradients of getVelocityGradients_State
// (navierstokes/VariableConversion_NS.f90:373-392).
HX_HD double les_viscosity(const Phys &ph, double delta, const double q[NV], const double gx[NV],
                           const double gy[NV], const double gz[NV])
{
    double ir = 1.0 / q[0], ir2 = ir * ir;
    double ud[3] = {q[1] * ir2, q[2] * ir2, q[3] * ir2};
    double G[3][3];   // G[a][b] = d u_b / d x_a
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        G[0][b] = ir * gx[1 + b] - ud[b] * gx[0];
        G[1][b] = ir * gy[1 + b] - ud[b] * gy[0];
        G[2][b] = ir * gz[1 + b] - ud[b] * gz[0];
    }
    double LS = ph.les_c * delta;
    if (ph.les == 1) {
        double ss = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double s = 0.5 * (G[b][a] + G[a][b]);
                ss += s * s;
            }
        return q[0] * LS * LS * sqrt(2.0 * ss);
    }
    if (ph.les == 2) {
        double G2[3][3], normS = 0.0, normSd = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double s = 0.5 * (G[i][j] + G[j][i]);
                normS += s * s;
                G2[i][j] = G[i][0] * G[0][j] + G[i][1] * G[1][j] + G[i][2] * G[2][j];
            }
        double div2 = G2[0][0] + G2[1][1] + G2[2][2];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double sd = 0.5 * (G2[i][j] + G2[j][i]) - (i == j ? div2 / 3.0 : 0.0);
                normSd += sd * sd;
            }
        if (normS < 1.0e-8 && normSd < 1.0e-8) return 0.0;
        return q[0] * LS * LS * (pow(normSd, 1.5) / (pow(normS, 2.5) + pow(normSd, 1.25)));
    }
    return 0.0;
}

// navierstokes/RiemannSolvers_NS.f90:1541-1656 RoeRiemannSolver
HX_HD void roe_flux(const Phys &ph, const double L[NV], const double R[NV], const double n[3], double f[NV])
{
    const double g = ph.gamma, gm1 = ph.gm1;
    double irL = 1.0 / L[0], irR = 1.0 / R[0];
    double ul = L[1] * irL, vl = L[2] * irL, wl = L[3] * irL;
    double ur = R[1] * irR, vr = R[2] * irR, wr = R[3] * irR;
    double pl = gm1 * (L[4] - 0.5 * irL * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]));
    double pr = gm1 * (R[4] - 0.5 * irR * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]));
    double ql = n[0] * ul + n[1] * vl + n[2] * wl;
    double qr = n[0] * ur + n[1] * vr + n[2] * wr;
    double gg = g / gm1;
    double hl = 0.5 * (ul * ul + vl * vl + wl * wl) + gg * pl * irL;
    double hr = 0.5 * (ur * ur + vr * vr + wr * wr) + gg * pr * irR;
    double rtd = sqrt(L[0] * R[0]);
    double bl = L[0] / (L[0] + rtd), br = 1.0 - bl;
    double utd = bl * ul + br * ur, vtd = bl * vl + br * vr, wtd = bl * wl + br * wr;
    double htd = bl * hl + br * hr;
    double atd2 = gm1 * (htd - 0.5 * (utd * utd + vtd * vtd + wtd * wtd));
    double atd = sqrt(atd2);
    double qtd = utd * n[0] + vtd * n[1] + wtd * n[2];
    double k = (g + 1.0) / 4.0 * atd / rtd;
    if (qtd >= 0.0) {
        double dw = 0.5 * ((pr - pl) / atd2 - (qr - ql) * rtd / atd);
        double sp = qtd - atd;
        double spm = sp < 0.0 ? sp : 0.0;
        double e = -fabs(sp) - k * dw;
        double eta = e > 0.0 ? e : 0.0;
        double udw = dw * (spm - 0.5 * eta);
        double rq = L[0] * ql;
        f[0] = rq + udw;
        f[1] = rq * ul + pl * n[0] + udw * (utd - atd * n[0]);
        f[2] = rq * vl + pl * n[1] + udw * (vtd - atd * n[1]);
        f[3] = rq * wl + pl * n[2] + udw * (wtd - atd * n[2]);
        f[4] = rq * hl + udw * (htd - qtd * atd);
    } else {
        double dw = 0.5 * ((pr - pl) / atd2 + (qr - ql) * rtd / atd);
        double sp = qtd + atd;
        double spp = sp > 0.0 ? sp : 0.0;
        double e = -fabs(sp) + k * dw;
        double eta = e > 0.0 ? e : 0.0;
        double udw = dw * (spp + 0.5 * eta);
        double rq = R[0] * qr;
        f[0] = rq - udw;
        f[1] = rq * ur + pr * n[0] - udw * (utd + atd * n[0]);
        f[2] = rq * vr + pr * n[1] - udw * (vtd + atd * n[1]);
        f[3] = rq * wr + pr * n[2] - udw * (wtd + atd * n[2]);
        f[4] = rq * hr - udw * (htd + qtd * atd);
    }
}

// navierstokes/RiemannSolvers_NS.f90:1251-1334 LxFRiemannSolver with AveragedStates
// (:1784-2077).  The reference rotates into (n, t1, t2), averages, and rotates the momentum
// back; every averaging is rotation-covariant, so the flux is evaluated directly in
// Cartesian components along n (identical in exact arithmetic, no t1/t2 needed).
HX_HD void lxf_flux(const Phys &ph, const double L[NV], const double R[NV], const double n[3], double f[NV])
{
    double irL = 1.0 / L[0], irR = 1.0 / R[0];
    double uL[3] = {L[1] * irL, L[2] * irL, L[3] * irL};
    double uR[3] = {R[1] * irR, R[2] * irR, R[3] * irR};
    double mnL = L[1] * n[0] + L[2] * n[1] + L[3] * n[2];   // rho u_n
    double mnR = R[1] * n[0] + R[2] * n[1] + R[3] * n[2];
    double pL = ph.gm1 * (L[4] - 0.5 * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3]) * irL);
    double pR = ph.gm1 * (R[4] - 0.5 * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3]) * irR);
    double aL = sqrt(ph.gamma * pL * irL), aR = sqrt(ph.gamma * pR * irR);
    double unL = mnL * irL, unR = mnR * irR;
    double l1 = fabs(unL) + aL, l2 = fabs(unR) + aR;
    double lam = l1 > l2 ? l1 : l2;
    switch (ph.averaging) {
    case 1: {  // Ducros
        double us = 0.25 * (unL + unR);
        double ps = 0.5 * (pL + pR);
        f[0] = (L[0] + R[0]) * us;
        for (int d = 0; d < 3; ++d) f[1 + d] = (L[1 + d] + R[1 + d]) * us + ps * n[d];
        f[4] = (L[4] + pL + R[4] + pR) * us;
        break; }
    case 2:    // Kennedy-Gruber
    case 3: {  // Pirozzoli
        double rho = 0.5 * (L[0] + R[0]), un = 0.5 * (unL + unR), p = 0.5 * (pL + pR);
        double m = rho * un;
        f[0] = m;
        for (int d = 0; d < 3; ++d) f[1 + d] = m * 0.5 * (uL[d] + uR[d]) + p * n[d];
        if (ph.averaging == 2) f[4] = m * 0.5 * (L[4] * irL + R[4] * irR) + p * un;
        else f[4] = m * 0.5 * ((L[4] + pL) * irL + (R[4] + pR) * irR);
        break; }
    case 4: {  // Morinishi
        double cp = ph.gamma / ph.gm1;
        double s = 0.25 * (mnL + mnR);
        f[0] = 0.5 * (mnL + mnR);
        for (int d = 0; d < 3; ++d) f[1 + d] = s * (uL[d] + uR[d]) + 0.5 * (pL + pR) * n[d];
        double uLuL = uL[0] * uL[0] + uL[1] * uL[1] + uL[2] * uL[2];
        double uRuR = uR[0] * uR[0] + uR[1] * uR[1] + uR[2] * uR[2];
        double uLuR = uL[0] * uR[0] + uL[1] * uR[1] + uL[2] * uR[2];
        // sum_c (mL uL_c + mR uR_c)(uL_c + uR_c)/4 - (mL uL_c^2 + mR uR_c^2)/4
        f[4] = 0.5 * (unL * cp * pL + unR * cp * pR) + 0.25 * ((mnL + mnR) * uLuR);
        (void)uLuL; (void)uRuR;
        break; }
    default: { // Standard
        double ps = 0.5 * (pL + pR);
        f[0] = 0.5 * (mnL + mnR);
        for (int d = 0; d < 3; ++d) f[1 + d] = 0.5 * (mnL * uL[d] + mnR * uR[d]) + ps * n[d];
        f[4] = 0.5 * (unL * (L[4] + pL) + unR * (R[4] + pR));
    }
    }
    double s = ph.lambda_stab * 0.5 * lam;
#pragma unroll
    for (int v = 0; v < NV; ++v) f[v] -= s * (R[v] - L[v]);
}

}  // namespace hx
