// HIP kernels of the DGSEM residual + low-storage RK stage for gfx950 (MI355X).
//
// Data layout in HBM (uniform polynomial order P = N+1 per handle):
//   Q, G      : FP64, reference layout  Q[(e*P*P + i + P*j)*4 + eq]
//   traces    : [(e*4 + f)*P + k][4]  element-local face orientation, FP64 or FP32
//   gradient  : dQ[(e*P*P + n)*8 + d*4 + eq], FP64 or FP32
//   metric    : five planes of e*P*P+n: Ja11, Ja12, Ja21, Ja22, jac (FP64 or FP32)
//
// One workgroup of 256 lanes (4 wavefronts) owns EPB = 256/(P*P) whole elements, one
// lane per node.  The element's state is staged in LDS; each element computes the
// Riemann fluxes of its own four faces (the neighbor reads the same trace values
// and evaluates the identical expression, so both sides see bit-identical fluxes,
// like the reference's FL = F, FR = -F), subtracts the face lifts in ascending
// global edge order (DGSpatialDiscretizationMethods.f90:196-220), divides by the
// Jacobian, applies the low-storage RK update and writes the next stage's traces
// from the updated state.  The residual max of the last stage is fused as well.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hg_device.h"

namespace hg {

namespace {

constexpr int kThreads = 256;

template <typename R>
struct Phys {
    R gamma, gm1, cp, cv, lambda, invSGM, mu, cpPr, sigma0, ip_eps;
    __device__ explicit Phys(const PhysParams &p)
        : gamma((R)p.gamma), gm1((R)p.gm1), cp((R)p.cp), cv((R)p.cv), lambda((R)p.lambda),
          invSGM((R)p.invSGM), mu((R)p.mu), cpPr((R)p.cpPr), sigma0((R)p.sigma0), ip_eps((R)p.ip_eps) {}
};

__device__ inline double dsqrt(double x) { return sqrt(x); }
__device__ inline float dsqrt(float x) { return sqrtf(x); }
__device__ inline double dpow(double a, double b) { return pow(a, b); }
__device__ inline float dpow(float a, float b) { return powf(a, b); }
__device__ inline double dabs(double x) { return fabs(x); }
__device__ inline float dabs(float x) { return fabsf(x); }

// ------------------------------------------------------------------ point fluxes
// InviscidFluxes.f90:104-147
template <typename R>
__device__ inline void inviscid_flux(const Phys<R> &ph, const R *u, R *f, R *g) {
    R invRho = R(1) / u[0];
    R vx = u[1] * invRho, vy = u[2] * invRho;
    R p = ph.gm1 * (u[3] - R(0.5) * u[1] * vx - R(0.5) * u[2] * vy);
    f[0] = u[1]; f[1] = u[1] * vx + p; f[2] = u[1] * vy; f[3] = (u[3] + p) * vx;
    g[0] = u[2]; g[1] = f[2]; g[2] = u[2] * vy + p; g[3] = (u[3] + p) * vy;
#pragma unroll
    for (int k = 0; k < 4; ++k) { f[k] *= ph.invSGM; g[k] *= ph.invSGM; }
}

// RiemannSolvers.f90:189-305
template <typename R>
__device__ void roe_flux(const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    R rhoL = qL[0], invRhoL = R(1) / rhoL, sqrtRhoL = dsqrt(rhoL);
    R uL = (qL[1] * nx + qL[2] * ny) * invRhoL, vL = (-qL[1] * ny + qL[2] * nx) * invRhoL;
    R pL = ph.gm1 * (qL[3] - R(0.5) * rhoL * (uL * uL + vL * vL));
    R HL = ph.cp * pL * invRhoL + R(0.5) * (uL * uL + vL * vL);
    R rhoR = qR[0], invRhoR = R(1) / rhoR, sqrtRhoR = dsqrt(rhoR);
    R uR = (qR[1] * nx + qR[2] * ny) * invRhoR, vR = (-qR[1] * ny + qR[2] * nx) * invRhoR;
    R pR = ph.gm1 * (qR[3] - R(0.5) * rhoR * (uR * uR + vR * vR));
    R HR = ph.cp * pR * invRhoR + R(0.5) * (uR * uR + vR * vR);
    R invrho = R(1) / (sqrtRhoL + sqrtRhoR);
    R u = (sqrtRhoL * uL + sqrtRhoR * uR) * invrho;
    R v = (sqrtRhoL * vL + sqrtRhoR * vR) * invrho;
    R H = (sqrtRhoL * HL + sqrtRhoR * HR) * invrho;
    R a = dsqrt(ph.gm1 * (H - R(0.5) * (u * u + v * v)));
    R d0 = rhoR - rhoL, d1 = rhoR * uR - rhoL * uL, d2 = rhoR * vR - rhoL * vL, d3 = qR[3] - qL[3];
    R al2 = d2 - v * d0;
    R al1 = ph.gm1 * (d0 * (H - u * u) + u * d1 - d3 + (d2 - v * d0) * v) / (a * a);
    R al0 = (d0 * (u + a) - d1 - a * al1) / (R(2) * a);
    R al3 = d0 - (al0 + al1);
    R Fs0 = rhoL * uL, Fs1 = Fs0 * uL + pL, Fs2 = Fs0 * vL, Fs3 = HL * Fs0;
    // waves with negative speed, counted like RiemannSolvers.f90:253-261
    if (!(u - a > R(0))) {
        R c = al0 * (u - a);
        Fs0 += c; Fs1 += c * (u - a); Fs2 += c * v; Fs3 += c * (H - u * a);
        if (!(u > R(0))) {
            R c1 = al1 * u;
            Fs0 += c1; Fs1 += c1 * u; Fs2 += c1 * v; Fs3 += c1 * (R(0.5) * (u * u + v * v));
            R c2 = al2 * u;
            Fs2 += c2; Fs3 += c2 * v;
            if (!(u + a > R(0))) {
                R c3 = al3 * (u + a);
                Fs0 += c3; Fs1 += c3 * (u + a); Fs2 += c3 * v; Fs3 += c3 * (H + u * a);
            }
        }
    }
    R Ft = Fs1 * nx - Fs2 * ny;
    Fs2 = Fs1 * ny + Fs2 * nx;
    F[0] = Fs0 * ph.invSGM; F[1] = Ft * ph.invSGM; F[2] = Fs2 * ph.invSGM; F[3] = Fs3 * ph.invSGM;
}

// RiemannSolvers.f90:307-399
template <typename R>
__device__ void hll_flux(const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    R rhoL = qL[0], invRhoL = R(1) / rhoL;
    R uL = (qL[1] * nx + qL[2] * ny) * invRhoL, vL = (-qL[1] * ny + qL[2] * nx) * invRhoL;
    R pL = ph.gm1 * (qL[3] - R(0.5) * rhoL * (uL * uL + vL * vL));
    R HL = ph.cp * pL * invRhoL + R(0.5) * (uL * uL + vL * vL);
    R aL = dsqrt(ph.gamma * pL * invRhoL), rhoeL = qL[3];
    R rhoR = qR[0], invRhoR = R(1) / rhoR;
    R uR = (qR[1] * nx + qR[2] * ny) * invRhoR, vR = (-qR[1] * ny + qR[2] * nx) * invRhoR;
    R pR = ph.gm1 * (qR[3] - R(0.5) * rhoR * (uR * uR + vR * vR));
    R HR = ph.cp * pR * invRhoR + R(0.5) * (uR * uR + vR * vR);
    R aR = dsqrt(ph.gamma * pR * invRhoR), rhoeR = qR[3];
    R sL = dsqrt(rhoL), sR = dsqrt(rhoR), invrho = R(1) / (sL + sR);
    R u = (sL * uL + sR * uR) * invrho, v = (sL * vL + sR * vR) * invrho, H = (sL * HL + sR * HR) * invrho;
    R a = dsqrt(ph.gm1 * (H - R(0.5) * (u * u + v * v)));
    R SL = fmin(uL - aL, u - a), SR = fmax(uR + aR, u + a);
    R Fs[4] = {R(0), R(0), R(0), R(0)};
    if (SL >= R(0)) {
        Fs[0] = rhoL * uL; Fs[1] = Fs[0] * uL + pL; Fs[2] = Fs[0] * vL; Fs[3] = HL * Fs[0];
    } else if (SR <= R(0)) {
        Fs[0] = rhoR * uR; Fs[1] = Fs[0] * uR + pR; Fs[2] = Fs[0] * vR; Fs[3] = HR * Fs[0];
    } else if ((SL < R(0)) || (SR > R(0))) {
        R idS = R(1) / (SR - SL);
        Fs[0] = (SR * rhoL * uL - SL * rhoR * uR + SL * SR * (rhoR - rhoL)) * idS;
        Fs[1] = (SR * (rhoL * uL * uL + pL) - SL * (rhoR * uR * uR + pR) + SL * SR * (rhoR * uR - rhoL * uL)) * idS;
        Fs[2] = (SR * rhoL * uL * vL - SL * rhoR * uR * vR + SR * SL * (rhoR * vR - rhoL * vL)) * idS;
        Fs[3] = (SR * uL * rhoL * HL - SL * uR * rhoR * HR + SR * SL * (rhoeR - rhoeL)) * idS;
    }
    R Ft = Fs[1] * nx - Fs[2] * ny;
    Fs[2] = Fs[1] * ny + Fs[2] * nx;
    F[0] = Fs[0] * ph.invSGM; F[1] = Ft * ph.invSGM; F[2] = Fs[2] * ph.invSGM; F[3] = Fs[3] * ph.invSGM;
}

// RiemannSolvers.f90:461-484
template <typename R>
__device__ inline void exact_F(const Phys<R> &ph, R p, const R *W, R &F, R &dF) {
    if (p > W[3]) {
        R A = R(2) / (W[0] * (ph.gamma + R(1)));
        R B = ph.gm1 * W[3] / (ph.gamma + R(1));
        R s = dsqrt(A / (p + B));
        F = (p - W[3]) * s;
        dF = s * (R(1) - R(0.5) * (p - W[3]) / (B + p));
    } else {
        F = R(2) * ph.cv * W[5] * (dpow(p / W[3], R(0.5) / ph.cp) - R(1));
        dF = R(1) / (W[0] * W[5]) * dpow(p / W[3], -R(0.5) * (ph.gamma + R(1)) / ph.gamma);
    }
}

// RiemannSolvers.f90:33-187 with ExactRiemann_ComputePStar :405-459 (at most 50 Newton
// iterations; the star velocity is taken from the last iterate when not converged).
template <typename R>
__device__ void exact_flux(const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    const R g = ph.gamma, gm1 = ph.gm1, cp = ph.cp, cv = ph.cv;
    R qnL[4] = {qL[0], qL[1] * nx + qL[2] * ny, -qL[1] * ny + qL[2] * nx, qL[3]};
    R qnR[4] = {qR[0], qR[1] * nx + qR[2] * ny, -qR[1] * ny + qR[2] * nx, qR[3]};
    R wL[6], wR[6];
    wL[0] = qnL[0]; wL[1] = qnL[1] / qnL[0]; wL[2] = qnL[2] / qnL[0];
    wL[3] = gm1 * (qnL[3] - R(0.5) * (qnL[1] * wL[1] + qnL[2] * wL[2])); wL[4] = R(0); wL[5] = dsqrt(g * wL[3] / wL[0]);
    wR[0] = qnR[0]; wR[1] = qnR[1] / qnR[0]; wR[2] = qnR[2] / qnR[0];
    wR[3] = gm1 * (qnR[3] - R(0.5) * (qnR[1] * wR[1] + qnR[2] * wR[2])); wR[4] = R(0); wR[5] = dsqrt(g * wR[3] / wR[0]);
    R FL, dFL, FR, dFR;
    R ps = dpow((wL[5] + wR[5] - R(0.5) * gm1 * (wR[1] - wL[1])) /
                    (wL[5] / dpow(wL[3], R(0.5) / cp) + wR[5] / dpow(wR[3], R(0.5) / cp)),
                R(2) * cp);
    for (int it = 1; it <= 50; ++it) {
        exact_F(ph, ps, wL, FL, dFL);
        exact_F(ph, ps, wR, FR, dFR);
        R Fv = FL + FR + wR[1] - wL[1], dF = dFL + dFR, pold = ps;
        ps = pold - Fv / dF;
        R cha = R(0.5) * dabs(ps - pold) / (ps + pold);
        if (dabs(cha) < R(1.0e-6)) break;
    }
    exact_F(ph, ps, wL, FL, dFL);
    exact_F(ph, ps, wR, FR, dFR);
    R us = R(0.5) * (wL[1] + wR[1] + FR - FL), pstar = ps, rs, uF, pF;
    R Fs[4] = {R(0), R(0), R(0), R(0)};
    if (us >= R(0)) {
        R sh = wL[5] * dsqrt((g + R(1)) / (R(2) * g) * pstar / wL[3] + R(0.5) / cp);
        if ((pstar <= wL[3]) && (wL[1] >= wL[5])) {
            Fs[0] = qnL[1]; Fs[1] = qnL[1] * wL[1] + wL[3]; Fs[2] = qnL[2] * wL[1];
            Fs[3] = wL[1] * (cp * wL[3] + R(0.5) * (qnL[1] * wL[1] + qnL[2] * wL[2]));
        } else if ((pstar <= wL[3]) && (wL[1] < wL[5]) && (us < wL[5] * dpow(pstar / wL[3], R(0.5) / cp))) {
            rs = wL[0] * dpow(pstar / wL[3], R(1) / g);
            Fs[0] = rs * us; Fs[1] = rs * us * us + pstar; Fs[2] = Fs[0] * wL[2];
            Fs[3] = us * (cp * pstar + R(0.5) * rs * (us * us + wL[2] * wL[2]));
        } else if ((pstar <= wL[3]) && (wL[1] < wL[5]) && (us >= wL[5] * dpow(pstar / wL[3], R(0.5) / cp))) {
            uF = R(2) * wL[5] / (g + R(1)) + (g - R(1)) / (g + R(1)) * wL[1];
            rs = wL[0] * dpow(uF / wL[5], R(2) * cv);
            pF = wL[3] * dpow(rs / wL[0], g);
            Fs[0] = rs * uF; Fs[1] = rs * uF * uF + pF; Fs[2] = rs * uF * wL[2];
            Fs[3] = uF * (cp * pF + R(0.5) * rs * (uF * uF + wL[2] * wL[2]));
        } else if ((pstar > wL[3]) && (wL[1] >= sh)) {
            Fs[0] = qnL[1]; Fs[1] = qnL[1] * wL[1] + wL[3]; Fs[2] = qnL[2] * wL[1];
            Fs[3] = wL[1] * (cp * wL[3] + R(0.5) * (qnL[1] * wL[1] + qnL[2] * wL[2]));
        } else if ((pstar > wL[3]) && (wL[1] < sh)) {
            rs = wL[0] * ((pstar / wL[3] + (g - R(1)) / (g + R(1))) / (pstar / wL[3] * (g - R(1)) / (g + R(1)) + R(1)));
            Fs[0] = rs * us; Fs[1] = rs * us * us + pstar; Fs[2] = rs * us * wL[2];
            Fs[3] = us * (cp * pstar + R(0.5) * rs * (us * us + wL[2] * wL[2]));
        }
    } else {
        if ((pstar <= wR[3]) && (wR[1] + wR[5] <= R(0))) {
            Fs[0] = qnR[1]; Fs[1] = qnR[1] * wR[1] + wR[3]; Fs[2] = qnR[2] * wR[1];
            Fs[3] = wR[1] * (cp * wR[3] + R(0.5) * (qnR[1] * wR[1] + qnR[2] * wR[2]));
        } else if ((pstar <= wR[3]) && (wR[1] + wR[5] >= R(0)) && (us + wR[5] * dpow(pstar / wR[3], R(0.5) / cp) >= R(0))) {
            rs = wR[0] * dpow(pstar / wR[3], R(1) / g);
            Fs[0] = rs * us; Fs[1] = rs * us * us + pstar; Fs[2] = Fs[0] * wR[2];
            Fs[3] = us * (cp * pstar + R(0.5) * rs * (us * us + wR[2] * wR[2]));
        } else if ((pstar <= wR[3]) && (wR[1] + wR[5] > R(0)) && (us + wR[5] * dpow(pstar / wR[3], R(0.5) / cp) < R(0))) {
            uF = -R(2) * wR[5] / (g + R(1)) + (g - R(1)) / (g + R(1)) * wR[1];
            rs = wR[0] * dpow(-uF / wR[5], R(2) * cv);
            pF = wR[3] * dpow(rs / wR[0], g);
            Fs[0] = rs * uF; Fs[1] = rs * uF * uF + pF; Fs[2] = rs * uF * wR[2];
            Fs[3] = uF * (cp * pF + R(0.5) * rs * (uF * uF + wR[2] * wR[2]));
        } else if ((pstar > wR[3]) && (wR[1] + wR[5] * (R(0.5) * (g + R(1)) / g * pstar / wR[3] + R(0.5) / cp) <= R(0))) {
            Fs[0] = qnR[1]; Fs[1] = qnR[1] * wR[1] + wR[3]; Fs[2] = qnR[2] * wR[1];
            Fs[3] = wR[1] * (cp * wR[3] + R(0.5) * (qnR[1] * wR[1] + qnR[2] * wR[2]));
        } else if ((pstar > wR[3]) && (wR[1] + wR[5] * ((g + R(1)) / (R(2) * g) * pstar / wR[3] + R(0.5) / cp) > R(0))) {
            rs = wR[0] * ((pstar / wR[3] + (g - R(1)) / (g + R(1))) / (pstar / wR[3] * (g - R(1)) / (g + R(1)) + R(1)));
            Fs[0] = rs * us; Fs[1] = rs * us * us + pstar; Fs[2] = rs * us * wR[2];
            Fs[3] = us * (cp * pstar + R(0.5) * rs * (us * us + wR[2] * wR[2]));
        }
    }
    R Ft = Fs[1] * nx - Fs[2] * ny;
    Fs[2] = Fs[1] * ny + Fs[2] * nx;
    F[0] = Fs[0] * ph.invSGM; F[1] = Ft * ph.invSGM; F[2] = Fs[2] * ph.invSGM; F[3] = Fs[3] * ph.invSGM;
}

// Lax-Friedrichs (Rusanov) flux; not present in the reference, defined as
// 1/2 (Fn(qL) + Fn(qR)) - 1/2 max(|un|+a) (qR - qL), scaled by 1/(sqrt(gamma) M).
template <typename R>
__device__ void lf_flux(const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    R fL[4], gL[4], fR[4], gR[4];
    R iL = R(1) / qL[0], iR = R(1) / qR[0];
    R uL = qL[1] * iL, vL = qL[2] * iL, uR = qR[1] * iR, vR = qR[2] * iR;
    R pL = ph.gm1 * (qL[3] - R(0.5) * qL[1] * uL - R(0.5) * qL[2] * vL);
    R pR = ph.gm1 * (qR[3] - R(0.5) * qR[1] * uR - R(0.5) * qR[2] * vR);
    fL[0] = qL[1]; fL[1] = qL[1] * uL + pL; fL[2] = qL[1] * vL; fL[3] = (qL[3] + pL) * uL;
    gL[0] = qL[2]; gL[1] = fL[2]; gL[2] = qL[2] * vL + pL; gL[3] = (qL[3] + pL) * vL;
    fR[0] = qR[1]; fR[1] = qR[1] * uR + pR; fR[2] = qR[1] * vR; fR[3] = (qR[3] + pR) * uR;
    gR[0] = qR[2]; gR[1] = fR[2]; gR[2] = qR[2] * vR + pR; gR[3] = (qR[3] + pR) * vR;
    R lL = dabs(uL * nx + vL * ny) + dsqrt(ph.gamma * pL * iL);
    R lR = dabs(uR * nx + vR * ny) + dsqrt(ph.gamma * pR * iR);
    R lam = lL > lR ? lL : lR;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        R fnL = fL[k] * nx + gL[k] * ny, fnR = fR[k] * nx + gR[k] * ny;
        F[k] = (R(0.5) * (fnL + fnR) - R(0.5) * lam * (qR[k] - qL[k])) * ph.invSGM;
    }
}

template <typename R>
__device__ inline void riemann(int id, const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    switch (id) {
    case 1: hll_flux(ph, qL, qR, nx, ny, F); break;
    case 2: exact_flux(ph, qL, qR, nx, ny, F); break;
    case 3: lf_flux(ph, qL, qR, nx, ny, F); break;
    default: roe_flux(ph, qL, qR, nx, ny, F); break;
    }
}

// ViscousFluxes.f90:39-80; dq[d*4 + eq]
template <typename R>
__device__ inline void viscous_flux(const Phys<R> &ph, const R *q, const R *dq, R *f, R *g) {
    R invRho = R(1) / q[0], u = q[1] * invRho, v = q[2] * invRho, uDR = u * invRho, vDR = v * invRho;
    const R *dx = dq, *dy = dq + 4;
    R divV = -uDR * dx[0] + invRho * dx[1] - vDR * dy[0] + invRho * dy[2];
    R dxT = ph.gm1 * invRho * (-q[3] * invRho * dx[0] + dx[3] + u * (u * dx[0] - dx[1]) + v * (v * dx[0] - dx[2]));
    R dyT = ph.gm1 * invRho * (-q[3] * invRho * dy[0] + dy[3] + u * (u * dy[0] - dy[1]) + v * (v * dy[0] - dy[2]));
    f[0] = R(0);
    f[1] = R(2) * (-uDR * dx[0] + invRho * dx[1]) + ph.lambda * divV;
    f[2] = -vDR * dx[0] - uDR * dy[0] + invRho * dy[1] + invRho * dx[2];
    f[3] = u * f[1] + v * f[2] + ph.cpPr * dxT;
    g[0] = R(0);
    g[1] = f[2];
    g[2] = R(2) * (-vDR * dy[0] + invRho * dy[2]) + ph.lambda * divV;
    g[3] = u * g[1] + v * g[2] + ph.cpPr * dyT;
}

// ViscousFluxes.f90:279-318 with qB = q (the BR1/IP adiabatic wall uses the interior state)
template <typename R>
__device__ inline void adiabatic_flux(const Phys<R> &ph, const R *q, const R *dq, R *f, R *g) {
    R invRho = R(1) / q[0], u = q[1] * invRho, v = q[2] * invRho, uDR = u * invRho, vDR = v * invRho;
    const R *dx = dq, *dy = dq + 4;
    R divV = -uDR * dx[0] + invRho * dx[1] - vDR * dy[0] + invRho * dy[2];
    f[0] = R(0);
    f[1] = R(2) * (-uDR * dx[0] + invRho * dx[1]) + ph.lambda * divV;
    f[2] = -vDR * dx[0] - uDR * dy[0] + invRho * dy[1] + invRho * dx[2];
    g[0] = R(0);
    g[1] = f[2];
    g[2] = R(2) * (-vDR * dy[0] + invRho * dy[2]) + ph.lambda * divV;
    f[3] = (q[1] * f[1] + q[2] * g[1]) / q[0];
    g[3] = (q[1] * f[2] + q[2] * g[2]) / q[0];
}

// Ghost states of the non-periodic boundaries, computed from the own trace.
// Farfield.incf:138-376 (uB = uSB; viscous type DIRICHLET, or NEUMANN at supersonic
// outflow -- the reference leaves it unset) and ViscousWall.incf:158-254.
template <typename R>
__device__ inline void ghost_state(int kind, int flags, const double *zd, const Phys<R> &ph, const R *Qt, R nx, R ny,
                                   R *uB, R *uSB, int &neumann) {
    neumann = 0;
    if (kind == FK_FARFIELD) {
        R q[4] = {(R)zd[0], (R)zd[1], (R)zd[2], (R)zd[3]};
        R rhoI = q[0], uI = q[1] / rhoI, vI = q[2] / rhoI;
        R pI = ph.gm1 * (q[3] - R(0.5) * q[1] * uI - R(0.5) * q[2] * vI);
        R rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
        R vnL = uL * nx + vL * ny;
        R pL = ph.gm1 * (Qt[3] - R(0.5) * Qt[1] * uL - R(0.5) * Qt[2] * vL);
        R aL = dsqrt(ph.gamma * pL / rhoL);
        R ML = dsqrt(uL * uL + vL * vL) / aL;
        if (vnL < R(0)) {
            if (ML > R(1)) {
                for (int k = 0; k < 4; ++k) uB[k] = q[k];
            } else {
                R pR = R(0.5) * (pL + pI - rhoL * aL * (nx * (uL - uI) + ny * (vL - vI)));
                R rhoR = rhoI + (pR - pI) / (aL * aL);
                R uR = uI + nx * (pR - pI) / (aL * rhoL), vR = vI + ny * (pR - pI) / (aL * rhoL);
                uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
                uB[3] = ph.cv * pR + R(0.5) * (uB[1] * uR + uB[2] * vR);
            }
        } else {
            if (ML > R(1)) {
                for (int k = 0; k < 4; ++k) uB[k] = Qt[k];
                neumann = 1;
            } else {
                R pR = pI;
                R rhoR = rhoL + (pI - pL) / (aL * aL);
                R uR = uL - nx * (pI - pL) / (rhoL * aL), vR = vL - ny * (pI - pL) / (rhoL * aL);
                uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
                uB[3] = ph.cv * pR + R(0.5) * (uB[1] * uR + uB[2] * vR);
            }
        }
        for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
    } else {  // viscous wall
        R Tw = (R)zd[0], vwx = (R)zd[1], vwy = (R)zd[2];
        bool iso = (flags & FF_ISOTHERMAL) != 0;
        uB[0] = Qt[0];
        uB[1] = R(2) * uB[0] * vwx - Qt[1];
        uB[2] = R(2) * uB[0] * vwy - Qt[2];
        R Ek = R(0.5) * (uB[1] * uB[1] + uB[2] * uB[2]) / uB[0];
        uB[3] = iso ? ph.cv * uB[0] * R(1) * Tw + Ek : Qt[3];
        uSB[0] = Qt[0]; uSB[1] = uSB[0] * vwx; uSB[2] = uSB[0] * vwy;
        uSB[3] = iso ? ph.cv * uSB[0] * R(1) * Tw + R(0.5) * (uSB[1] * uSB[1] + uSB[2] * uSB[2]) / uSB[0]
                     : Qt[3] - Ek + R(0.5) * uSB[0] * (vwx * vwx + vwy * vwy);
    }
}

// ------------------------------------------------------------------ spectral helpers
// Trace of a face from the staged element state (DGSpatialDiscretizationMethods.f90:1520-1574).
// Always accumulated in FP64 with explicit fma in l order so that every kernel that forms
// a trace (prolong, stage epilogue, stage recomputation) produces the same bits.
template <int P, int STRIDE, int NV>
__device__ inline void face_trace(const double *sX /* element block, [n*STRIDE + c] */, const double *sLb, int f, int k,
                                  double *out) {
    const double *lb = (f == 0 || f == 3) ? sLb : sLb + P;
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = 0.0;
#pragma unroll
    for (int l = 0; l < P; ++l) {
        int n = (f == 0 || f == 2) ? (k + P * l) : (l + P * k);
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = fma(sX[n * STRIDE + c], lb[l], out[c]);
    }
}

template <typename T>
__device__ inline T ld(const void *p, int64_t i) { return ((const T *)p)[i]; }

template <int P>
struct Cfg {
    static constexpr int PP = P * P;
    static constexpr int EPB = kThreads / PP;
    static constexpr int NF = EPB * 4 * P;
};

// ------------------------------------------------------------------ prolongation
template <int P, typename TR>
__global__ __launch_bounds__(kThreads) void k_prolong(StageArgs a) {
    using C = Cfg<P>;
    __shared__ double sQ[C::EPB * C::PP * 4];
    __shared__ double sLb[2 * P];
    const int t = threadIdx.x, e0 = blockIdx.x * C::EPB;
    if (t < 2 * P) sLb[t] = a.lb[t];
    const int64_t base = (int64_t)e0 * C::PP * 4;
    const int nval = min(C::EPB, a.K - e0) * C::PP * 4;
    for (int x = t; x < nval; x += kThreads) sQ[x] = a.Q[base + x];
    __syncthreads();
    for (int w = t; w < C::NF; w += kThreads) {
        int s = w / (4 * P), f = (w / P) % 4, k = w % P, e = e0 + s;
        if (e >= a.K) continue;
        double tq[4];
        face_trace<P, 4, 4>(sQ + s * C::PP * 4, sLb, f, k, tq);
        TR *o = (TR *)a.tr_out + ((int64_t)(e * 4 + f) * P + k) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] = (TR)tq[c];
    }
}

// ------------------------------------------------------------------ gradient kernels
// BR1: dQ = ( -Vector_StdVolumeGreen(Q) + sum_faces Vector_StdFace(uStar n dS) ) / jac
// (DGViscous_BR1.f90:48-111, DGWeakIntegrals.f90:503-622, DGViscousMethods.f90:538-917).
// IP: interior gradient (QuadElement_QuadProcedures.incf:42-88).
// Both write dQ at the nodes and the gradient traces of the four faces.
template <int P, typename R, typename TR>
__global__ __launch_bounds__(kThreads) void k_gradient(StageArgs a) {
    using C = Cfg<P>;
    constexpr int PP = C::PP, EPB = C::EPB;
    __shared__ double sQ[EPB * PP * 4];
    __shared__ R sJa[EPB * PP * 4];
    __shared__ R sFu[EPB * 4 * P * 8];
    __shared__ double sDQ[EPB * PP * 8];
    __shared__ double sLb[2 * P];
    __shared__ R sHat[PP], sLbw[2 * P], sD[PP];
    const int t = threadIdx.x, e0 = blockIdx.x * EPB;
    const Phys<R> ph(a.ph);
    const bool br1 = a.ph.viscous == 1;
    if (t < 2 * P) { sLb[t] = a.lb[t]; sLbw[t] = (R)a.lbw[t]; }
    for (int x = t; x < PP; x += kThreads) sHat[x] = (R)a.hatD[x];
    if (!br1)
        for (int x = t; x < PP; x += kThreads) sD[x] = (R)a.D[x];
    const int ne = min(EPB, a.K - e0);
    const int64_t nb0 = (int64_t)e0 * PP;
    for (int x = t; x < ne * PP * 4; x += kThreads) sQ[x] = a.Q[nb0 * 4 + x];
    for (int x = t; x < ne * PP * 4; x += kThreads) {
        int s = x / (PP * 4), c = (x / PP) % 4, n = x % PP;
        sJa[(s * PP + n) * 4 + c] = ld<R>(a.metric, c * a.nodes + nb0 + s * PP + n);
    }
    __syncthreads();
    // face phase (BR1 only): solution Riemann flux uStar * dS * n
    if (br1) {
        for (int w = t; w < C::NF; w += kThreads) {
            int s = w / (4 * P), f = (w / P) % 4, k = w % P, e = e0 + s;
            if (e >= a.K) continue;
            int4 fr = a.face[e * 4 + f];
            int kind = fr.x & 15;
            bool curved = fr.x & FF_CURVED, rev = fr.x & FF_REV, selfrev = fr.x & FF_SELFREV;
            int gi = curved ? (selfrev ? P - 1 - k : k) : 0;
            const double *geo = a.egeo + 3 * ((int64_t)fr.z + gi);
            R nx = (R)geo[0], ny = (R)geo[1], dS = (R)geo[2];
            double tq[4];
            face_trace<P, 4, 4>(sQ + s * PP * 4, sLb, f, k, tq);
            R qs[4] = {(R)tq[0], (R)tq[1], (R)tq[2], (R)tq[3]};
            R us[4];
            if (kind <= FK_PERIODIC) {
                const TR *nbt = (const TR *)a.tr_in + ((int64_t)fr.y * P + (rev ? P - 1 - k : k)) * 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) us[c] = R(0.5) * (qs[c] + (R)nbt[c]);
                if (kind == FK_INT_R) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) us[c] = -us[c];
                }
            } else {
                R uB[4];
                int neu;
                ghost_state<R>(kind, fr.x, a.zone + 8 * fr.y, ph, qs, nx, ny, uB, us, neu);
            }
            R *o = sFu + (s * 4 + f) * P * 8 + k * 8;
#pragma unroll
            for (int c = 0; c < 4; ++c) { o[c] = us[c] * dS * nx; o[4 + c] = us[c] * dS * ny; }
        }
        __syncthreads();
    }
    const int s = t / PP, n = t % PP, e = e0 + s;
    if (s < EPB && e < a.K) {
        const int i = n % P, j = n / P;
        const double *q = sQ + s * PP * 4;
        R jac = ld<R>(a.metric, 4 * a.nodes + nb0 + s * PP + n);
        R dq[8];
        if (br1) {
            const R *Ja = sJa + s * PP * 4;
#pragma unroll
            for (int d = 0; d < 2; ++d)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    R acc = R(0);
#pragma unroll
                    for (int l = 0; l < P; ++l) acc += (R)q[(l + P * j) * 4 + c] * Ja[(l + P * j) * 4 + d * 2 + 0] * sHat[i * P + l];
#pragma unroll
                    for (int l = 0; l < P; ++l) acc += (R)q[(i + P * l) * 4 + c] * Ja[(i + P * l) * 4 + d * 2 + 1] * sHat[j * P + l];
                    dq[d * 4 + c] = -acc;
                }
            int ord = a.forder[e];
#pragma unroll
            for (int o = 0; o < 4; ++o) {
                int f = (ord >> (2 * o)) & 3;
                const R *Fu = sFu + (s * 4 + f) * P * 8;
                int idx = (f == 0 || f == 2) ? i : j;
                R lw = (f == 0) ? sLbw[j] : (f == 1) ? sLbw[P + i] : (f == 2) ? sLbw[P + j] : sLbw[i];
#pragma unroll
                for (int c = 0; c < 8; ++c) dq[c] += Fu[idx * 8 + c] * lw;
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) dq[c] = dq[c] / jac;
        } else {
            const R *Ja = sJa + s * PP * 4;
            R j11 = Ja[n * 4 + 0], j12 = Ja[n * 4 + 1], j21 = Ja[n * 4 + 2], j22 = Ja[n * 4 + 3];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                R dxi = R(0), deta = R(0);
#pragma unroll
                for (int l = 0; l < P; ++l) dxi += sD[i * P + l] * (R)q[(l + P * j) * 4 + c];
#pragma unroll
                for (int l = 0; l < P; ++l) deta += sD[j * P + l] * (R)q[(i + P * l) * 4 + c];
                dq[c] = (dxi * j11 + deta * j12) / jac;
                dq[4 + c] = (dxi * j21 + deta * j22) / jac;
            }
        }
        R *go = (R *)a.dQ_out + (nb0 + s * PP + n) * 8;
#pragma unroll
        for (int c = 0; c < 8; ++c) { go[c] = dq[c]; sDQ[(s * PP + n) * 8 + c] = (double)dq[c]; }
    }
    __syncthreads();
    for (int w = t; w < C::NF; w += kThreads) {
        int s2 = w / (4 * P), f = (w / P) % 4, k = w % P, e2 = e0 + s2;
        if (e2 >= a.K) continue;
        double tg[8];
        face_trace<P, 8, 8>(sDQ + s2 * PP * 8, sLb, f, k, tg);
        TR *o = (TR *)a.dtr_out + ((int64_t)(e2 * 4 + f) * P + k) * 8;
#pragma unroll
        for (int c = 0; c < 8; ++c) o[c] = (TR)tg[c];
    }
}

// ------------------------------------------------------------------ residual helpers
__device__ inline void atomic_max_abs(unsigned long long *slot, double v) {
    // |v| >= 0: the IEEE bit pattern orders like the value; NaN patterns sort above +inf
    unsigned long long bits = __double_as_longlong(fabs(v));
    atomicMax(slot, bits);
}

// ------------------------------------------------------------------ the stage kernel
template <int P, bool NS, typename R, typename TR>
__global__ __launch_bounds__(kThreads) void k_stage(StageArgs a) {
    using C = Cfg<P>;
    constexpr int PP = C::PP, EPB = C::EPB, NF = C::NF;
    __shared__ double sQ[EPB * PP * 4];
    __shared__ R sCF[EPB * PP * 4];
    __shared__ R sCG[EPB * PP * 4];
    __shared__ R sF[EPB * 4 * P * 4];
    __shared__ R sU[NS ? EPB * 4 * P * 4 : 1];
    __shared__ R sV[NS ? EPB * 4 * P * 4 : 1];
    __shared__ R sU2[NS ? EPB * 4 * P * 4 : 1];
    __shared__ R sV2[NS ? EPB * 4 * P * 4 : 1];
    __shared__ double sLb[2 * P];
    __shared__ R sHat[PP], sLbw[2 * P], sDlbw[2 * P];
    const int t = threadIdx.x, e0 = blockIdx.x * EPB;
    const Phys<R> ph(a.ph);
    const bool ipgrad = NS && a.ph.viscous == 2 && a.ph.ip_gradflux;
    if (t < 2 * P) { sLb[t] = a.lb[t]; sLbw[t] = (R)a.lbw[t]; sDlbw[t] = (R)a.dlbw[t]; }
    for (int x = t; x < PP; x += kThreads) sHat[x] = (R)a.hatD[x];
    const int ne = min(EPB, a.K - e0);
    const int64_t nb0 = (int64_t)e0 * PP;
    {
        const double2 *src = (const double2 *)(a.Q + nb0 * 4);
        double2 *dst = (double2 *)sQ;
        for (int x = t; x < ne * PP * 2; x += kThreads) dst[x] = src[x];
    }
    __syncthreads();

    // ---- node phase: contravariant fluxes (DGSpatial_QDotVolumeLoop, DGWeakIntegrals.f90:111-171)
    const int s = t / PP, n = t % PP, e = e0 + s;
    const bool node_active = (s < EPB) && (e < a.K);
    R jac = R(1);
    if (node_active) {
        const int64_t g = nb0 + s * PP + n;
        R q[4], f[4], gg[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) q[c] = (R)sQ[(s * PP + n) * 4 + c];
        inviscid_flux(ph, q, f, gg);
        if (NS) {
            R dq[8], fv[4], gv[4];
            const TR *dqp = (const TR *)a.dQ + g * 8;
#pragma unroll
            for (int c = 0; c < 8; ++c) dq[c] = (R)dqp[c];
            viscous_flux(ph, q, dq, fv, gv);
#pragma unroll
            for (int c = 0; c < 4; ++c) { f[c] = (f[c] - ph.mu * fv[c]) - R(0); gg[c] = (gg[c] - ph.mu * gv[c]) - R(0); }
        }
        R j11 = ld<R>(a.metric, 0 * a.nodes + g), j12 = ld<R>(a.metric, 1 * a.nodes + g);
        R j21 = ld<R>(a.metric, 2 * a.nodes + g), j22 = ld<R>(a.metric, 3 * a.nodes + g);
        jac = ld<R>(a.metric, 4 * a.nodes + g);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sCF[(s * PP + n) * 4 + c] = f[c] * j11 + gg[c] * j21;
            sCG[(s * PP + n) * 4 + c] = f[c] * j12 + gg[c] * j22;
        }
    }

    // ---- face phase: Riemann fluxes of the element's own four faces
    for (int w = t; w < NF; w += kThreads) {
        const int fs = w / (4 * P), f = (w / P) % 4, k = w % P, fe = e0 + fs;
        if (fe >= a.K) continue;
        const int4 fr = a.face[fe * 4 + f];
        const int kind = fr.x & 15;
        const bool curved = fr.x & FF_CURVED, rev = fr.x & FF_REV, selfrev = fr.x & FF_SELFREV;
        const int gi = curved ? (selfrev ? P - 1 - k : k) : 0;
        const double *geo = a.egeo + 3 * ((int64_t)fr.z + gi);
        const R nx = (R)geo[0], ny = (R)geo[1], dS = (R)geo[2];
        double tq[4];
        face_trace<P, 4, 4>(sQ + fs * PP * 4, sLb, f, k, tq);
        R qs[4], qo[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) qs[c] = (R)(TR)tq[c];   // the same rounding the neighbor reads
        int rs = a.ph.riemann;
        R uSB[4];
        int neumann = 0;
        if (kind <= FK_PERIODIC) {
            const TR *nbt = (const TR *)a.tr_in + ((int64_t)fr.y * P + (rev ? P - 1 - k : k)) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) qo[c] = (R)nbt[c];
        } else {
            ghost_state<R>(kind, fr.x, a.zone + 8 * fr.y, ph, qs, nx, ny, qo, uSB, neumann);
            int zr = (fr.x >> FF_RS_SHIFT) & 15;
            if (zr) rs = zr - 1;
        }
        R Fi[4];
        if (kind == FK_INT_R) riemann<R>(rs, ph, qo, qs, nx, ny, Fi);
        else riemann<R>(rs, ph, qs, qo, nx, ny, Fi);
        R F[4];
        if (!NS) {
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = Fi[c] * dS;
        } else {
            // gradient traces: own and outer
            const int64_t oslot = ((int64_t)(fe * 4 + f) * P + k) * 8;
            R ds[8], dn[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) ds[c] = (R)((const TR *)a.dtr)[oslot + c];
            R Fv[4] = {R(0), R(0), R(0), R(0)};
            R G8[8] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
            const R pen = ph.sigma0 * ph.mu * (R)(P - 1) * (R)(P - 1) * (R)a.einvh[fr.w];
            if (kind <= FK_PERIODIC) {
                const int64_t nslot = ((int64_t)fr.y * P + (rev ? P - 1 - k : k)) * 8;
#pragma unroll
                for (int c = 0; c < 8; ++c) dn[c] = (R)((const TR *)a.dtr)[nslot + c];
                const R *uL = (kind == FK_INT_R) ? qo : qs, *uR = (kind == FK_INT_R) ? qs : qo;
                const R *dL = (kind == FK_INT_R) ? dn : ds, *dR = (kind == FK_INT_R) ? ds : dn;
                R fL[4], gL[4], fR[4], gR[4];
                viscous_flux(ph, uL, dL, fL, gL);
                viscous_flux(ph, uR, dR, fR, gR);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    R acc = R(0);
                    acc = acc + R(0.5) * (ph.mu * fL[c] + ph.mu * fR[c]) * nx;
                    acc = acc + R(0.5) * (ph.mu * gL[c] + ph.mu * gR[c]) * ny;
                    Fv[c] = acc;
                }
                if (a.ph.viscous == 2) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) Fv[c] = Fv[c] - pen * (uL[c] - uR[c]);
                }
                if (ipgrad) {
                    // IP_GradientRiemannSolver (DGViscous_IP.f90:216-249): this side's share
                    R fgr[8], ff[4], fg4[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c) { fgr[c] = (uL[c] - uR[c]) * nx; fgr[4 + c] = (uL[c] - uR[c]) * ny; }
                    const R *own = (kind == FK_INT_R) ? uR : uL;
                    viscous_flux(ph, own, fgr, ff, fg4);
                    R cf = (kind == FK_INT_R ? -R(0.5) : R(0.5)) * ph.ip_eps * ph.mu;
#pragma unroll
                    for (int c = 0; c < 4; ++c) { G8[c] = cf * ff[c] * dS; G8[4 + c] = cf * fg4[c] * dS; }
                }
            } else if (!neumann) {
                R fv4[4], gv4[4];
                if (kind == FK_WALL && !(fr.x & FF_ISOTHERMAL)) adiabatic_flux(ph, qs, ds, fv4, gv4);
                else viscous_flux(ph, qs, ds, fv4, gv4);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    R acc = R(0);
                    acc = acc + ph.mu * fv4[c] * nx;
                    acc = acc + ph.mu * gv4[c] * ny;
                    Fv[c] = acc;
                }
                if (a.ph.viscous == 2) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) Fv[c] = Fv[c] - pen * (qs[c] - uSB[c]);
                }
                if (ipgrad) {
                    R fgr[8], ff[4], fg4[4];
#pragma unroll
                    for (int c = 0; c < 4; ++c) { fgr[c] = (qs[c] - uSB[c]) * nx; fgr[4 + c] = (qs[c] - uSB[c]) * ny; }
                    if (kind == FK_WALL && !(fr.x & FF_ISOTHERMAL)) adiabatic_flux(ph, qs, fgr, ff, fg4);
                    else viscous_flux(ph, qs, fgr, ff, fg4);
                    R cf = ph.ip_eps * ph.mu;
#pragma unroll
                    for (int c = 0; c < 4; ++c) { G8[c] = cf * ff[c] * dS; G8[4 + c] = cf * fg4[c] * dS; }
                }
            }
            if (kind <= FK_INT_R) {
#pragma unroll
                for (int c = 0; c < 4; ++c) F[c] = ((Fi[c] - Fv[c]) - R(0)) * dS;
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) F[c] = ((Fi[c] * dS - Fv[c] * dS) - R(0));
            }
            if (ipgrad) {
                // Contravariant gradient-flux components at the face (DGWeakIntegrals.f90:268-495)
                const R *fm = (const R *)a.fmet + ((int64_t)(fe * 4 + f) * P + k) * 5;
                R *u = sU + ((fs * 4 + f) * P + k) * 4;   // temporarily: Fc
                R *v = sV + ((fs * 4 + f) * P + k) * 4;   // temporarily: Gc
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    R fc = (G8[c] * fm[0] + G8[4 + c] * fm[1]) / fm[4];
                    R gc = (G8[c] * fm[2] + G8[4 + c] * fm[3]) / fm[4];
                    u[c] = fc; v[c] = gc;
                }
            }
        }
        if (kind == FK_INT_R) {
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = -F[c];
        }
        R *o = sF + ((fs * 4 + f) * P + k) * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] = F[c];
    }
    __syncthreads();

    // ---- IP gradient face terms: U (lbw coefficient) and V (dlbw coefficient) per face node;
    // horizontal faces: U = hatD . Fc, V = Gc ; vertical faces: U = hatD . Gc, V = Fc
    if (ipgrad) {
        for (int w = t; w < NF; w += kThreads) {
            const int ws = w / (4 * P), wf = (w / P) % 4, wk = w % P;
            if (e0 + ws >= a.K) continue;
            const R *Fc = sU + (ws * 4 + wf) * P * 4, *Gc = sV + (ws * 4 + wf) * P * 4;
            const bool horiz = (wf == 0 || wf == 2);
            const R *src = horiz ? Fc : Gc;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                R acc = R(0);
#pragma unroll
                for (int l = 0; l < P; ++l) acc += sHat[wk * P + l] * src[l * 4 + c];
                sU2[((ws * 4 + wf) * P + wk) * 4 + c] = acc;
                sV2[((ws * 4 + wf) * P + wk) * 4 + c] = horiz ? Gc[wk * 4 + c] : Fc[wk * 4 + c];
            }
        }
        __syncthreads();
    }

    // ---- node phase 2: volume contraction, ordered face lifts, /jac, RK update
    double qd[4] = {0.0, 0.0, 0.0, 0.0};
    if (node_active) {
        const int i = n % P, j = n / P;
        R acc[4] = {R(0), R(0), R(0), R(0)};
        const R *cF = sCF + s * PP * 4, *cG = sCG + s * PP * 4;
#pragma unroll
        for (int l = 0; l < P; ++l) {
            const R hi = sHat[i * P + l], hj = sHat[j * P + l];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = acc[c] + hi * cF[(l + P * j) * 4 + c] + hj * cG[(i + P * l) * 4 + c];
        }
        const int ord = a.forder[e];
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            const int f = (ord >> (2 * o)) & 3;
            const int idx = (f == 0 || f == 2) ? i : j;
            const int oth = (f == 0 || f == 2) ? j : i;
            const int side = (f == 0 || f == 3) ? 0 : P;
            const R lw = sLbw[side + oth];
            const R *Ff = sF + ((s * 4 + f) * P + idx) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = acc[c] - Ff[c] * lw;
            if (ipgrad) {
                const R dlw = sDlbw[side + oth];
                const R *Uf = sU2 + ((s * 4 + f) * P + idx) * 4, *Vf = sV2 + ((s * 4 + f) * P + idx) * 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = acc[c] - (Uf[c] * lw + Vf[c] * dlw);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) qd[c] = (double)(acc[c] / jac);
        const int64_t g = nb0 + s * PP + n;
        if (a.mode == 1) {
            double2 *o = (double2 *)(a.QDot + g * 4);
            o[0] = make_double2(qd[0], qd[1]);
            o[1] = make_double2(qd[2], qd[3]);
        } else {
            double gv[4];
            if (a.first_stage) {
#pragma unroll
                for (int c = 0; c < 4; ++c) gv[c] = a.cq * qd[c];
            } else {
                const double2 *gp = (const double2 *)(a.G + g * 4);
                double2 g01 = gp[0], g23 = gp[1];
                gv[0] = fma(a.am, g01.x, a.cq * qd[0]); gv[1] = fma(a.am, g01.y, a.cq * qd[1]);
                gv[2] = fma(a.am, g23.x, a.cq * qd[2]); gv[3] = fma(a.am, g23.y, a.cq * qd[3]);
            }
            double2 *go = (double2 *)(a.G + g * 4);
            go[0] = make_double2(gv[0], gv[1]);
            go[1] = make_double2(gv[2], gv[3]);
            double qn[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) qn[c] = fma(a.cg, gv[c], sQ[(s * PP + n) * 4 + c]);
            double2 *qo = (double2 *)(a.Q + g * 4);
            qo[0] = make_double2(qn[0], qn[1]);
            qo[1] = make_double2(qn[2], qn[3]);
#pragma unroll
            for (int c = 0; c < 4; ++c) sQ[(s * PP + n) * 4 + c] = qn[c];
        }
    }
    if (a.do_resid) {
        double m[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double v = node_active ? fabs(qd[c]) : 0.0;
            if (v != v) v = __longlong_as_double(0x7ff8000000000000ll);
            for (int off = 32; off > 0; off >>= 1) {
                double o = __shfl_xor(v, off, 64);
                v = (o > v || o != o) ? o : v;
            }
            m[c] = v;
        }
        if ((t & 63) == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) atomic_max_abs(a.resid + c, m[c]);
        }
    }
    if (a.mode == 0) {
        __syncthreads();
        for (int w = t; w < NF; w += kThreads) {
            const int fs = w / (4 * P), f = (w / P) % 4, k = w % P, fe = e0 + fs;
            if (fe >= a.K) continue;
            double tq[4];
            face_trace<P, 4, 4>(sQ + fs * PP * 4, sLb, f, k, tq);
            TR *o = (TR *)a.tr_out + ((int64_t)(fe * 4 + f) * P + k) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = (TR)tq[c];
        }
    }
}

__global__ void k_clear_resid(unsigned long long *r) {
    if (threadIdx.x < 4) r[threadIdx.x] = 0ull;
}

// ------------------------------------------------------------------ time step estimate
// TimeIntegrator_EstimateTimeStep (DGTimeIntegrator.f90:462-511): per element maxima of
// |u| and a, then a prefix-max scan in element order starting from the saved maxima.
template <int P>
__global__ __launch_bounds__(kThreads) void k_dt_elem(const double *Q, int K, double gamma, double gm1, double *umax,
                                                      double *amax) {
    constexpr int PP = P * P, EPB = kThreads / PP;
    __shared__ double su[kThreads], sa[kThreads];
    const int t = threadIdx.x, s = t / PP, n = t % PP, e = blockIdx.x * EPB + s;
    double u = 0.0, aa = 0.0;
    if (s < EPB && e < K) {
        const double *q = Q + ((int64_t)e * PP + n) * 4;
        u = sqrt(q[1] * q[1] + q[2] * q[2]) / q[0];
        double ur = q[1] / q[0], vr = q[2] / q[0];
        aa = sqrt(gamma * gm1 * (q[3] / q[0] - 0.5 * (ur * ur) - 0.5 * (vr * vr)));
    }
    su[t] = u; sa[t] = aa;
    __syncthreads();
    if (s < EPB && e < K && n == 0) {
        double mu = 0.0, ma = 0.0;
        for (int x = 0; x < PP; ++x) { mu = fmax(mu, su[s * PP + x]); ma = fmax(ma, sa[s * PP + x]); }
        umax[e] = mu; amax[e] = ma;
    }
}

__global__ __launch_bounds__(1024) void k_dt_scan(int K, const double *umax, const double *amax, const double *dx,
                                                  double coef, double N1, double dt_max, double *state, double *dt_out) {
    __shared__ double su[1024], sa[1024], sd[1024];
    const int t = threadIdx.x, T = blockDim.x;
    const int chunk = (K + T - 1) / T, b = t * chunk, end = min(K, b + chunk);
    double mu = 0.0, ma = 0.0;
    for (int e = b; e < end; ++e) { mu = fmax(mu, umax[e]); ma = fmax(ma, amax[e]); }
    su[t] = mu; sa[t] = ma;
    __syncthreads();
    if (t == 0) {  // exclusive prefix max seeded with the saved running maxima
        double ru = state[0], ra = state[1];
        for (int x = 0; x < T; ++x) {
            double cu = su[x], ca = sa[x];
            su[x] = ru; sa[x] = ra;
            ru = fmax(ru, cu); ra = fmax(ra, ca);
        }
        state[0] = ru; state[1] = ra;
    }
    __syncthreads();
    double ru = su[t], ra = sa[t], dt = dt_max;
    for (int e = b; e < end; ++e) {
        ru = fmax(ru, umax[e]); ra = fmax(ra, amax[e]);
        double de = coef * dx[e] / (ru + ra) / N1;
        dt = fmin(dt, de);
    }
    sd[t] = dt;
    __syncthreads();
    for (int w = T / 2; w > 0; w >>= 1) {
        if (t < w) sd[t] = fmin(sd[t], sd[t + w]);
        __syncthreads();
    }
    if (t == 0) *dt_out = sd[0];
}

template <int P>
__global__ __launch_bounds__(kThreads) void k_cons_partial(const double *Q, const double *jac, const double *w, int K,
                                                           double *partial) {
    constexpr int PP = P * P, EPB = kThreads / PP;
    __shared__ double sv[kThreads][4];
    const int t = threadIdx.x, s = t / PP, n = t % PP, e = blockIdx.x * EPB + s;
    double v[4] = {0, 0, 0, 0};
    if (s < EPB && e < K) {
        const int i = n % P, j = n / P;
        const int64_t g = (int64_t)e * PP + n;
        double c = w[i] * w[j] * jac[g];
        for (int k = 0; k < 4; ++k) v[k] = c * Q[g * 4 + k];
    }
    for (int k = 0; k < 4; ++k) sv[t][k] = v[k];
    __syncthreads();
    for (int wd = kThreads / 2; wd > 0; wd >>= 1) {
        if (t < wd)
            for (int k = 0; k < 4; ++k) sv[t][k] += sv[t + wd][k];
        __syncthreads();
    }
    if (t < 4) partial[blockIdx.x * 4 + t] = sv[0][t];
}

__global__ void k_cons_final(const double *partial, int nb, double *out) {
    if (threadIdx.x < 4) {
        double acc = 0.0;
        for (int b = 0; b < nb; ++b) acc += partial[b * 4 + threadIdx.x];
        out[threadIdx.x] = acc;
    }
}

template <int P>
int nblocks(int K) {
    constexpr int EPB = kThreads / (P * P);
    return (K + EPB - 1) / EPB;
}

}  // namespace

bool supported_order(int N) { return N >= 1 && N <= 10; }

#define HG_DISPATCH_P(P_, BODY)                       \
    switch (P_) {                                     \
    case 2: { constexpr int PV = 2; BODY; } break;    \
    case 3: { constexpr int PV = 3; BODY; } break;    \
    case 4: { constexpr int PV = 4; BODY; } break;    \
    case 5: { constexpr int PV = 5; BODY; } break;    \
    case 6: { constexpr int PV = 6; BODY; } break;    \
    case 7: { constexpr int PV = 7; BODY; } break;    \
    case 8: { constexpr int PV = 8; BODY; } break;    \
    case 9: { constexpr int PV = 9; BODY; } break;    \
    case 10: { constexpr int PV = 10; BODY; } break;  \
    case 11: { constexpr int PV = 11; BODY; } break;  \
    default: return hipErrorInvalidValue;             \
    }

hipError_t launch_prolong(int P, bool mixed, const StageArgs &a, hipStream_t st) {
    HG_DISPATCH_P(P, {
        int nb = nblocks<PV>(a.K);
        if (mixed) hipLaunchKernelGGL((k_prolong<PV, float>), dim3(nb), dim3(kThreads), 0, st, a);
        else hipLaunchKernelGGL((k_prolong<PV, double>), dim3(nb), dim3(kThreads), 0, st, a);
    });
    return hipGetLastError();
}

hipError_t launch_stage(int P, bool mixed, const StageArgs &a, hipStream_t st) {
    const bool ns = a.ph.ns != 0;
    HG_DISPATCH_P(P, {
        int nb = nblocks<PV>(a.K);
        if (mixed) {
            if (ns) hipLaunchKernelGGL((k_stage<PV, true, float, float>), dim3(nb), dim3(kThreads), 0, st, a);
            else hipLaunchKernelGGL((k_stage<PV, false, float, float>), dim3(nb), dim3(kThreads), 0, st, a);
        } else {
            if (ns) hipLaunchKernelGGL((k_stage<PV, true, double, double>), dim3(nb), dim3(kThreads), 0, st, a);
            else hipLaunchKernelGGL((k_stage<PV, false, double, double>), dim3(nb), dim3(kThreads), 0, st, a);
        }
    });
    return hipGetLastError();
}

hipError_t launch_gradient(int P, bool mixed, const StageArgs &a, hipStream_t st) {
    HG_DISPATCH_P(P, {
        int nb = nblocks<PV>(a.K);
        if (mixed) hipLaunchKernelGGL((k_gradient<PV, float, float>), dim3(nb), dim3(kThreads), 0, st, a);
        else hipLaunchKernelGGL((k_gradient<PV, double, double>), dim3(nb), dim3(kThreads), 0, st, a);
    });
    return hipGetLastError();
}

hipError_t launch_clear_resid(unsigned long long *r, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_resid, dim3(1), dim3(64), 0, st, r);
    return hipGetLastError();
}

hipError_t launch_dt_elem(int P, const double *Q, int K, double gamma, double gm1, double *umax, double *amax,
                          hipStream_t st) {
    HG_DISPATCH_P(P, {
        int nb = nblocks<PV>(K);
        hipLaunchKernelGGL((k_dt_elem<PV>), dim3(nb), dim3(kThreads), 0, st, Q, K, gamma, gm1, umax, amax);
    });
    return hipGetLastError();
}

hipError_t launch_dt_scan(int K, const double *umax, const double *amax, const double *dx, double coef, double N1,
                          double dt_max, double *state, double *dt_out, hipStream_t st) {
    hipLaunchKernelGGL(k_dt_scan, dim3(1), dim3(1024), 0, st, K, umax, amax, dx, coef, N1, dt_max, state, dt_out);
    return hipGetLastError();
}

hipError_t launch_cons_sums(int P, const double *Q, const double *jac, const double *w, int K, double *partial,
                            int nbk, double *out, hipStream_t st) {
    HG_DISPATCH_P(P, {
        int nb = nblocks<PV>(K);
        if (nb > nbk) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_cons_partial<PV>), dim3(nb), dim3(kThreads), 0, st, Q, jac, w, K, partial);
        hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, st, partial, nb, out);
    });
    return hipGetLastError();
}

}  // namespace hg
