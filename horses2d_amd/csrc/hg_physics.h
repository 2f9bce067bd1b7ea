// Point-wise physics on the device: fluxes, Riemann solvers, boundary ghost states and the
// face-trace helper shared by every kernel.  Templated on the compute type R (double for the
// FP64 form, float for the mixed form).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hg_device.h"

namespace hg {

// In the FP32 form (R = float) the energy slot of every state handed to these functions holds
// rho*e - Eref, with Eref = pref/(gamma-1) for a reference pressure pref of the handle, so that
// gm1*(rho*e - KE) is the pressure perturbation p - pref: the momentum fluxes carry
// rho*u*u + (p - pref), whose FP32 rounding is relative to the perturbation instead of to p.
// The constant part pref*n has zero discrete divergence (free-stream preservation), and the
// energy flux, sound speed, enthalpy and temperature add the constants back.  The FP64 form
// (R = double) uses the reference's expressions unchanged.
template <typename R>
struct Phys {
    static constexpr bool shift = sizeof(R) == 4;
    R gamma, gm1, cp, cv, lambda, invSGM, mu, cpPr, sigma0, ip_eps, pref, Eref, Href;
    double Eref_d;
    __device__ explicit Phys(const PhysParams &p)
        : gamma((R)p.gamma), gm1((R)p.gm1), cp((R)p.cp), cv((R)p.cv), lambda((R)p.lambda),
          invSGM((R)p.invSGM), mu((R)p.mu), cpPr((R)p.cpPr), sigma0((R)p.sigma0), ip_eps((R)p.ip_eps),
          pref((R)p.pref), Eref((R)p.Eref), Href((R)(p.pref + p.Eref)), Eref_d(p.Eref) {}
};

// thermodynamic pressure from the pressure carried by the flux
template <typename R>
__device__ inline R p_full(const Phys<R> &ph, R pm) {
    if constexpr (Phys<R>::shift) return pm + ph.pref;
    else return pm;
}
// total energy from the energy slot
template <typename R>
__device__ inline R e_full(const Phys<R> &ph, R e) {
    if constexpr (Phys<R>::shift) return e + ph.Eref;
    else return e;
}
// rho*e + p from the energy slot and the flux pressure
template <typename R>
__device__ inline R e_plus_p(const Phys<R> &ph, R e, R pm) {
    if constexpr (Phys<R>::shift) return (e + pm) + ph.Href;
    else return e + pm;
}
// FP64 trace or node state -> the energy slot of the compute type
template <typename R>
__device__ inline double slot_energy(const PhysParams &p, double e) {
    if constexpr (sizeof(R) == 4) return e - p.Eref;
    else return e;
}

// Square root and reciprocal: IEEE in FP64 (the form that matches the oracle); the hardware's
// 1-ulp instructions in FP32, whose fluxes carry FP32 rounding anyway (the correctly rounded
// FP32 sequences cost about five and ten instructions).
__device__ inline double dsqrt(double x) { return sqrt(x); }
__device__ inline float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double rcp(double x) { return 1.0 / x; }
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double dpow(double a, double b) { return pow(a, b); }
__device__ inline float dpow(float a, float b) { return powf(a, b); }
__device__ inline double dabs(double x) { return fabs(x); }
__device__ inline float dabs(float x) { return fabsf(x); }

// ------------------------------------------------------------------ point fluxes
// InviscidFluxes.f90:104-147
template <typename R>
__device__ inline void inviscid_flux(const Phys<R> &ph, const R *u, R *f, R *g) {
    R invRho = rcp(u[0]);
    R vx = u[1] * invRho, vy = u[2] * invRho;
    R p = ph.gm1 * (u[3] - R(0.5) * u[1] * vx - R(0.5) * u[2] * vy);
    const R h = e_plus_p(ph, u[3], p);
    f[0] = u[1]; f[1] = u[1] * vx + p; f[2] = u[1] * vy; f[3] = h * vx;
    g[0] = u[2]; g[1] = f[2]; g[2] = u[2] * vy + p; g[3] = h * vy;
#pragma unroll
    for (int k = 0; k < 4; ++k) { f[k] *= ph.invSGM; g[k] *= ph.invSGM; }
}

// RiemannSolvers.f90:189-305
template <typename R>
__device__ void roe_flux(const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    R rhoL = qL[0], invRhoL = rcp(rhoL), sqrtRhoL = dsqrt(rhoL);
    R uL = (qL[1] * nx + qL[2] * ny) * invRhoL, vL = (-qL[1] * ny + qL[2] * nx) * invRhoL;
    R pL = ph.gm1 * (qL[3] - R(0.5) * rhoL * (uL * uL + vL * vL));
    R HL = ph.cp * p_full(ph, pL) * invRhoL + R(0.5) * (uL * uL + vL * vL);
    R rhoR = qR[0], invRhoR = rcp(rhoR), sqrtRhoR = dsqrt(rhoR);
    R uR = (qR[1] * nx + qR[2] * ny) * invRhoR, vR = (-qR[1] * ny + qR[2] * nx) * invRhoR;
    R pR = ph.gm1 * (qR[3] - R(0.5) * rhoR * (uR * uR + vR * vR));
    R HR = ph.cp * p_full(ph, pR) * invRhoR + R(0.5) * (uR * uR + vR * vR);
    R invrho = rcp(sqrtRhoL + sqrtRhoR);
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
    R rhoL = qL[0], invRhoL = rcp(rhoL);
    R uL = (qL[1] * nx + qL[2] * ny) * invRhoL, vL = (-qL[1] * ny + qL[2] * nx) * invRhoL;
    R pL = ph.gm1 * (qL[3] - R(0.5) * rhoL * (uL * uL + vL * vL));
    R HL = ph.cp * p_full(ph, pL) * invRhoL + R(0.5) * (uL * uL + vL * vL);
    R aL = dsqrt(ph.gamma * p_full(ph, pL) * invRhoL), rhoeL = qL[3];
    R rhoR = qR[0], invRhoR = rcp(rhoR);
    R uR = (qR[1] * nx + qR[2] * ny) * invRhoR, vR = (-qR[1] * ny + qR[2] * nx) * invRhoR;
    R pR = ph.gm1 * (qR[3] - R(0.5) * rhoR * (uR * uR + vR * vR));
    R HR = ph.cp * p_full(ph, pR) * invRhoR + R(0.5) * (uR * uR + vR * vR);
    R aR = dsqrt(ph.gamma * p_full(ph, pR) * invRhoR), rhoeR = qR[3];
    R sL = dsqrt(rhoL), sR = dsqrt(rhoR), invrho = rcp(sL + sR);
    R u = (sL * uL + sR * uR) * invrho, v = (sL * vL + sR * vR) * invrho, H = (sL * HL + sR * HR) * invrho;
    R a = dsqrt(ph.gm1 * (H - R(0.5) * (u * u + v * v)));
    R SL = fmin(uL - aL, u - a), SR = fmax(uR + aR, u + a);
    R Fs[4] = {R(0), R(0), R(0), R(0)};
    if (SL >= R(0)) {
        Fs[0] = rhoL * uL; Fs[1] = Fs[0] * uL + pL; Fs[2] = Fs[0] * vL; Fs[3] = HL * Fs[0];
    } else if (SR <= R(0)) {
        Fs[0] = rhoR * uR; Fs[1] = Fs[0] * uR + pR; Fs[2] = Fs[0] * vR; Fs[3] = HR * Fs[0];
    } else if ((SL < R(0)) || (SR > R(0))) {
        R idS = rcp(SR - SL);
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
// iterations; the star velocity is taken from the last iterate when not converged; a guess or
// iterate that is not a positive number -- an overshoot, where the reference takes the root of
// a negative pressure -- is raised to 1e-6 min(pL, pR); vacuum generation is sampled exactly).
template <typename R>
__device__ void exact_flux(const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    const R g = ph.gamma, gm1 = ph.gm1, cp = ph.cp, cv = ph.cv;
    R qnL[4] = {qL[0], qL[1] * nx + qL[2] * ny, -qL[1] * ny + qL[2] * nx, qL[3]};
    R qnR[4] = {qR[0], qR[1] * nx + qR[2] * ny, -qR[1] * ny + qR[2] * nx, qR[3]};
    R wL[6], wR[6];
    wL[0] = qnL[0]; wL[1] = qnL[1] / qnL[0]; wL[2] = qnL[2] / qnL[0];
    wL[3] = p_full(ph, gm1 * (qnL[3] - R(0.5) * (qnL[1] * wL[1] + qnL[2] * wL[2]))); wL[4] = R(0); wL[5] = dsqrt(g * wL[3] / wL[0]);
    wR[0] = qnR[0]; wR[1] = qnR[1] / qnR[0]; wR[2] = qnR[2] / qnR[0];
    wR[3] = p_full(ph, gm1 * (qnR[3] - R(0.5) * (qnR[1] * wR[1] + qnR[2] * wR[2]))); wR[4] = R(0); wR[5] = dsqrt(g * wR[3] / wR[0]);
    if (R(2) * cv * (wL[5] + wR[5]) <= wR[1] - wL[1]) {
        // Vacuum generation (the pressure positivity condition fails): the reference's
        // two-rarefaction guess is negative and its iteration has no root; the exact solution is
        // two rarefactions around a vacuum (Toro, Riemann Solvers and Numerical Methods for
        // Fluid Dynamics, sec. 4.6), sampled at x/t = 0.
        R Fs[4] = {R(0), R(0), R(0), R(0)};
        if (wL[1] - wL[5] >= R(0)) {
            Fs[0] = qnL[1]; Fs[1] = qnL[1] * wL[1] + wL[3]; Fs[2] = qnL[2] * wL[1];
            Fs[3] = wL[1] * (cp * wL[3] + R(0.5) * (qnL[1] * wL[1] + qnL[2] * wL[2]));
        } else if (wL[1] + R(2) * cv * wL[5] > R(0)) {
            const R uF = R(2) * wL[5] / (g + R(1)) + (g - R(1)) / (g + R(1)) * wL[1];
            const R rs = wL[0] * dpow(uF / wL[5], R(2) * cv), pF = wL[3] * dpow(rs / wL[0], g);
            Fs[0] = rs * uF; Fs[1] = rs * uF * uF + pF; Fs[2] = rs * uF * wL[2];
            Fs[3] = uF * (cp * pF + R(0.5) * rs * (uF * uF + wL[2] * wL[2]));
        } else if (wR[1] - R(2) * cv * wR[5] < R(0) && wR[1] + wR[5] > R(0)) {
            const R uF = -R(2) * wR[5] / (g + R(1)) + (g - R(1)) / (g + R(1)) * wR[1];
            const R rs = wR[0] * dpow(-uF / wR[5], R(2) * cv), pF = wR[3] * dpow(rs / wR[0], g);
            Fs[0] = rs * uF; Fs[1] = rs * uF * uF + pF; Fs[2] = rs * uF * wR[2];
            Fs[3] = uF * (cp * pF + R(0.5) * rs * (uF * uF + wR[2] * wR[2]));
        } else if (wR[1] + wR[5] <= R(0)) {
            Fs[0] = qnR[1]; Fs[1] = qnR[1] * wR[1] + wR[3]; Fs[2] = qnR[2] * wR[1];
            Fs[3] = wR[1] * (cp * wR[3] + R(0.5) * (qnR[1] * wR[1] + qnR[2] * wR[2]));
        }   // else x/t = 0 is in the vacuum: no flux
        if constexpr (Phys<R>::shift) Fs[1] -= ph.pref;
        R Ft = Fs[1] * nx - Fs[2] * ny;
        Fs[2] = Fs[1] * ny + Fs[2] * nx;
        F[0] = Fs[0] * ph.invSGM; F[1] = Ft * ph.invSGM; F[2] = Fs[2] * ph.invSGM; F[3] = Fs[3] * ph.invSGM;
        return;
    }
    R FL, dFL, FR, dFR;
    R ps = dpow((wL[5] + wR[5] - R(0.5) * gm1 * (wR[1] - wL[1])) /
                    (wL[5] / dpow(wL[3], R(0.5) / cp) + wR[5] / dpow(wR[3], R(0.5) / cp)),
                R(2) * cp);
    const R pfloor = R(1.0e-6) * (wL[3] < wR[3] ? wL[3] : wR[3]);
    if (!(ps > R(0))) ps = pfloor;
    for (int it = 1; it <= 50; ++it) {
        exact_F(ph, ps, wL, FL, dFL);
        exact_F(ph, ps, wR, FR, dFR);
        R Fv = FL + FR + wR[1] - wL[1], dF = dFL + dFR, pold = ps;
        ps = pold - Fv / dF;
        if (!(ps > R(0))) ps = pfloor;
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
    if constexpr (Phys<R>::shift) {
        if (Fs[0] != R(0) || Fs[1] != R(0) || Fs[3] != R(0)) Fs[1] -= ph.pref;   // a sampled state was taken
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
    R iL = rcp(qL[0]), iR = rcp(qR[0]);
    R uL = qL[1] * iL, vL = qL[2] * iL, uR = qR[1] * iR, vR = qR[2] * iR;
    R pL = ph.gm1 * (qL[3] - R(0.5) * qL[1] * uL - R(0.5) * qL[2] * vL);
    R pR = ph.gm1 * (qR[3] - R(0.5) * qR[1] * uR - R(0.5) * qR[2] * vR);
    const R hL = e_plus_p(ph, qL[3], pL), hR = e_plus_p(ph, qR[3], pR);
    fL[0] = qL[1]; fL[1] = qL[1] * uL + pL; fL[2] = qL[1] * vL; fL[3] = hL * uL;
    gL[0] = qL[2]; gL[1] = fL[2]; gL[2] = qL[2] * vL + pL; gL[3] = hL * vL;
    fR[0] = qR[1]; fR[1] = qR[1] * uR + pR; fR[2] = qR[1] * vR; fR[3] = hR * uR;
    gR[0] = qR[2]; gR[1] = fR[2]; gR[2] = qR[2] * vR + pR; gR[3] = hR * vR;
    R lL = dabs(uL * nx + vL * ny) + dsqrt(ph.gamma * p_full(ph, pL) * iL);
    R lR = dabs(uR * nx + vR * ny) + dsqrt(ph.gamma * p_full(ph, pR) * iR);
    R lam = lL > lR ? lL : lR;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        R fnL = fL[k] * nx + gL[k] * ny, fnR = fR[k] * nx + gR[k] * ny;
        F[k] = (R(0.5) * (fnL + fnR) - R(0.5) * lam * (qR[k] - qL[k])) * ph.invSGM;
    }
}

// Riemann solver RS fixed at compile time (0 Roe, 1 HLL, 2 exact, 3 Lax-Friedrichs), or
// RS = -1: chosen at run time by id.  A kernel that contains every solver allocates registers
// for the largest of them (the exact solver), so each handle runs the instantiation of its
// own solver and the run-time form is kept for meshes whose zones override the solver.
template <int RS, typename R>
__device__ inline void riemann(int id, const Phys<R> &ph, const R *qL, const R *qR, R nx, R ny, R *F) {
    if constexpr (RS == 0) roe_flux(ph, qL, qR, nx, ny, F);
    else if constexpr (RS == 1) hll_flux(ph, qL, qR, nx, ny, F);
    else if constexpr (RS == 2) exact_flux(ph, qL, qR, nx, ny, F);
    else if constexpr (RS == 3) lf_flux(ph, qL, qR, nx, ny, F);
    else {
        switch (id) {
        case 1: hll_flux(ph, qL, qR, nx, ny, F); break;
        case 2: exact_flux(ph, qL, qR, nx, ny, F); break;
        case 3: lf_flux(ph, qL, qR, nx, ny, F); break;
        default: roe_flux(ph, qL, qR, nx, ny, F); break;
        }
    }
}

// ViscousFluxes.f90:39-80; dq[d*4 + eq]
template <typename R>
__device__ inline void viscous_flux(const Phys<R> &ph, const R *q, const R *dq, R *f, R *g) {
    R invRho = rcp(q[0]), u = q[1] * invRho, v = q[2] * invRho, uDR = u * invRho, vDR = v * invRho;
    const R *dx = dq, *dy = dq + 4;
    R divV = -uDR * dx[0] + invRho * dx[1] - vDR * dy[0] + invRho * dy[2];
    const R E = e_full(ph, q[3]);
    R dxT = ph.gm1 * invRho * (-E * invRho * dx[0] + dx[3] + u * (u * dx[0] - dx[1]) + v * (v * dx[0] - dx[2]));
    R dyT = ph.gm1 * invRho * (-E * invRho * dy[0] + dy[3] + u * (u * dy[0] - dy[1]) + v * (v * dy[0] - dy[2]));
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
    R invRho = rcp(q[0]), u = q[1] * invRho, v = q[2] * invRho, uDR = u * invRho, vDR = v * invRho;
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

// Boundary faces whose viscous flux takes the adiabatic form (flux from the interior state and
// gradient, the boundary state entering only the interior-penalty jump).
__device__ inline bool is_adiabatic(int flags) {
    const int kind = flags & 15;
    return (kind == FK_WALL && !(flags & FF_ISOTHERMAL)) || kind == FK_EULERWALL;
}

// Euler wall, Dirichlet, pressure outlet and pressure inlet from the full-energy trace Qt, in
// the compute type T (the FP32 form: on the full energy, so its boundary pressures carry FP32
// rounding of O(1) values, as its fluxes do).  uB: the Riemann
// solver's outer state or, where the function returns 1 (the reference's WEAK_PRESCRIBED),
// the state whose inviscid normal flux is the face flux; uSB: the viscous boundary state;
// neumann: zero viscous face flux at this node.
template <typename T>
__device__ inline int ghost_state_ext(int kind, int flags, const double *zdd, const PhysParams &p, const T *Qt, T nx,
                                      T ny, T *uB, T *uSB, int &neumann) {
    const T gamma = (T)p.gamma, gm1 = (T)p.gm1, cv = (T)p.cv, cp = (T)p.cp;
    const T zd[4] = {(T)zdd[0], (T)zdd[1], (T)zdd[2], (T)zdd[3]};
    const T rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
    const T vnL = uL * nx + vL * ny;
    neumann = 0;
    if (kind == FK_EULERWALL) {
        // EulerWall_UpdateSolution (EulerWall.incf:96-159): wall state without the normal
        // momentum, energy from p and the tangential momentum; viscous type ADIABATIC
        const T rhovn = Qt[1] * nx + Qt[2] * ny, rhovt = -Qt[1] * ny + Qt[2] * nx;
        const T pw = gm1 * (Qt[3] - T(0.5) * (Qt[1] * Qt[1] + Qt[2] * Qt[2]) / Qt[0]);
        uB[0] = Qt[0];
        uB[1] = Qt[1] - rhovn * nx;
        uB[2] = Qt[2] - rhovn * ny;
        uB[3] = cv * pw + T(0.5) * rhovt * rhovt / Qt[0];
        for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
        return 1;
    }
    const T pL = gm1 * (Qt[3] - T(0.5) * (Qt[1] * uL + Qt[2] * vL));
    const T aL = sqrt(gamma * pL / rhoL);
    const T ML = sqrt(uL * uL + vL * vL) / aL;
    if (kind == FK_DIRICHLET) {
        // Dirichlet.incf:78-104 (uB = the zone state), :106-337 (uSB and the viscous type)
        for (int k = 0; k < 4; ++k) uB[k] = zd[k];
        if (vnL <= T(0)) {
            if (ML > T(1)) {
                for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
            } else {
                uSB[0] = uB[0]; uSB[1] = uB[1]; uSB[2] = uB[2];
                uSB[3] = cv * pL + T(0.5) * (uSB[1] * uSB[1] + uSB[2] * uSB[2]) / uSB[0];
            }
        } else if (ML > T(1)) {
            for (int k = 0; k < 4; ++k) uSB[k] = Qt[k];
            neumann = 1;
        } else {
            const T pB = gm1 * (uB[3] - T(0.5) * (uB[1] * uB[1] + uB[2] * uB[2]) / uB[0]);
            uSB[0] = Qt[0]; uSB[1] = Qt[1]; uSB[2] = Qt[2];
            uSB[3] = cv * pB + T(0.5) * (uSB[1] * uL + uSB[2] * vL);
        }
        return 0;
    }
    if (kind == FK_POUTLET) {
        // PressureOutletBC_UpdateSolution (PressureOutlet.incf:151-450)
        const T rhoI = zd[0], uI = zd[1] / rhoI, vI = zd[2] / rhoI;
        const T pI = gm1 * (zd[3] - T(0.5) * zd[1] * uI - T(0.5) * zd[2] * vI);
        if (ML > T(1)) {
            const T *src = vnL < T(0) ? zd : Qt;
            for (int k = 0; k < 4; ++k) { uB[k] = src[k]; uSB[k] = src[k]; }
            neumann = vnL < T(0) ? 0 : 1;
            return 0;
        }
        T rhoR, uR, vR, pR;
        if (flags & FF_REFLECTIVE) {
            const T R0 = pL - aL * aL * rhoL, Rplus = vnL + T(2) * cv * aL;
            rhoR = -(R0 - pI) / (aL * aL);
            const T aR = sqrt(gamma * pI / rhoR), vnR = Rplus - T(2) * aR * cv;
            uR = uL + (vnR - vnL) * nx;
            vR = vL + (vnR - vnL) * ny;
            pR = pI;
        } else {
            const T nrm = sqrt(uI * uI + vI * vI);
            pR = T(2) * pI - pL;
            uR = vnL < T(0) ? fabs(vnL) * (uI / nrm) : uL;
            vR = vnL < T(0) ? fabs(vnL) * (vI / nrm) : vL;
            rhoR = rhoL;
        }
        uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
        uB[3] = cv * pR + T(0.5) * (uB[1] * uR + uB[2] * vR);
        if (vnL < T(0)) {
            for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
        } else {
            uSB[0] = rhoL; uSB[1] = rhoL * uL; uSB[2] = rhoL * vL;
            uSB[3] = cv * pR + T(0.5) * rhoL * (uL * uL + vL * vL);
        }
        return 0;
    }
    // FK_PINLET and FK_NEWDIRICHLET share the external total conditions.
    // FK_PINLET: PressureInletBC_UpdateSolution (PressureInlet.incf:134-410); uB is the state
    // of the prescribed flux (the curved-edge supersonic outflow reads the trace transposed in
    // the reference, :344; the straight-edge form, :247, is used for both)
    const T uE = zd[1] / zd[0], vE = zd[2] / zd[0];
    const T pE = gm1 * (zd[3] - T(0.5) * (zd[1] * uE + zd[2] * vE));
    const T MExtSq = (uE * uE + vE * vE) / (gamma * pE) * zd[0];
    const T atExtSq = gamma * pE / zd[0] * (T(1) + T(0.5) * gm1 * MExtSq);
    const T sExt = pE / pow(zd[0], gamma);
    const T ptExt = pE * pow(T(1) + T(0.5) * gm1 * MExtSq, cp);
    if (kind == FK_NEWDIRICHLET) {
        // newDirichletBC_UpdateSolution (newDirichlet.incf:146-433): the Riemann solver's outer
        // state; uSB is the same array (:308-310), NEUMANN only at supersonic outflow
        if (vnL <= T(0) && ML > T(1)) {
            for (int k = 0; k < 4; ++k) uB[k] = zd[k];
        } else if (vnL <= T(0)) {
            const T pR = pL, rhoR = pow(pR / sExt, T(1) / gamma), aR = sqrt(gamma * pR / rhoR);
            if (flags & FF_TOTAL_INFLOW) {
                const T nrm = sqrt(uE * uE + vE * vE);
                // M_R^2 = 2/(gamma-1) ((p_hi/p_lo)^((gamma-1)/gamma) - 1) as expm1(k log1p(d)),
                // d = (p_hi - p_lo)/p_lo: the same value without the cancellation of x^k - 1
                // near x = 1 (at Mach 0.3, x^k - 1 = 0.018 would lose 6 bits of FP32)
                const bool rev = pR >= ptExt;
                const T d = rev ? (pR - ptExt) / ptExt : (ptExt - pR) / pR;
                const T m2 = T(2) / gm1 * expm1(gm1 / gamma * log1p(d));
                const T MR = rev ? -sqrt(m2) : sqrt(m2);
                uB[0] = rhoR;
                uB[1] = rhoR * aR * MR * (uE / nrm);
                uB[2] = rhoR * aR * MR * (vE / nrm);
                uB[3] = cv * pR + T(0.5) * rhoR * aR * aR * MR * MR;
            } else {
                uB[0] = zd[0]; uB[1] = zd[1]; uB[2] = zd[2];
                uB[3] = cv * pR + T(0.5) * (zd[1] * zd[1] + zd[2] * zd[2]) / zd[0];
            }
        } else if (ML > T(1)) {
            for (int k = 0; k < 4; ++k) uB[k] = Qt[k];
            neumann = 1;
        } else {
            const T pO = (flags & FF_TOTAL_OUTFLOW) ? ptExt : pE;
            uB[0] = Qt[0]; uB[1] = Qt[1]; uB[2] = Qt[2];
            uB[3] = cv * (T(2) * pO - pL) + T(0.5) * (uB[1] * uL + uB[2] * vL);
        }
        for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
        return 0;
    }
    const T Rplus = vnL + T(2) * cv * aL;
    if (vnL <= T(0) && ML > T(1)) {
        for (int k = 0; k < 4; ++k) { uB[k] = zd[k]; uSB[k] = zd[k]; }
    } else if (vnL <= T(0)) {
        const T a = (gamma + T(1)) / gm1, b = -T(2) * Rplus, c = T(0.5) * gm1 * (Rplus * Rplus) - atExtSq;
        const T aR = (-b + sqrt(b * b - T(4) * a * c)) / (T(2) * a);
        const T rhoR = pow(aR * aR / (gamma * sExt), cv);
        const T pR = aR * aR * rhoR / gamma, vnR = Rplus - T(2) * cv * aR;
        const T uR = vnR * nx, vR = vnR * ny;
        uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
        uB[3] = cv * pR + T(0.5) * (uB[1] * uR + uB[2] * vR);
        uSB[0] = rhoR; uSB[1] = rhoR * uR; uSB[2] = rhoR * vR;
        uSB[3] = cv * pL + T(0.5) * rhoR * (uR * uR + vR * vR);
    } else if (ML > T(1)) {
        for (int k = 0; k < 4; ++k) { uB[k] = Qt[k]; uSB[k] = Qt[k]; }
        neumann = 1;
    } else {
        uB[0] = Qt[0]; uB[1] = Qt[1]; uB[2] = Qt[2];
        uB[3] = cv * (T(2) * ptExt - pL) + T(0.5) * (uB[1] * uL + uB[2] * vL);
        uSB[0] = uB[0]; uSB[1] = uB[1]; uSB[2] = uB[2];
        uSB[3] = cv * ptExt + T(0.5) * (uB[1] * uL + uB[2] * vL);
    }
    return 1;
}

// Ghost states of the non-periodic boundaries, computed from the own trace.
// Farfield.incf:138-376 (uB = uSB; viscous type DIRICHLET, or NEUMANN at supersonic
// outflow -- the reference leaves it unset) and ViscousWall.incf:158-254 in the compute
// type; the other kinds through ghost_state_ext, compiled in only where XB is set (their FP64
// code would otherwise set the register budget of every FP32 kernel).  Returns 1 where uB is
// the state of a prescribed flux instead of the Riemann solver's outer state.
template <typename R, bool XB>
__device__ inline int ghost_state(int kind, int flags, const double *zd, const PhysParams &pp, const Phys<R> &ph,
                                  const R *Qt, R nx, R ny, R *uB, R *uSB, int &neumann) {
    neumann = 0;
    if constexpr (XB) {
      if (kind >= FK_EULERWALL) {
        const R es = (R)(Phys<R>::shift ? pp.Eref : 0.0);
        const R Qd[4] = {Qt[0], Qt[1], Qt[2], Qt[3] + es};
        R uBd[4], uSd[4];
        const int pres = ghost_state_ext<R>(kind, flags, zd, pp, Qd, nx, ny, uBd, uSd, neumann);
#pragma unroll
        for (int k = 0; k < 3; ++k) { uB[k] = uBd[k]; uSB[k] = uSd[k]; }
        uB[3] = uBd[3] - es;
        uSB[3] = uSd[3] - es;
        return pres;
      }
    }
    if (kind == FK_FARFIELD) {
        R q[4] = {(R)zd[0], (R)zd[1], (R)zd[2], (R)(Phys<R>::shift ? zd[3] - ph.Eref_d : zd[3])};
        R rhoI = q[0], uI = q[1] / rhoI, vI = q[2] / rhoI;
        R pI = ph.gm1 * (q[3] - R(0.5) * q[1] * uI - R(0.5) * q[2] * vI);
        R rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
        R vnL = uL * nx + vL * ny;
        R pL = ph.gm1 * (Qt[3] - R(0.5) * Qt[1] * uL - R(0.5) * Qt[2] * vL);
        R aL = dsqrt(ph.gamma * p_full(ph, pL) / rhoL);
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
        uB[3] = iso ? (ph.cv * uB[0] * R(1) * Tw + Ek) - (Phys<R>::shift ? ph.Eref : R(0)) : Qt[3];
        uSB[0] = Qt[0]; uSB[1] = uSB[0] * vwx; uSB[2] = uSB[0] * vwy;
        uSB[3] = iso ? (ph.cv * uSB[0] * R(1) * Tw + R(0.5) * (uSB[1] * uSB[1] + uSB[2] * uSB[2]) / uSB[0]) -
                           (Phys<R>::shift ? ph.Eref : R(0))
                     : Qt[3] - Ek + R(0.5) * uSB[0] * (vwx * vwx + vwy * vwy);
    }
    return 0;
}

// ------------------------------------------------------------------ spectral helpers
// Trace of a face from the staged element state (DGSpatialDiscretizationMethods.f90:1520-1574).
// Always accumulated in FP64 with explicit fma in l order so that every kernel that forms
// a trace (prolong, stage epilogue, stage recomputation) produces the same bits.
template <int P, int STRIDE, int NV, typename T>
__device__ inline void face_trace(const T *sX /* element block, [n*STRIDE + c] */, const double *sLb, int f, int k,
                                  double *out) {
    const double *lb = (f == 0 || f == 3) ? sLb : sLb + P;
    // nodes (k, l) of a horizontal face or (l, k) of a vertical one: a base and a stride, so
    // the loop body has no face-dependent branch
    const bool horiz = (f == 0 || f == 2);
    const int base = horiz ? k : P * k, step = horiz ? P : 1;
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = 0.0;
    // a rolled loop keeps one node's values live (unrolled, all P nodes are loaded up front)
#pragma unroll 1
    for (int l = 0; l < P; ++l) {
        const int n = base + step * l;
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = fma((double)sX[n * STRIDE + c], lb[l], out[c]);
    }
}

// face_trace accumulated in the tile's own type.  The FP32 form uses it for the gradient
// traces, which only the kernel that forms them ever evaluates (the viscous face fluxes are
// stored), so no other kernel needs to reproduce these bits.
template <int P, int STRIDE, int NV, typename T>
__device__ inline void face_trace_r(const T *sX, const T *sLb, int f, int k, T *out) {
    const T *lb = (f == 0 || f == 3) ? sLb : sLb + P;
    const bool horiz = (f == 0 || f == 2);
    const int base = horiz ? k : P * k, step = horiz ? P : 1;
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = T(0);
#pragma unroll 1
    for (int l = 0; l < P; ++l) {
        const int n = base + step * l;
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = fma(sX[n * STRIDE + c], lb[l], out[c]);
    }
}

}  // namespace hg
