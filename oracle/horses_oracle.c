/*
 * horses_oracle.c -- CPU restatement of HORSES2D's DGSEM residual and explicit
 * time integration, used ONLY as the test oracle for the HIP path.
 *
 * TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only to check results or as the
 * CPU baseline. The product path (horses2d_amd) never links or calls it.
 *
 * Every function cites the reference file:line (paths relative to the
 * HORSES2D source tree, Solver/src/...) whose algorithm it restates.  The
 * operation order of the reference is kept (serial face loop in global edge
 * order, volume sums in the reference's index order), and the file is
 * compiled with -ffp-contract=off so that no fused multiply-adds appear,
 * like the reference's gfortran/x86-64 build.
 *
 * Parity status (what pins each part) is listed in the module docstring of
 * oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

#define NCONS 4
#define MAXP 17

/* Reference constants Generic/Defines.h:108-127 */
#define FACE_INTERIOR 0
#define BC_PERIODIC 1
#define BC_DIRICHLET 2
#define BC_EULERWALL 3
#define BC_VISCOUSWALL 4
#define BC_FARFIELD 5
#define BC_PRESSUREOUTLET 6
#define BC_PRESSUREINLET 7
#define BC_NEWDIRICHLET 9
#define WEAK_RIEMANN 1
#define WEAK_PRESCRIBED 2
#define VBC_UNDEFINED 0
#define VBC_DIRICHLET 1
#define VBC_NEUMANN 2
#define VBC_PERIODIC 3
#define VBC_ADIABATIC 4
/* Pressure outlet "Outflow type" (BoundaryConditions/PressureOutlet.incf:66-72) */
#define PARTIALLY_NON_REFLECTIVE_OUTFLOW 0
#define REFLECTIVE_OUTFLOW 1
/* newDirichlet "Mode" and "Outflow pressure" (BoundaryConditions/DGBoundaryConditions.f90:47-54) */
#define SPECIFY_SPEED 1
#define SPECIFY_TOTAL_PRESSURE 2
#define STATIC_PRESSURE 1
#define TOTAL_PRESSURE 2

#define EBOTTOM 1
#define ERIGHT 2
#define ETOP 3
#define ELEFT 4
#define FORWARD 1
#define BACKWARD (-1)

enum { RS_ROE = 0, RS_HLL = 1, RS_EXACT = 2, RS_LF = 3 };
enum { VISC_NONE = 0, VISC_BR1 = 1, VISC_IP = 2 };
/* artificial dissipation type (DGArtificialDissipation.f90:100-137) */
enum { AD_NONE = 0, AD_PHYSICAL = 1, AD_LAPLACIAN = 2 };

static char g_err[512];
static int fail(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); return -1; }
const char *or_last_error(void) { return g_err; }

/* ------------------------------------------------------------------------ */
/* Utilities.f90:37-64  AlmostEqual (Kopriva)                                */
static int almost_equal(double a, double b) {
    if (a == 0.0 || b == 0.0) return fabs(a - b) <= 2.0 * DBL_EPSILON;
    double m = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
    return fabs(b - a) <= 2.0 * DBL_EPSILON * m;
}

/* ======================================================================== */
/* Spectral operators                                                        */
/* ======================================================================== */
typedef struct {
    int N, P;
    double xi[MAXP], w[MAXP], wb[MAXP];
    double D[MAXP][MAXP];     /* D[i][j] = l_j'(x_i) */
    double hatD[MAXP][MAXP];  /* hatD = Minv (M D)^T */
    double lb[2][MAXP];       /* [0]: at x=0 (reference LEFT=2), [1]: at x=1 (RIGHT=1) */
    double lbw[2][MAXP], dlb[2][MAXP], dlbw[2][MAXP];
} Spec;

/* LegendreAlgorithms.f90:55-102 */
static void legendre_poly_and_derivative(int N, double x, double *L, double *Lp) {
    if (N == 0) { *L = 1.0; *Lp = 0.0; return; }
    if (N == 1) { *L = x; *Lp = 1.0; return; }
    double Lm2 = 1.0, Lpm2 = 0.0, Lm1 = x, Lpm1 = 1.0, Ln = 0, Lpn = 0;
    for (int k = 2; k <= N; ++k) {
        Ln = ((2.0 * k - 1.0) * x * Lm1 - (k - 1.0) * Lm2) / k;
        Lpn = Lpm2 + (2.0 * k - 1.0) * Lm1;
        Lm2 = Lm1; Lm1 = Ln; Lpm2 = Lpm1; Lpm1 = Lpn;
    }
    *L = Ln; *Lp = Lpn;
}

/* LegendreAlgorithms.f90:106-182 */
static void gauss_legendre(int N, double *x, double *w) {
    const double PI = 3.141592653589793238462643;
    double tol = 4.0 * DBL_EPSILON, L, Lp;
    if (N == 0) { x[0] = 0.0; w[0] = 2.0; return; }
    if (N == 1) {
        x[0] = -sqrt(1.0 / 3.0); w[0] = 1.0; x[1] = -x[0]; w[1] = w[0];
    } else {
        int ND2 = (N + 1) / 2;
        for (int j = 0; j < ND2; ++j) {
            double xj = -cos((2 * j + 1) * PI / (2 * N + 2));
            for (int k = 0; k <= 100; ++k) {
                legendre_poly_and_derivative(N + 1, xj, &L, &Lp);
                double delta = -L / Lp;
                xj = xj + delta;
                if (fabs(delta) <= tol * fabs(xj)) break;
            }
            legendre_poly_and_derivative(N + 1, xj, &L, &Lp);
            x[j] = xj;
            w[j] = 2.0 / ((1.0 - xj * xj) * Lp * Lp);
            x[N - j] = -xj;
            w[N - j] = w[j];
        }
    }
    if (N % 2 == 0) {
        legendre_poly_and_derivative(N + 1, 0.0, &L, &Lp);
        x[N / 2] = 0.0;
        w[N / 2] = 2.0 / (Lp * Lp);
    }
}

/* InterpolationAndDerivatives.f90:206-231 */
static void barycentric_weights(int N, const double *x, double *w) {
    for (int j = 0; j <= N; ++j) w[j] = 1.0;
    for (int j = 1; j <= N; ++j)
        for (int k = 0; k < j; ++k) {
            w[k] = w[k] * (x[k] - x[j]);
            w[j] = w[j] * (x[j] - x[k]);
        }
    for (int j = 0; j <= N; ++j) w[j] = 1.0 / w[j];
}

/* InterpolationAndDerivatives.f90:106-158 */
static void lagrange_barycentric(double x, int N, const double *nodes, const double *wb, double *p) {
    int match = 0;
    for (int j = 0; j <= N; ++j) {
        p[j] = 0.0;
        if (almost_equal(x, nodes[j])) { p[j] = 1.0; match = 1; }
    }
    if (match) return;
    double d = 0.0;
    for (int j = 0; j <= N; ++j) { double t = wb[j] / (x - nodes[j]); p[j] = t; d = d + t; }
    for (int j = 0; j <= N; ++j) p[j] = p[j] / d;
}

/* InterpolationAndDerivatives.f90:440-469 */
static void derivative_matrix(int N, const double *x, double D[MAXP][MAXP]) {
    double bw[MAXP];
    barycentric_weights(N, x, bw);
    for (int i = 0; i <= N; ++i) {
        D[i][i] = 0.0;
        for (int j = 0; j <= N; ++j)
            if (j != i) {
                D[i][j] = bw[j] / (bw[i] * (x[i] - x[j]));
                D[i][i] = D[i][i] - D[i][j];
            }
    }
}

/* InterpolationAndDerivatives.f90:349-395 (T[k][j], k over new nodes) */
static void interp_matrix(int N, int M, const double *oldx, const double *wb, const double *newx,
                          double *T /* (M+1)*(N+1) */) {
    for (int k = 0; k <= M; ++k) {
        int row = 0;
        for (int j = 0; j <= N; ++j) {
            T[k * (N + 1) + j] = 0.0;
            if (almost_equal(newx[k], oldx[j])) { row = 1; T[k * (N + 1) + j] = 1.0; }
        }
        if (!row) {
            double s = 0.0;
            for (int j = 0; j <= N; ++j) { double t = wb[j] / (newx[k] - oldx[j]); T[k * (N + 1) + j] = t; s = s + t; }
            for (int j = 0; j <= N; ++j) T[k * (N + 1) + j] = T[k * (N + 1) + j] / s;
        }
    }
}

/* NodesAndWeightsClass.f90:251-332 (Legendre-Gauss nodes only) */
static void spec_init(Spec *s, int N) {
    memset(s, 0, sizeof *s);
    s->N = N; s->P = N + 1;
    gauss_legendre(N, s->xi, s->w);
    for (int i = 0; i <= N; ++i) { s->xi[i] = 0.5 + 0.5 * s->xi[i]; s->w[i] = 0.5 * s->w[i]; }
    barycentric_weights(N, s->xi, s->wb);
    derivative_matrix(N, s->xi, s->D);
    /* MD = M*D -> MD(i,j) = w_i D(i,j); trMD = MD^T; hatD = Minv*trMD: hatD(i,l) = (1/w_i)*(w_l D(l,i)) */
    for (int i = 0; i <= N; ++i)
        for (int l = 0; l <= N; ++l) {
            double md = s->w[l] * s->D[l][i];
            /* matmul(Minv, trMD) with diagonal Minv: sum_k Minv(i,k) trMD(k,l); only k=i is nonzero */
            s->hatD[i][l] = (1.0 / s->w[i]) * md;
        }
    lagrange_barycentric(0.0, N, s->xi, s->wb, s->lb[0]);
    lagrange_barycentric(1.0, N, s->xi, s->wb, s->lb[1]);
    for (int sd = 0; sd < 2; ++sd) {
        for (int i = 0; i <= N; ++i) s->lbw[sd][i] = s->lb[sd][i] / s->w[i];
        /* dlb = DT * lb -> dlb(i) = sum_j D(j,i) lb(j) */
        for (int i = 0; i <= N; ++i) {
            double acc = 0.0;
            for (int j = 0; j <= N; ++j) acc = acc + s->D[j][i] * s->lb[sd][j];
            s->dlb[sd][i] = acc;
        }
        for (int i = 0; i <= N; ++i) s->dlbw[sd][i] = s->dlb[sd][i] / s->w[i];
    }
}

static Spec *g_spec[MAXP];
static const Spec *spec_get(int N) {
    if (N < 1 || N >= MAXP - 1) return NULL;
    if (!g_spec[N]) { g_spec[N] = (Spec *)malloc(sizeof(Spec)); spec_init(g_spec[N], N); }
    return g_spec[N];
}

/* Exported for the spectral-operator tests. */
int or_spectral(int N, double *xi, double *w, double *D, double *hatD, double *lb, double *lbw,
                double *dlb, double *dlbw) {
    const Spec *s = spec_get(N);
    if (!s) return fail("N out of range");
    int P = N + 1;
    for (int i = 0; i < P; ++i) {
        xi[i] = s->xi[i]; w[i] = s->w[i];
        for (int j = 0; j < P; ++j) { D[i * P + j] = s->D[i][j]; hatD[i * P + j] = s->hatD[i][j]; }
        for (int sd = 0; sd < 2; ++sd) {
            lb[sd * P + i] = s->lb[sd][i]; lbw[sd * P + i] = s->lbw[sd][i];
            dlb[sd * P + i] = s->dlb[sd][i]; dlbw[sd * P + i] = s->dlbw[sd][i];
        }
    }
    return 0;
}

/* ======================================================================== */
/* Physics (PhysicsNS.f90:97-98, 405-481)                                    */
/* ======================================================================== */
typedef struct {
    double gamma, gm1, gogm1, invgm1, lambda;  /* thermodynamicsAir */
    double Mach, Re, Pr;
    double mu, kappa, cp, cv, invSqrtGammaMach, sqrtGammaMach;
} Phys;

static void phys_init(Phys *ph, double Mach, double Re, double Pr) {
    ph->gamma = 1.4; ph->gm1 = 0.4; ph->gogm1 = 3.5; ph->invgm1 = 2.5; ph->lambda = -2.0 / 3.0;
    ph->Mach = Mach; ph->Re = Re; ph->Pr = Pr;
    ph->sqrtGammaMach = sqrt(ph->gamma) * Mach;
    ph->invSqrtGammaMach = 1.0 / (sqrt(ph->gamma) * Mach);
    if (Re == 0.0) { ph->mu = 0.0; ph->kappa = 0.0; }
    else { ph->mu = 1.0 / Re; ph->kappa = ph->gogm1 / (Pr * Re); }
    ph->cp = ph->gogm1; ph->cv = ph->invgm1;
}

/* InviscidFluxes.f90:104-147 (one node): f = val(:,IX), g = val(:,IY) */
static void inviscid_flux(const Phys *ph, const double *u, double *f, double *g) {
    double invRho = 1.0 / u[0];
    double vx = u[1] * invRho, vy = u[2] * invRho;
    double p = ph->gm1 * (u[3] - 0.5 * u[1] * vx - 0.5 * u[2] * vy);
    f[0] = u[1]; f[1] = u[1] * vx + p; f[2] = u[1] * vy; f[3] = (u[3] + p) * vx;
    g[0] = u[2]; g[1] = f[2]; g[2] = u[2] * vy + p; g[3] = (u[3] + p) * vy;
    for (int k = 0; k < 4; ++k) { f[k] = f[k] * ph->invSqrtGammaMach; g[k] = g[k] * ph->invSqrtGammaMach; }
}

/* RiemannSolvers.f90:189-305 */
static void roe_flux(const Phys *ph, const double *qL, const double *qR, const double *n, double *F) {
    double rhoL = qL[0], invRhoL = 1.0 / rhoL, sqrtRhoL = sqrt(rhoL);
    double uL = (qL[1] * n[0] + qL[2] * n[1]) * invRhoL;
    double vL = (-qL[1] * n[1] + qL[2] * n[0]) * invRhoL;
    double pL = ph->gm1 * (qL[3] - 0.5 * rhoL * (uL * uL + vL * vL));
    double HL = ph->cp * pL * invRhoL + 0.5 * (uL * uL + vL * vL);
    double rhoR = qR[0], invRhoR = 1.0 / rhoR, sqrtRhoR = sqrt(rhoR);
    double uR = (qR[1] * n[0] + qR[2] * n[1]) * invRhoR;
    double vR = (-qR[1] * n[1] + qR[2] * n[0]) * invRhoR;
    double pR = ph->gm1 * (qR[3] - 0.5 * rhoR * (uR * uR + vR * vR));
    double HR = ph->cp * pR * invRhoR + 0.5 * (uR * uR + vR * vR);
    double invrho = 1.0 / (sqrtRhoL + sqrtRhoR);
    double u = (sqrtRhoL * uL + sqrtRhoR * uR) * invrho;
    double v = (sqrtRhoL * vL + sqrtRhoR * vR) * invrho;
    double H = (sqrtRhoL * HL + sqrtRhoR * HR) * invrho;
    double a = sqrt(ph->gm1 * (H - 0.5 * (u * u + v * v)));
    double lam[4] = {u - a, u, u, u + a};
    int neg;
    if (lam[0] > 0.0) neg = 0; else if (lam[1] > 0.0) neg = 1; else if (lam[3] > 0.0) neg = 3; else neg = 4;
    double K[4][4] = {{1.0, u - a, v, H - u * a}, {1.0, u, v, 0.5 * (u * u + v * v)},
                      {0.0, 0.0, 1.0, v}, {1.0, u + a, v, H + u * a}};
    double dq[4] = {rhoR - rhoL, rhoR * uR - rhoL * uL, rhoR * vR - rhoL * vL, qR[3] - qL[3]};
    double al[4];
    al[2] = dq[2] - v * dq[0];
    al[1] = ph->gm1 * (dq[0] * (H - u * u) + u * dq[1] - dq[3] + (dq[2] - v * dq[0]) * v) / (a * a);
    al[0] = (dq[0] * (u + a) - dq[1] - a * al[1]) / (2.0 * a);
    al[3] = dq[0] - (al[0] + al[1]);
    /* F_inviscidFlux(rho,u,v,p,H): InviscidFluxes.f90:149-163 */
    double Fs[4];
    Fs[0] = rhoL * uL; Fs[1] = Fs[0] * uL + pL; Fs[2] = Fs[0] * vL; Fs[3] = HL * Fs[0];
    for (int w = 0; w < neg; ++w)
        for (int k = 0; k < 4; ++k) Fs[k] = Fs[k] + al[w] * lam[w] * K[w][k];
    double Ft = Fs[1] * n[0] - Fs[2] * n[1];
    Fs[2] = Fs[1] * n[1] + Fs[2] * n[0];
    Fs[1] = Ft;
    for (int k = 0; k < 4; ++k) F[k] = Fs[k] * ph->invSqrtGammaMach;
}

/* RiemannSolvers.f90:307-399 */
static void hll_flux(const Phys *ph, const double *qL, const double *qR, const double *n, double *F) {
    double rhoL = qL[0], invRhoL = 1.0 / rhoL;
    double uL = (qL[1] * n[0] + qL[2] * n[1]) * invRhoL;
    double vL = (-qL[1] * n[1] + qL[2] * n[0]) * invRhoL;
    double pL = ph->gm1 * (qL[3] - 0.5 * rhoL * (uL * uL + vL * vL));
    double HL = ph->cp * pL * invRhoL + 0.5 * (uL * uL + vL * vL);
    double aL = sqrt(ph->gamma * pL * invRhoL), rhoeL = qL[3];
    double rhoR = qR[0], invRhoR = 1.0 / rhoR;
    double uR = (qR[1] * n[0] + qR[2] * n[1]) * invRhoR;
    double vR = (-qR[1] * n[1] + qR[2] * n[0]) * invRhoR;
    double pR = ph->gm1 * (qR[3] - 0.5 * rhoR * (uR * uR + vR * vR));
    double HR = ph->cp * pR * invRhoR + 0.5 * (uR * uR + vR * vR);
    double aR = sqrt(ph->gamma * pR * invRhoR), rhoeR = qR[3];
    double sqrtRhoL = sqrt(rhoL), sqrtRhoR = sqrt(rhoR);
    double invrho = 1.0 / (sqrtRhoL + sqrtRhoR);
    double u = (sqrtRhoL * uL + sqrtRhoR * uR) * invrho;
    double v = (sqrtRhoL * vL + sqrtRhoR * vR) * invrho;
    double H = (sqrtRhoL * HL + sqrtRhoR * HR) * invrho;
    double a = sqrt(ph->gm1 * (H - 0.5 * (u * u + v * v)));
    double SL = fmin(uL - aL, u - a), SR = fmax(uR + aR, u + a);
    double Fs[4] = {0, 0, 0, 0};
    if (SL >= 0.0) {
        Fs[0] = rhoL * uL; Fs[1] = Fs[0] * uL + pL; Fs[2] = Fs[0] * vL; Fs[3] = HL * Fs[0];
    } else if (SR <= 0.0) {
        Fs[0] = rhoR * uR; Fs[1] = Fs[0] * uR + pR; Fs[2] = Fs[0] * vR; Fs[3] = HR * Fs[0];
    } else if ((SL < 0.0) || (SR > 0.0)) {
        double invdS = 1.0 / (SR - SL);
        Fs[0] = (SR * rhoL * uL - SL * rhoR * uR + SL * SR * (rhoR - rhoL)) * invdS;
        Fs[1] = (SR * (rhoL * uL * uL + pL) - SL * (rhoR * uR * uR + pR) + SL * SR * (rhoR * uR - rhoL * uL)) * invdS;
        Fs[2] = (SR * rhoL * uL * vL - SL * rhoR * uR * vR + SR * SL * (rhoR * vR - rhoL * vL)) * invdS;
        Fs[3] = (SR * uL * rhoL * HL - SL * uR * rhoR * HR + SR * SL * (rhoeR - rhoeL)) * invdS;
    }
    double Ft = Fs[1] * n[0] - Fs[2] * n[1];
    Fs[2] = Fs[1] * n[1] + Fs[2] * n[0];
    Fs[1] = Ft;
    for (int k = 0; k < 4; ++k) F[k] = Fs[k] * ph->invSqrtGammaMach;
}

/* RiemannSolvers.f90:461-484 */
static void exact_F(const Phys *ph, double p, const double *W, double *F, double *dF) {
    /* W = (rho, u, v, p, T, a) */
    if (p > W[3]) {
        double A = 2.0 / (W[0] * (ph->gamma + 1.0));
        double B = ph->gm1 * W[3] / (ph->gamma + 1.0);
        double s = sqrt(A / (p + B));
        *F = (p - W[3]) * s;
        *dF = s * (1.0 - 0.5 * (p - W[3]) / (B + p));
    } else {
        *F = 2.0 * ph->cv * W[5] * (pow(p / W[3], 0.5 / ph->cp) - 1.0);
        *dF = 1.0 / (W[0] * W[5]) * pow(p / W[3], -0.5 * (ph->gamma + 1.0) / ph->gamma);
    }
}

/* RiemannSolvers.f90:405-459.  Two defined treatments where the reference has none:
 * when the Newton iteration does not converge in 50 iterations the reference leaves
 * ustar undefined, here it comes from the last iterate; and where the two-rarefaction
 * guess or an iterate is not a positive number (an overshoot), from which the
 * reference's next step takes the root of a negative pressure, the pressure is raised to
 * 1e-6 min(pL, pR) instead of going to NaN (vacuum generation is sampled exactly by the
 * caller).  Where the reference's iterates stay positive, nothing changes. */
static void exact_pstar(const Phys *ph, const double *WL, const double *WR, double *pstar, double *ustar) {
    double FL, dFL, FR, dFR;
    const double pfloor = 1.0e-6 * (WL[3] < WR[3] ? WL[3] : WR[3]);
    double ps = pow((WL[5] + WR[5] - 0.5 * ph->gm1 * (WR[1] - WL[1])) /
                        (WL[5] / pow(WL[3], 0.5 / ph->cp) + WR[5] / pow(WR[3], 0.5 / ph->cp)),
                    2.0 * ph->cp);
    if (!(ps > 0.0)) ps = pfloor;
    for (int it = 1; it <= 50; ++it) {
        exact_F(ph, ps, WL, &FL, &dFL);
        exact_F(ph, ps, WR, &FR, &dFR);
        double Fv = FL + FR + WR[1] - WL[1], dF = dFL + dFR;
        double pold = ps;
        ps = pold - Fv / dF;
        if (!(ps > 0.0)) ps = pfloor;
        double cha = 0.5 * fabs(ps - pold) / (ps + pold);
        if (fabs(cha) < 1.0e-6) break;
    }
    exact_F(ph, ps, WL, &FL, &dFL);
    exact_F(ph, ps, WR, &FR, &dFR);
    *ustar = 0.5 * (WL[1] + WR[1] + FR - FL);
    *pstar = ps;
}

/* RiemannSolvers.f90:33-187 */
static void exact_flux(const Phys *ph, const double *qL, const double *qR, const double *n, double *F) {
    double qnL[4], qnR[4], wL[6], wR[6], Fs[4] = {0, 0, 0, 0};
    qnL[0] = qL[0]; qnL[1] = qL[1] * n[0] + qL[2] * n[1]; qnL[2] = -qL[1] * n[1] + qL[2] * n[0]; qnL[3] = qL[3];
    qnR[0] = qR[0]; qnR[1] = qR[1] * n[0] + qR[2] * n[1]; qnR[2] = -qR[1] * n[1] + qR[2] * n[0]; qnR[3] = qR[3];
    double g = ph->gamma, gm1 = ph->gm1, cp = ph->cp, cv = ph->cv;
    wL[0] = qnL[0]; wL[1] = qnL[1] / qnL[0]; wL[2] = qnL[2] / qnL[0];
    wL[3] = gm1 * (qnL[3] - 0.5 * (qnL[1] * wL[1] + qnL[2] * wL[2])); wL[4] = 0; wL[5] = sqrt(g * wL[3] / wL[0]);
    wR[0] = qnR[0]; wR[1] = qnR[1] / qnR[0]; wR[2] = qnR[2] / qnR[0];
    wR[3] = gm1 * (qnR[3] - 0.5 * (qnR[1] * wR[1] + qnR[2] * wR[2])); wR[4] = 0; wR[5] = sqrt(g * wR[3] / wR[0]);
    double pstar, ustar, rhostar, uFan, pFan;
    if (2.0 * cv * (wL[5] + wR[5]) <= wR[1] - wL[1]) {
        /* Vacuum generation (pressure positivity condition fails): the reference's
         * two-rarefaction guess is negative and its iteration has no root; the exact solution
         * is two rarefactions around a vacuum (Toro, Riemann Solvers and Numerical Methods for
         * Fluid Dynamics, sec. 4.6), sampled at x/t = 0 (vacuum itself: no flux). */
        if (wL[1] - wL[5] >= 0.0) {
            Fs[0] = qnL[1]; Fs[1] = qnL[1] * wL[1] + wL[3]; Fs[2] = qnL[2] * wL[1];
            Fs[3] = wL[1] * (cp * wL[3] + 0.5 * (qnL[1] * wL[1] + qnL[2] * wL[2]));
        } else if (wL[1] + 2.0 * cv * wL[5] > 0.0) {
            uFan = 2.0 * wL[5] / (g + 1.0) + (g - 1.0) / (g + 1.0) * wL[1];
            rhostar = wL[0] * pow(uFan / wL[5], 2.0 * cv);
            pFan = wL[3] * pow(rhostar / wL[0], g);
            Fs[0] = rhostar * uFan; Fs[1] = rhostar * uFan * uFan + pFan; Fs[2] = rhostar * uFan * wL[2];
            Fs[3] = uFan * (cp * pFan + 0.5 * rhostar * (uFan * uFan + wL[2] * wL[2]));
        } else if (wR[1] - 2.0 * cv * wR[5] < 0.0 && wR[1] + wR[5] > 0.0) {
            uFan = -2.0 * wR[5] / (g + 1.0) + (g - 1.0) / (g + 1.0) * wR[1];
            rhostar = wR[0] * pow(-uFan / wR[5], 2.0 * cv);
            pFan = wR[3] * pow(rhostar / wR[0], g);
            Fs[0] = rhostar * uFan; Fs[1] = rhostar * uFan * uFan + pFan; Fs[2] = rhostar * uFan * wR[2];
            Fs[3] = uFan * (cp * pFan + 0.5 * rhostar * (uFan * uFan + wR[2] * wR[2]));
        } else if (wR[1] + wR[5] <= 0.0) {
            Fs[0] = qnR[1]; Fs[1] = qnR[1] * wR[1] + wR[3]; Fs[2] = qnR[2] * wR[1];
            Fs[3] = wR[1] * (cp * wR[3] + 0.5 * (qnR[1] * wR[1] + qnR[2] * wR[2]));
        }
        ustar = 1.0; pstar = -1.0;   /* the sampling below is skipped */
    } else {
        exact_pstar(ph, wL, wR, &pstar, &ustar);
    }
    if (pstar < 0.0) {
        /* vacuum: sampled above */
    } else if (ustar >= 0.0) {
        double sh = wL[5] * sqrt((g + 1.0) / (2.0 * g) * pstar / wL[3] + 0.5 / cp);
        if ((pstar <= wL[3]) && (wL[1] >= wL[5])) {
            Fs[0] = qnL[1]; Fs[1] = qnL[1] * wL[1] + wL[3]; Fs[2] = qnL[2] * wL[1];
            Fs[3] = wL[1] * (cp * wL[3] + 0.5 * (qnL[1] * wL[1] + qnL[2] * wL[2]));
        } else if ((pstar <= wL[3]) && (wL[1] < wL[5]) && (ustar < wL[5] * pow(pstar / wL[3], 0.5 / cp))) {
            rhostar = wL[0] * pow(pstar / wL[3], 1.0 / g);
            Fs[0] = rhostar * ustar; Fs[1] = rhostar * ustar * ustar + pstar; Fs[2] = Fs[0] * wL[2];
            Fs[3] = ustar * (cp * pstar + 0.5 * rhostar * (ustar * ustar + wL[2] * wL[2]));
        } else if ((pstar <= wL[3]) && (wL[1] < wL[5]) && (ustar >= wL[5] * pow(pstar / wL[3], 0.5 / cp))) {
            uFan = 2.0 * wL[5] / (g + 1.0) + (g - 1.0) / (g + 1.0) * wL[1];
            rhostar = wL[0] * pow(uFan / wL[5], 2.0 * cv);
            pFan = wL[3] * pow(rhostar / wL[0], g);
            Fs[0] = rhostar * uFan; Fs[1] = rhostar * uFan * uFan + pFan; Fs[2] = rhostar * uFan * wL[2];
            Fs[3] = uFan * (cp * pFan + 0.5 * rhostar * (uFan * uFan + wL[2] * wL[2]));
        } else if ((pstar > wL[3]) && (wL[1] >= sh)) {
            Fs[0] = qnL[1]; Fs[1] = qnL[1] * wL[1] + wL[3]; Fs[2] = qnL[2] * wL[1];
            Fs[3] = wL[1] * (cp * wL[3] + 0.5 * (qnL[1] * wL[1] + qnL[2] * wL[2]));
        } else if ((pstar > wL[3]) && (wL[1] < sh)) {
            rhostar = wL[0] * ((pstar / wL[3] + (g - 1.0) / (g + 1.0)) / (pstar / wL[3] * (g - 1.0) / (g + 1.0) + 1.0));
            Fs[0] = rhostar * ustar; Fs[1] = rhostar * ustar * ustar + pstar; Fs[2] = rhostar * ustar * wL[2];
            Fs[3] = ustar * (cp * pstar + 0.5 * rhostar * (ustar * ustar + wL[2] * wL[2]));
        }
    } else {
        if ((pstar <= wR[3]) && (wR[1] + wR[5] <= 0.0)) {
            Fs[0] = qnR[1]; Fs[1] = qnR[1] * wR[1] + wR[3]; Fs[2] = qnR[2] * wR[1];
            Fs[3] = wR[1] * (cp * wR[3] + 0.5 * (qnR[1] * wR[1] + qnR[2] * wR[2]));
        } else if ((pstar <= wR[3]) && (wR[1] + wR[5] >= 0.0) && (ustar + wR[5] * pow(pstar / wR[3], 0.5 / cp) >= 0.0)) {
            rhostar = wR[0] * pow(pstar / wR[3], 1.0 / g);
            Fs[0] = rhostar * ustar; Fs[1] = rhostar * ustar * ustar + pstar; Fs[2] = Fs[0] * wR[2];
            Fs[3] = ustar * (cp * pstar + 0.5 * rhostar * (ustar * ustar + wR[2] * wR[2]));
        } else if ((pstar <= wR[3]) && (wR[1] + wR[5] > 0.0) && (ustar + wR[5] * pow(pstar / wR[3], 0.5 / cp) < 0.0)) {
            uFan = -2.0 * wR[5] / (g + 1.0) + (g - 1.0) / (g + 1.0) * wR[1];
            rhostar = wR[0] * pow(-uFan / wR[5], 2.0 * cv);
            pFan = wR[3] * pow(rhostar / wR[0], g);
            Fs[0] = rhostar * uFan; Fs[1] = rhostar * uFan * uFan + pFan; Fs[2] = rhostar * uFan * wR[2];
            Fs[3] = uFan * (cp * pFan + 0.5 * rhostar * (uFan * uFan + wR[2] * wR[2]));
        } else if ((pstar > wR[3]) && (wR[1] + wR[5] * (0.5 * (g + 1.0) / g * pstar / wR[3] + 0.5 / cp) <= 0.0)) {
            Fs[0] = qnR[1]; Fs[1] = qnR[1] * wR[1] + wR[3]; Fs[2] = qnR[2] * wR[1];
            Fs[3] = wR[1] * (cp * wR[3] + 0.5 * (qnR[1] * wR[1] + qnR[2] * wR[2]));
        } else if ((pstar > wR[3]) && (wR[1] + wR[5] * ((g + 1.0) / (2.0 * g) * pstar / wR[3] + 0.5 / cp) > 0.0)) {
            rhostar = wR[0] * ((pstar / wR[3] + (g - 1.0) / (g + 1.0)) / (pstar / wR[3] * (g - 1.0) / (g + 1.0) + 1.0));
            Fs[0] = rhostar * ustar; Fs[1] = rhostar * ustar * ustar + pstar; Fs[2] = rhostar * ustar * wR[2];
            Fs[3] = ustar * (cp * pstar + 0.5 * rhostar * (ustar * ustar + wR[2] * wR[2]));
        }
    }
    double Ft = Fs[1] * n[0] - Fs[2] * n[1];
    Fs[2] = Fs[1] * n[1] + Fs[2] * n[0];
    Fs[1] = Ft;
    for (int k = 0; k < 4; ++k) F[k] = Fs[k] * ph->invSqrtGammaMach;
}

/* Lax-Friedrichs (Rusanov) flux.  ABSENT from the reference (only Roe, HLL and
 * Exact are selectable, DGInviscidMethods.f90:143-173); defined here as
 *   F* = 1/2 (Fn(qL) + Fn(qR)) - 1/2 lambda (qR - qL),
 *   lambda = max(|unL| + aL, |unR| + aR),  Fn = (f nx + g ny),
 * scaled by 1/(sqrt(gamma) Mach) like every reference Riemann solver.  Parity
 * against the reference is impossible; it is pinned by consistency,
 * antisymmetry and the dissipation sign. */
static void lf_flux(const Phys *ph, const double *qL, const double *qR, const double *n, double *F) {
    double fL[4], gL[4], fR[4], gR[4];
    double iL = 1.0 / qL[0], iR = 1.0 / qR[0];
    double uL = qL[1] * iL, vL = qL[2] * iL, uR = qR[1] * iR, vR = qR[2] * iR;
    double pL = ph->gm1 * (qL[3] - 0.5 * qL[1] * uL - 0.5 * qL[2] * vL);
    double pR = ph->gm1 * (qR[3] - 0.5 * qR[1] * uR - 0.5 * qR[2] * vR);
    fL[0] = qL[1]; fL[1] = qL[1] * uL + pL; fL[2] = qL[1] * vL; fL[3] = (qL[3] + pL) * uL;
    gL[0] = qL[2]; gL[1] = fL[2]; gL[2] = qL[2] * vL + pL; gL[3] = (qL[3] + pL) * vL;
    fR[0] = qR[1]; fR[1] = qR[1] * uR + pR; fR[2] = qR[1] * vR; fR[3] = (qR[3] + pR) * uR;
    gR[0] = qR[2]; gR[1] = fR[2]; gR[2] = qR[2] * vR + pR; gR[3] = (qR[3] + pR) * vR;
    double unL = uL * n[0] + vL * n[1], unR = uR * n[0] + vR * n[1];
    double aL = sqrt(ph->gamma * pL * iL), aR = sqrt(ph->gamma * pR * iR);
    double lL = fabs(unL) + aL, lR = fabs(unR) + aR;
    double lam = lL > lR ? lL : lR;
    for (int k = 0; k < 4; ++k) {
        double fnL = fL[k] * n[0] + gL[k] * n[1];
        double fnR = fR[k] * n[0] + gR[k] * n[1];
        F[k] = (0.5 * (fnL + fnR) - 0.5 * lam * (qR[k] - qL[k])) * ph->invSqrtGammaMach;
    }
}

typedef void (*RiemannFn)(const Phys *, const double *, const double *, const double *, double *);
static RiemannFn riemann_fn(int id) {
    switch (id) {
    case RS_ROE: return roe_flux;
    case RS_HLL: return hll_flux;
    case RS_EXACT: return exact_flux;
    case RS_LF: return lf_flux;
    }
    return NULL;
}

/* Exported point-wise Riemann flux (tests). */
int or_riemann(int solver, double Mach, int n, const double *qL, const double *qR, const double *nrm, double *F) {
    Phys ph; phys_init(&ph, Mach, 0.0, 0.72);
    RiemannFn fn = riemann_fn(solver);
    if (!fn) return fail("unknown Riemann solver");
    for (int i = 0; i < n; ++i) fn(&ph, qL + 4 * i, qR + 4 * i, nrm + 2 * i, F + 4 * i);
    return 0;
}

/* ViscousFluxes.f90:39-80 (one node), dq[k][dim] -> dq[k + 4*dim] */
static void viscous_flux(const Phys *ph, const double *q, const double *dq, double *f, double *g) {
    double invRho = 1.0 / q[0], u = q[1] * invRho, v = q[2] * invRho;
    double uDR = u * invRho, vDR = v * invRho;
    const double *dx = dq, *dy = dq + 4;
    double divV = -uDR * dx[0] + invRho * dx[1] - vDR * dy[0] + invRho * dy[2];
    double dxT = ph->gm1 * invRho * (-q[3] * invRho * dx[0] + dx[3] + u * (u * dx[0] - dx[1]) + v * (v * dx[0] - dx[2]));
    double dyT = ph->gm1 * invRho * (-q[3] * invRho * dy[0] + dy[3] + u * (u * dy[0] - dy[1]) + v * (v * dy[0] - dy[2]));
    double cpPr = ph->cp / ph->Pr;
    f[0] = 0.0;
    f[1] = 2.0 * (-uDR * dx[0] + invRho * dx[1]) + ph->lambda * divV;
    f[2] = -vDR * dx[0] - uDR * dy[0] + invRho * dy[1] + invRho * dx[2];
    f[3] = u * f[1] + v * f[2] + cpPr * dxT;
    g[0] = 0.0;
    g[1] = f[2];
    g[2] = 2.0 * (-vDR * dy[0] + invRho * dy[2]) + ph->lambda * divV;
    g[3] = u * g[1] + v * g[2] + cpPr * dyT;
}

/* ViscousFluxes.f90:279-318 (adiabatic; qB enters only the energy row) */
static void adiabatic_viscous_flux(const Phys *ph, const double *q, const double *qB, const double *dq, double *f, double *g) {
    double invRho = 1.0 / q[0], u = q[1] * invRho, v = q[2] * invRho;
    double uDR = u * invRho, vDR = v * invRho;
    const double *dx = dq, *dy = dq + 4;
    double divV = -uDR * dx[0] + invRho * dx[1] - vDR * dy[0] + invRho * dy[2];
    f[0] = 0.0;
    f[1] = 2.0 * (-uDR * dx[0] + invRho * dx[1]) + ph->lambda * divV;
    f[2] = -vDR * dx[0] - uDR * dy[0] + invRho * dy[1] + invRho * dx[2];
    g[0] = 0.0;
    g[1] = f[2];
    g[2] = 2.0 * (-vDR * dy[0] + invRho * dy[2]) + ph->lambda * divV;
    f[3] = (qB[1] * f[1] + qB[2] * g[1]) / qB[0];
    g[3] = (qB[1] * f[2] + qB[2] * g[2]) / qB[0];
}

/* ======================================================================== */
/* Mesh                                                                      */
/* ======================================================================== */
/* A subdivided edge (SubdividedEdge_t): its two mortars (NORTH: hanging node to node 2, SOUTH:
 * node 1 to the hanging node), their orders N_N, N_S, and the eight transformation matrices
 * (SubdividedEdge_ComputeMortarsTransformationMatrices, QuadElement_EdgeProcedures.incf:530-634),
 * row-major [rows][cols]; the RIGHT_SOUTH element's traces (the third storage). */
typedef struct {
    int NN, NS;
    double *TLN_F, *TLS_F, *TLN_B, *TLS_B, *TRN_F, *TRN_B, *TRS_F, *TRS_B;
    /* the mortars' normals and surface Jacobians per mortar node (QuadElement_EdgeMappings.incf:450-565):
     * straight, the edge's n and 0.5 dS at every node; curved, from the edge's dS n interpolated at
     * 0.5 + 0.5 xi_N and 0.5 xi_S */
    double nN[2 * MAXP], nS[2 * MAXP], dSN[MAXP], dSS[MAXP];
    double tq2[NCONS * MAXP];
    double tdq2[2 * NCONS * MAXP];
} SubData;

typedef struct {
    int id, bdry, curved;
    int sub;               /* subdivided edge: nodes[2] the hanging node, quads LEFT / RIGHT_NORTH / RIGHT_SOUTH */
    SubData *sd;
    int nodes[3];          /* node ids after linking inversions */
    int N, NLow, NCurve;
    /* p-non-conforming edge (Edge_LinkWithElements, QuadElement_EdgeProcedures.incf:105-124,
     * 154-168): the side (0 LEFT, 1 RIGHT) whose order NLow is below the edge's N, or -1;
     * T_forward [i*(NLow+1) + l] (edge node i from the low side's node l) and
     * T_backward [i*(N+1) + l] = Minv_low T_forward^T M_edge (low node i from edge node l) */
    int low;
    double *Tf, *Tb;
    int quads[3];          /* [0] LEFT element (lowest id) / boundary element, [1] RIGHT (subdivided: [1] NORTH, [2] SOUTH) */
    int loc[3];            /* edge location 1..4 in quads[0..2] */
    int inverse;
    int marker;
    double *X, *dX, *n, *dS; /* X,dX,n: 2*P (straight: dX,n,dS only [0]) */
    double Area, invh;
    double mu_a;            /* artificial viscosity of the edge (ComputeEdgeViscosity) */
    double lap;             /* Laplacian dissipation: the edge's EdgeViscosity (mean of its elements') */
    /* boundary data */
    int inviscidBCType, rs, associated, partner;
    int vbc[MAXP];
    double uB[NCONS * MAXP], uSB[NCONS * MAXP];
    double FB[NCONS * MAXP];   /* WEAK_PRESCRIBED: F.n per node (times dS when applied) */
    double *fq;                /* boundary-state function values per node [4 P] (or NULL: the zone state) */
    /* traces: storage(LEFT)=tq[0], storage(RIGHT)=tq[1] (boundary: tq[0]) */
    double tq[2][NCONS * MAXP];
    double tdq[2][2 * NCONS * MAXP]; /* [eq + 4*i + 4*P*dim] */
} Edge;

typedef struct {
    int N, P;
    int nodes[4];
    int edges[4];   /* edge index per position 1..4 */
    int dir[4];     /* edgesDirection */
    int qpos[4];    /* 0 = LEFT/boundary, 1 = RIGHT (storage index of this element on the edge) */
    double *X;      /* X[(i + P*j)*2 + d] */
    double *Ja;     /* Ja[(r*2+c)*P*P + i + P*j], r,c in {0,1} = (IX,IY) */
    double *jac;
    double Volume;
    int64_t off;    /* node offset */
} Elem;

typedef struct {
    int type;       /* BC_* */
    int direction;  /* periodic: 1=x, 2=y */
    int partner;    /* periodic connected marker */
    int wall_iso;   /* viscous wall: 0 adiabatic, 1 isothermal */
    int outflow;    /* pressure outlet: PARTIALLY_NON_REFLECTIVE_OUTFLOW or REFLECTIVE_OUTFLOW */
    int nd_mode;    /* newDirichlet: SPECIFY_SPEED or SPECIFY_TOTAL_PRESSURE */
    int nd_outflow; /* newDirichlet: STATIC_PRESSURE or TOTAL_PRESSURE */
    double q[4];    /* far field, Dirichlet, pressure outlet / inlet: the zone state (dimensionless) */
    double Tw, vw[2];
    int rs;         /* zone Riemann solver (-1 = interior method) */
    int associated;
    int nedges; int *edges;
} Zone;

typedef struct {
    Phys ph;
    int ns, riemann, viscous;
    double sigma0, ip_eps; int ip_gradflux;
    int ad; double ad_ceps; double *mu_e;   /* artificial dissipation: type (AD_*), Ceps, per element mu_a */
    double *lap_e;                          /* Laplacian dissipation: per element ElementViscosity */
    int nnodes; double *pts;
    int K; Elem *el;
    int E; Edge *ed;
    int nmarkers; Zone *zones;
    int64_t nodes_total, ndof;
    double *Q, *QDot, *dQ, *G;
    double dt_umax, dt_amax;  /* SAVEd running maxima of EstimateTimeStep */
    int traces_valid;         /* the edge traces are those of the last derivative evaluation */
} Mesh;

/* Mesh input (the content of a HiOMesh file, coordinates already divided by
 * the Reynolds length, ids 0-based). */
typedef struct {
    int nnodes; const double *points;
    int K; const int *quads; const int *order;
    int nbedges; const int *bedges; const int *bmarker;
    int ncurved; int np1; const int *cedges; const double *cxy; /* cxy[(c*np1 + k)*2 + d] */
    int nmarkers;
    const int *bc_type; const int *bc_int; const double *bc_real; const int *bc_rs;
} MeshIn;

static int same_face(const int *a, const int *b) {
    return (a[0] == b[0] && a[1] == b[1]) || (a[0] == b[1] && a[1] == b[0]);
}

/* QuadElement_Auxiliars.incf:27-117 searchEdge/compareEdges */
static void search_edge(const int *el_nodes, const int *edge_nodes, int *needsInv, int *pos, int *dirn) {
    for (int e = 1; e <= 4; ++e) {
        int c0 = el_nodes[e - 1], c1 = el_nodes[e % 4];
        int fwd_pos = (e == EBOTTOM || e == ERIGHT);
        if (c0 == edge_nodes[0] && c1 == edge_nodes[1]) {
            *pos = e; *dirn = fwd_pos ? FORWARD : BACKWARD; *needsInv = (*dirn != FORWARD); return;
        }
        if (c0 == edge_nodes[1] && c1 == edge_nodes[0]) {
            *pos = e; *dirn = fwd_pos ? BACKWARD : FORWARD; *needsInv = (*dirn != FORWARD); return;
        }
    }
    *pos = -1; *dirn = 0; *needsInv = 0;
}

/* QuadElement_EdgeProcedures.incf:433-521 Edge_Invert */
static void edge_invert(Edge *ed, int changeSign) {
    int P = ed->N + 1;
    int t = ed->nodes[0]; ed->nodes[0] = ed->nodes[1]; ed->nodes[1] = t;
    for (int i = 0; i < P / 2; ++i)
        for (int d = 0; d < 2; ++d) {
            double a = ed->X[2 * i + d]; ed->X[2 * i + d] = ed->X[2 * (P - 1 - i) + d]; ed->X[2 * (P - 1 - i) + d] = a;
        }
    if (!ed->curved) {
        ed->dX[0] = -ed->dX[0]; ed->dX[1] = -ed->dX[1];
        if (changeSign) { ed->n[0] = -ed->n[0]; ed->n[1] = -ed->n[1]; }
    } else {
        double tdX[2 * MAXP], tn[2 * MAXP], tdS[MAXP];
        memcpy(tdX, ed->dX, sizeof(double) * 2 * P); memcpy(tn, ed->n, sizeof(double) * 2 * P); memcpy(tdS, ed->dS, sizeof(double) * P);
        for (int i = 0; i < P; ++i) {
            int r = P - 1 - i;
            ed->dX[2 * i] = -tdX[2 * r]; ed->dX[2 * i + 1] = -tdX[2 * r + 1];
            ed->dS[i] = tdS[r];
            double s = changeSign ? -1.0 : 1.0;
            ed->n[2 * i] = s * tn[2 * r]; ed->n[2 * i + 1] = s * tn[2 * r + 1];
        }
    }
}

/* QuadElement_EdgeMappings.incf:22-50 (straight) and :185-260 (curved) */
static void edge_set_curve(Mesh *m, Edge *ed, const double *cpts /* curved points (np1 x 2) or NULL */, int order) {
    const Spec *s = spec_get(ed->N);
    int P = ed->N + 1;
    const double *n1 = m->pts + 2 * ed->nodes[0], *n2 = m->pts + 2 * ed->nodes[1];
    if (!ed->curved) {
        for (int p = 0; p < P; ++p)
            for (int d = 0; d < 2; ++d) ed->X[2 * p + d] = n1[d] * (1.0 - s->xi[p]) + n2[d] * s->xi[p];
        double dS0 = n2[1] - n1[1], dS1 = n1[0] - n2[0];
        ed->dS[0] = sqrt(dS0 * dS0 + dS1 * dS1);
        ed->n[0] = dS0 / ed->dS[0]; ed->n[1] = dS1 / ed->dS[0];
        ed->dX[0] = n2[0] - n1[0]; ed->dX[1] = n2[1] - n1[1];
        return;
    }
    const double PI = 3.141592653589793238462643;
    int NC = ed->NCurve;
    double cgl[MAXP], cgli[MAXP], wb[MAXP], T[MAXP * MAXP], Xi[2 * MAXP];
    for (int k = 0; k <= order; ++k) cgl[k] = 0.5 + 0.5 * cos(PI * (order - k) / (1.0 * order));
    for (int k = 0; k <= NC; ++k) cgli[k] = 0.5 + 0.5 * cos(PI * (NC - k) / (1.0 * NC));
    barycentric_weights(order, cgl, wb);
    interp_matrix(order, NC, cgl, wb, cgli, T);
    /* Xintermediate = points * T^T : Xi(d,k) = sum_j points(d,j) T(k,j) */
    for (int k = 0; k <= NC; ++k)
        for (int d = 0; d < 2; ++d) {
            double acc = 0.0;
            for (int j = 0; j <= order; ++j) acc = acc + cpts[2 * j + d] * T[k * (order + 1) + j];
            Xi[2 * k + d] = acc;
        }
    barycentric_weights(NC, cgli, wb);
    interp_matrix(NC, ed->N, cgli, wb, s->xi, T);
    for (int i = 0; i < P; ++i)
        for (int d = 0; d < 2; ++d) {
            double acc = 0.0;
            for (int k = 0; k <= NC; ++k) acc = acc + Xi[2 * k + d] * T[i * (NC + 1) + k];
            ed->X[2 * i + d] = acc;
        }
    /* dX = X * D^T : dX(d,i) = sum_j X(d,j) D(i,j) */
    for (int i = 0; i < P; ++i)
        for (int d = 0; d < 2; ++d) {
            double acc = 0.0;
            for (int j = 0; j < P; ++j) acc = acc + ed->X[2 * j + d] * s->D[i][j];
            ed->dX[2 * i + d] = acc;
        }
    for (int i = 0; i < P; ++i) {
        ed->dS[i] = sqrt(ed->dX[2 * i] * ed->dX[2 * i] + ed->dX[2 * i + 1] * ed->dX[2 * i + 1]);
        ed->n[2 * i] = ed->dX[2 * i + 1] / ed->dS[i];
        ed->n[2 * i + 1] = -ed->dX[2 * i] / ed->dS[i];
    }
}

/* EvaluateX of an edge at x, forwards: Edge_AnalyticalX from the nodes, CurvedEdge_InterpolantX
 * from the edge's points (QuadElement_EdgeMappings.incf:52-78, 355-380). */
static void edge_point_at(const Mesh *m, const Edge *ed, double x, double *X) {
    if (!ed->curved) {
        const double *n1 = m->pts + 2 * ed->nodes[0], *n2 = m->pts + 2 * ed->nodes[1];
        for (int d = 0; d < 2; ++d) X[d] = n1[d] * (1.0 - x) + n2[d] * x;
        return;
    }
    const Spec *s = spec_get(ed->N);
    double lj[MAXP];
    lagrange_barycentric(x, ed->N, s->xi, s->wb, lj);
    for (int d = 0; d < 2; ++d) {
        double a = 0.0;
        for (int j = 0; j <= ed->N; ++j) a = a + ed->X[2 * j + d] * lj[j];
        X[d] = a;
    }
}

/* Edge geometry seen by an element of order eN at its local coordinate index iXi
 * (QuadElement_EdgeMappings.incf:52-177, 262-442): interior edges are evaluated at the
 * element's node xi(iXi) (QuadElement_QuadMappings.incf:55-62), which is not one of the edge's
 * nodes when the edge takes the higher order of its two elements. */
static void edge_eval_X(Mesh *m, Edge *ed, int eN, int iXi, int dirn, int qpos, double *X, double *dX) {
    const Spec *s = spec_get(ed->N);
    int P = ed->N + 1;
    if (ed->bdry) { /* getX/getdX */
        int ci = (dirn == FORWARD) ? iXi : ed->N - iXi;
        X[0] = ed->X[2 * ci]; X[1] = ed->X[2 * ci + 1];
        if (!ed->curved) { dX[0] = (double)dirn * ed->dX[0]; dX[1] = (double)dirn * ed->dX[1]; }
        else { dX[0] = (double)dirn * ed->dX[2 * ci]; dX[1] = (double)dirn * ed->dX[2 * ci + 1]; }
        return;
    }
    if (ed->sub) {
        /* a subdivided edge, always evaluated forwards: the big element on [0, 1], the NORTH
         * element on [0.5, 1], the SOUTH one on [0, 0.5], a backward small element at its nodes
         * reversed, dX scaled by the half and the direction (QuadElement_QuadMappings.incf:66-115) */
        const double *xs = spec_get(eN)->xi;
        double xv = xs[iXi], scal = 1.0;
        if (qpos == 1) { xv = 0.5 + 0.5 * (dirn == BACKWARD ? xs[eN - iXi] : xs[iXi]); scal = dirn == BACKWARD ? -0.5 : 0.5; }
        else if (qpos == 2) { xv = 0.5 * (dirn == BACKWARD ? xs[eN - iXi] : xs[iXi]); scal = dirn == BACKWARD ? -0.5 : 0.5; }
        edge_point_at(m, ed, xv, X);
        if (!ed->curved) {
            const double *n1 = m->pts + 2 * ed->nodes[0], *n2 = m->pts + 2 * ed->nodes[1];
            for (int d = 0; d < 2; ++d) dX[d] = scal * (n2[d] - n1[d]);
        } else {
            double lj[MAXP];
            lagrange_barycentric(xv, ed->N, s->xi, s->wb, lj);
            for (int d = 0; d < 2; ++d) {
                double b = 0.0;
                for (int j = 0; j < P; ++j) b = b + ed->dX[2 * j + d] * lj[j];
                dX[d] = scal * b;
            }
        }
        return;
    }
    double xi = spec_get(eN)->xi[iXi];
    double cxi = (dirn == BACKWARD) ? 1.0 - xi : xi;
    if (!ed->curved) {
        const double *n1 = m->pts + 2 * ed->nodes[0], *n2 = m->pts + 2 * ed->nodes[1];
        for (int d = 0; d < 2; ++d) {
            X[d] = n1[d] * (1.0 - cxi) + n2[d] * cxi;
            dX[d] = (double)dirn * (n2[d] - n1[d]);
        }
        return;
    }
    double lj[MAXP];
    lagrange_barycentric(cxi, ed->N, s->xi, s->wb, lj);
    for (int d = 0; d < 2; ++d) {
        double a = 0.0, b = 0.0;
        for (int j = 0; j < P; ++j) { a = a + ed->X[2 * j + d] * lj[j]; b = b + ed->dX[2 * j + d] * lj[j]; }
        X[d] = a; dX[d] = (double)dirn * b;
    }
}

/* QuadElement_QuadMappings.incf:22-367 (transfinite map) */
static void elem_set_mappings(Mesh *m, Elem *e) {
    const Spec *s = spec_get(e->N);
    int P = e->P;
    double *Xe = (double *)malloc(sizeof(double) * 2 * P * P * 4), *dXe = (double *)malloc(sizeof(double) * 2 * P * P * 4);
#define XE(d, i, j, ed) Xe[(((ed) * P + (j)) * P + (i)) * 2 + (d)]
#define DXE(d, i, j, ed) dXe[(((ed) * P + (j)) * P + (i)) * 2 + (d)]
    for (int pos = EBOTTOM; pos <= ETOP; pos += 2) {
        Edge *ed = &m->ed[e->edges[pos - 1]];
        for (int i = 0; i < P; ++i) {
            double X[2], dX[2];
            edge_eval_X(m, ed, e->N, i, e->dir[pos - 1], e->qpos[pos - 1], X, dX);
            for (int j = 0; j < P; ++j)
                for (int d = 0; d < 2; ++d) { XE(d, i, j, pos - 1) = X[d]; DXE(d, i, j, pos - 1) = dX[d]; }
        }
    }
    for (int pos = ERIGHT; pos <= ELEFT; pos += 2) {
        Edge *ed = &m->ed[e->edges[pos - 1]];
        for (int j = 0; j < P; ++j) {
            double X[2], dX[2];
            edge_eval_X(m, ed, e->N, j, e->dir[pos - 1], e->qpos[pos - 1], X, dX);
            for (int i = 0; i < P; ++i)
                for (int d = 0; d < 2; ++d) { XE(d, i, j, pos - 1) = X[d]; DXE(d, i, j, pos - 1) = dX[d]; }
        }
    }
    const double *n1 = m->pts + 2 * e->nodes[0], *n2 = m->pts + 2 * e->nodes[1];
    const double *n3 = m->pts + 2 * e->nodes[2], *n4 = m->pts + 2 * e->nodes[3];
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < P; ++j) {
            double xi = s->xi[i], eta = s->xi[j];
            double dx[2][2];
            for (int d = 0; d < 2; ++d) {
                e->X[(i + P * j) * 2 + d] = (1.0 - eta) * XE(d, i, j, 0) + xi * XE(d, i, j, 1) + eta * XE(d, i, j, 2) +
                                            (1.0 - xi) * XE(d, i, j, 3) - n1[d] * (1.0 - xi) * (1.0 - eta) -
                                            n2[d] * xi * (1.0 - eta) - n3[d] * xi * eta - n4[d] * (1.0 - xi) * eta;
                dx[d][0] = (1.0 - eta) * DXE(d, i, j, 0) + XE(d, i, j, 1) + eta * DXE(d, i, j, 2) - XE(d, i, j, 3) +
                           n1[d] * (1.0 - eta) - n2[d] * (1.0 - eta) - n3[d] * eta + n4[d] * eta;
                dx[d][1] = -XE(d, i, j, 0) + xi * DXE(d, i, j, 1) + XE(d, i, j, 2) + (1.0 - xi) * DXE(d, i, j, 3) +
                           n1[d] * (1.0 - xi) + n2[d] * xi - n3[d] * xi - n4[d] * (1.0 - xi);
            }
            int k = i + P * j, PP = P * P;
            e->jac[k] = dx[0][0] * dx[1][1] - dx[0][1] * dx[1][0];
            e->Ja[0 * PP + k] = dx[1][1];   /* Ja(IX,IX) =  Y_eta */
            e->Ja[1 * PP + k] = -dx[1][0];  /* Ja(IX,IY) = -Y_xi  */
            e->Ja[2 * PP + k] = -dx[0][1];  /* Ja(IY,IX) = -X_eta */
            e->Ja[3 * PP + k] = dx[0][0];   /* Ja(IY,IY) =  X_xi  */
        }
#undef XE
#undef DXE
    free(Xe); free(dXe);
}

static void mesh_free(Mesh *m) {
    if (!m) return;
    for (int e = 0; e < m->K; ++e) { free(m->el[e].X); free(m->el[e].Ja); free(m->el[e].jac); }
    for (int i = 0; i < m->E; ++i) {
        free(m->ed[i].X); free(m->ed[i].dX); free(m->ed[i].n); free(m->ed[i].dS);
        free(m->ed[i].Tf); free(m->ed[i].Tb); free(m->ed[i].fq);
        if (m->ed[i].sd) {
            SubData *d = m->ed[i].sd;
            free(d->TLN_F); free(d->TLS_F); free(d->TLN_B); free(d->TLS_B);
            free(d->TRN_F); free(d->TRN_B); free(d->TRS_F); free(d->TRS_B); free(d);
        }
    }
    for (int z = 0; z < m->nmarkers; ++z) free(m->zones[z].edges);
    free(m->el); free(m->ed); free(m->zones); free(m->pts);
    free(m->Q); free(m->QDot); free(m->dQ); free(m->G); free(m->mu_e); free(m->lap_e);
    free(m);
}

/* The points_of_edges list (IntegerArrayLinkedList.f90) as arrays: entry k holds fnn[k] node
 * ids fn[3k..] (2, or 3 for a merged subdivided edge) and a boundary attribute fat[k]. */
static int entry_has(const int *fn, const int *fnn, int k, int node) {
    for (int c = 0; c < fnn[k]; ++c)
        if (fn[3 * k + c] == node) return 1;
    return 0;
}

/* IntegerArrayLinkedList_SearchIfPresent (:489-567): the first `how` entries holding `node`,
 * -1 padded; their attributes. */
static void search_if_present(const int *fn, const int *fnn, const int *fat, int E, int node, int how, int *pos,
                              int *att) {
    int c = 0;
    for (int k = 0; k < E && c < how; ++k)
        if (entry_has(fn, fnn, k, node)) { pos[c] = k; if (att) att[c] = fat[k]; ++c; }
    for (; c < how; ++c) { pos[c] = -1; if (att) att[c] = 0; }
}

/* IntegerArrayLinkedList_Search (:346-412) for a two-node entry, either order. */
static int search_pair(const int *fn, const int *fnn, int E, int a, int b) {
    for (int k = 0; k < E; ++k)
        if (fnn[k] == 2 && ((fn[3 * k] == a && fn[3 * k + 1] == b) || (fn[3 * k] == b && fn[3 * k + 1] == a))) return k;
    return -1;
}

static void remove_entry(int *fn, int *fnn, int *fat, int *E, int k) {
    if (k < 0) return;
    memmove(fn + 3 * k, fn + 3 * (k + 1), sizeof(int) * 3 * (*E - k - 1));
    memmove(fnn + k, fnn + k + 1, sizeof(int) * (*E - k - 1));
    memmove(fat + k, fat + k + 1, sizeof(int) * (*E - k - 1));
    --*E;
}

/* mergeDividedFaces (MeshFile.f90:364-490), nodes in ascending order: a node with exactly three
 * edges, none of them a boundary edge, whose neighbours are joined by a fourth edge is a hanging
 * node; that edge and the two halves are removed and [n1, n2, hanging] appended. */
static void merge_divided(int nnodes, int *fn, int *fnn, int *fat, int *E) {
    for (int node = 0; node < nnodes; ++node) {
        int e4[4], at[4];
        search_if_present(fn, fnn, fat, *E, node, 4, e4, at);
        if (e4[3] != -1 || e4[2] == -1 || at[0] || at[1] || at[2] || at[3]) continue;
        int pts[3];
        for (int j = 0; j < 3; ++j) {
            const int *v = fn + 3 * e4[j];
            pts[j] = v[0] != node ? v[0] : (v[1] != node ? v[1] : -1);
        }
        int done = 0;
        for (int j = 0; j < 3 && !done; ++j) {
            int ne[8];
            search_if_present(fn, fnn, fat, *E, pts[j], 8, ne, NULL);
            for (int k = 0; k < 8 && !done; ++k) {
                if (ne[k] == -1) continue;
                const int p0 = fn[3 * ne[k]], p1 = fn[3 * ne[k] + 1];
                int point = -1;
                if (p0 != pts[j] && p0 != node) point = p0;
                else if (p1 != pts[j] && p1 != node) point = p1;
                if (point != -1 && (point == pts[0] || point == pts[1] || point == pts[2])) {
                    remove_entry(fn, fnn, fat, E, search_pair(fn, fnn, *E, p0, p1));
                    remove_entry(fn, fnn, fat, E, search_pair(fn, fnn, *E, p0, node));
                    remove_entry(fn, fnn, fat, E, search_pair(fn, fnn, *E, p1, node));
                    fn[3 * *E] = p0; fn[3 * *E + 1] = p1; fn[3 * *E + 2] = node;
                    fnn[*E] = 3; fat[*E] = 0; ++*E;
                    done = 1;
                }
            }
        }
    }
}

/* SubdividedEdge_ConstructMortars / CurvedSubdividedEdge_ConstructMortars
 * (QuadElement_EdgeMappings.incf:450-565), the edge's n and dS already oriented by the linking. */
static void mortar_geometry(const Edge *ed, SubData *d) {
    const Spec *sN = spec_get(d->NN), *sS = spec_get(d->NS), *s = spec_get(ed->N);
    for (int half = 0; half < 2; ++half) {
        const int NM = half == 0 ? d->NN : d->NS;
        const Spec *sm = half == 0 ? sN : sS;
        double *nrm = half == 0 ? d->nN : d->nS, *dS = half == 0 ? d->dSN : d->dSS;
        for (int p = 0; p <= NM; ++p) {
            if (!ed->curved) {
                nrm[2 * p] = ed->n[0]; nrm[2 * p + 1] = ed->n[1];
                dS[p] = 0.5 * ed->dS[0];
                continue;
            }
            /* CurvedEdge_InterpolantdS: (dS n) interpolated at the mortar node's edge coordinate */
            const double x = half == 0 ? 0.5 + 0.5 * sm->xi[p] : 0.5 * sm->xi[p];
            double lj[MAXP], v[2];
            lagrange_barycentric(x, ed->N, s->xi, s->wb, lj);
            for (int c = 0; c < 2; ++c) {
                double acc = 0.0;
                for (int j = 0; j <= ed->N; ++j) acc = acc + (ed->dS[j] * ed->n[2 * j + c]) * lj[j];
                v[c] = acc;
            }
            dS[p] = 0.5 * sqrt(v[0] * v[0] + v[1] * v[1]);
            nrm[2 * p] = 0.5 * v[0] / dS[p];
            nrm[2 * p + 1] = 0.5 * v[1] / dS[p];
        }
    }
}

/* SubdividedEdge_LinkWithElements (QuadElement_EdgeProcedures.incf:201-355) with
 * SubdividedEdge_ConstructMortars (QuadElement_EdgeMappings.incf:450-497) and
 * SubdividedEdge_ComputeMortarsTransformationMatrices (QuadElement_EdgeProcedures.incf:530-634).
 * The big element is LEFT and forward; R1 (nodes 1 and h) and R2 (nodes 2 and h) become NORTH and
 * SOUTH according to the inversion the LEFT element needs. */
static int link_subdivided(Mesh *m, Edge *ed, int k) {
    int needs, pos, dirn;
    int ids[3] = {ed->nodes[0], ed->nodes[1], ed->nodes[2]};
    Elem *L = &m->el[ed->quads[0]];
    search_edge(L->nodes, ids, &needs, &pos, &dirn);
    if (pos < 0) return fail("subdivided edge not found in its element");
    L->edges[pos - 1] = k; L->dir[pos - 1] = FORWARD; L->qpos[pos - 1] = 0; ed->loc[0] = pos;
    int rn, rs;
    if (needs) {
        rn = ed->quads[1]; rs = ed->quads[2];
        edge_invert(ed, pos == EBOTTOM || pos == ERIGHT);
        int t = ids[0]; ids[0] = ids[1]; ids[1] = t;
    } else {
        rn = ed->quads[2]; rs = ed->quads[1];
        if (pos == ETOP || pos == ELEFT) { ed->n[0] = -ed->n[0]; ed->n[1] = -ed->n[1]; }
    }
    ed->quads[1] = rn; ed->quads[2] = rs;
    SubData *d = (SubData *)calloc(1, sizeof(SubData));
    ed->sd = d;
    const int NL = L->N, NRN = m->el[rn].N, NRS = m->el[rs].N;
    d->NN = NL > NRN ? NL : NRN;
    d->NS = NL > NRS ? NL : NRS;
    mortar_geometry(ed, d);
    /* NORTH: [h, node 2]; SOUTH: [node 1, h] */
    int sub_n[2] = {ids[2], ids[1]}, sub_s[2] = {ids[0], ids[2]};
    Elem *RN = &m->el[rn], *RS = &m->el[rs];
    int dn, ds;
    search_edge(RN->nodes, sub_n, &needs, &pos, &dn);
    if (pos < 0) return fail("subdivided edge not found in its north element");
    RN->edges[pos - 1] = k; RN->dir[pos - 1] = dn; RN->qpos[pos - 1] = 1; ed->loc[1] = pos;
    search_edge(RS->nodes, sub_s, &needs, &pos, &ds);
    if (pos < 0) return fail("subdivided edge not found in its south element");
    RS->edges[pos - 1] = k; RS->dir[pos - 1] = ds; RS->qpos[pos - 1] = 2; ed->loc[2] = pos;
    /* the eight transformation matrices */
    const Spec *sL = spec_get(NL), *sN = spec_get(d->NN), *sS = spec_get(d->NS), *sRN = spec_get(NRN), *sRS = spec_get(NRS);
    const int PL = NL + 1, PN = d->NN + 1, PS = d->NS + 1, PRN = NRN + 1, PRS = NRS + 1;
    double xs[MAXP], xn[MAXP];
    for (int p = 0; p < PS; ++p) xs[p] = 0.5 * sS->xi[p];
    for (int p = 0; p < PN; ++p) xn[p] = 0.5 + 0.5 * sN->xi[p];
    d->TLS_F = (double *)malloc(sizeof(double) * PS * PL); interp_matrix(NL, d->NS, sL->xi, sL->wb, xs, d->TLS_F);
    d->TLN_F = (double *)malloc(sizeof(double) * PN * PL); interp_matrix(NL, d->NN, sL->xi, sL->wb, xn, d->TLN_F);
    d->TRS_F = (double *)malloc(sizeof(double) * PS * PRS); interp_matrix(NRS, d->NS, sRS->xi, sRS->wb, sS->xi, d->TRS_F);
    d->TRN_F = (double *)malloc(sizeof(double) * PN * PRN); interp_matrix(NRN, d->NN, sRN->xi, sRN->wb, sN->xi, d->TRN_F);
    d->TLS_B = (double *)malloc(sizeof(double) * PL * PS);
    d->TLN_B = (double *)malloc(sizeof(double) * PL * PN);
    d->TRS_B = (double *)malloc(sizeof(double) * PRS * PS);
    d->TRN_B = (double *)malloc(sizeof(double) * PRN * PN);
    for (int i = 0; i < PL; ++i) {
        for (int l = 0; l < PS; ++l) d->TLS_B[i * PS + l] = ((1.0 / sL->w[i]) * d->TLS_F[l * PL + i]) * sS->w[l];
        for (int l = 0; l < PN; ++l) d->TLN_B[i * PN + l] = ((1.0 / sL->w[i]) * d->TLN_F[l * PL + i]) * sN->w[l];
    }
    for (int i = 0; i < PRS; ++i)
        for (int l = 0; l < PS; ++l) d->TRS_B[i * PS + l] = ((1.0 / sRS->w[i]) * d->TRS_F[l * PRS + i]) * sS->w[l];
    for (int i = 0; i < PRN; ++i)
        for (int l = 0; l < PN; ++l) d->TRN_B[i * PN + l] = ((1.0 / sRN->w[i]) * d->TRN_F[l * PRN + i]) * sN->w[l];
    /* a small element running its half backwards: T_FWD times the reversal (its columns
     * reversed), the reversal times T_BKW (its rows reversed) */
    if (ds == BACKWARD) {
        for (int i = 0; i < PS; ++i)
            for (int l = 0; l < PRS / 2; ++l) {
                double t = d->TRS_F[i * PRS + l]; d->TRS_F[i * PRS + l] = d->TRS_F[i * PRS + PRS - 1 - l]; d->TRS_F[i * PRS + PRS - 1 - l] = t;
            }
        for (int i = 0; i < PRS / 2; ++i)
            for (int l = 0; l < PS; ++l) {
                double t = d->TRS_B[i * PS + l]; d->TRS_B[i * PS + l] = d->TRS_B[(PRS - 1 - i) * PS + l]; d->TRS_B[(PRS - 1 - i) * PS + l] = t;
            }
    }
    if (dn == BACKWARD) {
        for (int i = 0; i < PN; ++i)
            for (int l = 0; l < PRN / 2; ++l) {
                double t = d->TRN_F[i * PRN + l]; d->TRN_F[i * PRN + l] = d->TRN_F[i * PRN + PRN - 1 - l]; d->TRN_F[i * PRN + PRN - 1 - l] = t;
            }
        for (int i = 0; i < PRN / 2; ++i)
            for (int l = 0; l < PN; ++l) {
                double t = d->TRN_B[i * PN + l]; d->TRN_B[i * PN + l] = d->TRN_B[(PRN - 1 - i) * PN + l]; d->TRN_B[(PRN - 1 - i) * PN + l] = t;
            }
    }
    return 0;
}

/* MeshFile.f90:77-673 (connectivity), QuadMesh_Construct.incf:22-305
 * (elements/edges/linking/geometry), QuadMesh.f90:215-309 + QuadMesh_Zones.incf
 * (zones and periodic pairing), BoundaryConditions/ *.incf Associate. */
static Mesh *mesh_build(const MeshIn *in) {
    Mesh *m = (Mesh *)calloc(1, sizeof(Mesh));
    m->nnodes = in->nnodes;
    m->pts = (double *)malloc(sizeof(double) * 2 * in->nnodes);
    memcpy(m->pts, in->points, sizeof(double) * 2 * in->nnodes);
    m->K = in->K;
    m->el = (Elem *)calloc(in->K, sizeof(Elem));
    /* computeFaces: first appearance order over (element, face 1..4), O(E^2) search */
    int cap = 4 * in->K + 4;
    int *fn = (int *)malloc(sizeof(int) * 3 * cap), *fnn = (int *)malloc(sizeof(int) * cap);
    int *fat = (int *)calloc(cap, sizeof(int));
    int E = 0;
    for (int e = 0; e < in->K; ++e)
        for (int f = 0; f < 4; ++f) {
            int face[2] = {in->quads[4 * e + f], in->quads[4 * e + (f + 1) % 4]};
            int found = 0;
            for (int k = 0; k < E; ++k)
                if (same_face(fn + 3 * k, face)) { found = 1; break; }
            if (!found) { fn[3 * E] = face[0]; fn[3 * E + 1] = face[1]; fnn[E] = 2; ++E; }
        }
    /* the entries that are boundary edges of the file carry the attribute (:337-360) */
    for (int b = 0; b < in->nbedges; ++b) {
        int k = search_pair(fn, fnn, E, in->bedges[2 * b], in->bedges[2 * b + 1]);
        if (k >= 0) fat[k] = 1;
    }
    merge_divided(in->nnodes, fn, fnn, fat, &E);
    m->E = E;
    m->ed = (Edge *)calloc(E, sizeof(Edge));
    /* computeElementOfEdges (:500-570): elements containing both nodes, in element order; a
     * subdivided edge [n1, n2, h]: the element with n1 and n2, then n1 and h, then n2 and h
     * (searchDividedEdgeInElements, :785-824) */
    for (int k = 0; k < E; ++k) {
        Edge *ed = &m->ed[k];
        ed->id = k; ed->nodes[0] = fn[3 * k]; ed->nodes[1] = fn[3 * k + 1]; ed->nodes[2] = -1;
        ed->quads[0] = ed->quads[1] = ed->quads[2] = -1;
        ed->partner = -1;
        if (fnn[k] == 3) {
            ed->sub = 1;
            ed->nodes[2] = fn[3 * k + 2];
            for (int e = 0; e < in->K; ++e) {
                int h0 = 0, h1 = 0, h2 = 0;
                for (int c = 0; c < 4; ++c) {
                    const int q = in->quads[4 * e + c];
                    h0 |= q == ed->nodes[0]; h1 |= q == ed->nodes[1]; h2 |= q == ed->nodes[2];
                }
                if (h0 && h1) ed->quads[0] = e;
                else if (h0 && h2) ed->quads[1] = e;
                else if (h1 && h2) ed->quads[2] = e;
                if (ed->quads[0] >= 0 && ed->quads[1] >= 0 && ed->quads[2] >= 0) break;
            }
            continue;
        }
        int cnt = 0;
        for (int e = 0; e < in->K && cnt < 2; ++e) {
            int h0 = 0, h1 = 0;
            for (int c = 0; c < 4; ++c) { h0 |= in->quads[4 * e + c] == ed->nodes[0]; h1 |= in->quads[4 * e + c] == ed->nodes[1]; }
            if (h0 && h1) ed->quads[cnt++] = e;
        }
        ed->bdry = (cnt == 1);
    }
    free(fn); free(fnn); free(fat);
    /* computeFaceMarkers */
    for (int b = 0; b < in->nbedges; ++b)
        for (int k = 0; k < E; ++k)
            if (same_face(in->bedges + 2 * b, m->ed[k].nodes)) { m->ed[k].marker = in->bmarker[b]; break; }
    /* computeCurvedEdges */
    int *curve_of = (int *)malloc(sizeof(int) * E);
    for (int k = 0; k < E; ++k) curve_of[k] = -1;
    double *cxy = NULL;
    if (in->ncurved > 0) {
        cxy = (double *)malloc(sizeof(double) * 2 * in->np1 * in->ncurved);
        memcpy(cxy, in->cxy, sizeof(double) * 2 * in->np1 * in->ncurved);
        for (int c = 0; c < in->ncurved; ++c)
            for (int k = 0; k < E; ++k) {
                Edge *ed = &m->ed[k];
                const int *cp = in->cedges + 2 * c;
                if (!same_face(cp, ed->nodes)) continue;
                curve_of[k] = c;
                if (ed->nodes[0] == cp[1] && ed->nodes[1] == cp[0]) { /* reverse the curve points */
                    for (int a = 0; a < in->np1 / 2; ++a)
                        for (int d = 0; d < 2; ++d) {
                            double t = cxy[(c * in->np1 + a) * 2 + d];
                            cxy[(c * in->np1 + a) * 2 + d] = cxy[(c * in->np1 + in->np1 - 1 - a) * 2 + d];
                            cxy[(c * in->np1 + in->np1 - 1 - a) * 2 + d] = t;
                        }
                }
                break;
            }
    }
    /* elements */
    int64_t off = 0;
    for (int e = 0; e < in->K; ++e) {
        Elem *el = &m->el[e];
        el->N = in->order[e]; el->P = el->N + 1;
        if (!spec_get(el->N)) { mesh_free(m); free(curve_of); free(cxy); fail("polynomial order out of range"); return NULL; }
        for (int c = 0; c < 4; ++c) el->nodes[c] = in->quads[4 * e + c];
        el->off = off; off += (int64_t)el->P * el->P;
        el->X = (double *)calloc(2 * el->P * el->P, sizeof(double));
        el->Ja = (double *)calloc(4 * el->P * el->P, sizeof(double));
        el->jac = (double *)calloc(el->P * el->P, sizeof(double));
    }
    m->nodes_total = off; m->ndof = 4 * off;
    /* edges: order, curves, construct */
    for (int k = 0; k < E; ++k) {
        Edge *ed = &m->ed[k];
        ed->low = -1;
        if (ed->sub) {
            /* divided interior edge: the big (LEFT) element's order (QuadMesh_Construct.incf:121-139) */
            if (ed->quads[0] < 0 || ed->quads[1] < 0 || ed->quads[2] < 0) {
                mesh_free(m); free(curve_of); free(cxy); fail("subdivided edge without its three elements"); return NULL;
            }
            int a = m->el[ed->quads[0]].N, b = m->el[ed->quads[1]].N, c = m->el[ed->quads[2]].N;
            ed->N = a;
            ed->NLow = a < b ? (a < c ? a : c) : (b < c ? b : c);
            ed->NCurve = ed->NLow > 1 ? ed->NLow / 2 : ed->NLow;
            if (b != c) {
                /* the reference's inversion matrix of the NORTH mortar is sized by the SOUTH element's
                 * order (QuadElement_EdgeProcedures.incf:556-559): defined only when they agree */
                mesh_free(m); free(curve_of); free(cxy);
                fail("subdivided edges need the two small elements of one order"); return NULL;
            }
        } else if (ed->bdry) { ed->N = m->el[ed->quads[0]].N; ed->NLow = ed->N; ed->NCurve = ed->N; }
        else {
            int a = m->el[ed->quads[0]].N, b = m->el[ed->quads[1]].N;
            ed->N = a > b ? a : b; ed->NLow = a < b ? a : b;
            ed->NCurve = ed->NLow > 1 ? ed->NLow / 2 : ed->NLow;
            if (a != b) {
                /* PolynomialInterpolationMatrix(NLow, N, ...) and TripleMatrixProduct(Minv_low,
                 * T_forward^T, M_edge): (1/w_low(i) * T_forward(l, i)) * w_edge(l) */
                const Spec *sl = spec_get(ed->NLow), *sh = spec_get(ed->N);
                int Pl = ed->NLow + 1, Ph = ed->N + 1;
                ed->low = a < b ? 0 : 1;
                ed->Tf = (double *)malloc(sizeof(double) * Pl * Ph);
                ed->Tb = (double *)malloc(sizeof(double) * Pl * Ph);
                interp_matrix(ed->NLow, ed->N, sl->xi, sl->wb, sh->xi, ed->Tf);
                for (int i = 0; i < Pl; ++i)
                    for (int l = 0; l < Ph; ++l) ed->Tb[i * Ph + l] = ((1.0 / sl->w[i]) * ed->Tf[l * Pl + i]) * sh->w[l];
            }
        }
        ed->curved = curve_of[k] >= 0;
        int P = ed->N + 1;
        ed->X = (double *)calloc(2 * P, sizeof(double));
        ed->dX = (double *)calloc(2 * P, sizeof(double));
        ed->n = (double *)calloc(2 * P, sizeof(double));
        ed->dS = (double *)calloc(P, sizeof(double));
        edge_set_curve(m, ed, ed->curved ? cxy + 2 * in->np1 * curve_of[k] : NULL, in->np1 - 1);
        if (ed->sub) {
            /* the hanging node moves onto the edge at 0.5 (QuadMesh_Construct.incf:168-173) */
            double X[2];
            edge_point_at(m, ed, 0.5, X);
            m->pts[2 * ed->nodes[2]] = X[0];
            m->pts[2 * ed->nodes[2] + 1] = X[1];
        }
    }
    free(curve_of); free(cxy);
    /* straight conforming interior edges again, with the moved nodes (:177-185) */
    for (int k = 0; k < E; ++k) {
        Edge *ed = &m->ed[k];
        if (!ed->sub && !ed->bdry && !ed->curved) edge_set_curve(m, ed, NULL, 0);
    }
    /* linking (QuadElement_EdgeProcedures.incf:30-199, 357-424) */
    for (int k = 0; k < E; ++k) {
        Edge *ed = &m->ed[k];
        int needs, pos, dirn;
        if (ed->sub) {
            if (link_subdivided(m, ed, k)) { mesh_free(m); return NULL; }
            continue;
        }
        Elem *L = &m->el[ed->quads[0]];
        search_edge(L->nodes, ed->nodes, &needs, &pos, &dirn);
        if (pos < 0) { mesh_free(m); fail("edge not found in element"); return NULL; }
        L->edges[pos - 1] = k; L->dir[pos - 1] = FORWARD; L->qpos[pos - 1] = 0; ed->loc[0] = pos;
        if (needs) edge_invert(ed, pos == EBOTTOM || pos == ERIGHT);
        else if (pos == ETOP || pos == ELEFT) {
            int P = ed->N + 1, np = ed->curved ? P : 1;
            for (int i = 0; i < 2 * np; ++i) ed->n[i] = -ed->n[i];
        }
        if (!ed->bdry) {
            Elem *R = &m->el[ed->quads[1]];
            search_edge(R->nodes, ed->nodes, &needs, &pos, &dirn);
            if (pos < 0) { mesh_free(m); fail("edge not found in right element"); return NULL; }
            R->edges[pos - 1] = k; R->dir[pos - 1] = dirn; R->qpos[pos - 1] = 1; ed->loc[1] = pos;
            ed->inverse = (dirn != FORWARD);
        }
    }
    /* geometry */
    for (int e = 0; e < m->K; ++e) {
        Elem *el = &m->el[e];
        const Spec *s = spec_get(el->N);
        elem_set_mappings(m, el);
        /* BilinearForm_F(jac, w, w): sum_i sum_j w_i jac(i,j) w_j */
        double vol = 0.0;
        for (int j = 0; j < el->P; ++j) {
            double acc = 0.0;
            for (int i = 0; i < el->P; ++i) acc = acc + s->w[i] * el->jac[i + el->P * j];
            vol = vol + acc * s->w[j];
        }
        el->Volume = vol;
    }
    for (int k = 0; k < E; ++k) {
        Edge *ed = &m->ed[k];
        const Spec *s = spec_get(ed->N);
        double A = 0.0;
        for (int i = 0; i <= ed->N; ++i) A = A + (ed->curved ? ed->dS[i] : ed->dS[0]) * s->w[i];
        ed->Area = A;
        if (ed->bdry) ed->invh = A / m->el[ed->quads[0]].Volume;
        else if (ed->sub) {
            double v0 = m->el[ed->quads[0]].Volume, v1 = m->el[ed->quads[1]].Volume, v2 = m->el[ed->quads[2]].Volume;
            double vm = v0 < v1 ? v0 : v1;
            ed->invh = A / (vm < v2 ? vm : v2);
        } else {
            double vl = m->el[ed->quads[0]].Volume, vr = m->el[ed->quads[1]].Volume;
            ed->invh = A / (vl < vr ? vl : vr);
        }
    }
    /* zones and boundary conditions */
    m->nmarkers = in->nmarkers;
    m->zones = (Zone *)calloc(in->nmarkers + 1, sizeof(Zone));
    for (int z = 1; z <= in->nmarkers; ++z) {
        Zone *Z = &m->zones[z - 1];
        Z->type = in->bc_type[z - 1];
        Z->direction = in->bc_int[2 * (z - 1)];
        Z->partner = in->bc_int[2 * (z - 1) + 1];
        Z->wall_iso = in->bc_int[2 * (z - 1)];
        Z->outflow = in->bc_int[2 * (z - 1)];
        /* newDirichletBC_Construct (newDirichlet.incf:79-105): anything else is the default */
        Z->nd_mode = in->bc_int[2 * (z - 1)] == SPECIFY_TOTAL_PRESSURE ? SPECIFY_TOTAL_PRESSURE : SPECIFY_SPEED;
        Z->nd_outflow = in->bc_int[2 * (z - 1) + 1] == TOTAL_PRESSURE ? TOTAL_PRESSURE : STATIC_PRESSURE;
        for (int k = 0; k < 4; ++k) Z->q[k] = in->bc_real[4 * (z - 1) + k];
        Z->Tw = in->bc_real[4 * (z - 1)]; Z->vw[0] = in->bc_real[4 * (z - 1) + 1]; Z->vw[1] = in->bc_real[4 * (z - 1) + 2];
        Z->rs = in->bc_rs ? in->bc_rs[z - 1] : -1;
        Z->edges = (int *)malloc(sizeof(int) * (E + 1));
        for (int k = 0; k < E; ++k)
            if (m->ed[k].bdry && m->ed[k].marker == z) Z->edges[Z->nedges++] = k;
        for (int a = 0; a < Z->nedges; ++a) {
            Edge *ed = &m->ed[Z->edges[a]];
            /* BoundaryConditions_construct, DGBoundaryConditions.f90:244-357 */
            ed->inviscidBCType = (Z->type == BC_EULERWALL || Z->type == BC_PRESSUREINLET) ? WEAK_PRESCRIBED : WEAK_RIEMANN;
            ed->rs = Z->rs;
            int P = ed->N + 1;
            for (int i = 0; i < P; ++i) ed->vbc[i] = VBC_UNDEFINED;
            if (Z->type == BC_EULERWALL)   /* EulerWall.incf:44-94 */
                for (int i = 0; i < P; ++i) ed->vbc[i] = VBC_ADIABATIC;
            if (Z->type == BC_VISCOUSWALL)
                for (int i = 0; i < P; ++i) ed->vbc[i] = Z->wall_iso ? VBC_DIRICHLET : VBC_ADIABATIC;
            if (Z->type == BC_FARFIELD) /* defined treatment: the reference leaves it unset */
                for (int i = 0; i < P; ++i) ed->vbc[i] = VBC_DIRICHLET;
        }
    }
    for (int z = 1; z <= in->nmarkers; ++z) {
        Zone *Z1 = &m->zones[z - 1];
        if (Z1->type != BC_PERIODIC || Z1->associated) continue;
        if (Z1->partner < 1 || Z1->partner > in->nmarkers) { mesh_free(m); fail("bad periodic partner"); return NULL; }
        Zone *Z2 = &m->zones[Z1->partner - 1];
        int coord = (Z1->direction == 1 && Z2->direction == 1) ? 1 : 0;
        for (int a = 0; a < Z1->nedges; ++a) {
            Edge *e1 = &m->ed[Z1->edges[a]];
            if (e1->curved) { mesh_free(m); fail("periodic curved edges are not supported"); return NULL; }
            for (int b = 0; b < Z2->nedges; ++b) {
                Edge *e2 = &m->ed[Z2->edges[b]];
                if (e2->associated) continue;
                const double *p11 = m->pts + 2 * e1->nodes[0], *p12 = m->pts + 2 * e1->nodes[1];
                const double *p21 = m->pts + 2 * e2->nodes[0], *p22 = m->pts + 2 * e2->nodes[1];
                int same = (almost_equal(p11[coord], p21[coord]) && almost_equal(p12[coord], p22[coord])) ||
                           (almost_equal(p11[coord], p22[coord]) && almost_equal(p12[coord], p21[coord]));
                if (!same) continue;
                double dot = e1->dX[0] * e2->dX[0] + e1->dX[1] * e2->dX[1];
                int inv = !(dot >= 0.0);
                e1->inverse = e2->inverse = inv;
                e1->associated = e2->associated = 1;
                e1->partner = Z2->edges[b]; e2->partner = Z1->edges[a];
                int P = e1->N + 1;
                for (int i = 0; i < P; ++i) { e1->vbc[i] = VBC_PERIODIC; e2->vbc[i] = VBC_PERIODIC; }
            }
        }
        Z1->associated = Z2->associated = 1;
    }
    for (int k = 0; k < E; ++k)
        if (m->ed[k].bdry && m->ed[k].marker == 0) {
            /* an element edge with no neighbour and no boundary marker, which the merging of
             * subdivided edges (MeshFile.f90:364-490) did not absorb */
            mesh_free(m); fail("an edge with neither a neighbour nor a boundary marker"); return NULL;
        }
    m->Q = (double *)calloc(m->ndof, sizeof(double));
    m->QDot = (double *)calloc(m->ndof, sizeof(double));
    m->dQ = (double *)calloc(2 * m->ndof, sizeof(double));
    m->G = (double *)calloc(m->ndof, sizeof(double));
    return m;
}

/* ======================================================================== */
/* Residual                                                                  */
/* ======================================================================== */
#define QIDX(e, eq, i, j) ((e)->off * 4 + (eq) + 4 * ((i) + (e)->P * (j)))
#define DQIDX(e, eq, i, j, d) ((e)->off * 8 + (eq) + 4 * ((i) + (e)->P * (j)) + 4 * (e)->P * (e)->P * (d))

/* DGSpatialDiscretizationMethods.f90:1520-1574 (mode 0: Q, mode 1: dQ) */
static void prolong(Mesh *m, int grad) {
#pragma omp parallel for schedule(static)
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        const Spec *s = spec_get(e->N);
        int P = e->P;
        for (int pos = 1; pos <= 4; ++pos) {
            Edge *ed = &m->ed[e->edges[pos - 1]];
            int side = e->qpos[pos - 1];
            const double *lb = (pos == EBOTTOM || pos == ELEFT) ? s->lb[0] : s->lb[1];
            int ndim = grad ? 2 : 1;
            for (int d = 0; d < ndim; ++d)
                for (int i = 0; i < P; ++i)
                    for (int eq = 0; eq < 4; ++eq) {
                        double acc = 0.0;
                        for (int l = 0; l < P; ++l) {
                            double v;
                            if (pos == EBOTTOM || pos == ETOP) v = grad ? m->dQ[DQIDX(e, eq, i, l, d)] : m->Q[QIDX(e, eq, i, l)];
                            else v = grad ? m->dQ[DQIDX(e, eq, l, i, d)] : m->Q[QIDX(e, eq, l, i)];
                            acc = acc + v * lb[l];
                        }
                        double *tq = side == 2 ? ed->sd->tq2 : ed->tq[side];
                        double *tdq = side == 2 ? ed->sd->tdq2 : ed->tdq[side];
                        if (grad) tdq[eq + 4 * i + 4 * P * d] = acc;
                        else tq[eq + 4 * i] = acc;
                    }
        }
    }
}

/* BoundaryConditions: FarfieldBC_UpdateSolution (Farfield.incf:138-376),
 * ViscousWall_UpdateSolution (ViscousWall.incf:158-254). Periodic edges read
 * the partner trace directly (QuadMesh_Zones.incf:111-117). */
/* The boundary states of edge ed's nodes i0 .. i1-1 from zone Z's state Z->q. */
static void update_edge_bc(Mesh *m, const Zone *Z, Edge *ed, int i0, int i1) {
    const Phys *ph = &m->ph;
    {
        {
            if (Z->type == BC_FARFIELD) {
                const double *q = Z->q;
                double rhoI = q[0], uI = q[1] / rhoI, vI = q[2] / rhoI;
                double pI = ph->gm1 * (q[3] - 0.5 * q[1] * uI - 0.5 * q[2] * vI);
                for (int i = i0; i < i1; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    const double *Qt = ed->tq[0] + 4 * i;
                    double rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
                    double vnL = uL * nn[0] + vL * nn[1];
                    double pL = ph->gm1 * (Qt[3] - 0.5 * Qt[1] * uL - 0.5 * Qt[2] * vL);
                    double aL = sqrt(ph->gamma * pL / rhoL);
                    double ML = sqrt(uL * uL + vL * vL) / aL;
                    double *uB = ed->uB + 4 * i;
                    int supersonic_out = 0;
                    if (vnL < 0.0) {
                        if (ML > 1.0) { for (int k = 0; k < 4; ++k) uB[k] = q[k]; }
                        else {
                            double pR = 0.5 * (pL + pI - rhoL * aL * (nn[0] * (uL - uI) + nn[1] * (vL - vI)));
                            double rhoR = rhoI + (pR - pI) / (aL * aL);
                            double uR = uI + nn[0] * (pR - pI) / (aL * rhoL);
                            double vR = vI + nn[1] * (pR - pI) / (aL * rhoL);
                            uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
                            uB[3] = ph->cv * pR + 0.5 * (uB[1] * uR + uB[2] * vR);
                        }
                    } else {
                        if (ML > 1.0) {
                            /* the straight-edge branch transposes indices (Farfield.incf:246);
                             * restated here in the curved-edge form (:344) */
                            for (int k = 0; k < 4; ++k) uB[k] = Qt[k];
                            supersonic_out = 1;
                        } else {
                            double pR = pI;
                            double rhoR = rhoL + (pI - pL) / (aL * aL);
                            double uR = uL - nn[0] * (pI - pL) / (rhoL * aL);
                            double vR = vL - nn[1] * (pI - pL) / (rhoL * aL);
                            uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
                            uB[3] = ph->cv * pR + 0.5 * (uB[1] * uR + uB[2] * vR);
                        }
                    }
                    for (int k = 0; k < 4; ++k) ed->uSB[4 * i + k] = uB[k];
                    ed->vbc[i] = supersonic_out ? VBC_NEUMANN : VBC_DIRICHLET;
                }
            } else if (Z->type == BC_EULERWALL) {
                /* EulerWall_UpdateSolution (EulerWall.incf:96-159): the wall state q (normal
                 * momentum removed, energy from p and the tangential momentum) gives the
                 * prescribed flux FB = sum_d F_d(q) n_d dS; uB mirrors the normal momentum */
                for (int i = i0; i < i1; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    const double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                    const double *Qt = ed->tq[0] + 4 * i;
                    double rhovn = Qt[1] * nn[0] + Qt[2] * nn[1];
                    double rhovt = -Qt[1] * nn[1] + Qt[2] * nn[0];
                    double p = ph->gm1 * (Qt[3] - 0.5 * (Qt[1] * Qt[1] + Qt[2] * Qt[2]) / Qt[0]);
                    double q[4] = {Qt[0], Qt[1] - rhovn * nn[0], Qt[2] - rhovn * nn[1],
                                   ph->cv * p + 0.5 * rhovt * rhovt / Qt[0]};
                    double f[4], g[4];
                    inviscid_flux(ph, q, f, g);
                    for (int k = 0; k < 4; ++k) {
                        double fb = 0.0;
                        fb = fb + f[k] * nn[0] * dS;
                        fb = fb + g[k] * nn[1] * dS;
                        ed->FB[4 * i + k] = fb;
                    }
                    double *uB = ed->uB + 4 * i;
                    uB[0] = Qt[0];
                    uB[1] = Qt[1] - 2.0 * rhovn * nn[0];
                    uB[2] = Qt[2] - 2.0 * rhovn * nn[1];
                    uB[3] = Qt[3];
                    for (int k = 0; k < 4; ++k) ed->uSB[4 * i + k] = q[k];
                }
            } else if (Z->type == BC_DIRICHLET) {
                /* DirichletBC_Associate (Dirichlet.incf:78-104): uB = the zone state;
                 * DirichletBC_UpdateSolution (:106-337): uSB and the viscous type per node */
                for (int i = i0; i < i1; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    const double *Qt = ed->tq[0] + 4 * i;
                    double *uB = ed->uB + 4 * i, *uSB = ed->uSB + 4 * i;
                    for (int k = 0; k < 4; ++k) uB[k] = Z->q[k];
                    double rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
                    double vnL = uL * nn[0] + vL * nn[1];
                    double pL = ph->gm1 * (Qt[3] - 0.5 * (Qt[1] * uL + Qt[2] * vL));
                    double aL = sqrt(ph->gamma * pL / rhoL);
                    double ML = sqrt(uL * uL + vL * vL) / aL;
                    if (vnL <= 0.0) {
                        if (ML > 1.0) {
                            for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
                        } else {
                            double pR = pL;
                            uSB[0] = uB[0]; uSB[1] = uB[1]; uSB[2] = uB[2];
                            uSB[3] = ph->cv * pR + 0.5 * (uSB[1] * uSB[1] + uSB[2] * uSB[2]) / uSB[0];
                        }
                        ed->vbc[i] = VBC_DIRICHLET;
                    } else if (ML > 1.0) {
                        for (int k = 0; k < 4; ++k) uSB[k] = Qt[k];
                        ed->vbc[i] = VBC_NEUMANN;
                    } else {
                        uSB[0] = Qt[0]; uSB[1] = Qt[1]; uSB[2] = Qt[2];
                        double pB = ph->gm1 * (uB[3] - 0.5 * (uB[1] * uB[1] + uB[2] * uB[2]) / uB[0]);
                        uSB[3] = ph->cv * pB + 0.5 * (uSB[1] * uL + uSB[2] * vL);
                        ed->vbc[i] = VBC_DIRICHLET;
                    }
                }
            } else if (Z->type == BC_PRESSUREOUTLET) {
                /* PressureOutletBC_UpdateSolution (PressureOutlet.incf:151-450) */
                const double *q = Z->q;
                double rhoI = q[0], uI = q[1] / rhoI, vI = q[2] / rhoI;
                double pI = ph->gm1 * (q[3] - 0.5 * q[1] * uI - 0.5 * q[2] * vI);
                double nrm = sqrt(uI * uI + vI * vI), nIx = uI / nrm, nIy = vI / nrm;
                for (int i = i0; i < i1; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    const double *Qt = ed->tq[0] + 4 * i;
                    double *uB = ed->uB + 4 * i, *uSB = ed->uSB + 4 * i;
                    double rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
                    double vnL = uL * nn[0] + vL * nn[1];
                    double pL = ph->gm1 * (Qt[3] - 0.5 * Qt[1] * uL - 0.5 * Qt[2] * vL);
                    double aL = sqrt(ph->gamma * pL / rhoL);
                    double ML = sqrt(uL * uL + vL * vL) / aL;
                    double rhoR, uR, vR, pR;
                    if (ML > 1.0) {   /* supersonic: inflow takes the zone state, outflow the interior */
                        const double *src = vnL < 0.0 ? q : Qt;
                        for (int k = 0; k < 4; ++k) { uB[k] = src[k]; uSB[k] = src[k]; }
                        ed->vbc[i] = vnL < 0.0 ? VBC_DIRICHLET : VBC_NEUMANN;
                        continue;
                    }
                    if (Z->outflow == REFLECTIVE_OUTFLOW) {
                        double R0 = pL - aL * aL * rhoL;
                        double Rplus = vnL + 2.0 * ph->cv * aL;
                        rhoR = -(R0 - pI) / (aL * aL);
                        double aR = sqrt(ph->gamma * pI / rhoR);
                        double vnR = Rplus - 2.0 * aR * ph->cv;
                        uR = uL + (vnR - vnL) * nn[0];
                        vR = vL + (vnR - vnL) * nn[1];
                        pR = pI;
                    } else {
                        pR = 2.0 * pI - pL;
                        uR = vnL < 0.0 ? fabs(vnL) * nIx : uL;
                        vR = vnL < 0.0 ? fabs(vnL) * nIy : vL;
                        rhoR = rhoL;
                    }
                    uB[0] = rhoR; uB[1] = rhoR * uR; uB[2] = rhoR * vR;
                    uB[3] = ph->cv * pR + 0.5 * (uB[1] * uR + uB[2] * vR);
                    if (vnL < 0.0) {
                        for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
                    } else {
                        uSB[0] = rhoL; uSB[1] = rhoL * uL; uSB[2] = rhoL * vL;
                        uSB[3] = ph->cv * pR + 0.5 * rhoL * (uL * uL + vL * vL);
                    }
                    ed->vbc[i] = VBC_DIRICHLET;
                }
            } else if (Z->type == BC_NEWDIRICHLET) {
                /* newDirichletBC_UpdateSolution (newDirichlet.incf:146-433): the outer state of
                 * the Riemann flux from the characteristics against the zone state; uSB is
                 * uB itself (:308-310, :429-431), the viscous type Neumann only for
                 * supersonic outflow.  Straight edges use their one normal (:201). */
                const double *q = Z->q;
                double uE = q[1] / q[0], vE = q[2] / q[0];
                double nrm = sqrt(uE * uE + vE * vE), nIx = uE / nrm, nIy = vE / nrm;
                double pE = ph->gm1 * (q[3] - 0.5 * (q[1] * uE + q[2] * vE));
                double MExtSq = (uE * uE + vE * vE) / (ph->gamma * pE) * q[0];
                double sExt = pE / pow(q[0], ph->gamma);
                double ptExt = pE * pow(1.0 + 0.5 * ph->gm1 * MExtSq, ph->cp);
                for (int i = i0; i < i1; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    const double *Qt = ed->tq[0] + 4 * i;
                    double *uB = ed->uB + 4 * i, *uSB = ed->uSB + 4 * i;
                    double rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
                    double vnL = uL * nn[0] + vL * nn[1];
                    double pL = ph->gm1 * (Qt[3] - 0.5 * (Qt[1] * uL + Qt[2] * vL));
                    double aL = sqrt(ph->gamma * pL / rhoL);
                    double ML = sqrt(uL * uL + vL * vL) / aL;
                    ed->vbc[i] = VBC_DIRICHLET;
                    if (vnL <= 0.0 && ML > 1.0) {
                        for (int k = 0; k < 4; ++k) uB[k] = q[k];
                    } else if (vnL <= 0.0) {
                        double pR = pL, rhoR = pow(pR / sExt, 1.0 / ph->gamma);
                        double aR = sqrt(ph->gamma * pR / rhoR);
                        if (Z->nd_mode == SPECIFY_SPEED) {
                            uB[0] = q[0]; uB[1] = q[1]; uB[2] = q[2];
                            uB[3] = ph->cv * pR + 0.5 * (uB[1] * uB[1] + uB[2] * uB[2]) / uB[0];
                        } else {
                            double MR = pR >= ptExt ? -sqrt(2.0 / ph->gm1 * (pow(pR / ptExt, ph->gm1 / ph->gamma) - 1.0))
                                                    : sqrt(2.0 / ph->gm1 * (pow(ptExt / pR, ph->gm1 / ph->gamma) - 1.0));
                            uB[0] = rhoR;
                            uB[1] = rhoR * aR * MR * nIx;
                            uB[2] = rhoR * aR * MR * nIy;
                            uB[3] = ph->cv * pR + 0.5 * rhoR * aR * aR * MR * MR;
                        }
                    } else if (ML > 1.0) {
                        for (int k = 0; k < 4; ++k) uB[k] = Qt[k];
                        ed->vbc[i] = VBC_NEUMANN;
                    } else {
                        double pO = Z->nd_outflow == STATIC_PRESSURE ? pE : ptExt;
                        uB[0] = Qt[0]; uB[1] = Qt[1]; uB[2] = Qt[2];
                        uB[3] = ph->cv * (2.0 * pO - pL) + 0.5 * (uB[1] * uL + uB[2] * vL);
                    }
                    for (int k = 0; k < 4; ++k) uSB[k] = uB[k];
                }
            } else if (Z->type == BC_PRESSUREINLET) {
                /* PressureInletBC_UpdateSolution (PressureInlet.incf:134-410): prescribed flux
                 * FB = (F_x(q) n_x + F_y(q) n_y) dS of the state q each branch forms.  The
                 * curved-edge supersonic outflow reads the trace transposed (:344); the
                 * straight-edge form (:247) is used for both. */
                const double *qe = Z->q;
                double uE = qe[1] / qe[0], vE = qe[2] / qe[0];
                double pE = ph->gm1 * (qe[3] - 0.5 * (qe[1] * uE + qe[2] * vE));
                double MExtSq = (uE * uE + vE * vE) / (ph->gamma * pE) * qe[0];
                double atExtSq = ph->gamma * pE / qe[0] * (1.0 + 0.5 * ph->gm1 * MExtSq);
                double sExt = pE / pow(qe[0], ph->gamma);
                double ptExt = pE * pow(1.0 + 0.5 * ph->gm1 * MExtSq, ph->cp);
                for (int i = i0; i < i1; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    const double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                    const double *Qt = ed->tq[0] + 4 * i;
                    double *uSB = ed->uSB + 4 * i;
                    double rhoL = Qt[0], uL = Qt[1] / rhoL, vL = Qt[2] / rhoL;
                    double vnL = uL * nn[0] + vL * nn[1];
                    double pL = ph->gm1 * (Qt[3] - 0.5 * (Qt[1] * uL + Qt[2] * vL));
                    double aL = sqrt(ph->gamma * pL / rhoL);
                    double ML = sqrt(uL * uL + vL * vL) / aL;
                    double Rplus = vnL + 2.0 * ph->cv * aL;
                    double q[4];
                    if (vnL <= 0.0 && ML > 1.0) {
                        for (int k = 0; k < 4; ++k) q[k] = qe[k];
                        for (int k = 0; k < 4; ++k) uSB[k] = q[k];
                        ed->vbc[i] = VBC_DIRICHLET;
                    } else if (vnL <= 0.0) {
                        double a = (ph->gamma + 1.0) / ph->gm1, b = -2 * Rplus;
                        double c = 0.5 * ph->gm1 * (Rplus * Rplus) - atExtSq;
                        double aR = (-b + sqrt(b * b - 4.0 * a * c)) / (2.0 * a);
                        double rhoR = pow(aR * aR / (ph->gamma * sExt), ph->cv);
                        double pR = aR * aR * rhoR / ph->gamma;
                        double vnR = Rplus - 2.0 * ph->cv * aR;
                        double uR = vnR * nn[0], vR = vnR * nn[1];
                        q[0] = rhoR; q[1] = rhoR * uR; q[2] = rhoR * vR;
                        q[3] = ph->cv * pR + 0.5 * (q[1] * uR + q[2] * vR);
                        uSB[0] = rhoR; uSB[1] = rhoR * uR; uSB[2] = rhoR * vR;
                        uSB[3] = ph->cv * pL + 0.5 * rhoR * (uR * uR + vR * vR);
                        ed->vbc[i] = VBC_DIRICHLET;
                    } else if (ML > 1.0) {
                        for (int k = 0; k < 4; ++k) q[k] = Qt[k];
                        for (int k = 0; k < 4; ++k) uSB[k] = q[k];
                        ed->vbc[i] = VBC_NEUMANN;
                    } else {
                        q[0] = Qt[0]; q[1] = Qt[1]; q[2] = Qt[2];
                        q[3] = ph->cv * (2.0 * ptExt - pL) + 0.5 * (q[1] * uL + q[2] * vL);
                        uSB[0] = q[0]; uSB[1] = q[1]; uSB[2] = q[2];
                        uSB[3] = ph->cv * ptExt + 0.5 * (q[1] * uL + q[2] * vL);
                        ed->vbc[i] = VBC_DIRICHLET;
                    }
                    double f[4], g[4];
                    inviscid_flux(ph, q, f, g);
                    for (int k = 0; k < 4; ++k) ed->FB[4 * i + k] = (f[k] * nn[0] + g[k] * nn[1]) * dS;
                    for (int k = 0; k < 4; ++k) ed->uB[4 * i + k] = q[k];
                }
            } else if (Z->type == BC_VISCOUSWALL) {
                for (int i = i0; i < i1; ++i) {
                    const double *Qt = ed->tq[0] + 4 * i;
                    double *uB = ed->uB + 4 * i, *uSB = ed->uSB + 4 * i;
                    uB[0] = Qt[0];
                    uB[1] = 2.0 * uB[0] * Z->vw[0] - Qt[1];
                    uB[2] = 2.0 * uB[0] * Z->vw[1] - Qt[2];
                    double Ek = 0.5 * (uB[1] * uB[1] + uB[2] * uB[2]) / uB[0];
                    if (Z->wall_iso) uB[3] = ph->cv * uB[0] * 1.0 * Z->Tw + Ek;
                    else uB[3] = Qt[3];
                    uSB[0] = Qt[0];
                    uSB[1] = uSB[0] * Z->vw[0];
                    uSB[2] = uSB[0] * Z->vw[1];
                    if (Z->wall_iso)
                        uSB[3] = ph->cv * uSB[0] * 1.0 * Z->Tw + 0.5 * (uSB[1] * uSB[1] + uSB[2] * uSB[2]) / uSB[0];
                    else
                        uSB[3] = Qt[3] - Ek + 0.5 * uSB[0] * (Z->vw[0] * Z->vw[0] + Z->vw[1] * Z->vw[1]);
                }
            }
        }
    }
}

/* Every boundary edge's states (DGSpatial_newTimeStep, DGSpatialDiscretizationMethods.f90:
 * 118-150).  A zone whose state comes from a boundary-state function ("Use function",
 * BoundaryConditionFunctionDrivers.incf:22-63) has it per edge node (ed->fq, set by the
 * caller from the function at the edge's node coordinates): each node then sees the zone with
 * its own state, as the reference evaluates q = BoundaryState(x(iXi), time) node by node. */
static void update_bcs(Mesh *m) {
    for (int z = 0; z < m->nmarkers; ++z) {
        const Zone *Z = &m->zones[z];
        for (int a = 0; a < Z->nedges; ++a) {
            Edge *ed = &m->ed[Z->edges[a]];
            const int P = ed->N + 1;
            if (!ed->fq) {
                update_edge_bc(m, Z, ed, 0, P);
                continue;
            }
            for (int i = 0; i < P; ++i) {
                Zone Zi = *Z;
                for (int k = 0; k < 4; ++k) Zi.q[k] = ed->fq[4 * i + k];
                update_edge_bc(m, &Zi, ed, i, i + 1);
            }
        }
    }
}

/* Partner trace of a periodic edge at local node i (reversed when inverse). */
static const double *partner_q(Mesh *m, Edge *ed, int i) {
    Edge *pe = &m->ed[ed->partner];
    int P = ed->N + 1;
    return pe->tq[0] + 4 * (ed->inverse ? (P - 1 - i) : i);
}
static const double *partner_dq(Mesh *m, Edge *ed, int i, double *buf) {
    Edge *pe = &m->ed[ed->partner];
    int P = ed->N + 1, ii = ed->inverse ? (P - 1 - i) : i;
    for (int d = 0; d < 2; ++d)
        for (int k = 0; k < 4; ++k) buf[k + 4 * d] = pe->tdq[0][k + 4 * ii + 4 * P * d];
    return buf;
}
static const double *bdry_state(Mesh *m, Edge *ed, int i) {
    return ed->associated ? partner_q(m, ed, i) : ed->uB + 4 * i;
}

/* Edge_ProjectSolution / Edge_ProjectSolutionAndGradient, Types 1-6
 * (QuadElement_EdgeProcedures.incf:749-1079): side `side`'s trace values of one plane
 * (src[eq + 4 l], the side's own nodes) at the edge's nodes i, in the edge's (LEFT) orientation.
 * The low side is interpolated with T_forward; a RIGHT side of an inverse edge is read (Type 2)
 * or written (Type 6) backwards. */
static void edge_side_values(const Edge *ed, int side, const double *src, double *out /* 4 (N+1) */) {
    int P = ed->N + 1, Pl = ed->NLow + 1;
    for (int i = 0; i < P; ++i) {
        for (int eq = 0; eq < 4; ++eq) {
            double v;
            if (side != ed->low) v = src[eq + 4 * ((side == 1 && ed->inverse) ? P - 1 - i : i)];
            else {
                int row = (side == 1 && ed->inverse) ? P - 1 - i : i;
                double acc = 0.0;
                for (int l = 0; l < Pl; ++l) acc = acc + ed->Tf[row * Pl + l] * src[eq + 4 * l];
                v = acc;
            }
            out[eq + 4 * i] = v;
        }
    }
}

/* Edge_ProjectFluxes, Types 1-6 (QuadElement_EdgeProcedures.incf:1081-1206): a flux at the
 * edge's nodes (F[eq + 4 i]) to side `side`'s own nodes, negated for the RIGHT side; the low
 * side takes T_backward F. */
static void edge_flux_to_side(const Edge *ed, int side, const double *F, double *out) {
    int P = ed->N + 1, Pl = ed->NLow + 1;
    if (side != ed->low) {
        for (int i = 0; i < P; ++i) {
            int o = (side == 1 && ed->inverse) ? P - 1 - i : i;
            for (int eq = 0; eq < 4; ++eq) out[eq + 4 * o] = side == 0 ? F[eq + 4 * i] : -F[eq + 4 * i];
        }
        return;
    }
    for (int i = 0; i < Pl; ++i) {
        int o = (side == 1 && ed->inverse) ? Pl - 1 - i : i;
        for (int eq = 0; eq < 4; ++eq) {
            double acc = 0.0;
            for (int l = 0; l < P; ++l) {
                if (side == 0) acc = acc + ed->Tb[i * P + l] * F[eq + 4 * l];
                else acc = acc - ed->Tb[i * P + l] * F[eq + 4 * l];
            }
            out[eq + 4 * o] = acc;
        }
    }
}

/* Edge_ProjectGradientFluxes, Types 1-6 (QuadElement_EdgeProcedures.incf:1208-1339): one
 * direction of a side's gradient flux at the edge's nodes (G[eq + 4 i]) to that side's own
 * nodes.  Unlike ProjectFluxes the RIGHT side is not negated (each side's flux was formed from
 * its own state); the low side takes T_backward G, a RIGHT side of an inverse edge reversed. */
static void edge_grad_to_side(const Edge *ed, int side, const double *G, double *out) {
    int P = ed->N + 1, Pl = ed->NLow + 1;
    if (side != ed->low) {
        for (int i = 0; i < P; ++i) {
            int o = (side == 1 && ed->inverse) ? P - 1 - i : i;
            for (int eq = 0; eq < 4; ++eq) out[eq + 4 * o] = G[eq + 4 * i];
        }
        return;
    }
    for (int i = 0; i < Pl; ++i) {
        int o = (side == 1 && ed->inverse) ? Pl - 1 - i : i;
        for (int eq = 0; eq < 4; ++eq) {
            double acc = 0.0;
            for (int l = 0; l < P; ++l) acc = acc + ed->Tb[i * P + l] * G[eq + 4 * l];
            out[eq + 4 * o] = acc;
        }
    }
}

/* A subdivided edge's mortar values from one side's traces, out[eq + 4 i] = sum_l T(i, l) src[eq + 4 l]
 * (ComputeRiemannSolver_SubdividedEdge, DGSpatialDiscretizationMethods.f90:809-828). */
static void mortar_in(const double *T, int PM, int Pe, const double *src, double *out) {
    for (int i = 0; i < PM; ++i)
        for (int eq = 0; eq < 4; ++eq) {
            double acc = 0.0;
            for (int l = 0; l < Pe; ++l) acc = acc + T[i * Pe + l] * src[eq + 4 * l];
            out[eq + 4 * i] = acc;
        }
}

/* Mortar values back to a side, accumulated into out (sign +1 or -1 per term):
 * out[eq + 4 i] (+/-)= T(i, l) F[eq + 4 l] for l in order (:887-907). */
static void mortar_out(const double *T, int Pe, int PM, const double *F, double sign, double *out) {
    for (int i = 0; i < Pe; ++i)
        for (int eq = 0; eq < 4; ++eq) {
            double acc = out[eq + 4 * i];
            for (int l = 0; l < PM; ++l) {
                if (sign > 0.0) acc = acc + T[i * PM + l] * F[eq + 4 * l];
                else acc = acc - T[i * PM + l] * F[eq + 4 * l];
            }
            out[eq + 4 * i] = acc;
        }
}

/* The three traces of a subdivided edge on its NORTH and SOUTH mortars: q[0..3] = LN, LS, RN, RS. */
static void subdivided_mortar_values(const Mesh *m, const Edge *ed, int grad, int d, double q[4][4 * MAXP]) {
    const SubData *sd = ed->sd;
    const int PL = m->el[ed->quads[0]].P, PN = m->el[ed->quads[1]].P, PS = m->el[ed->quads[2]].P;
    const double *tl = grad ? ed->tdq[0] + 4 * PL * d : ed->tq[0];
    const double *tn = grad ? ed->tdq[1] + 4 * PN * d : ed->tq[1];
    const double *ts = grad ? sd->tdq2 + 4 * PS * d : sd->tq2;
    mortar_in(sd->TLN_F, sd->NN + 1, PL, tl, q[0]);
    mortar_in(sd->TLS_F, sd->NS + 1, PL, tl, q[1]);
    mortar_in(sd->TRN_F, sd->NN + 1, PN, tn, q[2]);
    mortar_in(sd->TRS_F, sd->NS + 1, PS, ts, q[3]);
}

/* DGWeakIntegrals.f90:503-553 Vector_StdVolumeGreen + BR1 face loop
 * (DGViscous_BR1.f90:48-280) + DGViscousMethods.f90:538-616,786-917. */
static void gradient_br1(Mesh *m) {
#pragma omp parallel for schedule(static)
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        const Spec *s = spec_get(e->N);
        int P = e->P, PP = P * P;
        for (int d = 0; d < 2; ++d)
            for (int j = 0; j < P; ++j)
                for (int i = 0; i < P; ++i)
                    for (int eq = 0; eq < 4; ++eq) {
                        double acc = 0.0;
                        for (int l = 0; l < P; ++l)
                            acc = acc + m->Q[QIDX(e, eq, l, j)] * e->Ja[(d * 2 + 0) * PP + l + P * j] * s->hatD[i][l];
                        for (int l = 0; l < P; ++l)
                            acc = acc + m->Q[QIDX(e, eq, i, l)] * e->Ja[(d * 2 + 1) * PP + i + P * l] * s->hatD[j][l];
                        m->dQ[DQIDX(e, eq, i, j, d)] = -acc;
                    }
    }
    for (int k = 0; k < m->E; ++k) {
        Edge *ed = &m->ed[k];
        int P = ed->N + 1;
        double FuL[2][4 * MAXP], FuR[2][4 * MAXP], FuS[2][4 * MAXP];
        if (ed->sub) {
            /* BaseClass_ComputeSolutionRiemann_Subdivided / _CurvedSubdivided
             * (DGViscousMethods.f90:618-780): uStar on each mortar; straight edges return
             * 0.5 uStar to the big side and scale by the edge's dS n, curved ones scale each
             * mortar node first and return the products */
            const SubData *sd = ed->sd;
            const int PL = m->el[ed->quads[0]].P, PN = m->el[ed->quads[1]].P, PS = m->el[ed->quads[2]].P;
            const int PMN = sd->NN + 1, PMS = sd->NS + 1;
            double q[4][4 * MAXP], uN[4 * MAXP], uS[4 * MAXP];
            subdivided_mortar_values(m, ed, 0, 0, q);
            for (int c = 0; c < 4 * PMN; ++c) uN[c] = 0.5 * (q[0][c] + q[2][c]);
            for (int c = 0; c < 4 * PMS; ++c) uS[c] = 0.5 * (q[1][c] + q[3][c]);
            if (!ed->curved) {
                double TN[MAXP * MAXP], TS[MAXP * MAXP], uL[4 * MAXP], uRN[4 * MAXP], uRS[4 * MAXP];
                for (int c = 0; c < PL * PMN; ++c) TN[c] = 0.5 * sd->TLN_B[c];
                for (int c = 0; c < PL * PMS; ++c) TS[c] = 0.5 * sd->TLS_B[c];
                for (int c = 0; c < 4 * PL; ++c) uL[c] = 0.0;
                for (int c = 0; c < 4 * PN; ++c) uRN[c] = 0.0;
                for (int c = 0; c < 4 * PS; ++c) uRS[c] = 0.0;
                mortar_out(TN, PL, PMN, uN, 1.0, uL);
                mortar_out(TS, PL, PMS, uS, 1.0, uL);
                mortar_out(sd->TRN_B, PN, PMN, uN, 1.0, uRN);
                mortar_out(sd->TRS_B, PS, PMS, uS, 1.0, uRS);
                for (int d = 0; d < 2; ++d) {
                    for (int c = 0; c < 4 * PL; ++c) FuL[d][c] = uL[c] * ed->dS[0] * ed->n[d];
                    for (int c = 0; c < 4 * PN; ++c) FuR[d][c] = (-uRN[c]) * sd->dSN[0] * sd->nN[d];
                    for (int c = 0; c < 4 * PS; ++c) FuS[d][c] = (-uRS[c]) * sd->dSS[0] * sd->nS[d];
                }
            } else {
                for (int d = 0; d < 2; ++d) {
                    double fN[4 * MAXP], fS[4 * MAXP];
                    for (int i = 0; i < PMN; ++i)
                        for (int eq = 0; eq < 4; ++eq) fN[eq + 4 * i] = uN[eq + 4 * i] * sd->dSN[i] * sd->nN[2 * i + d];
                    for (int i = 0; i < PMS; ++i)
                        for (int eq = 0; eq < 4; ++eq) fS[eq + 4 * i] = uS[eq + 4 * i] * sd->dSS[i] * sd->nS[2 * i + d];
                    for (int c = 0; c < 4 * PL; ++c) FuL[d][c] = 0.0;
                    for (int c = 0; c < 4 * PN; ++c) FuR[d][c] = 0.0;
                    for (int c = 0; c < 4 * PS; ++c) FuS[d][c] = 0.0;
                    mortar_out(sd->TLN_B, PL, PMN, fN, 1.0, FuL[d]);
                    mortar_out(sd->TLS_B, PL, PMS, fS, 1.0, FuL[d]);
                    mortar_out(sd->TRN_B, PN, PMN, fN, -1.0, FuR[d]);
                    mortar_out(sd->TRS_B, PS, PMS, fS, -1.0, FuS[d]);
                }
            }
        } else if (!ed->bdry && ed->low >= 0) {
            /* p-non-conforming: BaseClass_ComputeSolutionRiemann_Interior / _CurvedInterior
             * (DGViscousMethods.f90:538-616): uStar on the edge's nodes from the projected
             * traces; straight edges project uStar and then scale by dS n, curved edges scale
             * first and project each direction */
            double qL[4 * MAXP], qR[4 * MAXP], us[4 * MAXP];
            edge_side_values(ed, 0, ed->tq[0], qL);
            edge_side_values(ed, 1, ed->tq[1], qR);
            for (int c = 0; c < 4 * P; ++c) us[c] = 0.5 * (qL[c] + qR[c]);
            int PL = m->el[ed->quads[0]].P, PR = m->el[ed->quads[1]].P;
            if (!ed->curved) {
                double uL[4 * MAXP], uR[4 * MAXP];
                edge_flux_to_side(ed, 0, us, uL);
                edge_flux_to_side(ed, 1, us, uR);
                for (int d = 0; d < 2; ++d) {
                    for (int c = 0; c < 4 * PL; ++c) FuL[d][c] = uL[c] * ed->dS[0] * ed->n[d];
                    for (int c = 0; c < 4 * PR; ++c) FuR[d][c] = uR[c] * ed->dS[0] * ed->n[d];
                }
            } else {
                for (int d = 0; d < 2; ++d) {
                    double fs[4 * MAXP];
                    for (int i = 0; i < P; ++i)
                        for (int eq = 0; eq < 4; ++eq) fs[eq + 4 * i] = us[eq + 4 * i] * ed->dS[i] * ed->n[2 * i + d];
                    edge_flux_to_side(ed, 0, fs, FuL[d]);
                    edge_flux_to_side(ed, 1, fs, FuR[d]);
                }
            }
        } else if (!ed->bdry) {
            for (int i = 0; i < P; ++i) {
                const double *qL = ed->tq[0] + 4 * i;
                const double *qR = ed->tq[1] + 4 * (ed->inverse ? P - 1 - i : i);
                int ir = ed->inverse ? P - 1 - i : i;
                double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                for (int eq = 0; eq < 4; ++eq) {
                    double us = 0.5 * (qL[eq] + qR[eq]);
                    if (!ed->curved) {
                        /* uStarL = uStar, uStarR = -uStar (projected); FuStar = u * dS * n */
                        for (int d = 0; d < 2; ++d) {
                            FuL[d][eq + 4 * i] = us * dS * nn[d];
                            FuR[d][eq + 4 * ir] = (-us) * dS * nn[d];
                        }
                    } else {
                        for (int d = 0; d < 2; ++d) {
                            double f = us * dS * nn[d];
                            FuL[d][eq + 4 * i] = f;
                            FuR[d][eq + 4 * ir] = -f;
                        }
                    }
                }
            }
        } else {
            for (int i = 0; i < P; ++i) {
                double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                for (int eq = 0; eq < 4; ++eq) {
                    double us;
                    if (ed->associated) us = 0.5 * (ed->tq[0][eq + 4 * i] + partner_q(m, ed, i)[eq]);
                    else us = ed->uSB[eq + 4 * i];
                    for (int d = 0; d < 2; ++d) FuL[d][eq + 4 * i] = us * dS * nn[d];
                }
            }
        }
        for (int side = 0; side < (ed->bdry ? 1 : (ed->sub ? 3 : 2)); ++side) {
            Elem *e = &m->el[ed->quads[side]];
            const Spec *s = spec_get(e->N);
            int loc = ed->loc[side], Pe = e->P;
            double(*Fu)[4 * MAXP] = side == 0 ? FuL : (side == 1 ? FuR : FuS);
            for (int d = 0; d < 2; ++d)
                for (int jj = 0; jj < Pe; ++jj)
                    for (int ii = 0; ii < Pe; ++ii)
                        for (int eq = 0; eq < 4; ++eq) {
                            double v;
                            switch (loc) {
                            case EBOTTOM: v = Fu[d][eq + 4 * ii] * s->lbw[0][jj]; break;
                            case ERIGHT: v = Fu[d][eq + 4 * jj] * s->lbw[1][ii]; break;
                            case ETOP: v = Fu[d][eq + 4 * ii] * s->lbw[1][jj]; break;
                            default: v = Fu[d][eq + 4 * jj] * s->lbw[0][ii]; break;
                            }
                            m->dQ[DQIDX(e, eq, ii, jj, d)] = m->dQ[DQIDX(e, eq, ii, jj, d)] + v;
                        }
        }
    }
#pragma omp parallel for schedule(static)
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        int P = e->P;
        for (int d = 0; d < 2; ++d)
            for (int i = 0; i < P; ++i)
                for (int j = 0; j < P; ++j)
                    for (int eq = 0; eq < 4; ++eq) m->dQ[DQIDX(e, eq, i, j, d)] = m->dQ[DQIDX(e, eq, i, j, d)] / e->jac[i + P * j];
    }
}

/* QuadElement_QuadProcedures.incf:42-88 (IP interior gradient) */
/* QuadElement_ComputeInteriorGradient (QuadElement_QuadProcedures.incf:42-88) at node (i, j):
 * dq[eq + 4 d], the broken gradient (D Q, Q D^T) mapped with Ja and divided by jac. */
static void interior_gradient(const Elem *e, const Spec *s, const double *Q, int i, int j, double *dq) {
    int P = e->P, PP = P * P, k = i + P * j;
    for (int eq = 0; eq < 4; ++eq) {
        double dxi = 0.0, deta = 0.0;
        for (int l = 0; l < P; ++l) dxi = dxi + s->D[i][l] * Q[QIDX(e, eq, l, j)];
        for (int l = 0; l < P; ++l) deta = deta + s->D[j][l] * Q[QIDX(e, eq, i, l)];
        dq[eq] = (dxi * e->Ja[0 * PP + k] + deta * e->Ja[1 * PP + k]) / e->jac[k];
        dq[eq + 4] = (dxi * e->Ja[2 * PP + k] + deta * e->Ja[3 * PP + k]) / e->jac[k];
    }
}

static void gradient_ip(Mesh *m) {
#pragma omp parallel for schedule(static)
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        const Spec *s = spec_get(e->N);
        int P = e->P, PP = P * P;
        for (int i = 0; i < P; ++i)
            for (int j = 0; j < P; ++j)
                for (int eq = 0; eq < 4; ++eq) {
                    double dxi = 0.0, deta = 0.0;
                    for (int l = 0; l < P; ++l) dxi = dxi + s->D[i][l] * m->Q[QIDX(e, eq, l, j)];
                    for (int l = 0; l < P; ++l) deta = deta + s->D[j][l] * m->Q[QIDX(e, eq, i, l)];
                    int k = i + P * j;
                    m->dQ[DQIDX(e, eq, i, j, 0)] = (dxi * e->Ja[0 * PP + k] + deta * e->Ja[1 * PP + k]) / e->jac[k];
                    m->dQ[DQIDX(e, eq, i, j, 1)] = (dxi * e->Ja[2 * PP + k] + deta * e->Ja[3 * PP + k]) / e->jac[k];
                }
    }
}

/* Viscous Riemann solvers: BR1 (DGViscous_BR1.f90:330-441), IP
 * (DGViscous_IP.f90:88-214), with the viscosity mu + edge % mu_a. */
/* Nn: the order of the nodes the flux is evaluated on (the edge's, a subdivided edge's mortar's),
 * which sets the IP penalty sigma0 mu Nn^2 / h. */
static void visc_interior_n(Mesh *m, Edge *ed, int Nn, const double *uL, const double *uR, const double *dL,
                            const double *dR, const double *nn, double *Fs) {
    const Phys *ph = &m->ph;
    const double mu = ph->mu + ed->mu_a;
    double fL[4], gL[4], fR[4], gR[4];
    viscous_flux(ph, uL, dL, fL, gL);
    viscous_flux(ph, uR, dR, fR, gR);
    for (int k = 0; k < 4; ++k) {
        double FL0 = mu * fL[k], FL1 = mu * gL[k], FR0 = mu * fR[k], FR1 = mu * gR[k];
        double acc = 0.0;
        acc = acc + 0.5 * (FL0 + FR0) * nn[0];
        acc = acc + 0.5 * (FL1 + FR1) * nn[1];
        Fs[k] = acc;
    }
    if (m->viscous == VISC_IP) {
        double pen = m->sigma0 * mu * (double)Nn * (double)Nn * ed->invh;
        for (int k = 0; k < 4; ++k) Fs[k] = Fs[k] - pen * (uL[k] - uR[k]);
    }
}
static void visc_interior(Mesh *m, Edge *ed, const double *uL, const double *uR, const double *dL, const double *dR,
                          const double *nn, double *Fs) {
    visc_interior_n(m, ed, ed->N, uL, uR, dL, dR, nn, Fs);
}

/* DGWeakIntegrals.f90:268-495 Scalar_StdGradientFace, subtracted from QDot */
static void std_gradient_face(Mesh *m, Elem *e, int loc, const double *G /* [eq + 4*i + 4*P*d] */) {
    const Spec *s = spec_get(e->N);
    int P = e->P, PP = P * P;
    const double *lb = (loc == EBOTTOM || loc == ELEFT) ? s->lb[0] : s->lb[1];
    const double *lbw = (loc == EBOTTOM || loc == ELEFT) ? s->lbw[0] : s->lbw[1];
    const double *dlbw = (loc == EBOTTOM || loc == ELEFT) ? s->dlbw[0] : s->dlbw[1];
    int horiz = (loc == EBOTTOM || loc == ETOP);
    double Jxi[2][MAXP], Jeta[2][MAXP], jac[MAXP];
    for (int i = 0; i < P; ++i) {
        for (int d = 0; d < 2; ++d) {
            double a = 0.0, b = 0.0;
            for (int l = 0; l < P; ++l) {
                int k = horiz ? (i + P * l) : (l + P * i);
                a = a + e->Ja[(d * 2 + 0) * PP + k] * lb[l];
                b = b + e->Ja[(d * 2 + 1) * PP + k] * lb[l];
            }
            Jxi[d][i] = a; Jeta[d][i] = b;
        }
        double c = 0.0;
        for (int l = 0; l < P; ++l) c = c + e->jac[horiz ? (i + P * l) : (l + P * i)] * lb[l];
        jac[i] = c;
    }
    double Fc[4 * MAXP], Gc[4 * MAXP], dF[4 * MAXP];
    for (int i = 0; i < P; ++i)
        for (int eq = 0; eq < 4; ++eq) {
            double a = 0.0, b = 0.0;
            for (int d = 0; d < 2; ++d) a = a + G[eq + 4 * i + 4 * P * d] * Jxi[d][i];
            a = a / jac[i];
            for (int d = 0; d < 2; ++d) b = b + G[eq + 4 * i + 4 * P * d] * Jeta[d][i];
            b = b / jac[i];
            Fc[eq + 4 * i] = a; Gc[eq + 4 * i] = b;
        }
    const double *src = horiz ? Fc : Gc;
    for (int i = 0; i < P; ++i)
        for (int eq = 0; eq < 4; ++eq) {
            double acc = 0.0;
            for (int l = 0; l < P; ++l) acc = acc + s->hatD[i][l] * src[eq + 4 * l];
            dF[eq + 4 * i] = acc;
        }
    for (int jj = 0; jj < P; ++jj)
        for (int ii = 0; ii < P; ++ii)
            for (int eq = 0; eq < 4; ++eq) {
                double v;
                if (horiz) v = dF[eq + 4 * ii] * lbw[jj] + Gc[eq + 4 * ii] * dlbw[jj];
                else v = Fc[eq + 4 * jj] * dlbw[ii] + dF[eq + 4 * jj] * lbw[ii];
                m->QDot[QIDX(e, eq, ii, jj)] = m->QDot[QIDX(e, eq, ii, jj)] - v;
            }
}

/* DGWeakIntegrals.f90:183-266 Scalar_StdFace, subtracted from QDot */
static void std_face(Mesh *m, Elem *e, int loc, const double *F) {
    const Spec *s = spec_get(e->N);
    int P = e->P;
    for (int jj = 0; jj < P; ++jj)
        for (int ii = 0; ii < P; ++ii)
            for (int eq = 0; eq < 4; ++eq) {
                double v;
                switch (loc) {
                case EBOTTOM: v = F[eq + 4 * ii] * s->lbw[0][jj]; break;
                case ERIGHT: v = F[eq + 4 * jj] * s->lbw[1][ii]; break;
                case ETOP: v = F[eq + 4 * ii] * s->lbw[1][jj]; break;
                default: v = F[eq + 4 * jj] * s->lbw[0][ii]; break;
                }
                m->QDot[QIDX(e, eq, ii, jj)] = m->QDot[QIDX(e, eq, ii, jj)] - v;
            }
}

/* IP gradient Riemann solvers (DGViscous_IP.f90:216-313): G[eq+4i+4P d] */
static void ip_grad_interior(Mesh *m, const Edge *ed, const double *uL, const double *uR, const double *nn,
                             double *GL, double *GR) {
    const Phys *ph = &m->ph;
    const double mu = ph->mu + ed->mu_a;
    double fg[8], f[4], g[4];
    for (int d = 0; d < 2; ++d)
        for (int k = 0; k < 4; ++k) fg[k + 4 * d] = (uL[k] - uR[k]) * nn[d];
    viscous_flux(ph, uL, fg, f, g);
    double cL = 0.5 * m->ip_eps * mu;
    for (int k = 0; k < 4; ++k) { GL[k] = cL * f[k]; GL[k + 4] = cL * g[k]; }
    viscous_flux(ph, uR, fg, f, g);
    double cR = -0.5 * m->ip_eps * mu;
    for (int k = 0; k < 4; ++k) { GR[k] = cR * f[k]; GR[k + 4] = cR * g[k]; }
}

/* Jumps of variable `var` across an edge (Edge_ComputeJumps and its variants,
 * QuadElement_EdgeProcedures.incf:1348-1619): nodal 2 (QL - QR) / (QL + QR) squared, weighted
 * by w dS.  Boundary edges take QR from uSB (periodic: the partner's trace, reversed when
 * inverse) and, straight or curved, dS(0). */
static double edge_jumps(Mesh *m, Edge *ed, int var) {
    const Spec *s = spec_get(ed->N);
    const int P = ed->N + 1;
    double acc = 0.0;
    for (int i = 0; i < P; ++i) {
        double qL = ed->tq[0][4 * i + var], qR, dS;
        if (!ed->bdry) {
            qR = ed->tq[1][4 * (ed->inverse ? P - 1 - i : i) + var];
            dS = ed->curved ? ed->dS[i] : ed->dS[0];
        } else {
            qR = ed->associated ? partner_q(m, ed, i)[var] : ed->uSB[4 * i + var];
            dS = ed->dS[0];
        }
        double nj = 2.0 * (qL - qR) / (qL + qR);
        acc = acc + s->w[i] * nj * nj * dS;
    }
    return acc;
}

/* Jumps-based artificial viscosity (DGArtificialDissipation.f90:277-363): per element
 * mu = Ceps sqrt(Volume)/(N+1) G(log10(max(sum J_rho, sum J_rhoE) / sum Area)), G the smooth
 * switch of half-width k = 5 around s0 = log10((N+1)^-4) + 1; per edge the mean of its
 * elements' values (one element for boundary edges, periodic ones included).  The "Physical"
 * type adds them to mu in every viscous flux (e % mu_a, ed % mu_a); the "Laplacian" type keeps
 * those at zero (it has no ComputeElementViscosity, :146-171) and uses the values in its own
 * volume and face fluxes (:207-249). */
static void artificial_viscosity(Mesh *m) {
    if (!m->ad) return;
    double *mu_e = m->ad == AD_LAPLACIAN ? m->lap_e : m->mu_e;
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        double gr = 0.0, ge = 0.0, area = 0.0;
        for (int f = 0; f < 4; ++f) {
            Edge *ed = &m->ed[e->edges[f]];
            gr = gr + edge_jumps(m, ed, 0);
            ge = ge + edge_jumps(m, ed, 3);
            area = area + ed->Area;
        }
        double gk = (gr > ge ? gr : ge) / area;
        const double k = 5.0, s0 = log10(1.0 / pow((double)(e->N + 1), 4.0)) + 1.0;
        double G;
        if (!(gk > 1.0e-8)) G = 0.0;
        else {
            gk = log10(gk);
            if (gk < s0 - k) G = 0.0;
            else if (gk < s0 + k) G = 0.5 * sin(0.5 * 3.141592653589793 * (gk - s0) / k) + 0.5;
            else G = 1.0;
        }
        mu_e[eid] = m->ad_ceps * (sqrt(e->Volume) / (double)(e->N + 1)) * G;
    }
    for (int k = 0; k < m->E; ++k) {
        Edge *ed = &m->ed[k];
        double mu = ed->bdry ? mu_e[ed->quads[0]] / 1.0 : (mu_e[ed->quads[0]] + mu_e[ed->quads[1]]) / 2.0;
        if (m->ad == AD_LAPLACIAN) ed->lap = mu;
        else ed->mu_a = mu;
    }
}

/* LaplaceDissipation_ComputeFaceFluxes (DGArtificialDissipation.f90:224-249): -mu Area (QL - QR)
 * with mu the edge's viscosity; zero without the Laplacian type. */
static void laplace_face(const Mesh *m, const Edge *ed, const double *qL, const double *qR, double *Fa) {
    for (int eq = 0; eq < 4; ++eq)
        Fa[eq] = m->ad == AD_LAPLACIAN ? -((ed->lap * ed->Area) * (qL[eq] - qR[eq])) : 0.0;
}

/* DGSpatialDiscretizationMethods.f90:160-235 computeQDot (+ newTimeStep :118-158) */
static void time_derivative(Mesh *m) {
    const Phys *ph = &m->ph;
    prolong(m, 0);
    update_bcs(m);
    artificial_viscosity(m);
    if (m->ns) {
        if (m->viscous == VISC_BR1) gradient_br1(m);
        else gradient_ip(m);
        prolong(m, 1);
    }
    RiemannFn rdefault = riemann_fn(m->riemann);
    /* volume loop: DGSpatial_QDotVolumeLoop :243-268 + Scalar_StdVolumeGreen DGWeakIntegrals.f90:111-181 */
#pragma omp parallel for schedule(static)
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        const Spec *s = spec_get(e->N);
        int P = e->P, PP = P * P;
        double *cF = (double *)malloc(sizeof(double) * 4 * PP), *cG = (double *)malloc(sizeof(double) * 4 * PP);
        for (int j = 0; j < P; ++j)
            for (int i = 0; i < P; ++i) {
                int k = i + P * j;
                const double *q = m->Q + QIDX(e, 0, i, j);
                double f[4], g[4];
                inviscid_flux(ph, q, f, g);
                if (m->ns) {
                    double dq[8], fv[4], gv[4];
                    for (int d = 0; d < 2; ++d)
                        for (int eq = 0; eq < 4; ++eq) dq[eq + 4 * d] = m->dQ[DQIDX(e, eq, i, j, d)];
                    viscous_flux(ph, q, dq, fv, gv);
                    const double mu = ph->mu + m->mu_e[eid];   /* e % mu_a, DGViscous_BR1.f90:307 */
                    /* Laplacian: Fa = mu * the element's interior gradient (:207-222, ComputeInteriorGradient,
                     * QuadElement_QuadProcedures.incf:42-88: the IP gradient itself) */
                    double fa[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    if (m->ad == AD_LAPLACIAN) {
                        double dqi[8];
                        if (m->viscous == VISC_IP) for (int c = 0; c < 8; ++c) dqi[c] = dq[c];
                        else interior_gradient(e, s, m->Q, i, j, dqi);
                        for (int c = 0; c < 8; ++c) fa[c] = m->lap_e[eid] * dqi[c];
                    }
                    for (int eq = 0; eq < 4; ++eq) { f[eq] = (f[eq] - mu * fv[eq]) - fa[eq]; g[eq] = (g[eq] - mu * gv[eq]) - fa[eq + 4]; }
                }
                for (int eq = 0; eq < 4; ++eq) {
                    cF[eq + 4 * k] = f[eq] * e->Ja[0 * PP + k] + g[eq] * e->Ja[2 * PP + k];
                    cG[eq + 4 * k] = f[eq] * e->Ja[1 * PP + k] + g[eq] * e->Ja[3 * PP + k];
                }
            }
        for (int j = 0; j < P; ++j)
            for (int i = 0; i < P; ++i)
                for (int eq = 0; eq < 4; ++eq) {
                    double acc = 0.0;
                    for (int l = 0; l < P; ++l)
                        acc = acc + s->hatD[i][l] * cF[eq + 4 * (l + P * j)] + s->hatD[j][l] * cG[eq + 4 * (i + P * l)];
                    m->QDot[QIDX(e, eq, i, j)] = acc;
                }
        free(cF); free(cG);
    }
    /* face loop in global edge order (serial, like the reference's !$omp master) */
    for (int k = 0; k < m->E; ++k) {
        Edge *ed = &m->ed[k];
        int P = ed->N + 1;
        double FL[4 * MAXP], FR[4 * MAXP];
        double GL[8 * MAXP], GR[8 * MAXP];
        int gradfl = m->ns && m->viscous == VISC_IP && m->ip_gradflux;
        if (ed->sub) {
            /* ComputeRiemannSolver_SubdividedEdge / _CurvedSubdividedEdge (:761-1140): the Riemann
             * problem on each mortar, FStar = (Fi - Fv - Fa) dS_mortar, returned with T_BKW: the big
             * side takes NORTH then SOUTH, each small side its own mortar, negated */
            const SubData *sd = ed->sd;
            const int PL = m->el[ed->quads[0]].P, PN = m->el[ed->quads[1]].P, PS = m->el[ed->quads[2]].P;
            const int PM[2] = {sd->NN + 1, sd->NS + 1};
            double q[4][4 * MAXP], dq[2][4][4 * MAXP], Fs[2][4 * MAXP], FRS[4 * MAXP];
            subdivided_mortar_values(m, ed, 0, 0, q);
            if (m->ns)
                for (int d = 0; d < 2; ++d) subdivided_mortar_values(m, ed, 1, d, dq[d]);
            for (int h = 0; h < 2; ++h) {
                const double *nrm = h == 0 ? sd->nN : sd->nS, *dSm = h == 0 ? sd->dSN : sd->dSS;
                const double *qa = q[h], *qb = q[2 + h];
                for (int i = 0; i < PM[h]; ++i) {
                    double Fi[4];
                    rdefault(ph, qa + 4 * i, qb + 4 * i, nrm + 2 * i, Fi);
                    if (m->ns) {
                        double a[8], b[8], Fv[4];
                        for (int d = 0; d < 2; ++d)
                            for (int eq = 0; eq < 4; ++eq) {
                                a[eq + 4 * d] = dq[d][h][eq + 4 * i];
                                b[eq + 4 * d] = dq[d][2 + h][eq + 4 * i];
                            }
                        visc_interior_n(m, ed, PM[h] - 1, qa + 4 * i, qb + 4 * i, a, b, nrm + 2 * i, Fv);
                        for (int eq = 0; eq < 4; ++eq) Fs[h][eq + 4 * i] = (Fi[eq] - Fv[eq]) * dSm[i];
                    } else {
                        for (int eq = 0; eq < 4; ++eq) Fs[h][eq + 4 * i] = Fi[eq] * dSm[i];
                    }
                }
            }
            for (int c = 0; c < 4 * PL; ++c) FL[c] = 0.0;
            for (int c = 0; c < 4 * PN; ++c) FR[c] = 0.0;
            for (int c = 0; c < 4 * PS; ++c) FRS[c] = 0.0;
            mortar_out(sd->TLN_B, PL, PM[0], Fs[0], 1.0, FL);
            mortar_out(sd->TLS_B, PL, PM[1], Fs[1], 1.0, FL);
            mortar_out(sd->TRN_B, PN, PM[0], Fs[0], -1.0, FR);
            mortar_out(sd->TRS_B, PS, PM[1], Fs[1], -1.0, FRS);
            std_face(m, &m->el[ed->quads[0]], ed->loc[0], FL);
            std_face(m, &m->el[ed->quads[1]], ed->loc[1], FR);
            std_face(m, &m->el[ed->quads[2]], ed->loc[2], FRS);
            if (gradfl) {
                /* IP_GradientRiemannSolver on each mortar, times the mortar's dS (straight: its
                 * node 0's, curved: each node's), returned with T_BKW to all three sides, the
                 * small sides not negated (:906-940, :1093-1136) */
                double Gm[2][2][8 * MAXP];   /* [mortar][big, small][eq + 4 i + 4 PM d] */
                for (int h = 0; h < 2; ++h) {
                    const double *nrm = h == 0 ? sd->nN : sd->nS, *dSm = h == 0 ? sd->dSN : sd->dSS;
                    for (int i = 0; i < PM[h]; ++i) {
                        double gl[8], gr[8];
                        ip_grad_interior(m, ed, q[h] + 4 * i, q[2 + h] + 4 * i, nrm + 2 * i, gl, gr);
                        const double sc = ed->curved ? dSm[i] : dSm[0];
                        for (int d = 0; d < 2; ++d)
                            for (int eq = 0; eq < 4; ++eq) {
                                Gm[h][0][eq + 4 * i + 4 * PM[h] * d] = gl[eq + 4 * d] * sc;
                                Gm[h][1][eq + 4 * i + 4 * PM[h] * d] = gr[eq + 4 * d] * sc;
                            }
                    }
                }
                double GRS[8 * MAXP];
                for (int c = 0; c < 8 * PL; ++c) GL[c] = 0.0;
                for (int c = 0; c < 8 * PN; ++c) GR[c] = 0.0;
                for (int c = 0; c < 8 * PS; ++c) GRS[c] = 0.0;
                for (int d = 0; d < 2; ++d) {
                    mortar_out(sd->TLN_B, PL, PM[0], Gm[0][0] + 4 * PM[0] * d, 1.0, GL + 4 * PL * d);
                    mortar_out(sd->TLS_B, PL, PM[1], Gm[1][0] + 4 * PM[1] * d, 1.0, GL + 4 * PL * d);
                    mortar_out(sd->TRN_B, PN, PM[0], Gm[0][1] + 4 * PM[0] * d, 1.0, GR + 4 * PN * d);
                    mortar_out(sd->TRS_B, PS, PM[1], Gm[1][1] + 4 * PM[1] * d, 1.0, GRS + 4 * PS * d);
                }
                std_gradient_face(m, &m->el[ed->quads[0]], ed->loc[0], GL);
                std_gradient_face(m, &m->el[ed->quads[1]], ed->loc[1], GR);
                std_gradient_face(m, &m->el[ed->quads[2]], ed->loc[2], GRS);
            }
        } else if (!ed->bdry && ed->low >= 0) {
            /* p-non-conforming (ComputeRiemannSolver_InteriorEdge / _InteriorCurvedEdge
             * :545-759): the Riemann problem on the edge's nodes from the projected traces and
             * gradients, FStar = (Fi - Fv - Fa) dS, returned to each side by ProjectFluxes */
            double qL[4 * MAXP], qR[4 * MAXP], dL[2][4 * MAXP], dR[2][4 * MAXP], Fs[4 * MAXP];
            int PL = m->el[ed->quads[0]].P, PR = m->el[ed->quads[1]].P;
            edge_side_values(ed, 0, ed->tq[0], qL);
            edge_side_values(ed, 1, ed->tq[1], qR);
            if (m->ns)
                for (int d = 0; d < 2; ++d) {
                    edge_side_values(ed, 0, ed->tdq[0] + 4 * PL * d, dL[d]);
                    edge_side_values(ed, 1, ed->tdq[1] + 4 * PR * d, dR[d]);
                }
            for (int i = 0; i < P; ++i) {
                const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                double Fi[4];
                rdefault(ph, qL + 4 * i, qR + 4 * i, nn, Fi);
                if (m->ns) {
                    double a[8], b[8], Fv[4], Fa[4];
                    for (int d = 0; d < 2; ++d)
                        for (int eq = 0; eq < 4; ++eq) { a[eq + 4 * d] = dL[d][eq + 4 * i]; b[eq + 4 * d] = dR[d][eq + 4 * i]; }
                    visc_interior(m, ed, qL + 4 * i, qR + 4 * i, a, b, nn, Fv);
                    laplace_face(m, ed, qL + 4 * i, qR + 4 * i, Fa);
                    for (int eq = 0; eq < 4; ++eq) Fs[eq + 4 * i] = ((Fi[eq] - Fv[eq]) - Fa[eq]) * dS;
                } else {
                    for (int eq = 0; eq < 4; ++eq) Fs[eq + 4 * i] = Fi[eq] * dS;
                }
            }
            edge_flux_to_side(ed, 0, Fs, FL);
            edge_flux_to_side(ed, 1, Fs, FR);
            std_face(m, &m->el[ed->quads[0]], ed->loc[0], FL);
            std_face(m, &m->el[ed->quads[1]], ed->loc[1], FR);
            if (gradfl) {
                /* IP_GradientRiemannSolver on the projected traces, each side's flux returned by
                 * ProjectGradientFluxes: straight edges project and then multiply by dS(0),
                 * curved edges multiply by dS(i) first (:638-651, :744-757) */
                double Ge[2][8 * MAXP];   /* [side][eq + 4 i + 4 P d] on the edge's nodes */
                for (int i = 0; i < P; ++i) {
                    const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                    double gl[8], gr[8];
                    ip_grad_interior(m, ed, qL + 4 * i, qR + 4 * i, nn, gl, gr);
                    for (int d = 0; d < 2; ++d)
                        for (int eq = 0; eq < 4; ++eq) {
                            Ge[0][eq + 4 * i + 4 * P * d] = ed->curved ? gl[eq + 4 * d] * ed->dS[i] : gl[eq + 4 * d];
                            Ge[1][eq + 4 * i + 4 * P * d] = ed->curved ? gr[eq + 4 * d] * ed->dS[i] : gr[eq + 4 * d];
                        }
                }
                for (int side = 0; side < 2; ++side) {
                    const int Ps = side == 0 ? PL : PR;
                    double *Gs = side == 0 ? GL : GR;
                    for (int d = 0; d < 2; ++d) edge_grad_to_side(ed, side, Ge[side] + 4 * P * d, Gs + 4 * Ps * d);
                    if (!ed->curved)
                        for (int c = 0; c < 8 * Ps; ++c) Gs[c] = Gs[c] * ed->dS[0];
                }
                std_gradient_face(m, &m->el[ed->quads[0]], ed->loc[0], GL);
                std_gradient_face(m, &m->el[ed->quads[1]], ed->loc[1], GR);
            }
        } else if (!ed->bdry) {
            /* ComputeRiemannSolver_InteriorEdge :545-653 / InteriorCurvedEdge :655-759 */
            for (int i = 0; i < P; ++i) {
                int ir = ed->inverse ? P - 1 - i : i;
                const double *qL = ed->tq[0] + 4 * i, *qR = ed->tq[1] + 4 * ir;
                const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                double Fi[4], Fs[4];
                rdefault(ph, qL, qR, nn, Fi);
                if (m->ns) {
                    double dL[8], dR[8], Fv[4];
                    for (int d = 0; d < 2; ++d)
                        for (int eq = 0; eq < 4; ++eq) {
                            dL[eq + 4 * d] = ed->tdq[0][eq + 4 * i + 4 * P * d];
                            dR[eq + 4 * d] = ed->tdq[1][eq + 4 * ir + 4 * P * d];
                        }
                    visc_interior(m, ed, qL, qR, dL, dR, nn, Fv);
                    double Fa[4];
                    laplace_face(m, ed, qL, qR, Fa);
                    for (int eq = 0; eq < 4; ++eq) Fs[eq] = ((Fi[eq] - Fv[eq]) - Fa[eq]) * dS;
                    if (gradfl) {
                        double gl[8], gr[8];
                        ip_grad_interior(m, ed, qL, qR, nn, gl, gr);
                        for (int d = 0; d < 2; ++d)
                            for (int eq = 0; eq < 4; ++eq) {
                                /* straight: project then *dS(0); curved: *dS(i) then project */
                                GL[eq + 4 * i + 4 * P * d] = gl[eq + 4 * d] * dS;
                                GR[eq + 4 * ir + 4 * P * d] = gr[eq + 4 * d] * dS;
                            }
                    }
                } else {
                    for (int eq = 0; eq < 4; ++eq) Fs[eq] = Fi[eq] * dS;
                }
                for (int eq = 0; eq < 4; ++eq) { FL[eq + 4 * i] = Fs[eq]; FR[eq + 4 * ir] = -Fs[eq]; }
            }
            Elem *eL = &m->el[ed->quads[0]], *eR = &m->el[ed->quads[1]];
            std_face(m, eL, ed->loc[0], FL);
            std_face(m, eR, ed->loc[1], FR);
            if (gradfl) {
                std_gradient_face(m, eL, ed->loc[0], GL);
                std_gradient_face(m, eR, ed->loc[1], GR);
            }
        } else {
            /* ComputeRiemannSolver_StraightBdryEdge :1142-1339 / CurvedBdryEdge :1341-1511 */
            RiemannFn rs = ed->rs >= 0 ? riemann_fn(ed->rs) : rdefault;
            for (int i = 0; i < P; ++i) {
                const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
                double dS = ed->curved ? ed->dS[i] : ed->dS[0];
                const double *q = ed->tq[0] + 4 * i;
                const double *qb = bdry_state(m, ed, i);
                double Fi[4], Fv[4] = {0, 0, 0, 0}, Fa[4] = {0, 0, 0, 0}, Gv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                if (ed->inviscidBCType == WEAK_PRESCRIBED) {   /* Fi = FB (:1194-1198) */
                    for (int eq = 0; eq < 4; ++eq) Fi[eq] = ed->FB[4 * i + eq];
                } else {
                    rs(ph, q, qb, nn, Fi);
                    for (int eq = 0; eq < 4; ++eq) Fi[eq] = Fi[eq] * dS;
                }
                if (m->ns) {
                    double dq[8];
                    for (int d = 0; d < 2; ++d)
                        for (int eq = 0; eq < 4; ++eq) dq[eq + 4 * d] = ed->tdq[0][eq + 4 * i + 4 * P * d];
                    int vbc = ed->vbc[i];
                    if (ed->associated) vbc = VBC_PERIODIC;
                    if (vbc == VBC_PERIODIC) {
                        double dqb[8];
                        partner_dq(m, ed, i, dqb);
                        visc_interior(m, ed, q, qb, dq, dqb, nn, Fv);
                        laplace_face(m, ed, q, qb, Fa);
                        for (int eq = 0; eq < 4; ++eq) { Fv[eq] = Fv[eq] * dS; Fa[eq] = Fa[eq] * dS; }
                        if (gradfl) {
                            double gl[8], gr[8];
                            ip_grad_interior(m, ed, q, qb, nn, gl, gr);
                            for (int c = 0; c < 8; ++c) Gv[c] = gl[c] * dS;
                        }
                    } else if (vbc == VBC_ADIABATIC) {
                        const double *usb = ed->uSB + 4 * i;
                        double f[4], g[4];
                        adiabatic_viscous_flux(ph, q, q, dq, f, g);
                        for (int eq = 0; eq < 4; ++eq) {
                            double acc = 0.0;
                            acc = acc + (ph->mu + ed->mu_a) * f[eq] * nn[0];
                            acc = acc + (ph->mu + ed->mu_a) * g[eq] * nn[1];
                            Fv[eq] = acc;
                        }
                        if (m->viscous == VISC_IP) {
                            double pen = m->sigma0 * (ph->mu + ed->mu_a) * (double)ed->N * (double)ed->N * ed->invh;
                            for (int eq = 0; eq < 4; ++eq) Fv[eq] = Fv[eq] - pen * (q[eq] - usb[eq]);
                        }
                        laplace_face(m, ed, q, usb, Fa);
                        for (int eq = 0; eq < 4; ++eq) { Fv[eq] = Fv[eq] * dS; Fa[eq] = Fa[eq] * dS; }
                        if (gradfl) {
                            double fg[8];
                            for (int d = 0; d < 2; ++d)
                                for (int eq = 0; eq < 4; ++eq) fg[eq + 4 * d] = (q[eq] - usb[eq]) * nn[d];
                            adiabatic_viscous_flux(ph, q, q, fg, f, g);
                            double c = m->ip_eps * (ph->mu + ed->mu_a);
                            for (int eq = 0; eq < 4; ++eq) { Gv[eq] = c * f[eq] * dS; Gv[eq + 4] = c * g[eq] * dS; }
                        }
                    } else if (vbc == VBC_DIRICHLET) {
                        const double *usb = ed->uSB + 4 * i;
                        double f[4], g[4];
                        viscous_flux(ph, q, dq, f, g);
                        for (int eq = 0; eq < 4; ++eq)
                            Fv[eq] = (ph->mu + ed->mu_a) * f[eq] * nn[0] + (ph->mu + ed->mu_a) * g[eq] * nn[1];
                        if (m->viscous == VISC_IP) {
                            double pen = m->sigma0 * (ph->mu + ed->mu_a) * (double)ed->N * (double)ed->N * ed->invh;
                            for (int eq = 0; eq < 4; ++eq) Fv[eq] = Fv[eq] - pen * (q[eq] - usb[eq]);
                        }
                        laplace_face(m, ed, q, usb, Fa);
                        for (int eq = 0; eq < 4; ++eq) { Fv[eq] = Fv[eq] * dS; Fa[eq] = Fa[eq] * dS; }
                        if (gradfl) {
                            double fg[8];
                            for (int d = 0; d < 2; ++d)
                                for (int eq = 0; eq < 4; ++eq) fg[eq + 4 * d] = (q[eq] - usb[eq]) * nn[d];
                            viscous_flux(ph, q, fg, f, g);
                            double c = 1.0 * m->ip_eps * (ph->mu + ed->mu_a);
                            for (int eq = 0; eq < 4; ++eq) { Gv[eq] = c * f[eq] * dS; Gv[eq + 4] = c * g[eq] * dS; }
                        }
                    } /* NEUMANN: Fv = 0, Gv = 0 */
                }
                for (int eq = 0; eq < 4; ++eq) FL[eq + 4 * i] = m->ns ? ((Fi[eq] - Fv[eq]) - Fa[eq]) : Fi[eq];
                for (int d = 0; d < 2; ++d)
                    for (int eq = 0; eq < 4; ++eq) GL[eq + 4 * i + 4 * P * d] = Gv[eq + 4 * d];
            }
            Elem *e = &m->el[ed->quads[0]];
            std_face(m, e, ed->loc[0], FL);
            if (gradfl) std_gradient_face(m, e, ed->loc[0], GL);
        }
    }
#pragma omp parallel for schedule(static)
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        int P = e->P;
        for (int j = 0; j < P; ++j)
            for (int i = 0; i < P; ++i)
                for (int eq = 0; eq < 4; ++eq) m->QDot[QIDX(e, eq, i, j)] = m->QDot[QIDX(e, eq, i, j)] / e->jac[i + P * j];
    }
    m->traces_valid = 1;
}

/* ======================================================================== */
/* Public API                                                                */
/* ======================================================================== */
typedef struct {
    double Mach, Re, Pr;
    int ns, riemann, viscous, ip_method; /* ip_method: 0 SIPG, 1 NIPG, 2 IIPG */
    double sigma0;
    int ad;            /* artificial dissipation (AD_*): none, "Physical" or "Laplacian", jumps-based indicator */
    double ad_ceps;    /* "Artificial dissipation intensity" */
} OrOptions;

void *or_create(const MeshIn *in, const OrOptions *o) {
    Mesh *m = mesh_build(in);
    if (!m) return NULL;
    phys_init(&m->ph, o->Mach, o->Re, o->Pr);
    m->ns = o->ns; m->riemann = o->riemann; m->viscous = o->viscous; m->sigma0 = o->sigma0;
    if (m->ns && m->viscous == VISC_NONE) m->viscous = VISC_BR1;
    if (o->ip_method == 0) { m->ip_eps = -1.0; m->ip_gradflux = 1; }
    else if (o->ip_method == 1) { m->ip_eps = 1.0; m->ip_gradflux = 1; }
    else { m->ip_eps = 0.0; m->ip_gradflux = 0; }
    if (!riemann_fn(m->riemann)) { mesh_free(m); fail("unknown Riemann solver"); return NULL; }
    m->ad = m->ns ? o->ad : 0;   /* the module is Navier-Stokes only */
    m->ad_ceps = o->ad_ceps;
    for (int k = 0; k < m->E; ++k) {
        const int sub = m->ed[k].sub != 0, pnc = m->ed[k].low >= 0;
        if ((sub || pnc) && m->ad) {
            fail(sub ? "subdivided edges: no artificial dissipation" : "p-non-conforming edges: no artificial dissipation");
            mesh_free(m);
            return NULL;
        }
    }
    m->mu_e = (double *)calloc(m->K + 1, sizeof(double));
    m->lap_e = (double *)calloc(m->K + 1, sizeof(double));
    for (int k = 0; k < m->E; ++k) m->ed[k].mu_a = 0.0;
    return m;
}
void or_destroy(void *h) { mesh_free((Mesh *)h); }
int64_t or_ndof(void *h) { return ((Mesh *)h)->ndof; }
int or_num_edges(void *h) { return ((Mesh *)h)->E; }

int or_set_state(void *h, const double *Q) {
    Mesh *m = h;
    memcpy(m->Q, Q, sizeof(double) * m->ndof);
    m->traces_valid = 0;
    return 0;
}
int or_get_state(void *h, double *Q) { Mesh *m = h; memcpy(Q, m->Q, sizeof(double) * m->ndof); return 0; }
int or_get_qdot(void *h, double *Q) { Mesh *m = h; memcpy(Q, m->QDot, sizeof(double) * m->ndof); return 0; }
int or_get_grad(void *h, double *dQ) { Mesh *m = h; memcpy(dQ, m->dQ, sizeof(double) * 2 * m->ndof); return 0; }

/* Node coordinates X[node*2 + d] (node = off_e + i + P j) for evaluating ICs. */
int or_get_coords(void *h, double *X) {
    Mesh *m = h;
    for (int e = 0; e < m->K; ++e)
        memcpy(X + 2 * m->el[e].off, m->el[e].X, sizeof(double) * 2 * m->el[e].P * m->el[e].P);
    return 0;
}
/* Geometry: Ja[node*4 + (r*2+c)], jac[node] */
int or_get_metrics(void *h, double *Ja, double *jac) {
    Mesh *m = h;
    for (int e = 0; e < m->K; ++e) {
        Elem *el = &m->el[e];
        int PP = el->P * el->P;
        for (int k = 0; k < PP; ++k) {
            for (int c = 0; c < 4; ++c) Ja[(el->off + k) * 4 + c] = el->Ja[c * PP + k];
            jac[el->off + k] = el->jac[k];
        }
    }
    return 0;
}

/* Connectivity export (for bit-exact comparison with the product builder):
 *  edges[k*8 + ...] = {node0, node1, quadL, quadR, locL, locR, inverse, marker}
 *  elem[e*12 + ...] = {edges[4], dir[4], qpos[4]}
 *  periodic partner per edge (or -1) */
int or_get_connectivity(void *h, int *edges, int *elem, int *partner) {
    Mesh *m = h;
    for (int k = 0; k < m->E; ++k) {
        Edge *ed = &m->ed[k];
        int *o = edges + 8 * k;
        o[0] = ed->nodes[0]; o[1] = ed->nodes[1]; o[2] = ed->quads[0]; o[3] = ed->quads[1];
        o[4] = ed->loc[0]; o[5] = ed->bdry ? 0 : ed->loc[1]; o[6] = ed->inverse; o[7] = ed->marker;
        partner[k] = ed->partner;
    }
    for (int e = 0; e < m->K; ++e) {
        int *o = elem + 12 * e;
        for (int c = 0; c < 4; ++c) { o[c] = m->el[e].edges[c]; o[4 + c] = m->el[e].dir[c]; o[8 + c] = m->el[e].qpos[c]; }
    }
    return 0;
}
/* Edge geometry export: geo[k*(3*MAXPE)+...] normals and dS in edge node order, per node. */
/* Node count of zone `marker`'s edges (ascending edge id, each edge's N + 1 nodes in its own
 * order), or -1. */
static int64_t zone_nodes(const Mesh *m, int marker) {
    if (marker < 1 || marker > m->nmarkers) return -1;
    const Zone *Z = &m->zones[marker - 1];
    int64_t n = 0;
    for (int a = 0; a < Z->nedges; ++a) n += m->ed[Z->edges[a]].N + 1;
    return n;
}
/* The coordinates x[2 node + d] of those nodes (edge % x, where the reference evaluates a
 * boundary-state function); x = NULL returns their number. */
int64_t or_boundary_nodes(void *h, int marker, int64_t n, double *x) {
    Mesh *m = h;
    if (!x) return zone_nodes(m, marker) < 0 ? fail("or_boundary_nodes: marker out of range") : zone_nodes(m, marker);
    if (zone_nodes(m, marker) != n) return fail("or_boundary_nodes: marker or node count mismatch");
    const Zone *Z = &m->zones[marker - 1];
    int64_t o = 0;
    for (int a = 0; a < Z->nedges; ++a) {
        const Edge *ed = &m->ed[Z->edges[a]];
        for (int i = 0; i <= ed->N; ++i, ++o) { x[2 * o] = ed->X[2 * i]; x[2 * o + 1] = ed->X[2 * i + 1]; }
    }
    return 0;
}
/* Boundary-state function values q[4 node + eq] at those nodes (dimensionless), replacing the
 * zone state of a far-field, Dirichlet, newDirichlet or pressure zone; q = NULL drops them. */
int or_set_boundary_state(void *h, int marker, int64_t n, const double *q) {
    Mesh *m = h;
    if (zone_nodes(m, marker) < 0 || (q && zone_nodes(m, marker) != n))
        return fail("or_set_boundary_state: marker or node count mismatch");
    const Zone *Z = &m->zones[marker - 1];
    if (q && Z->type != BC_FARFIELD && Z->type != BC_DIRICHLET && Z->type != BC_NEWDIRICHLET &&
        Z->type != BC_PRESSUREOUTLET && Z->type != BC_PRESSUREINLET)
        return fail("or_set_boundary_state: the zone's kind has no boundary state");
    int64_t o = 0;
    for (int a = 0; a < Z->nedges; ++a) {
        Edge *ed = &m->ed[Z->edges[a]];
        const int P = ed->N + 1;
        if (!q) { free(ed->fq); ed->fq = NULL; continue; }
        if (!ed->fq) ed->fq = (double *)malloc(sizeof(double) * 4 * P);
        memcpy(ed->fq, q + 4 * o, sizeof(double) * 4 * P);
        o += P;
    }
    m->traces_valid = 0;
    return 0;
}

int or_get_edge_geometry(void *h, int maxP, double *nx, double *ny, double *dS, double *invh) {
    Mesh *m = h;
    for (int k = 0; k < m->E; ++k) {
        Edge *ed = &m->ed[k];
        int P = ed->N + 1;
        for (int i = 0; i < P && i < maxP; ++i) {
            int c = ed->curved ? i : 0;
            nx[k * maxP + i] = ed->n[2 * c]; ny[k * maxP + i] = ed->n[2 * c + 1]; dS[k * maxP + i] = ed->dS[c];
        }
        invh[k] = ed->invh;
    }
    return 0;
}

/* The p-non-conforming data of edge k: low side (-1 conforming), orders, T_forward and T_backward. */
int or_edge_mortar(void *h, int k, int *low, int *N, int *NLow, double *Tf, double *Tb) {
    Mesh *m = (Mesh *)h;
    if (k < 0 || k >= m->E) return fail("bad edge");
    Edge *ed = &m->ed[k];
    *low = ed->low; *N = ed->N; *NLow = ed->NLow;
    if (ed->low >= 0) {
        int n = (ed->N + 1) * (ed->NLow + 1);
        memcpy(Tf, ed->Tf, sizeof(double) * n);
        memcpy(Tb, ed->Tb, sizeof(double) * n);
    }
    return 0;
}

/* A subdivided edge's linking and mortar matrices: info = {N_left, N_north, N_south, N of the
 * NORTH and SOUTH elements, their directions, the three elements, their edge locations}; mats =
 * TLN_F, TLS_F, TLN_B, TLS_B, TRN_F, TRS_F, TRN_B, TRS_B, each row-major in a MAXP*MAXP slot.
 * Returns 1 for an edge that is not subdivided. */
int or_edge_subdivided(void *h, int k, int *info, double *mats) {
    Mesh *m = h;
    if (k < 0 || k >= m->E) return fail("edge index out of range");
    const Edge *ed = &m->ed[k];
    if (!ed->sub) return 1;
    const SubData *d = ed->sd;
    const Elem *L = &m->el[ed->quads[0]], *RN = &m->el[ed->quads[1]], *RS = &m->el[ed->quads[2]];
    info[0] = L->N; info[1] = d->NN; info[2] = d->NS; info[3] = RN->N; info[4] = RS->N;
    info[5] = RN->dir[ed->loc[1] - 1]; info[6] = RS->dir[ed->loc[2] - 1];
    for (int c = 0; c < 3; ++c) { info[7 + c] = ed->quads[c]; info[10 + c] = ed->loc[c]; }
    const double *src[8] = {d->TLN_F, d->TLS_F, d->TLN_B, d->TLS_B, d->TRN_F, d->TRS_F, d->TRN_B, d->TRS_B};
    const int rows[8] = {d->NN + 1, d->NS + 1, L->P, L->P, d->NN + 1, d->NS + 1, RN->P, RS->P};
    const int cols[8] = {L->P, L->P, d->NN + 1, d->NS + 1, RN->P, RS->P, d->NN + 1, d->NS + 1};
    for (int t = 0; t < 8; ++t)
        memcpy(mats + (size_t)t * MAXP * MAXP, src[t], sizeof(double) * rows[t] * cols[t]);
    return 0;
}

int or_time_derivative(void *h, double t) { (void)t; time_derivative((Mesh *)h); return 0; }

/* Residual max|QDot| per equation (QuadMesh_Auxiliars.incf:22-38) */
int or_residuals(void *h, double *res) {
    Mesh *m = h;
    for (int k = 0; k < 4; ++k) res[k] = 0.0;
    for (int64_t i = 0; i < m->ndof; ++i) { double a = fabs(m->QDot[i]); if (a > res[i & 3]) res[i & 3] = a; }
    return 0;
}

/* DGTimeIntegrator.f90:263-365: scheme 0 explicit Euler, 1 RK3, 2 RK5 */
int or_step(void *h, int scheme, double t, double dt) {
    Mesh *m = h;
    if (scheme == 0) {
        time_derivative(m);
        for (int64_t i = 0; i < m->ndof; ++i) m->Q[i] = m->Q[i] + dt * m->QDot[i];
        return 0;
    }
    if (scheme == 1) {
        const double am[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
        const double gm[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
        const double cm[3] = {0.0, 1.0 / 3.0, 3.0 / 4.0};
        for (int s = 0; s < 3; ++s) {
            (void)cm; (void)t;
            time_derivative(m);
            if (s == 0) for (int64_t i = 0; i < m->ndof; ++i) m->G[i] = m->QDot[i];
            else for (int64_t i = 0; i < m->ndof; ++i) m->G[i] = am[s] * m->G[i] + m->QDot[i];
            double c = gm[s] * dt;
            for (int64_t i = 0; i < m->ndof; ++i) m->Q[i] = m->Q[i] + c * m->G[i];
        }
        return 0;
    }
    if (scheme == 2) {
        const double am[5] = {0.0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
        const double gm[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681};
        for (int s = 0; s < 5; ++s) {
            time_derivative(m);
            if (s == 0) {
                for (int64_t i = 0; i < m->ndof; ++i) { m->G[i] = dt * m->QDot[i]; m->Q[i] = m->Q[i] + gm[s] * m->G[i]; }
            } else {
                for (int64_t i = 0; i < m->ndof; ++i) {
                    m->G[i] = am[s] * m->G[i] + dt * m->QDot[i];
                    m->Q[i] = m->Q[i] + gm[s] * m->G[i];
                }
            }
        }
        return 0;
    }
    return fail("unknown scheme");
}

/* DGTimeIntegrator.f90:462-511 with the SAVEd running maxima umax/amax */
int or_estimate_dt(void *h, double dt_max, double cfl, double *dt_out) {
    Mesh *m = h;
    const Phys *ph = &m->ph;
    double dt = dt_max;
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        for (int i = 0; i < e->P; ++i)
            for (int j = 0; j < e->P; ++j) {
                const double *q = m->Q + QIDX(e, 0, i, j);
                double u = sqrt(q[1] * q[1] + q[2] * q[2]) / q[0];
                if (u > m->dt_umax) m->dt_umax = u;
                double ur = q[1] / q[0], vr = q[2] / q[0];
                double a = sqrt(ph->gamma * ph->gm1 * (q[3] / q[0] - 0.5 * (ur * ur) - 0.5 * (vr * vr)));
                if (a > m->dt_amax) m->dt_amax = a;
            }
        const double *x1 = m->pts + 2 * e->nodes[0], *x2 = m->pts + 2 * e->nodes[1];
        const double *x3 = m->pts + 2 * e->nodes[2], *x4 = m->pts + 2 * e->nodes[3];
#define DIST(a, b) sqrt(((a)[0] - (b)[0]) * ((a)[0] - (b)[0]) + ((a)[1] - (b)[1]) * ((a)[1] - (b)[1]))
        double dx = DIST(x1, x2), d2 = DIST(x2, x3), d3 = DIST(x3, x4), d4 = DIST(x1, x4);
#undef DIST
        if (d2 < dx) dx = d2;
        if (d3 < dx) dx = d3;
        if (d4 < dx) dx = d4;
        double de = ph->Mach * sqrt(ph->gamma) * cfl * dx / (m->dt_umax + m->dt_amax) / (e->N + 1);
        if (de < dt) dt = de;
    }
    *dt_out = dt;
    return 0;
}
void or_reset_dt_state(void *h) { ((Mesh *)h)->dt_umax = 0.0; ((Mesh *)h)->dt_amax = 0.0; }

/* Conservation sums sum_e sum_ij w_i w_j J Q (FP64) */
int or_conservation_sums(void *h, double *out) {
    Mesh *m = h;
    for (int k = 0; k < 4; ++k) out[k] = 0.0;
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        const Spec *s = spec_get(e->N);
        for (int j = 0; j < e->P; ++j)
            for (int i = 0; i < e->P; ++i)
                for (int k = 0; k < 4; ++k) out[k] += s->w[i] * s->w[j] * e->jac[i + e->P * j] * m->Q[QIDX(e, k, i, j)];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Monitor integrals of the current state (QuadMesh_Integrals.incf:22-346)   */
/* ------------------------------------------------------------------------ */
/* VariableConversion.f90:48-55 */
static double pressure0(const Phys *ph, const double *q) {
    return ph->gm1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
}

/* Compute_volumeIntegral (:22-65): out[0] "One" = sum_e w^T jac w; out[1] "dSnorm" =
 * sqrt(sum_e w^T ((p / rho^gamma - 1)^2 jac) w) (BilinearForm_F, MatrixOperations.f90:211-234). */
int or_volume_integrals(void *h, double *out) {
    Mesh *m = h;
    const Phys *ph = &m->ph;
    double one = 0.0, sn = 0.0;
    for (int eid = 0; eid < m->K; ++eid) {
        Elem *e = &m->el[eid];
        const Spec *s = spec_get(e->N);
        const int P = e->P;
        double b1 = 0.0, b2 = 0.0;
        for (int j = 0; j < P; ++j) {
            double r1 = 0.0, r2 = 0.0;   /* (A w)(i) summed over i with w(i): X^T (A Y) */
            for (int i = 0; i < P; ++i) {
                const double jc = e->jac[i + P * j];
                const double *q = m->Q + QIDX(e, 0, i, j);
                const double v = pressure0(ph, q) / pow(q[0], ph->gamma) - 1.0;
                r1 = r1 + s->w[i] * jc;
                r2 = r2 + s->w[i] * (v * v * jc);
            }
            b1 = b1 + r1 * s->w[j];
            b2 = b2 + r2 * s->w[j];
        }
        one = one + b1;
        sn = sn + b2;
    }
    out[0] = one;
    out[1] = sqrt(sn);
    return 0;
}

/* Surface integrals over the edges of zone `marker` (1-based), in the zone's edge order, with
 * the element-side traces storage(1) (interpolated from the current state; for Navier-Stokes
 * also the gradient and its traces, as the residual evaluation forms them):
 *   out[0] "Surface"  sum w dS                         (Compute_ScalarScalarSurfaceIntegral :83-99)
 *   out[1] "pressure" sum w p dS                       (:101-126)
 *   out[2] "density"  sum w rho dS                     (:128-152; the reference indexes
 *                     storage(1)%Q(0:N, IRHO) of a (1:NCONS, 0:N) array, read here as rho)
 *   out[3..4] "pressure" sum w p n dS                  (Compute_ScalarVectorSurfaceIntegral :174-204)
 *   out[5] "mass-flow" sum w (rho u . n) dS            (Compute_VectorVectorSurfaceIntegral :226-252)
 *   out[6] "flow"      sum w (rho u . n) dS / rho      (:254-281)
 *   out[7..8] "viscous" sum w (tau . n) dS              (Compute_TensorVectorSurfaceIntegral :305-342,
 *                     tau of ComputeViscousTensor, ViscousFluxes.f90:416-461; zero for Euler) */
/* The boundary states of zone `marker` from its last update_bcs (the last time derivative
 * or stage): per edge node, in the zone's edge order, the trace Q (4), the outward normal (2),
 * uB (4), uSB (4) and the viscous type.  Returns the node count; with NULL arrays only counts. */
int or_boundary_states(void *h, int marker, double *qt, double *nrm, double *ub, double *usb, int *vbc) {
    Mesh *m = h;
    if (marker < 1 || marker > m->nmarkers) return fail("or_boundary_states: bad marker");
    const Zone *Z = &m->zones[marker - 1];
    if (Z->type == BC_PERIODIC) return fail("or_boundary_states: periodic zone");
    int c = 0;
    for (int a = 0; a < Z->nedges; ++a) {
        const Edge *ed = &m->ed[Z->edges[a]];
        for (int i = 0; i <= ed->N; ++i, ++c) {
            if (!qt) continue;
            const double *nn = ed->curved ? ed->n + 2 * i : ed->n;
            for (int k = 0; k < 4; ++k) {
                qt[4 * c + k] = ed->tq[0][4 * i + k];
                ub[4 * c + k] = ed->uB[4 * i + k];
                usb[4 * c + k] = ed->uSB ? ed->uSB[4 * i + k] : 0.0;
            }
            nrm[2 * c] = nn[0]; nrm[2 * c + 1] = nn[1];
            vbc[c] = ed->vbc[i];
        }
    }
    return c;
}

int or_surface_integrals(void *h, int marker, double *out) {
    Mesh *m = h;
    const Phys *ph = &m->ph;
    if (marker < 1 || marker > m->nmarkers) return fail("or_surface_integrals: bad marker");
    /* The integrals read the edge storage (QuadMesh_Integrals.incf:109-314), i.e. the traces and
     * gradients of the last derivative evaluation: after a time step those of its last stage's
     * input state, not of the updated state.  Only when the state was set after the last
     * evaluation are they formed from the current state. */
    if (!m->traces_valid) {
        prolong(m, 0);
        if (m->ns) {
            update_bcs(m);
            if (m->viscous == VISC_BR1) gradient_br1(m);
            else gradient_ip(m);
            prolong(m, 1);
        }
        m->traces_valid = 1;
    }
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    const Zone *Z = &m->zones[marker - 1];
    for (int a = 0; a < Z->nedges; ++a) {
        const Edge *ed = &m->ed[Z->edges[a]];
        const Spec *s = spec_get(ed->N);
        const int P = ed->N + 1;
        for (int j = 0; j < P; ++j) {
            const double *q = ed->tq[0] + 4 * j;
            const double nx = ed->curved ? ed->n[2 * j] : ed->n[0];
            const double ny = ed->curved ? ed->n[2 * j + 1] : ed->n[1];
            const double dS = ed->curved ? ed->dS[j] : ed->dS[0];
            const double w = s->w[j], p = pressure0(ph, q);
            out[0] = out[0] + w * dS;
            out[1] = out[1] + p * dS * w;
            out[2] = out[2] + q[0] * dS * w;
            out[3] = out[3] + p * w * nx * dS;
            out[4] = out[4] + p * w * ny * dS;
            out[5] = out[5] + w * (q[1] * nx + q[2] * ny) * dS;
            out[6] = out[6] + w * (q[1] * nx + q[2] * ny) * dS / q[0];
            if (m->ns) {
                const double *dq = ed->tdq[0];
                const double dx0 = dq[0 + 4 * j], dx1 = dq[1 + 4 * j], dx2 = dq[2 + 4 * j];
                const double dy0 = dq[0 + 4 * j + 4 * P], dy1 = dq[1 + 4 * j + 4 * P], dy2 = dq[2 + 4 * j + 4 * P];
                const double mu = ph->mu, invRho = 1.0 / q[0], u = q[1] * invRho, v = q[2] * invRho;
                const double uDR = u * invRho, vDR = v * invRho;
                const double divV = -uDR * dx0 + invRho * dx1 - vDR * dy0 + invRho * dy2;
                const double txx = 2.0 * mu * (-uDR * dx0 + invRho * dx1) + ph->lambda * mu * divV;
                const double tyy = 2.0 * mu * (-vDR * dy0 + invRho * dy2) + ph->lambda * mu * divV;
                const double txy = -mu * vDR * dx0 - mu * uDR * dy0 + mu * invRho * dy1 + mu * invRho * dx2;
                out[7] = out[7] + w * (txx * nx + txy * ny) * dS;
                out[8] = out[8] + w * (txy * nx + tyy * ny) * dS;
            }
        }
    }
    return 0;
}

/* Probe_Update (IO/Monitors.f90:739-846) for n probes at reference coordinates (xi, eta) of
 * elements elem (0-based): the Lagrange interpolants of Probe_Initialization (:715-716,
 * barycentric as NodesAndWeightsClass.f90:334-342), B = X^T A Y with (A Y)(i) = sum_j A(i,j) Y(j)
 * (BilinearForm, MatrixOperations.f90:185-209) per conservative variable of Q (QDot for var
 * 4..7), then var: 0 rho 1 rhou 2 rhov 3 rhoe 4 rhot 5 rhout 6 rhovt 7 rhoet 8 u 9 v 10 p
 * 11 Mach, dimensionless (without the refValues factors). */
int or_probes(void *h, int n, const int *elem, const double *xi, const double *eta, const int *var, double *out) {
    Mesh *m = h;
    const Phys *ph = &m->ph;
    for (int k = 0; k < n; ++k) {
        if (elem[k] < 0 || elem[k] >= m->K || var[k] < 0 || var[k] > 11) return fail("or_probes: bad probe");
        Elem *e = &m->el[elem[k]];
        const Spec *s = spec_get(e->N);
        const int P = e->P;
        double lx[MAXP], ly[MAXP], v[4];
        lagrange_barycentric(xi[k], e->N, s->xi, s->wb, lx);
        lagrange_barycentric(eta[k], e->N, s->xi, s->wb, ly);
        const double *A = (var[k] >= 4 && var[k] <= 7) ? m->QDot : m->Q;
        for (int eq = 0; eq < 4; ++eq) {
            double b = 0.0;
            for (int i = 0; i < P; ++i) {
                double r = 0.0;
                for (int j = 0; j < P; ++j) r = r + A[QIDX(e, eq, i, j)] * ly[j];
                b = b + lx[i] * r;
            }
            v[eq] = b;
        }
        switch (var[k]) {
        case 8: out[k] = v[1] / v[0]; break;
        case 9: out[k] = v[2] / v[0]; break;
        case 10: out[k] = ph->gm1 * (v[3] - 0.5 * (v[1] * v[1] + v[2] * v[2]) / v[0]); break;
        case 11: out[k] = sqrt(v[1] * v[1] + v[2] * v[2]) / v[0] / sqrt(ph->gamma); break;
        default: out[k] = v[var[k] & 3]; break;
        }
    }
    return 0;
}

/* Point-wise flux exports for tests */
int or_inviscid_flux(double Mach, int n, const double *q, double *f, double *g) {
    Phys ph; phys_init(&ph, Mach, 0.0, 0.72);
    for (int i = 0; i < n; ++i) inviscid_flux(&ph, q + 4 * i, f + 4 * i, g + 4 * i);
    return 0;
}
int or_viscous_flux(double Re, double Pr, int n, const double *q, const double *dq, double *f, double *g) {
    Phys ph; phys_init(&ph, 1.0, Re, Pr);
    for (int i = 0; i < n; ++i) viscous_flux(&ph, q + 4 * i, dq + 8 * i, f + 4 * i, g + 4 * i);
    return 0;
}

/* Threads of the OpenMP element loops (for timing the restatement at a stated core count);
 * returns the count now in effect. */
int or_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}
