// p-non-conforming edges (mortars) of a mesh whose elements have several polynomial orders.
//
// An edge between elements of orders NLow < N takes the order N (QuadMesh_Construct.incf:104-118).
// Its Riemann problem is solved on its own N+1 nodes: the low side's traces are interpolated up
// with T_forward (Edge_ProjectSolution Types 3-6, QuadElement_EdgeProcedures.incf:749-871), and
// the result returns to the low side as T_backward = Minv_low T_forward^T M_edge times the
// edge values (Edge_ProjectFluxes Types 3-6, :1081-1206), negated for the RIGHT side.  One
// workgroup per edge: lane i forms edge node i, then lane j the projection to a side's node j.
// The results are written per element face node, where the stage kernel (flux) and the BR1
// gradient (uStar) of each side read them in place of evaluating the face.
#include <hip/hip_runtime.h>

#include "hg_device.h"
#include "hg_physics.h"

namespace hg {
namespace {

constexpr int kEdgeMax = 16;   // edge points per direction held in LDS (orders up to 15)

// Side `side`'s NV values at edge node i, in the edge's (LEFT) orientation, from `src` (that
// side's own face nodes): the low side interpolated with T_forward (rows in ascending l, as
// the reference accumulates them), a RIGHT side of an inverse edge read backwards (Types 2, 4)
// or interpolated into the reversed node (Type 6).
template <int NV, typename T>
__device__ void side_values(const MortarEdge &me, const double *mats, int side, const T *src, int i, double *out) {
    const int P = me.P, Pl = me.Ps[me.low];
    const bool rev = side == 1 && me.inverse;
    if (side != me.low) {
        const T *v = src + (rev ? P - 1 - i : i) * NV;
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = (double)v[c];
        return;
    }
    const double *Tf = mats + me.mat + (rev ? P - 1 - i : i) * Pl;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double acc = 0.0;
        for (int l = 0; l < Pl; ++l) acc = acc + Tf[l] * (double)src[l * NV + c];
        out[c] = acc;
    }
}

// A side's Q traces at its face slot, in the traces' type.
template <typename TR>
__device__ inline const TR *q_trace(const MortarArgs &m, int sub, int slot, int P) {
    return (const TR *)m.tr[sub] + (int64_t)slot * P * 4;
}

// The mixed form keeps the energy slot of its traces less Eref: the fluxes need it back.
template <typename TR>
__device__ inline void unshift(const MortarArgs &m, double *q) {
    if constexpr (sizeof(TR) == 4) q[3] += m.Eref;
}

// The mixed form's momentum fluxes carry the pressure less pref (a uniform pressure adds nothing
// to an element's balance, so its volume and face terms both leave it out): the FP64 Riemann
// flux of a mortar node, in the solver's scaling, follows that convention.
template <typename TR>
__device__ inline void shift_flux(const MortarArgs &m, const Phys<double> &ph, double nx, double ny, double *F) {
    if constexpr (sizeof(TR) == 4) {
        F[1] -= m.ph.pref * ph.invSGM * nx;
        F[2] -= m.ph.pref * ph.invSGM * ny;
    }
}

// The edge values sF[i*NV + c] at side `side`'s node j: the high side takes them as they are
// (negated for RIGHT when NEG, reversed for an inverse edge), the low side T_backward times them
// (the RIGHT side accumulating the negated terms, as Types 5 and 6 of ProjectFluxes do;
// ProjectGradientFluxes, :1208-1339, negates neither side: NEG = false).
template <int NV = 4, bool NEG = true>
__device__ void flux_to_side(const MortarEdge &me, const double *mats, int side, const double *sF, int j, double *out) {
    const int P = me.P, Pl = me.Ps[me.low];
    const bool rev = side == 1 && me.inverse;
    const bool neg = NEG && side == 1;
    if (side != me.low) {
        const int i = rev ? P - 1 - j : j;
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = neg ? -sF[i * NV + c] : sF[i * NV + c];
        return;
    }
    const int i = rev ? Pl - 1 - j : j;
    const double *Tb = mats + me.mat + P * Pl + i * P;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double acc = 0.0;
        if (!neg) {
            for (int l = 0; l < P; ++l) acc = acc + Tb[l] * sF[l * NV + c];
        } else {
            for (int l = 0; l < P; ++l) acc = acc - Tb[l] * sF[l * NV + c];
        }
        out[c] = acc;
    }
}

// Geometry of edge node i: a curved edge's own (from the pool), a straight edge's one normal and dS.
__device__ inline void edge_geo(const MortarArgs &m, const MortarEdge &me, int i, double &nx, double &ny, double &dS) {
    if (me.geo < 0) { nx = me.nx; ny = me.ny; dS = me.dS; return; }
    const double *g = m.sgeo + 3 * ((int64_t)me.geo + i);
    nx = g[0]; ny = g[1]; dS = g[2];
}

__device__ inline const double *side_trace(const MortarArgs &m, const MortarEdge &me, int side, const double *const *buf,
                                           int nv) {
    return buf[me.sub[side]] + (int64_t)me.slot[side] * me.Ps[side] * nv;
}
template <typename TR>
__device__ inline const TR *side_q(const MortarArgs &m, const MortarEdge &me, int side) {
    return q_trace<TR>(m, me.sub[side], me.slot[side], me.Ps[side]);
}

// BaseClass_ComputeSolutionRiemann_Interior / _CurvedInterior (DGViscousMethods.f90:538-616):
// uStar = (QL + QR)/2 on the edge's nodes (BR1_SolutionRiemannSolver, DGViscous_BR1.f90:311-328).
// A straight edge projects uStar to each side (the gradient kernels form uStar dS n from it); a
// curved edge scales each node's uStar by its dS n first and projects each direction, which the
// gradient lift of the side then takes as they are.
template <typename TR>
__global__ __launch_bounds__(64) void k_mortar_ustar(MortarArgs m) {
    __shared__ double sU[kEdgeMax * 8];
    const MortarEdge me = m.edges[blockIdx.x];
    const int t = threadIdx.x;
    const bool curved = me.geo >= 0;
    if (t < me.P) {
        double qL[4], qR[4];
        // (the mixed form's uStar stays in the traces' shifted energy, as its gradient kernels
        // lift it against the shifted state)
        side_values<4>(me, m.mats, 0, side_q<TR>(m, me, 0), t, qL);
        side_values<4>(me, m.mats, 1, side_q<TR>(m, me, 1), t, qR);
        if (!curved) {
#pragma unroll
            for (int c = 0; c < 4; ++c) sU[t * 4 + c] = 0.5 * (qL[c] + qR[c]);
        } else {
            double nx, ny, dS;
            edge_geo(m, me, t, nx, ny, dS);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double us = 0.5 * (qL[c] + qR[c]);
                sU[t * 8 + c] = us * dS * nx;
                sU[t * 8 + 4 + c] = us * dS * ny;
            }
        }
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        if (t < me.Ps[side]) {
            double *o = m.mustar[me.sub[side]] + ((int64_t)me.slot[side] * me.Ps[side] + t) * 8;
            if (!curved) {
                double u[4];
                flux_to_side<4>(me, m.mats, side, sU, t, u);
#pragma unroll
                for (int c = 0; c < 4; ++c) o[c] = u[c];
            } else {
                double u[8];
                flux_to_side<8>(me, m.mats, side, sU, t, u);
#pragma unroll
                for (int c = 0; c < 8; ++c) o[c] = u[c];
            }
        }
    }
}

// The viscous part of an interior mortar flux at one node: the average of the two sides' viscous
// fluxes (BR1_RiemannSolver, DGViscous_BR1.f90:330-371; IP_RiemannSolver, DGViscous_IP.f90:87-135)
// minus, for interior penalty, sigma0 mu Nn^2 invh (QL - QR), Nn the order of the nodes.
__device__ inline void mortar_viscous(const MortarArgs &m, const Phys<double> &ph, const double *qL, const double *qR,
                                      const double *dL, const double *dR, double nx, double ny, int Nn, double invh,
                                      double *Fv) {
    double fL[4], gL[4], fR[4], gR[4];
    viscous_flux(ph, qL, dL, fL, gL);
    viscous_flux(ph, qR, dR, fR, gR);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double FL0 = ph.mu * fL[c], FL1 = ph.mu * gL[c], FR0 = ph.mu * fR[c], FR1 = ph.mu * gR[c];
        double acc = 0.0;
        acc = acc + 0.5 * (FL0 + FR0) * nx;
        acc = acc + 0.5 * (FL1 + FR1) * ny;
        Fv[c] = acc;
    }
    if (m.ph.viscous == 2) {
        const double pen = m.ph.sigma0 * ph.mu * (double)Nn * (double)Nn * invh;
#pragma unroll
        for (int c = 0; c < 4; ++c) Fv[c] = Fv[c] - pen * (qL[c] - qR[c]);
    }
}

// IP_GradientRiemannSolver (DGViscous_IP.f90:216-249) at one node: the false gradient
// (QL - QR) n, then 0.5 eps mu F_v(QL, .) for the LEFT side and -0.5 eps mu F_v(QR, .) for the
// RIGHT one, each [f(4), g(4)].
__device__ inline void ip_gradient_flux(const MortarArgs &m, const Phys<double> &ph, const double *qL, const double *qR,
                                        double nx, double ny, double *gl, double *gr) {
    double fg[8], f[4], g[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) { fg[c] = (qL[c] - qR[c]) * nx; fg[4 + c] = (qL[c] - qR[c]) * ny; }
    viscous_flux(ph, qL, fg, f, g);
    const double cL = 0.5 * m.ph.ip_eps * ph.mu;
#pragma unroll
    for (int c = 0; c < 4; ++c) { gl[c] = cL * f[c]; gl[4 + c] = cL * g[c]; }
    viscous_flux(ph, qR, fg, f, g);
    const double cR = -0.5 * m.ph.ip_eps * ph.mu;
#pragma unroll
    for (int c = 0; c < 4; ++c) { gr[c] = cR * f[c]; gr[4 + c] = cR * g[c]; }
}

// ComputeRiemannSolver_InteriorEdge / _InteriorCurvedEdge (DGSpatialDiscretizationMethods.f90:
// 545-759): Fi from the interior Riemann solver on the projected traces, Fv the viscous flux from
// the projected traces and gradients, FStar = (Fi - Fv) dS (dS and n per node on a curved edge),
// projected to each side.  Interior penalty with gradient fluxes: each side's IP gradient flux,
// returned by ProjectGradientFluxes (straight: projected, then times dS; curved: times each
// node's dS, then projected).
template <typename TR>
__global__ __launch_bounds__(64) void k_mortar_flux(MortarArgs m) {
    __shared__ double sF[kEdgeMax * 4];
    __shared__ double sG[2][kEdgeMax * 8];
    const MortarEdge me = m.edges[blockIdx.x];
    const int t = threadIdx.x;
    const Phys<double> ph(m.ph);
    const bool gradfl = m.ph.ns && m.ph.viscous == 2 && m.ph.ip_gradflux;
    const bool curved = me.geo >= 0;
    if (t < me.P) {
        double qL[4], qR[4], Fi[4], F[4], nx, ny, dS;
        edge_geo(m, me, t, nx, ny, dS);
        side_values<4>(me, m.mats, 0, side_q<TR>(m, me, 0), t, qL);
        side_values<4>(me, m.mats, 1, side_q<TR>(m, me, 1), t, qR);
        unshift<TR>(m, qL);
        unshift<TR>(m, qR);
        riemann<-1, double>(m.ph.riemann, ph, qL, qR, nx, ny, Fi);
        shift_flux<TR>(m, ph, nx, ny, Fi);
        if (m.ph.ns) {
            double dL[8], dR[8], Fv[4];
            side_values<8>(me, m.mats, 0, side_trace(m, me, 0, m.mdq, 8), t, dL);
            side_values<8>(me, m.mats, 1, side_trace(m, me, 1, m.mdq, 8), t, dR);
            mortar_viscous(m, ph, qL, qR, dL, dR, nx, ny, me.P - 1, me.invh, Fv);
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = (Fi[c] - Fv[c]) * dS;
            if (gradfl) {
                double gl[8], gr[8];
                ip_gradient_flux(m, ph, qL, qR, nx, ny, gl, gr);
                // [direction][eq] per node: x components, then y
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    sG[0][t * 8 + c] = curved ? gl[c] * dS : gl[c];
                    sG[1][t * 8 + c] = curved ? gr[c] * dS : gr[c];
                }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = Fi[c] * dS;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) sF[t * 4 + c] = F[c];
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        if (t < me.Ps[side]) {
            double f[4];
            flux_to_side<4>(me, m.mats, side, sF, t, f);
            double *o = m.mflux[me.sub[side]] + ((int64_t)me.slot[side] * me.Ps[side] + t) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = f[c];
            if (gradfl) {
                double g[8];
                flux_to_side<8, false>(me, m.mats, side, sG[side], t, g);
                double *og = m.mgrad[me.sub[side]] + ((int64_t)me.slot[side] * me.Ps[side] + t) * 8;
#pragma unroll
                for (int c = 0; c < 8; ++c) og[c] = curved ? g[c] : g[c] * me.dS;
            }
        }
    }
}

// ---- subdivided edges: the big (LEFT) element's edge meets a NORTH and a SOUTH element; the
// Riemann problem is solved on two mortars (NORTH: the hanging node to node 2, SOUTH: node 1 to
// the hanging node) from the three sides' traces interpolated with the T_FWD matrices, and
// returned with the T_BKW matrices: the big side takes NORTH's terms then SOUTH's, each small
// side its own mortar's, negated (ComputeRiemannSolver_SubdividedEdge /
// _CurvedSubdividedEdge, DGSpatialDiscretizationMethods.f90:761-1140).  One workgroup per edge:
// lanes 0-31 the NORTH mortar's nodes, lanes 32-63 the SOUTH one's; then lane j a side's node j.

struct SubMats {
    const double *LF[2], *LB[2], *RF[2], *RB[2];   // [0] NORTH, [1] SOUTH
};

__device__ inline SubMats sub_mats(const MortarArgs &m, const SubMortarEdge &se) {
    const int PL = se.Ps[0], PN = se.PM[0], PS = se.PM[1];
    const double *p = m.mats + se.mat;
    SubMats t;
    t.LF[0] = p; p += PN * PL;
    t.LF[1] = p; p += PS * PL;
    t.LB[0] = p; p += PL * PN;
    t.LB[1] = p; p += PL * PS;
    t.RF[0] = p; p += PN * se.Ps[1];
    t.RF[1] = p; p += PS * se.Ps[2];
    t.RB[0] = p; p += se.Ps[1] * PN;
    t.RB[1] = p;
    return t;
}

__device__ inline const double *sub_trace(const double *const *buf, const SubMortarEdge &se, int side, int nv) {
    return buf[se.sub[side]] + (int64_t)se.slot[side] * se.Ps[side] * nv;
}
template <typename TR>
__device__ inline const TR *sub_q(const MortarArgs &m, const SubMortarEdge &se, int side) {
    return q_trace<TR>(m, se.sub[side], se.slot[side], se.Ps[side]);
}

// Mortar node i of half h: the big side's (a) and the small side's (b) NV values there, sums
// over the sides' nodes in ascending order (as the reference accumulates them).
template <int NV, typename T>
__device__ void sub_values(const SubMortarEdge &se, const SubMats &Tm, int h, int i, const T *big,
                           const T *small, double *a, double *b) {
    const int PL = se.Ps[0], Pe = se.Ps[1 + h];
    const double *Tl = Tm.LF[h] + i * PL, *Tr = Tm.RF[h] + i * Pe;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double acc = 0.0;
        for (int l = 0; l < PL; ++l) acc = acc + Tl[l] * (double)big[l * NV + c];
        a[c] = acc;
        acc = 0.0;
        for (int l = 0; l < Pe; ++l) acc = acc + Tr[l] * (double)small[l * NV + c];
        b[c] = acc;
    }
}

// Mortar values (sV[h][l*NV + c]) at side `side`'s node j: the big side sums NORTH's then SOUTH's
// T_BKW terms (scaled by `bigw`), a small side subtracts its own mortar's (adds, when !NEG: the
// gradient fluxes, :906-940).
template <int NV = 4, bool NEG = true>
__device__ void sub_to_side(const SubMortarEdge &se, const SubMats &T, int side, int j, const double (*sV)[kEdgeMax * 8],
                            double bigw, double *out) {
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double acc = 0.0;
        if (side == 0) {
            for (int h = 0; h < 2; ++h) {
                const double *Tb = T.LB[h] + j * se.PM[h];
                for (int l = 0; l < se.PM[h]; ++l) acc = acc + (bigw * Tb[l]) * sV[h][l * NV + c];
            }
        } else {
            const int h = side - 1;
            const double *Tb = T.RB[h] + j * se.PM[h];
            if (NEG) {
                for (int l = 0; l < se.PM[h]; ++l) acc = acc - Tb[l] * sV[h][l * NV + c];
            } else {
                for (int l = 0; l < se.PM[h]; ++l) acc = acc + Tb[l] * sV[h][l * NV + c];
            }
        }
        out[c] = acc;
    }
}

// BaseClass_ComputeSolutionRiemann_Subdivided / _CurvedSubdivided (DGViscousMethods.f90:618-780):
// uStar = (QL + QR)/2 on each mortar.  Straight edges: the big side takes 0.5 T_BKW uStar of both
// (its gradient lift multiplies by the whole edge's dS n), a small side -T_BKW uStar of its own
// (lifted with its mortar's dS / 2 and n).  Curved edges: each mortar node's uStar times its dS n,
// each direction returned (big side: T_BKW of both; small sides: -T_BKW of their own) and lifted
// as it is.
template <typename TR>
__global__ __launch_bounds__(64) void k_sub_ustar(MortarArgs m) {
    __shared__ double sU[2][kEdgeMax * 8];
    const SubMortarEdge se = m.subs[blockIdx.x];
    const SubMats T = sub_mats(m, se);
    const int t = threadIdx.x, h = t >> 5, i = t & 31;
    if (i < se.PM[h]) {
        double a[4], b[4];
        sub_values<4>(se, T, h, i, sub_q<TR>(m, se, 0), sub_q<TR>(m, se, 1 + h), a, b);
        if (!se.curved) {
#pragma unroll
            for (int c = 0; c < 4; ++c) sU[h][i * 4 + c] = 0.5 * (a[c] + b[c]);
        } else {
            const double *g = m.sgeo + 3 * ((int64_t)se.geo + (h ? se.PM[0] : 0) + i);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double us = 0.5 * (a[c] + b[c]);
                sU[h][i * 8 + c] = us * g[2] * g[0];
                sU[h][i * 8 + 4 + c] = us * g[2] * g[1];
            }
        }
    }
    __syncthreads();
    for (int side = 0; side < 3; ++side) {
        if (t < se.Ps[side]) {
            double *o = m.mustar[se.sub[side]] + ((int64_t)se.slot[side] * se.Ps[side] + t) * 8;
            if (!se.curved) {
                double u[4];
                sub_to_side<4>(se, T, side, t, sU, 0.5, u);
#pragma unroll
                for (int c = 0; c < 4; ++c) o[c] = u[c];
            } else {
                double u[8];
                sub_to_side<8>(se, T, side, t, sU, 1.0, u);
#pragma unroll
                for (int c = 0; c < 8; ++c) o[c] = u[c];
            }
        }
    }
}

// FStar = (Fi - Fv) dS_mortar at each mortar node (its own normal and surface Jacobian), Fi the
// interior Riemann solver, Fv the viscous flux (interior penalty: sigma0 mu N_mortar^2 invh);
// to each side.  Interior penalty with gradient fluxes: each side's IP gradient flux on each
// mortar times the mortar node's dS, returned with T_BKW to all three sides, not negated
// (:906-940, :1093-1136).
template <typename TR>
__global__ __launch_bounds__(64) void k_sub_flux(MortarArgs m) {
    __shared__ double sF[2][kEdgeMax * 8];
    __shared__ double sGa[2][kEdgeMax * 8], sGb[2][kEdgeMax * 8];   // [mortar]: big side, small side
    const SubMortarEdge se = m.subs[blockIdx.x];
    const SubMats T = sub_mats(m, se);
    const int t = threadIdx.x, h = t >> 5, i = t & 31;
    const Phys<double> ph(m.ph);
    const bool gradfl = m.ph.ns && m.ph.viscous == 2 && m.ph.ip_gradflux;
    if (i < se.PM[h]) {
        const double *g = m.sgeo + 3 * ((int64_t)se.geo + (h ? se.PM[0] : 0) + i);
        const double nx = g[0], ny = g[1], dS = g[2];
        double qa[4], qb[4], Fi[4], F[4];
        sub_values<4>(se, T, h, i, sub_q<TR>(m, se, 0), sub_q<TR>(m, se, 1 + h), qa, qb);
        unshift<TR>(m, qa);
        unshift<TR>(m, qb);
        riemann<-1, double>(m.ph.riemann, ph, qa, qb, nx, ny, Fi);
        shift_flux<TR>(m, ph, nx, ny, Fi);
        if (m.ph.ns) {
            double da[8], db[8], Fv[4];
            sub_values<8>(se, T, h, i, sub_trace(m.mdq, se, 0, 8), sub_trace(m.mdq, se, 1 + h, 8), da, db);
            mortar_viscous(m, ph, qa, qb, da, db, nx, ny, se.PM[h] - 1, se.invh, Fv);
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = (Fi[c] - Fv[c]) * dS;
            if (gradfl) {
                double gl[8], gr[8];
                ip_gradient_flux(m, ph, qa, qb, nx, ny, gl, gr);
#pragma unroll
                for (int c = 0; c < 8; ++c) { sGa[h][i * 8 + c] = gl[c] * dS; sGb[h][i * 8 + c] = gr[c] * dS; }
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = Fi[c] * dS;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) sF[h][i * 4 + c] = F[c];
    }
    __syncthreads();
    for (int side = 0; side < 3; ++side) {
        if (t < se.Ps[side]) {
            double f[4];
            sub_to_side<4>(se, T, side, t, sF, 1.0, f);
            double *o = m.mflux[se.sub[side]] + ((int64_t)se.slot[side] * se.Ps[side] + t) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = f[c];
            if (gradfl) {
                double gv[8];
                sub_to_side<8, false>(se, T, side, t, side == 0 ? sGa : sGb, 1.0, gv);
                double *og = m.mgrad[se.sub[side]] + ((int64_t)se.slot[side] * se.Ps[side] + t) * 8;
#pragma unroll
                for (int c = 0; c < 8; ++c) og[c] = gv[c];
            }
        }
    }
}

__global__ void k_scatter(const double *src, const int *map, int n, double *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[map[i]] = src[i];
}

}  // namespace

hipError_t launch_mortar_ustar(const MortarArgs &m, hipStream_t st) {
    if (m.nedges <= 0) return hipSuccess;
    if (m.mixed) hipLaunchKernelGGL(k_mortar_ustar<float>, dim3(m.nedges), dim3(64), 0, st, m);
    else hipLaunchKernelGGL(k_mortar_ustar<double>, dim3(m.nedges), dim3(64), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_mortar_flux(const MortarArgs &m, hipStream_t st) {
    if (m.nedges <= 0) return hipSuccess;
    if (m.mixed) hipLaunchKernelGGL(k_mortar_flux<float>, dim3(m.nedges), dim3(64), 0, st, m);
    else hipLaunchKernelGGL(k_mortar_flux<double>, dim3(m.nedges), dim3(64), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_sub_ustar(const MortarArgs &m, hipStream_t st) {
    if (m.nsubs <= 0) return hipSuccess;
    if (m.mixed) hipLaunchKernelGGL(k_sub_ustar<float>, dim3(m.nsubs), dim3(64), 0, st, m);
    else hipLaunchKernelGGL(k_sub_ustar<double>, dim3(m.nsubs), dim3(64), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_sub_flux(const MortarArgs &m, hipStream_t st) {
    if (m.nsubs <= 0) return hipSuccess;
    if (m.mixed) hipLaunchKernelGGL(k_sub_flux<float>, dim3(m.nsubs), dim3(64), 0, st, m);
    else hipLaunchKernelGGL(k_sub_flux<double>, dim3(m.nsubs), dim3(64), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_scatter(const double *src, const int *map, int n, double *dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, st, src, map, n, dst);
    return hipGetLastError();
}

}  // namespace hg
