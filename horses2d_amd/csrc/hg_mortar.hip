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
template <int NV>
__device__ void side_values(const MortarEdge &me, const double *mats, int side, const double *src, int i, double *out) {
    const int P = me.P, Pl = me.Ps[me.low];
    const bool rev = side == 1 && me.inverse;
    if (side != me.low) {
        const double *v = src + (rev ? P - 1 - i : i) * NV;
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = v[c];
        return;
    }
    const double *Tf = mats + me.mat + (rev ? P - 1 - i : i) * Pl;
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        double acc = 0.0;
        for (int l = 0; l < Pl; ++l) acc = acc + Tf[l] * src[l * NV + c];
        out[c] = acc;
    }
}

// The edge values sF[i*4 + c] at side `side`'s node j: the high side takes them as they are
// (negated for RIGHT, reversed for an inverse edge), the low side T_backward times them (the
// RIGHT side accumulating the negated terms, as Types 5 and 6 do).
__device__ void flux_to_side(const MortarEdge &me, const double *mats, int side, const double *sF, int j, double *out) {
    const int P = me.P, Pl = me.Ps[me.low];
    const bool rev = side == 1 && me.inverse;
    if (side != me.low) {
        const int i = rev ? P - 1 - j : j;
#pragma unroll
        for (int c = 0; c < 4; ++c) out[c] = side == 0 ? sF[i * 4 + c] : -sF[i * 4 + c];
        return;
    }
    const int i = rev ? Pl - 1 - j : j;
    const double *Tb = mats + me.mat + P * Pl + i * P;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double acc = 0.0;
        if (side == 0) {
            for (int l = 0; l < P; ++l) acc = acc + Tb[l] * sF[l * 4 + c];
        } else {
            for (int l = 0; l < P; ++l) acc = acc - Tb[l] * sF[l * 4 + c];
        }
        out[c] = acc;
    }
}

__device__ inline const double *side_trace(const MortarArgs &m, const MortarEdge &me, int side, const double *const *buf,
                                           int nv) {
    return buf[me.sub[side]] + (int64_t)me.slot[side] * me.Ps[side] * nv;
}

// BaseClass_ComputeSolutionRiemann_Interior (DGViscousMethods.f90:538-577) on a straight edge:
// uStar = (QL + QR)/2 on the edge's nodes (BR1_SolutionRiemannSolver, DGViscous_BR1.f90:311-328),
// projected to each side (the gradient kernels form uStar dS n from it).
__global__ __launch_bounds__(64) void k_mortar_ustar(MortarArgs m) {
    __shared__ double sU[kEdgeMax * 4];
    const MortarEdge me = m.edges[blockIdx.x];
    const int t = threadIdx.x;
    if (t < me.P) {
        double qL[4], qR[4];
        side_values<4>(me, m.mats, 0, side_trace(m, me, 0, m.tr, 4), t, qL);
        side_values<4>(me, m.mats, 1, side_trace(m, me, 1, m.tr, 4), t, qR);
#pragma unroll
        for (int c = 0; c < 4; ++c) sU[t * 4 + c] = 0.5 * (qL[c] + qR[c]);
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        if (t < me.Ps[side]) {
            double u[4];
            flux_to_side(me, m.mats, side, sU, t, u);
            double *o = m.mustar[me.sub[side]] + ((int64_t)me.slot[side] * me.Ps[side] + t) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = u[c];
        }
    }
}

// ComputeRiemannSolver_InteriorEdge (DGSpatialDiscretizationMethods.f90:545-653) on a straight
// edge: Fi from the interior Riemann solver on the projected traces, Fv the BR1 average of the
// two sides' viscous fluxes from the projected traces and gradients (BR1_RiemannSolver,
// DGViscous_BR1.f90:330-371), FStar = (Fi - Fv) dS, projected to each side.
__global__ __launch_bounds__(64) void k_mortar_flux(MortarArgs m) {
    __shared__ double sF[kEdgeMax * 4];
    const MortarEdge me = m.edges[blockIdx.x];
    const int t = threadIdx.x;
    const Phys<double> ph(m.ph);
    if (t < me.P) {
        double qL[4], qR[4], Fi[4], F[4];
        side_values<4>(me, m.mats, 0, side_trace(m, me, 0, m.tr, 4), t, qL);
        side_values<4>(me, m.mats, 1, side_trace(m, me, 1, m.tr, 4), t, qR);
        riemann<-1, double>(m.ph.riemann, ph, qL, qR, me.nx, me.ny, Fi);
        if (m.ph.ns) {
            double dL[8], dR[8], fL[4], gL[4], fR[4], gR[4];
            side_values<8>(me, m.mats, 0, side_trace(m, me, 0, m.mdq, 8), t, dL);
            side_values<8>(me, m.mats, 1, side_trace(m, me, 1, m.mdq, 8), t, dR);
            viscous_flux(ph, qL, dL, fL, gL);
            viscous_flux(ph, qR, dR, fR, gR);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double FL0 = ph.mu * fL[c], FL1 = ph.mu * gL[c], FR0 = ph.mu * fR[c], FR1 = ph.mu * gR[c];
                double acc = 0.0;
                acc = acc + 0.5 * (FL0 + FR0) * me.nx;
                acc = acc + 0.5 * (FL1 + FR1) * me.ny;
                F[c] = (Fi[c] - acc) * me.dS;
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) F[c] = Fi[c] * me.dS;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) sF[t * 4 + c] = F[c];
    }
    __syncthreads();
    for (int side = 0; side < 2; ++side) {
        if (t < me.Ps[side]) {
            double f[4];
            flux_to_side(me, m.mats, side, sF, t, f);
            double *o = m.mflux[me.sub[side]] + ((int64_t)me.slot[side] * me.Ps[side] + t) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = f[c];
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
    hipLaunchKernelGGL(k_mortar_ustar, dim3(m.nedges), dim3(64), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_mortar_flux(const MortarArgs &m, hipStream_t st) {
    if (m.nedges <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mortar_flux, dim3(m.nedges), dim3(64), 0, st, m);
    return hipGetLastError();
}

hipError_t launch_scatter(const double *src, const int *map, int n, double *dst, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter, dim3((n + 255) / 256), dim3(256), 0, st, src, map, n, dst);
    return hipGetLastError();
}

}  // namespace hg
