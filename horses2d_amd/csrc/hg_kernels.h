// Stage, gradient and trace kernels of the DGSEM residual + low-storage RK for gfx950.
//
// Data layout in HBM (uniform polynomial order P = N+1 per handle):
//   Q, G      : FP64, reference layout  Q[(e*P*P + i + P*j)*4 + eq]
//   traces    : [(e*4 + f)*P + k][4]  element-local face orientation, FP64 or FP32
//   viscous flux traces: [(e*4 + f)*P + k][VT], FP64 or FP32: mu times each side's viscous
//               flux at its face nodes (the gradient itself is never stored: the nodal
//               gradient is recomputed inside each stage)
//   metric    : five planes of e*P*P+n: Ja11, Ja12, Ja21, Ja22, 1/jac (FP64 or FP32; the FP32
//               form stores one record per element when the metric is element-constant); the
//               divisions by the Jacobian (DGSpatialDiscretizationMethods.f90:226-233) are
//               multiplications by its reciprocal (a last-bit difference, not bit-exact)
//
// Tiling: a workgroup owns EPB whole elements (one lane per node, one lane per face node);
// two tile shapes per order are compiled (the widest that fits one wavefront, and the widest
// that fits four) and the faster is chosen per handle at creation.  Every global load a lane
// needs is issued before the first barrier so that its latency overlaps with the other
// lanes' loads; the element state is staged in LDS for the tensor-product contractions.
//
// Each element computes the Riemann fluxes of its own four faces: the neighbor reads the
// same trace values and evaluates the identical expression, so both sides see
// bit-identical fluxes (the reference's FL = F, FR = -F), and the face lifts are subtracted
// in ascending global edge order (DGSpatialDiscretizationMethods.f90:196-220).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "hg_device.h"
#include "hg_physics.h"

namespace hg {

template <typename T>
__device__ inline T ld(const void *p, int64_t i) { return ((const T *)p)[i]; }

__device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// Tile handled by workgroup b of nb.  The dispatcher hands consecutive workgroups to the eight
// XCDs in turn; this gives each XCD a contiguous range of tiles (a band of the mesh), so that
// the traces a tile reads from its neighbours were mostly written through the same XCD's L2.
// A bijection of [0, nb): every tile is processed once, wherever it runs.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int x = b & 7, q = b >> 3, per = nb >> 3, rem = nb & 7;
    return x * per + (x < rem ? x : rem) + q;
}

template <int P, int EPB_>
struct Tile {
    static constexpr int PP = P * P;
    static constexpr int EPB = EPB_;
    static constexpr int NF = EPB * 4 * P;
    static constexpr int NN = EPB * PP;
    static constexpr int TPB = (((NN > NF ? NN : NF) + 63) / 64) * 64;
};

// Tile choices: 0 = widest tile within one wavefront, 1 = widest within four wavefronts whose
// Euler FP64 LDS tile (state, two contravariant fluxes, face fluxes: 96 P^2 + 128 P bytes per
// element) stays within 31 KB, so that five workgroups fit a CU's LDS (this caps P = 4 at 15
// elements instead of 16; the batched kernels' LDS is the largest tile over all orders).
template <int P>
struct Tiles {
    static constexpr int a64 = 64 / (P * P), b64 = 64 / (4 * P);
    static constexpr int a256 = 256 / (P * P), b256 = 256 / (4 * P);
    static constexpr int lds_cap = (31 * 1024) / (96 * P * P + 128 * P);
    static constexpr int small = (a64 < b64 ? a64 : b64) > 0 ? (a64 < b64 ? a64 : b64) : 1;
    static constexpr int wide = (a256 < b256 ? a256 : b256) > 0 ? (a256 < b256 ? a256 : b256) : 1;
    static constexpr int large = (wide < lds_cap ? wide : lds_cap) > 0 ? (wide < lds_cap ? wide : lds_cap) : 1;
};

// ------------------------------------------------------------------ prolongation
template <int P, int EPB, typename TR>
__global__ __launch_bounds__((Tile<P, EPB>::TPB)) void k_prolong(StageArgs a) {
    using T = Tile<P, EPB>;
    __shared__ double sQ[T::NN * 4];
    __shared__ double sLb[2 * P];
    const int t = threadIdx.x, e0 = blockIdx.x * EPB;
    if (t < 2 * P) sLb[t] = a.lb[t];
    const int64_t base = (int64_t)e0 * T::PP * 4;
    const int nval = min(EPB, a.K - e0) * T::PP * 4;
    for (int x = t; x < nval; x += T::TPB) sQ[x] = a.Q[base + x];
    __syncthreads();
    if (t < T::NF) {
        const int s = t / (4 * P), f = (t / P) % 4, k = t % P, e = e0 + s;
        if (e < a.K) {
            double tq[4];
            face_trace<P, 4, 4>(sQ + s * T::PP * 4, sLb, f, k, tq);
            tq[3] = slot_energy<TR>(a.ph, tq[3]);
            TR *o = (TR *)a.tr_out + ((int64_t)(e * 4 + f) * P + k) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = (TR)tq[c];
        }
    }
}

// ------------------------------------------------------------------ gradient helpers
// Both the trace kernel (k_gradient) and the stage kernel evaluate the element gradient with
// these functions, so the gradient a stage recomputes for its own element is the one whose
// face traces its neighbours read.

// The state slot the compute type sees (energy shifted in the FP32 form).
template <typename R>
__device__ __forceinline__ R qslot(const PhysParams &pp, double v, int c) {
    return (R)(c == 3 ? slot_energy<R>(pp, v) : v);
}

// The state a boundary face node's ghost state starts from: its zone's record (8 doubles per
// zone, the state in the first four), or, on a face flagged FF_FSTATE, the node's own value of
// the zone's boundary-state function (slot = (element*4 + face)*P + node).
__device__ __forceinline__ const double *zone_state(const StageArgs &a, int flags, int y, int64_t slot) {
    return (flags & FF_FSTATE) ? a.fstate + slot * 4 : a.zone + 8 * y;
}

// BR1 solution flux at a face node: uStar n dS in both directions (DGViscous_BR1.f90:313-328;
// boundary faces use the ghost state's uSB, DGViscousMethods.f90:786-917), stored as
// out[0..3] = uStar dS and out[4..5] = (nx, ny); the lift forms (uStar dS) nx and (uStar dS) ny.
// zd: the boundary's zone state (zone_state).
template <typename R, bool XB>
__device__ __forceinline__ void br1_face_flux(const StageArgs &a, const Phys<R> &ph, int flags, const double *zd,
                                              const R *qs, const R *qo, R nx, R ny, R dS, R *out) {
    const int kind = flags & 15;
    R us[4];
    if (kind <= FK_PERIODIC) {
#pragma unroll
        for (int c = 0; c < 4; ++c) us[c] = R(0.5) * (qs[c] + qo[c]);
        if (kind == FK_INT_R) {
#pragma unroll
            for (int c = 0; c < 4; ++c) us[c] = -us[c];
        }
    } else if (kind == FK_MORTAR) {
        // projected and signed by the mortar kernel (load_face put it in qo)
        if (flags & FF_MCURVED) {
            // a curved mortar edge: qo holds uStar dS nx already; the y terms follow separately
#pragma unroll
            for (int c = 0; c < 4; ++c) out[c] = qo[c];
            out[4] = R(1); out[5] = R(0);
            return;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) us[c] = qo[c];
    } else {
        R uB[4];
        int neu;
        ghost_state<R, XB>(kind, flags, zd, a.ph, ph, qs, nx, ny, uB, us, neu);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = us[c] * dS;
    out[4] = nx; out[5] = ny;
}

// BR1 gradient at node (i, j): ( -Vector_StdVolumeGreen(Q) + sum_faces Vector_StdFace(uStar) ) / jac
// (DGViscous_BR1.f90:48-111, DGWeakIntegrals.f90:503-622), faces lifted in edge-id order.
// q: element state in the compute type [n*4] (energy slot shifted in the FP32 form),
// Ja: element metric [n*4], Fu: element uStar fluxes [(f*P + k)*6] (see br1_face_flux).
// ROLLED keeps the two l loops rolled (one node's values live; the stage kernel, where the
// registers bound occupancy); the sums are the same either way.
// Faces whose bit 8 + f is set in ord are curved p-non-conforming or subdivided edges, whose
// lift terms come per direction from the mortar kernel (uStar dS n projected direction by
// direction): the face lane stores the x terms with the "normal" (1, 0), and mortar_y_lift adds
// the y terms afterwards.
template <int P, typename R, int QS = 4, int JS = 4, bool ROLLED = false>
__device__ __forceinline__ void br1_node_gradient(const R *q, const R *Ja, const R *sHat, const R *Fu, const R *sLbw,
                                                  int ord, int i, int j, R ijac, R *dq) {
    // Every accumulator receives its xi terms (l = 0..P-1) then its eta terms; each state value
    // is read once and used for both directions, and the metric-weighted derivative entry
    // Ja * hatD is formed once per node for all four variables.
    constexpr int UNROLL = ROLLED ? 1 : P;
    R acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = R(0);
#pragma unroll UNROLL
    for (int l = 0; l < P; ++l) {
        const int m = l + P * j;
        const R h = sHat[i * P + l];
        const R w0 = Ja[m * JS + 0] * h, w2 = Ja[m * JS + 2] * h;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const R v = q[m * QS + c];
            acc[c] += v * w0;
            acc[4 + c] += v * w2;
        }
    }
#pragma unroll UNROLL
    for (int l = 0; l < P; ++l) {
        const int m = i + P * l;
        const R h = sHat[j * P + l];
        const R w1 = Ja[m * JS + 1] * h, w3 = Ja[m * JS + 3] * h;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const R v = q[m * QS + c];
            acc[c] += v * w1;
            acc[4 + c] += v * w3;
        }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) dq[c] = -acc[c];
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const int f = (ord >> (2 * o)) & 3;
        const R *F = Fu + (f * P + ((f == 0 || f == 2) ? i : j)) * 6;
        const R lw = (f == 0) ? sLbw[j] : (f == 1) ? sLbw[P + i] : (f == 2) ? sLbw[P + j] : sLbw[i];
        const R fnx = F[4], fny = F[5];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const R ux = F[c] * fnx, uy = F[c] * fny;
            dq[c] += ux * lw;
            dq[4 + c] += uy * lw;
        }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) dq[c] = dq[c] * ijac;
}

// The y-direction lift terms of an element's curved mortar faces (bit 8 + f of ord), added to
// the gradient br1_node_gradient formed: (uStar dS ny projected) lw / jac from mu8 (the element's
// [(f*P + k)*8] block of StageArgs::mustar).  FP64 only (meshes with mortar edges run in it).
template <int P, typename R>
__device__ __forceinline__ void mortar_y_lift(const double *mu8, const R *sLbw, int ord, int i, int j, R ijac, R *dq) {
    if constexpr (sizeof(R) == 8) {
#pragma unroll 1
        for (int f = 0; f < 4; ++f) {
            if (!((ord >> (8 + f)) & 1)) continue;
            const R lw = (f == 0) ? sLbw[j] : (f == 1) ? sLbw[P + i] : (f == 2) ? sLbw[P + j] : sLbw[i];
            const double *M = mu8 + (f * P + ((f == 0 || f == 2) ? i : j)) * 8;
#pragma unroll
            for (int c = 0; c < 4; ++c) dq[4 + c] += (R)(M[4 + c] * lw) * ijac;
        }
    }
}

// Interior (broken) gradient of the IP method (QuadElement_QuadProcedures.incf:42-88); q as above.
template <int P, typename R, int QS = 4>
__device__ __forceinline__ void ip_node_gradient(const R *q, const R *sD, int i, int j, R j11, R j12, R j21, R j22,
                                                 R ijac, R *dq) {
    // l outermost (each accumulator still sums l = 0..P-1 in order): one node's values live
    R dxi[4] = {R(0), R(0), R(0), R(0)}, deta[4] = {R(0), R(0), R(0), R(0)};
#pragma unroll 1
    for (int l = 0; l < P; ++l) {
        const R di = sD[i * P + l], dj = sD[j * P + l];
        const R *qx = q + (l + P * j) * QS, *qy = q + (i + P * l) * QS;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            dxi[c] += di * qx[c];
            deta[c] += dj * qy[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        dq[c] = (dxi[c] * j11 + deta[c] * j12) * ijac;
        dq[4 + c] = (dxi[c] * j21 + deta[c] * j22) * ijac;
    }
}

// Stages an element tile's state from HBM into LDS: FP64 always, plus the compute-type copy
// (energy slot shifted) that the gradient contractions read when the compute type is FP32.
// The FP64 tile has a node stride of QS doubles (4, or 6 in the gradient kernel).
template <typename R, bool COPY, bool KEEP64 = true, int QS = 4>
__device__ __forceinline__ void load_tile(const StageArgs &a, const int64_t nb0, const int ne, const int PP, double *sQ,
                                          R *sQf, const int t, const int TPB) {
    const double2 *src = (const double2 *)(a.Q + nb0 * 4);
    double2 *dst = (double2 *)sQ;
    for (int x = t; x < ne * PP * 2; x += TPB) {
        const double2 v = src[x];
        if constexpr (KEEP64) dst[(x >> 1) * (QS / 2) + (x & 1)] = v;
        if constexpr (COPY) {
            // pairs (0,1) and (2,3) of a node: the odd pair holds the energy
            ((float2 *)sQf)[x] = make_float2((float)v.x, (float)((x & 1) ? slot_energy<R>(a.ph, v.y) : v.y));
        }
    }
}

// Metric of node g (element e): five planes per node, or per element when the handle stores an
// element-constant metric.
template <typename R>
__device__ __forceinline__ void load_metric(const StageArgs &a, int e, int64_t g, R &j11, R &j12, R &j21, R &j22,
                                            R &ijac) {
    const int64_t i = a.metric_elem ? (int64_t)e : g, st = a.metric_elem ? (int64_t)a.K : a.nodes;
    j11 = ld<R>(a.metric, i); j12 = ld<R>(a.metric, st + i);
    j21 = ld<R>(a.metric, 2 * st + i); j22 = ld<R>(a.metric, 3 * st + i);
    ijac = ld<R>(a.metric, 4 * st + i);
}

// Viscosity of a face: mu + edge % mu_a, the edge's artificial viscosity being the mean of its
// elements' (ComputeEdgeViscosity, DGArtificialDissipation.f90:340-363; boundary and periodic
// faces: the own element's).  Without artificial dissipation, mu.
template <typename R>
__device__ __forceinline__ R face_mu(const StageArgs &a, const Phys<R> &ph, int fe, int4 fr) {
    if (!a.mu_a) return ph.mu;
    const double ma = (fr.x & 15) <= FK_INT_R ? (a.mu_a[fe] + a.mu_a[fr.y >> 2]) / 2.0 : a.mu_a[fe];
    return (R)(a.ph.mu + ma);
}

// Edge viscosity of the Laplacian dissipation (JumpsBasedEdgeViscosity,
// DGArtificialDissipation.f90:340-363): the mean of the two elements' (boundary and periodic
// faces: the own element's).
__device__ __forceinline__ double face_lap(const StageArgs &a, int fe, int4 fr) {
    return (fr.x & 15) <= FK_INT_R ? (a.lap[fe] + a.lap[fr.y >> 2]) / 2.0 : a.lap[fe] / 1.0;
}

// Face-node geometry and the neighbour's Q trace (periodic and interior faces).
template <int P, typename R, typename TR>
__device__ __forceinline__ void load_face(const StageArgs &a, int fe, int ff, int fk, int4 &fr, R &nx, R &ny, R &dS,
                                          R *qo, int64_t &nslot) {
    fr = a.face[fe * 4 + ff];
    const bool curved = fr.x & FF_CURVED, selfrev = fr.x & FF_SELFREV, rev = fr.x & FF_REV;
    const int gi = curved ? (selfrev ? P - 1 - fk : fk) : 0;
    const double *geo = a.egeo + 3 * ((int64_t)fr.z + gi);
    nx = (R)geo[0]; ny = (R)geo[1]; dS = (R)geo[2];
    nslot = (int64_t)fr.y * P + (rev ? P - 1 - fk : fk);
    if ((fr.x & 15) <= FK_PERIODIC) {
        const TR *nbt = (const TR *)a.tr_in + nslot * 4;
#pragma unroll
        for (int c = 0; c < 4; ++c) qo[c] = (R)nbt[c];
    } else if ((fr.x & 15) == FK_MORTAR && a.mustar) {
        // p-non-conforming edge: this side's projected uStar in place of a neighbour trace
        const double *u = a.mustar + ((int64_t)(fe * 4 + ff) * P + fk) * 8;
#pragma unroll
        for (int c = 0; c < 4; ++c) qo[c] = (R)u[c];
    }
}

// mu (f, g) of one side of a face node from its viscous flux trace (layout of VT).
template <typename R>
__device__ __forceinline__ void viscous_trace_flux(const R *v, R *f, R *g) {
    f[0] = R(0); f[1] = v[0]; f[2] = v[1]; f[3] = v[2];
    g[0] = R(0); g[1] = v[1]; g[2] = v[3]; g[3] = v[4];
}

// Own Q trace of a face node as the previous stage (or the prolongation) stored it: the values
// the neighbour reads, in the compute type.
template <int P, typename R, typename TR>
__device__ __forceinline__ void stored_trace(const StageArgs &a, int fe, int ff, int fk, R *qs) {
    const TR *o = (const TR *)a.tr_in + ((int64_t)(fe * 4 + ff) * P + fk) * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) qs[c] = (R)o[c];
}

// Own Q trace of a face node from the LDS tile, in the compute type, rounded like the stored
// traces the neighbour reads.
template <int P, typename R, typename TR>
__device__ __forceinline__ void own_trace(const StageArgs &a, const double *sQe, const double *sLb, int ff, int fk,
                                          R *qs) {
    double tq[4];
    face_trace<P, 4, 4>(sQe, sLb, ff, fk, tq);
    tq[3] = slot_energy<TR>(a.ph, tq[3]);
#pragma unroll
    for (int c = 0; c < 4; ++c) qs[c] = (R)(TR)tq[c];
}

// This side's viscous flux at a face node from its Q trace qs and the element's nodal gradient
// tile (node stride DS), times mu, stored in the VT layout (ViscousFluxes.f90:39-80, adiabatic
// walls :279-318).  The gradient trace is accumulated in FP64 in the FP64 form and in FP32 in
// the FP32 form (lbR: lb in the compute type).
// mdq: where a p-non-conforming face keeps its gradient trace for the mortar kernel (or null).
template <int P, int DS, typename R, typename TR>
__device__ __forceinline__ void viscous_face_trace(const StageArgs &a, const Phys<R> &ph, int flags, const R *qs,
                                                   const R *sDQe, const double *sLb, const R *lbR, int ff, int fk,
                                                   TR *o, R muf, double *mdq = nullptr) {
    R ds[8], fv[4], gv[4];
    if constexpr (sizeof(R) == 4) {
        face_trace_r<P, DS, 8>(sDQe, lbR, ff, fk, ds);
    } else {
        double tg[8];
        face_trace<P, DS, 8>(sDQe, sLb, ff, fk, tg);
#pragma unroll
        for (int c = 0; c < 8; ++c) ds[c] = (R)(TR)tg[c];
    }
    if (mdq) {
#pragma unroll
        for (int c = 0; c < 8; ++c) mdq[c] = (double)ds[c];
    }
    if (is_adiabatic(flags)) adiabatic_flux(ph, qs, ds, fv, gv);
    else viscous_flux(ph, qs, ds, fv, gv);
    o[0] = (TR)(muf * fv[1]); o[1] = (TR)(muf * fv[2]); o[2] = (TR)(muf * fv[3]);
    o[3] = (TR)(muf * gv[2]); o[4] = (TR)(muf * gv[3]); o[5] = TR(0);
}

// ------------------------------------------------------------------ gradient trace kernel
// Evaluates the element gradient (BR1 or IP), its traces on the element's four faces and, from
// them and the Q traces, this side's viscous flux at every face node, which both elements of
// the face read in their stage kernels.  The gradient at the nodes is not stored: each stage
// kernel recomputes it for its own element.
// Node strides (in elements of the tile's type) of the gradient kernel's LDS tiles, chosen so
// that the P nodes of a row or a column of an element fall in distinct LDS banks for 16-byte
// reads: FP64 state and metric 48 B, FP64 gradient 80 B; the FP32 state copy and metric
// (16 B per node) need no padding, the FP32 gradient is padded to 48 B.
template <typename R>
struct GradStrides {
    static constexpr int QS = sizeof(R) == 8 ? 6 : 4;    // state (FP64 tile or FP32 copy)
    static constexpr int JS = sizeof(R) == 8 ? 6 : 4;    // metric
    static constexpr int DS = sizeof(R) == 8 ? 10 : 12;  // gradient
};

template <int P, int EPB, bool BR1, typename R>
struct GradSmem {
    using T = Tile<P, EPB>;
    using ST = GradStrides<R>;
    struct Inputs {
        alignas(16) double sQ[sizeof(R) == 4 ? 2 : T::NN * ST::QS];   // FP64 form: the state
        alignas(16) R sQf[sizeof(R) == 4 ? T::NN * ST::QS : 1];        // FP32 form: the compute-type state
        alignas(16) R sJa[BR1 ? T::NN * ST::JS : 1];
    };
    // the state and the metric are dead once every node's gradient is formed: the gradient
    // tile takes their place
    union {
        Inputs in;
        alignas(16) R sDQ[T::NN * ST::DS];
    } u;
    alignas(16) double sLb[2 * P];
    alignas(16) R sFu[BR1 ? T::NF * 6 : 1];
    alignas(16) R sHat[BR1 ? T::PP : 1];
    alignas(16) R sLbw[4 * P];   // lbw, then lb in the compute type
    alignas(16) R sD[BR1 ? 1 : T::PP];
};

// One element tile [e0, e0+EPB) of the gradient; any workgroup size >= Tile<P, EPB>::TPB.
// XB: the Euler-wall, Dirichlet and pressure boundary kinds are compiled in.
template <int P, int EPB, bool BR1, typename R, typename TR, bool XB = false>
__device__ __forceinline__ void gradient_body(const StageArgs &a, const int e0, GradSmem<P, EPB, BR1, R> &S) {
    using T = Tile<P, EPB>;
    constexpr int PP = T::PP, NN = T::NN, NF = T::NF;
    const int TPB = blockDim.x;
    double *sQ = S.u.in.sQ, *sLb = S.sLb;
    R *sDQ = S.u.sDQ, *sJa = S.u.in.sJa, *sFu = S.sFu, *sHat = S.sHat, *sLbw = S.sLbw, *sD = S.sD;
    const int t = threadIdx.x;
    const Phys<R> ph(a.ph);
    if (t < 2 * P) { sLb[t] = a.lb[t]; sLbw[t] = (R)a.lbw[t]; sLbw[2 * P + t] = (R)a.lb[t]; }
    for (int x = t; x < PP; x += TPB) {
        if (BR1) sHat[x] = (R)a.hatD[x];
        else sD[x] = (R)a.D[x];
    }
    const int ne = min(EPB, a.K - e0);
    const int64_t nb0 = (int64_t)e0 * PP;
    constexpr bool COPY = sizeof(R) == 4;
    using ST = GradStrides<R>;
    constexpr int QS = ST::QS, JS = ST::JS, DS = ST::DS;
    load_tile<R, COPY, !COPY, QS>(a, nb0, ne, PP, sQ, S.u.in.sQf, t, TPB);
    const R *sQr;
    if constexpr (COPY) sQr = S.u.in.sQf; else sQr = (const R *)sQ;
    // ---- hoisted global loads
    const int s = t / PP, n = t % PP, e = e0 + s;
    const bool node_active = (t < NN) && (e < a.K);
    R ijac = R(1), j11 = R(0), j12 = R(0), j21 = R(0), j22 = R(0);
    int ord = 0;
    if (node_active) {
        const int64_t g = nb0 + s * PP + n;
        load_metric<R>(a, e, g, j11, j12, j21, j22, ijac);
        if (BR1) {
            ord = a.forder[e];
            sJa[t * JS + 0] = j11; sJa[t * JS + 1] = j12; sJa[t * JS + 2] = j21; sJa[t * JS + 3] = j22;
        }
    }
    const int fs = t / (4 * P), ff = (t / P) % 4, fk = t % P, fe = e0 + fs;
    const bool face_active = (t < NF) && (fe < a.K);
    int4 fr = make_int4(0, 0, 0, 0);
    R nx = R(0), ny = R(0), dS = R(0), qo[4] = {R(0), R(0), R(0), R(0)};
    int64_t nslot = 0;
    R qs[4] = {R(0), R(0), R(0), R(0)};
    if (face_active) {
        if (BR1) load_face<P, R, TR>(a, fe, ff, fk, fr, nx, ny, dS, qo, nslot);
        else fr = a.face[fe * 4 + ff];
        stored_trace<P, R, TR>(a, fe, ff, fk, qs);
    }
    __syncthreads();
    if (BR1 && face_active)
        br1_face_flux<R, XB>(a, ph, fr.x, zone_state(a, fr.x, fr.y, (int64_t)(fe * 4 + ff) * P + fk), qs, qo, nx, ny,
                             dS, sFu + t * 6);
    if (BR1) __syncthreads();
    R dq[8];
    if (node_active) {
        const int i = n % P, j = n / P;
        if (BR1) {
            br1_node_gradient<P, R, QS, JS>(sQr + s * PP * QS, sJa + s * PP * JS, sHat, sFu + s * 4 * P * 6, sLbw, ord, i,
                                            j, ijac, dq);
            if (ord >> 8) mortar_y_lift<P, R>(a.mustar + (int64_t)e * 4 * P * 8, sLbw, ord, i, j, ijac, dq);
        }
        else ip_node_gradient<P, R, QS>(sQr + s * PP * QS, sD, i, j, j11, j12, j21, j22, ijac, dq);
    }
    __syncthreads();   // every lane is done with the state and the metric
    if (node_active) {
#pragma unroll
        for (int c = 0; c < 8; ++c) sDQ[t * DS + c] = dq[c];
    }
    __syncthreads();
    if (face_active) {
        const R muf = XB ? face_mu<R>(a, ph, fe, fr) : ph.mu;
        double *mdq = ((fr.x & 15) == FK_MORTAR && a.mdq) ? a.mdq + ((int64_t)(fe * 4 + ff) * P + fk) * 8 : nullptr;
        viscous_face_trace<P, DS, R, TR>(a, ph, fr.x, qs, sDQ + fs * PP * DS, sLb, sLbw + 2 * P, ff, fk,
                                         (TR *)a.dtr_out + ((int64_t)(fe * 4 + ff) * P + fk) * VT, muf, mdq);
    }
}

template <int P, int EPB, bool BR1, typename R, typename TR, bool XB = false>
__global__ __launch_bounds__((Tile<P, EPB>::TPB)) void k_gradient(StageArgs a) {
    __shared__ GradSmem<P, EPB, BR1, R> S;
    gradient_body<P, EPB, BR1, R, TR, XB>(a, xcd_tile(blockIdx.x, gridDim.x) * EPB, S);
}

// ------------------------------------------------------------------ the stage kernel
template <int P, int EPB, bool NS, bool IP, typename R>
struct StageSmem {
    using T = Tile<P, EPB>;
    alignas(16) double sQ[T::NN * 4];
    alignas(16) double sLb[2 * P];
    alignas(16) double sRes[4][4];
    struct Fluxes {
        alignas(16) R sCF[T::NN * 4];          // BR1: the metric Ja during the gradient phase
        alignas(16) R sCG[T::NN * 4];          // NS, FP32: the compute-type state until the node phase
    };
    // IP: the epilogue's gradient tile takes the place of the contravariant fluxes
    union {
        Fluxes f;
        alignas(16) R sDQ[IP ? T::NN * 8 : 1];
    } u;
    alignas(16) R sF[NS && !IP ? T::NF * 6 : T::NF * 4];  // BR1: uStar fluxes during the gradient phase
    // IP gradient-face terms: the face phase's contravariant fluxes, then (same arrays) the
    // U / V lift coefficients
    alignas(16) R sU[IP ? T::NF * 4 : 1];
    alignas(16) R sV[IP ? T::NF * 4 : 1];
    alignas(16) R sHat[T::PP];
    alignas(16) R sD[NS ? T::PP : 1];   // IP gradient; BR1: the Laplacian dissipation's interior gradient
    alignas(16) R sLbw[2 * P];
    alignas(16) R sDlbw[2 * P];
    alignas(16) R sLbR[IP ? 2 * P : 1];   // lb in the compute type (IP epilogue, FP32 traces)
};

// One element tile [e0, e0+EPB) of a stage; any workgroup size >= Tile<P, EPB>::TPB (at most
// 256).  rslot is the residual partial this tile writes; cq, cg are the stage's RK factors.
// Navier-Stokes: the element's nodal gradient is recomputed here from the LDS tile and the Q
// traces (it stays in registers); the viscous fluxes of both sides of a face are read from
// what the gradient kernel stored.  The own Q traces are read from the stored traces as well.
// Interior penalty: the gradient is element-local, so the RK update's epilogue also forms the
// next stage's viscous face-flux traces from the updated tile (the work of k_gradient, with
// the same device functions and the same bits), and no gradient kernel runs between stages.
// XBF: the Euler-wall, Dirichlet and pressure boundary kinds and artificial dissipation are
// compiled in (always with the run-time solver choice, RS = -1).
template <int P, int EPB, bool NS, bool IP, int RS, typename R, typename TR, bool XBF = false>
__device__ __forceinline__ void stage_body(const StageArgs &a, const int e0, const int64_t rslot, const double cq,
                                           const double cg, StageSmem<P, EPB, NS, IP, R> &S) {
    using T = Tile<P, EPB>;
    constexpr bool XB = RS < 0 || XBF;
    constexpr int PP = T::PP, NN = T::NN, NF = T::NF;
    constexpr bool BR1 = NS && !IP;
    const int TPB = blockDim.x;
    double *sQ = S.sQ, *sLb = S.sLb;
    R *sCF = S.u.f.sCF, *sCG = S.u.f.sCG, *sF = S.sF, *sU = S.sU, *sV = S.sV;
    R *sHat = S.sHat, *sD = S.sD, *sLbw = S.sLbw, *sDlbw = S.sDlbw;
    const int t = threadIdx.x;
    const Phys<R> ph(a.ph);
    const bool ipgrad = IP && a.ph.ip_gradflux;
    if (t < 2 * P) {
        sLb[t] = a.lb[t]; sLbw[t] = (R)a.lbw[t]; sDlbw[t] = (R)a.dlbw[t];
        if (IP) S.sLbR[t] = (R)a.lb[t];
    }
    for (int x = t; x < PP; x += TPB) {
        sHat[x] = (R)a.hatD[x];
        if (IP || (BR1 && XB)) sD[x] = (R)a.D[x];
    }
    const int ne = min(EPB, a.K - e0);
    const int64_t nb0 = (int64_t)e0 * PP;
    // NS in FP32: a compute-type copy of the state for the gradient contractions, in sCG (each
    // node lane reads its own entry once more before the node phase overwrites it)
    constexpr bool COPY = NS && sizeof(R) == 4;
    load_tile<R, COPY>(a, nb0, ne, PP, sQ, sCG, t, TPB);
    const R *sQr;
    if constexpr (COPY) sQr = sCG; else sQr = (const R *)sQ;

    // ---- hoisted global loads: node lane
    const int s = t / PP, n = t % PP, e = e0 + s;
    const bool node_active = (t < NN) && (e < a.K);
    const int64_t g = nb0 + s * PP + n;
    R j11 = R(0), j12 = R(0), j21 = R(0), j22 = R(0), ijac = R(1);
    int ord = 0;
    if (node_active) {
        load_metric<R>(a, e, g, j11, j12, j21, j22, ijac);
        ord = a.forder[e];
        if (BR1) {
            sCF[t * 4 + 0] = j11; sCF[t * 4 + 1] = j12; sCF[t * 4 + 2] = j21; sCF[t * 4 + 3] = j22;
        }
    }
    // ---- hoisted global loads: face lane
    const int fs = t / (4 * P), ff = (t / P) % 4, fk = t % P, fe = e0 + fs;
    const bool face_active = (t < NF) && (fe < a.K);
    int4 fr = make_int4(0, 0, 0, 0);
    R nx = R(0), ny = R(0), dS = R(0), qo[4] = {R(0), R(0), R(0), R(0)};
    int64_t nslot = 0;
    R qs[4] = {R(0), R(0), R(0), R(0)};
    if (face_active) {
        load_face<P, R, TR>(a, fe, ff, fk, fr, nx, ny, dS, qo, nslot);
        stored_trace<P, R, TR>(a, fe, ff, fk, qs);
    }
    __syncthreads();

    if (BR1) {
        if (face_active)
            br1_face_flux<R, XB>(a, ph, fr.x, zone_state(a, fr.x, fr.y, (int64_t)(fe * 4 + ff) * P + fk), qs, qo, nx,
                                 ny, dS, sF + t * 6);
        __syncthreads();
    }

    // ---- own nodal gradient (Navier-Stokes), kept in registers for the node phase
    R dq[NS ? 8 : 1];
    // BR1 with the Laplacian dissipation: the element's interior gradient as well (its volume
    // flux, DGArtificialDissipation.f90:207-222); interior penalty: that is dq itself
    constexpr bool LAPG = BR1 && XB;
    R dqi[LAPG ? 8 : 1];
    if (NS && node_active) {
        const int i = n % P, j = n / P;
        if (BR1) {
            br1_node_gradient<P, R, 4, 4, true>(sQr + s * PP * 4, sCF + s * PP * 4, sHat, sF + s * 4 * P * 6, sLbw, ord,
                                                i, j, ijac, dq);
            if (ord >> 8) mortar_y_lift<P, R>(a.mustar + (int64_t)e * 4 * P * 8, sLbw, ord, i, j, ijac, dq);
        }
        else ip_node_gradient<P, R>(sQr + s * PP * 4, sD, i, j, j11, j12, j21, j22, ijac, dq);
        if (LAPG && a.lap) ip_node_gradient<P, R>(sQr + s * PP * 4, sD, i, j, j11, j12, j21, j22, ijac, dqi);
    }
    // the neighbour's viscous flux trace, issued before the barrier and consumed in the face phase
    R vn[NS ? VT : 1];
    R invh = R(0);
    if (NS && face_active) {
        if ((fr.x & 15) <= FK_PERIODIC) {
#pragma unroll
            for (int c = 0; c < (NS ? VT : 1); ++c) vn[c] = (R)((const TR *)a.dtr)[nslot * VT + c];
            if (XB && a.mu_a && (fr.x & 15) == FK_PERIODIC) {
                // a periodic face is a boundary edge of each side, whose viscosity (its own
                // element's mu_a) both sides' fluxes take; the partner stored its trace with its own
                const double ms = a.ph.mu + a.mu_a[fe], mp = a.ph.mu + a.mu_a[fr.y >> 2];
                const R sc = (R)(ms / mp);
#pragma unroll
                for (int c = 0; c < (NS ? VT : 1); ++c) vn[c] = vn[c] * sc;
            }
        }
        if (IP) invh = (R)a.einvh[fr.w];
    }
    // the node phase overwrites sCF (the metric, BR1) and sCG (the FP32 state copy)
    if (BR1 || COPY) __syncthreads();

    // ---- node phase: contravariant fluxes (DGSpatial_QDotVolumeLoop, DGWeakIntegrals.f90:111-171)
    if (node_active) {
        R q[4], f[4], gg[4];
        if constexpr (COPY) {
#pragma unroll
            for (int c = 0; c < 4; ++c) q[c] = sQr[t * 4 + c];
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) q[c] = (R)sQ[t * 4 + c];
            q[3] = (R)slot_energy<R>(a.ph, sQ[t * 4 + 3]);
        }
        inviscid_flux(ph, q, f, gg);
        if (NS) {
            R fv[4], gv[4];
            viscous_flux(ph, q, dq, fv, gv);
            // e % mu_a (DGViscous_BR1.f90:307, DGViscous_IP.f90:84)
            const R mue = (XB && a.mu_a) ? (R)(a.ph.mu + a.mu_a[e]) : ph.mu;
            // Laplacian dissipation: F = Fi - Fv - Fa, Fa = mu_e times the interior gradient
            R fa[4] = {R(0), R(0), R(0), R(0)}, ga[4] = {R(0), R(0), R(0), R(0)};
            if (XB && a.lap) {
                const R lm = (R)a.lap[e];
                const R *di = LAPG ? dqi : dq;
#pragma unroll
                for (int c = 0; c < 4; ++c) { fa[c] = lm * di[c]; ga[c] = lm * di[4 + c]; }
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) { f[c] = (f[c] - mue * fv[c]) - fa[c]; gg[c] = (gg[c] - mue * gv[c]) - ga[c]; }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            sCF[t * 4 + c] = f[c] * j11 + gg[c] * j21;
            sCG[t * 4 + c] = f[c] * j12 + gg[c] * j22;
        }
    }

    // ---- face phase: Riemann fluxes of the element's own four faces
    if (face_active) {
        const int kind = fr.x & 15;
        if (kind == FK_MORTAR) {
            // p-non-conforming edge: (Fi - Fv) dS projected to this side by the mortar kernel
            const double *mf = a.mflux + ((int64_t)(fe * 4 + ff) * P + fk) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) sF[t * 4 + c] = (R)mf[c];
            if (ipgrad) {
                // this side's IP gradient flux, projected and times dS by the mortar kernel, in
                // contravariant components (DGWeakIntegrals.f90:268-495)
                const double *mg = a.mgrad + ((int64_t)(fe * 4 + ff) * P + fk) * 8;
                const R *fm = (const R *)a.fmet + ((int64_t)(fe * 4 + ff) * P + fk) * 5;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const R gx = (R)mg[c], gy = (R)mg[4 + c];
                    sU[t * 4 + c] = (gx * fm[0] + gy * fm[1]) * fm[4];
                    sV[t * 4 + c] = (gx * fm[2] + gy * fm[3]) * fm[4];
                }
            }
        } else {
            R vs[NS ? VT : 1];   // own viscous flux trace
            if (NS) {
                const TR *vt = (const TR *)a.dtr + ((int64_t)(fe * 4 + ff) * P + fk) * VT;
#pragma unroll
                for (int c = 0; c < (NS ? VT : 1); ++c) vs[c] = (R)vt[c];
            }
            int rs = a.ph.riemann;
            R uSB[4] = {R(0), R(0), R(0), R(0)};
            int neumann = 0, prescribed = 0;
            if (kind > FK_PERIODIC) {
                prescribed = ghost_state<R, XB>(kind, fr.x, zone_state(a, fr.x, fr.y, (int64_t)(fe * 4 + ff) * P + fk),
                                                a.ph, ph, qs, nx, ny, qo, uSB, neumann);
                const int zr = (fr.x >> FF_RS_SHIFT) & 15;
                if (zr) rs = zr - 1;
            }
            R Fi[4];
            if (XB && prescribed) {
                // WEAK_PRESCRIBED (DGSpatialDiscretizationMethods.f90:1194-1198): the inviscid
                // normal flux of the boundary's state
                R fp[4], gp[4];
                inviscid_flux(ph, qo, fp, gp);
#pragma unroll
                for (int c = 0; c < 4; ++c) Fi[c] = fp[c] * nx + gp[c] * ny;
            } else if (kind == FK_INT_R) riemann<RS, R>(rs, ph, qo, qs, nx, ny, Fi);
            else riemann<RS, R>(rs, ph, qs, qo, nx, ny, Fi);
            R F[4];
            if (!NS) {
#pragma unroll
                for (int c = 0; c < 4; ++c) F[c] = Fi[c] * dS;
            } else {
                R Fv[4] = {R(0), R(0), R(0), R(0)};
                R Fa[4] = {R(0), R(0), R(0), R(0)};   // Laplacian dissipation: -mu Area (QL - QR)
                R G8[8] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
                const R muf = XB ? face_mu<R>(a, ph, fe, fr) : ph.mu;
                const R lapA = (XB && a.lap) ? (R)(face_lap(a, fe, fr) * a.earea[fr.w]) : R(0);
                const R pen = IP ? ph.sigma0 * muf * (R)(P - 1) * (R)(P - 1) * invh : R(0);
                if (kind <= FK_PERIODIC) {
                    const R *uL = (kind == FK_INT_R) ? qo : qs, *uR = (kind == FK_INT_R) ? qs : qo;
                    const R *vL = (kind == FK_INT_R) ? vn : vs, *vR = (kind == FK_INT_R) ? vs : vn;
                    R fL[4], gL[4], fR[4], gR[4];
                    viscous_trace_flux(vL, fL, gL);
                    viscous_trace_flux(vR, fR, gR);
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        R acc = R(0);
                        acc = acc + R(0.5) * (fL[c] + fR[c]) * nx;
                        acc = acc + R(0.5) * (gL[c] + gR[c]) * ny;
                        Fv[c] = acc;
                    }
                    if (IP) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) Fv[c] = Fv[c] - pen * (uL[c] - uR[c]);
                    }
                    if (XB && a.lap) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) Fa[c] = -(lapA * (uL[c] - uR[c]));
                    }
                    if (ipgrad) {
                        // IP_GradientRiemannSolver (DGViscous_IP.f90:216-249): this side's share
                        R fgr[8], f4[4], g4[4];
#pragma unroll
                        for (int c = 0; c < 4; ++c) { fgr[c] = (uL[c] - uR[c]) * nx; fgr[4 + c] = (uL[c] - uR[c]) * ny; }
                        const R *own = (kind == FK_INT_R) ? uR : uL;
                        viscous_flux(ph, own, fgr, f4, g4);
                        const R cf = (kind == FK_INT_R ? -R(0.5) : R(0.5)) * ph.ip_eps * muf;
#pragma unroll
                        for (int c = 0; c < 4; ++c) { G8[c] = cf * f4[c] * dS; G8[4 + c] = cf * g4[c] * dS; }
                    }
                } else if (!neumann) {
                    const bool adiab = is_adiabatic(fr.x);
                    R fv4[4], gv4[4];
                    viscous_trace_flux(vs, fv4, gv4);   // adiabatic or not, as the gradient kernel chose
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        R acc = R(0);
                        acc = acc + fv4[c] * nx;
                        acc = acc + gv4[c] * ny;
                        Fv[c] = acc;
                    }
                    if (IP) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) Fv[c] = Fv[c] - pen * (qs[c] - uSB[c]);
                    }
                    if (XB && a.lap) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) Fa[c] = -(lapA * (qs[c] - uSB[c]));
                    }
                    if (ipgrad) {
                        R fgr[8], f4[4], g4[4];
#pragma unroll
                        for (int c = 0; c < 4; ++c) { fgr[c] = (qs[c] - uSB[c]) * nx; fgr[4 + c] = (qs[c] - uSB[c]) * ny; }
                        if (adiab) adiabatic_flux(ph, qs, fgr, f4, g4);
                        else viscous_flux(ph, qs, fgr, f4, g4);
                        const R cf = ph.ip_eps * muf;
#pragma unroll
                        for (int c = 0; c < 4; ++c) { G8[c] = cf * f4[c] * dS; G8[4 + c] = cf * g4[c] * dS; }
                    }
                }
                if (kind <= FK_INT_R) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) F[c] = ((Fi[c] - Fv[c]) - Fa[c]) * dS;
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c) F[c] = ((Fi[c] * dS - Fv[c] * dS) - Fa[c] * dS);
                }
                if (ipgrad) {
                    // contravariant gradient-flux components at the face (DGWeakIntegrals.f90:268-495)
                    const R *fm = (const R *)a.fmet + ((int64_t)(fe * 4 + ff) * P + fk) * 5;
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        sU[t * 4 + c] = (G8[c] * fm[0] + G8[4 + c] * fm[1]) * fm[4];   // fm[4] = 1/jac at the face node
                        sV[t * 4 + c] = (G8[c] * fm[2] + G8[4 + c] * fm[3]) * fm[4];
                    }
                }
            }
            if (kind == FK_INT_R) {
#pragma unroll
                for (int c = 0; c < 4; ++c) F[c] = -F[c];
            }
#pragma unroll
            for (int c = 0; c < 4; ++c) sF[t * 4 + c] = F[c];
        }
    }
    __syncthreads();

    // ---- IP gradient face terms: U (lbw coefficient) and V (dlbw coefficient) per face node;
    // horizontal faces: U = hatD . Fc, V = Gc ; vertical faces: U = hatD . Gc, V = Fc
    if (ipgrad) {
        R u2[4], v2[4];
        if (face_active) {
            const R *Fc = sU + (fs * 4 + ff) * P * 4, *Gc = sV + (fs * 4 + ff) * P * 4;
            const bool horiz = (ff == 0 || ff == 2);
            const R *src = horiz ? Fc : Gc;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                R acc = R(0);
#pragma unroll
                for (int l = 0; l < P; ++l) acc += sHat[fk * P + l] * src[l * 4 + c];
                u2[c] = acc;
                v2[c] = horiz ? Gc[fk * 4 + c] : Fc[fk * 4 + c];
            }
        }
        __syncthreads();   // every face has read its fluxes: U and V take their place
        if (face_active) {
#pragma unroll
            for (int c = 0; c < 4; ++c) { sU[t * 4 + c] = u2[c]; sV[t * 4 + c] = v2[c]; }
        }
        __syncthreads();
    }

    // ---- node phase 2: volume contraction, ordered face lifts, /jac, RK update
    double qd[4] = {0.0, 0.0, 0.0, 0.0};
    double gold[4] = {0.0, 0.0, 0.0, 0.0};
    if (node_active && a.mode == 0 && !a.first_stage) {
        const double2 *gp = (const double2 *)(a.G + g * 4);
        const double2 g01 = gp[0], g23 = gp[1];
        gold[0] = g01.x; gold[1] = g01.y; gold[2] = g23.x; gold[3] = g23.y;
    }
    if (node_active) {
        const int i = n % P, j = n / P;
        R acc[4] = {R(0), R(0), R(0), R(0)};
        const R *cF = sCF + s * PP * 4, *cG = sCG + s * PP * 4;
#pragma unroll 1
        for (int l = 0; l < P; ++l) {
            const R hi = sHat[i * P + l], hj = sHat[j * P + l];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = acc[c] + hi * cF[(l + P * j) * 4 + c] + hj * cG[(i + P * l) * 4 + c];
        }
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
                const R *Uf = sU + ((s * 4 + f) * P + idx) * 4, *Vf = sV + ((s * 4 + f) * P + idx) * 4;
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] = acc[c] - (Uf[c] * lw + Vf[c] * dlw);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) qd[c] = (double)(acc[c] * ijac);   // QDot / jac
        if (a.mode == 1) {
            double2 *o = (double2 *)(a.QDot + g * 4);
            o[0] = make_double2(qd[0], qd[1]);
            o[1] = make_double2(qd[2], qd[3]);
        } else {
            double gv[4];
            if (a.first_stage) {
#pragma unroll
                for (int c = 0; c < 4; ++c) gv[c] = cq * qd[c];
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) gv[c] = fma(a.am, gold[c], cq * qd[c]);
            }
            double2 *go = (double2 *)(a.G + g * 4);
            go[0] = make_double2(gv[0], gv[1]);
            go[1] = make_double2(gv[2], gv[3]);
            double qn[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) qn[c] = fma(cg, gv[c], sQ[t * 4 + c]);
            double2 *qo2 = (double2 *)(a.Q + g * 4);
            qo2[0] = make_double2(qn[0], qn[1]);
            qo2[1] = make_double2(qn[2], qn[3]);
#pragma unroll
            for (int c = 0; c < 4; ++c) sQ[t * 4 + c] = qn[c];
        }
    }
    if (a.do_resid) {
        // max|QDot| per equation: wavefront shuffle, then across the workgroup's wavefronts in
        // LDS; one partial per workgroup (reduced by k_resid_final).  NaN wins every compare.
        double(*sRes)[4] = S.sRes;
        double m[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double v = node_active ? fabs(qd[c]) : 0.0;
            for (int off = 32; off > 0; off >>= 1) v = nan_max(v, __shfl_xor(v, off, 64));
            m[c] = v;
        }
        if ((t & 63) == 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) sRes[t >> 6][c] = m[c];
        }
        __syncthreads();
        if (t < 4) {
            double v = sRes[0][t];
            for (int w = 1; w < TPB / 64; ++w) v = nan_max(v, sRes[w][t]);
            a.resid_part[rslot * 4 + t] = v;
        }
    }
    if (a.mode == 0) {
        __syncthreads();
        R qn[4];   // the next stage's own Q trace in the compute type (IP epilogue)
        if (face_active) {
            double tq[4];
            face_trace<P, 4, 4>(sQ + fs * PP * 4, sLb, ff, fk, tq);
            tq[3] = slot_energy<TR>(a.ph, tq[3]);
            TR *o = (TR *)a.tr_out + ((int64_t)(fe * 4 + ff) * P + fk) * 4;
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = (TR)tq[c];
#pragma unroll
            for (int c = 0; c < 4; ++c) qn[c] = (R)(TR)tq[c];
        }
        if constexpr (IP) {
          if (!XB || a.ip_epi) {   // uniform: the launch's flag
            // the updated state in the compute type (FP32: a fresh copy, the node phase
            // overwrote the first one), the broken gradient, its traces and this side's
            // viscous flux: what k_gradient evaluates for the updated state
            const R *sQn;
            if constexpr (COPY) {
                if (node_active) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) sCG[t * 4 + c] = qslot<R>(a.ph, sQ[t * 4 + c], c);
                }
                __syncthreads();
                sQn = sCG;
            } else {
                sQn = (const R *)sQ;
            }
            R dq[8];
            if (node_active) {
                // the metric again from memory: keeping it live from the node phase costs the
                // FP64 kernel a wavefront per SIMD
                R m11, m12, m21, m22, mij;
                load_metric<R>(a, e, g, m11, m12, m21, m22, mij);
                ip_node_gradient<P, R>(sQn + s * PP * 4, sD, n % P, n / P, m11, m12, m21, m22, mij, dq);
            }
            __syncthreads();   // the state copy and the fluxes' space are free
            R *sDQ = S.u.sDQ;
            if (node_active) {
#pragma unroll
                for (int c = 0; c < 8; ++c) sDQ[t * 8 + c] = dq[c];
            }
            __syncthreads();
            if (face_active) {
                double *mdq = ((fr.x & 15) == FK_MORTAR && a.mdq) ? a.mdq + ((int64_t)(fe * 4 + ff) * P + fk) * 8 : nullptr;
                viscous_face_trace<P, 8, R, TR>(a, ph, fr.x, qn, sDQ + fs * PP * 8, sLb, S.sLbR, ff, fk,
                                                (TR *)a.dtr_out + ((int64_t)(fe * 4 + ff) * P + fk) * VT, ph.mu, mdq);
            }
          }
        }
    }
}

template <int P, int EPB, bool NS, bool IP, int RS, typename R, typename TR, bool XBF = false>
__global__ __launch_bounds__((Tile<P, EPB>::TPB)) void k_stage(StageArgs a) {
    __shared__ StageSmem<P, EPB, NS, IP, R> S;
    double cq = a.cq, cg = a.cg;
    if (a.dt_dev) {   // graph replay: the step size lives in device memory
        const double d = *a.dt_dev;
        if (a.dt_cq) cq *= d;
        if (a.dt_cg) cg *= d;
    }
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    stage_body<P, EPB, NS, IP, RS, R, TR, XBF>(a, b * EPB, b, cq, cg, S);
}

// ------------------------------------------------------------------ time step and sums
template <int P>
__global__ __launch_bounds__(256) void k_dt_elem(const double *Q, int K, double gamma, double gm1, double *umax,
                                                 double *amax) {
    constexpr int PP = P * P, EPB = 256 / PP > 0 ? 256 / PP : 1;
    __shared__ double su[256], sa[256];
    const int t = threadIdx.x, s = t / PP, n = t % PP, e = blockIdx.x * EPB + s;
    double u = 0.0, aa = 0.0;
    if (t < EPB * PP && e < K) {
        const double *q = Q + ((int64_t)e * PP + n) * 4;
        u = sqrt(q[1] * q[1] + q[2] * q[2]) / q[0];
        const double ur = q[1] / q[0], vr = q[2] / q[0];
        aa = sqrt(gamma * gm1 * (q[3] / q[0] - 0.5 * (ur * ur) - 0.5 * (vr * vr)));
    }
    su[t] = u; sa[t] = aa;
    __syncthreads();
    if (t < EPB * PP && e < K && n == 0) {
        double mu = 0.0, ma = 0.0;
        for (int x = 0; x < PP; ++x) { mu = fmax(mu, su[s * PP + x]); ma = fmax(ma, sa[s * PP + x]); }
        umax[e] = mu; amax[e] = ma;
    }
}

// mode 0: sum_e sum_ij w_i w_j J Q (four conserved variables); mode 1: the volume monitors
// sum w_i w_j J ("One") and sum w_i w_j J (p / rho^gamma - 1)^2 ("dSnorm" squared),
// QuadMesh_Integrals.incf:22-65.
template <int P>
__global__ __launch_bounds__(256) void k_cons_partial(const double *Q, const double *jac, const double *w, int K,
                                                      int mode, double gamma, double gm1, double *partial) {
    constexpr int PP = P * P, EPB = 256 / PP > 0 ? 256 / PP : 1;
    __shared__ double sv[256][4];
    const int t = threadIdx.x, s = t / PP, n = t % PP, e = blockIdx.x * EPB + s;
    double v[4] = {0, 0, 0, 0};
    if (t < EPB * PP && e < K) {
        const int i = n % P, j = n / P;
        const int64_t g = (int64_t)e * PP + n;
        const double c = w[i] * w[j] * jac[g];
        if (mode == 0) {
            for (int k = 0; k < 4; ++k) v[k] = c * Q[g * 4 + k];
        } else {
            const double *q = Q + g * 4;
            const double p = gm1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
            const double d = p / pow(q[0], gamma) - 1.0;
            v[0] = c;
            v[1] = c * (d * d);
        }
    }
    for (int k = 0; k < 4; ++k) sv[t][k] = v[k];
    __syncthreads();
    for (int wd = 128; wd > 0; wd >>= 1) {
        if (t < wd)
            for (int k = 0; k < 4; ++k) sv[t][k] += sv[t + wd][k];
        __syncthreads();
    }
    if (t < 4) partial[blockIdx.x * 4 + t] = sv[0][t];
}

// ------------------------------------------------------------------ per-order launchers
template <int P, int EPB>
inline int tile_blocks(int K) { return (K + EPB - 1) / EPB; }

// PHYS: 0 Euler, 1 Navier-Stokes BR1, 2 Navier-Stokes IP.
template <int P, int EPB, int PHYS, int RS, bool XBF = false>
hipError_t stage_launch(bool mixed, const StageArgs &a, hipStream_t st) {
    constexpr int TPB = Tile<P, EPB>::TPB;
    constexpr bool NS = PHYS >= 1, IP = PHYS == 2;
    const int nb = tile_blocks<P, EPB>(a.K);
    if (mixed) hipLaunchKernelGGL((k_stage<P, EPB, NS, IP, RS, float, float, XBF>), dim3(nb), dim3(TPB), 0, st, a);
    else hipLaunchKernelGGL((k_stage<P, EPB, NS, IP, RS, double, double, XBF>), dim3(nb), dim3(TPB), 0, st, a);
    return hipGetLastError();
}

template <int P, int EPB, int PHYS>
hipError_t stage_tile(bool mixed, const StageArgs &a, hipStream_t st) {
    switch (a.ph.rs_kernel) {
    case 0: return stage_launch<P, EPB, PHYS, 0>(mixed, a, st);
    case 1: return stage_launch<P, EPB, PHYS, 1>(mixed, a, st);
    case 2: return stage_launch<P, EPB, PHYS, 2>(mixed, a, st);
    case 3: return stage_launch<P, EPB, PHYS, 3>(mixed, a, st);
    default: return hipErrorInvalidValue;
    }
}

template <int P, int EPB>
hipError_t gradient_tile(bool mixed, const StageArgs &a, hipStream_t st) {
    constexpr int TPB = Tile<P, EPB>::TPB;
    const int nb = tile_blocks<P, EPB>(a.K);
    const bool br1 = a.ph.viscous == 1;
    // BR1 lifts the boundary's uSB (the boundary kinds must be compiled in); with artificial
    // dissipation every viscous trace takes the face's mu + mu_a
    if (a.ph.xbc && (br1 || a.ph.ad)) {
        if (mixed) {
            if (br1) hipLaunchKernelGGL((k_gradient<P, EPB, true, float, float, true>), dim3(nb), dim3(TPB), 0, st, a);
            else hipLaunchKernelGGL((k_gradient<P, EPB, false, float, float, true>), dim3(nb), dim3(TPB), 0, st, a);
        } else {
            if (br1) hipLaunchKernelGGL((k_gradient<P, EPB, true, double, double, true>), dim3(nb), dim3(TPB), 0, st, a);
            else hipLaunchKernelGGL((k_gradient<P, EPB, false, double, double, true>), dim3(nb), dim3(TPB), 0, st, a);
        }
        return hipGetLastError();
    }
    if (mixed) {
        if (br1) hipLaunchKernelGGL((k_gradient<P, EPB, true, float, float>), dim3(nb), dim3(TPB), 0, st, a);
        else hipLaunchKernelGGL((k_gradient<P, EPB, false, float, float>), dim3(nb), dim3(TPB), 0, st, a);
    } else {
        if (br1) hipLaunchKernelGGL((k_gradient<P, EPB, true, double, double>), dim3(nb), dim3(TPB), 0, st, a);
        else hipLaunchKernelGGL((k_gradient<P, EPB, false, double, double>), dim3(nb), dim3(TPB), 0, st, a);
    }
    return hipGetLastError();
}

// The run-time-solver form (rs_kernel = -1) is compiled for the wide tile only.
template <int P, int PHYS>
hipError_t stage_phys_p(int tile, bool mixed, const StageArgs &a, hipStream_t st) {
    if (a.ph.rs_kernel < 0) return stage_launch<P, Tiles<P>::large, PHYS, -1>(mixed, a, st);
    if (a.ph.xbc) {   // one solver, the boundary kinds and artificial dissipation compiled in (wide tile)
        constexpr int EPB = Tiles<P>::large;
        switch (a.ph.rs_kernel) {
        case 0: return stage_launch<P, EPB, PHYS, 0, true>(mixed, a, st);
        case 1: return stage_launch<P, EPB, PHYS, 1, true>(mixed, a, st);
        case 2: return stage_launch<P, EPB, PHYS, 2, true>(mixed, a, st);
        case 3: return stage_launch<P, EPB, PHYS, 3, true>(mixed, a, st);
        default: return hipErrorInvalidValue;
        }
    }
    return tile == 0 ? stage_tile<P, Tiles<P>::small, PHYS>(mixed, a, st)
                     : stage_tile<P, Tiles<P>::large, PHYS>(mixed, a, st);
}

template <int P>
int tile_threads_p(int tile) {
    return tile == 0 ? Tile<P, Tiles<P>::small>::TPB : Tile<P, Tiles<P>::large>::TPB;
}

template <int P>
int tile_elems_p(int tile) {
    return tile == 0 ? Tiles<P>::small : Tiles<P>::large;
}

template <int P>
int stage_blocks_p(int tile, int K) {
    return tile == 0 ? tile_blocks<P, Tiles<P>::small>(K) : tile_blocks<P, Tiles<P>::large>(K);
}

template <int P>
hipError_t gradient_p(int tile, bool mixed, const StageArgs &a, hipStream_t st) {
    return tile == 0 ? gradient_tile<P, Tiles<P>::small>(mixed, a, st)
                     : gradient_tile<P, Tiles<P>::large>(mixed, a, st);
}

template <int P>
hipError_t prolong_p(bool mixed, const StageArgs &a, hipStream_t st) {
    constexpr int EPB = Tiles<P>::large, TPB = Tile<P, EPB>::TPB;
    const int nb = tile_blocks<P, EPB>(a.K);
    if (mixed) hipLaunchKernelGGL((k_prolong<P, EPB, float>), dim3(nb), dim3(TPB), 0, st, a);
    else hipLaunchKernelGGL((k_prolong<P, EPB, double>), dim3(nb), dim3(TPB), 0, st, a);
    return hipGetLastError();
}

template <int P>
hipError_t dt_elem_p(const double *Q, int K, double gamma, double gm1, double *umax, double *amax, hipStream_t st) {
    constexpr int EPB = 256 / (P * P) > 0 ? 256 / (P * P) : 1;
    hipLaunchKernelGGL((k_dt_elem<P>), dim3((K + EPB - 1) / EPB), dim3(256), 0, st, Q, K, gamma, gm1, umax, amax);
    return hipGetLastError();
}

template <int P>
int cons_blocks_p(int K) {
    constexpr int EPB = 256 / (P * P) > 0 ? 256 / (P * P) : 1;
    return (K + EPB - 1) / EPB;
}

template <int P>
hipError_t cons_partial_p(const double *Q, const double *jac, const double *w, int K, int mode, double gamma,
                          double gm1, double *partial, hipStream_t st) {
    hipLaunchKernelGGL((k_cons_partial<P>), dim3(cons_blocks_p<P>(K)), dim3(256), 0, st, Q, jac, w, K, mode, gamma,
                       gm1, partial);
    return hipGetLastError();
}

// ------------------------------------------------------------------ artificial viscosity
// Jumps-based artificial viscosity of each element (JumpsBasedElementViscosity,
// DGArtificialDissipation.f90:277-338), from the stage's Q traces: per
// face sum_k w_k (2 (qL - qR) / (qL + qR))^2 dS for rho and rho E (Edge_ComputeJumps and its
// variants, QuadElement_EdgeProcedures.incf:1348-1619: curved interior faces per-node dS, the
// others dS of the edge's first node; boundary faces the boundary state uSB), normalised by the
// four faces' area, through the smooth switch, times Ceps sqrt(Volume)/(N+1).  FP64 in both
// forms (the FP32 form's traces, energy shifted back).
template <int P, typename TR>
__global__ __launch_bounds__(256) void k_art_visc(StageArgs a, const double *evol, const double *w, double *mu_a) {
    // one lane per face node (EPB elements of 4P lanes per workgroup); the first lane of each
    // element sums its 4P terms face by face in node order, then applies the switch
    constexpr int FN = 4 * P, EPB = 256 / FN;
    __shared__ double sj[EPB * FN][3];
    const int t = threadIdx.x, s = t / FN, l = t % FN, e = blockIdx.x * EPB + s;
    const bool act = (t < EPB * FN) && (e < a.K);
    if (act) {
        const Phys<double> ph(a.ph);
        const double es = sizeof(TR) == 4 ? a.ph.Eref : 0.0;
        const int f = l / P, k = l % P;
        const int4 fr = a.face[e * 4 + f];
        const int kind = fr.x & 15;
        const bool curved = fr.x & FF_CURVED, selfrev = fr.x & FF_SELFREV, rev = fr.x & FF_REV;
        const TR *o = (const TR *)a.tr_in + ((int64_t)(e * 4 + f) * P + k) * 4;
        const double qs[4] = {(double)o[0], (double)o[1], (double)o[2], (double)o[3] + es};
        const int gi = curved ? (selfrev ? P - 1 - k : k) : 0;
        const double *geo = a.egeo + 3 * ((int64_t)fr.z + gi);
        double qo[4], dS = geo[2];
        if (kind <= FK_PERIODIC) {
            const TR *nb = (const TR *)a.tr_in + ((int64_t)fr.y * P + (rev ? P - 1 - k : k)) * 4;
            qo[0] = (double)nb[0]; qo[1] = (double)nb[1]; qo[2] = (double)nb[2]; qo[3] = (double)nb[3] + es;
            if (kind == FK_PERIODIC) dS = a.egeo[3 * (int64_t)fr.z + 2];
        } else {
            double uB[4];
            int neu;
            ghost_state<double, true>(kind, fr.x, zone_state(a, fr.x, fr.y, (int64_t)(e * 4 + f) * P + k), a.ph, ph, qs,
                                      geo[0], geo[1], uB, qo, neu);
            dS = a.egeo[3 * (int64_t)fr.z + 2];
        }
        const double nr = 2.0 * (qs[0] - qo[0]) / (qs[0] + qo[0]);
        const double ne = 2.0 * (qs[3] - qo[3]) / (qs[3] + qo[3]);
        sj[t][0] = w[k] * nr * nr * dS;
        sj[t][1] = w[k] * ne * ne * dS;
        sj[t][2] = w[k] * geo[2];
    }
    __syncthreads();
    if (!act || l != 0) return;
    double gr = 0.0, ge = 0.0, area = 0.0;
    for (int f = 0; f < 4; ++f) {
        double jr = 0.0, je = 0.0, ar = 0.0;
        for (int k = 0; k < P; ++k) {
            const double *v = sj[s * FN + f * P + k];
            jr = jr + v[0];
            je = je + v[1];
            ar = ar + v[2];
        }
        gr = gr + jr;
        ge = ge + je;
        area = area + ar;
    }
    double gk = (gr > ge ? gr : ge) / area;
    const double kw = 5.0, s0 = log10(1.0 / pow((double)P, 4.0)) + 1.0;
    double G;
    if (!(gk > 1.0e-8)) {
        G = 0.0;
    } else {
        gk = log10(gk);
        if (gk < s0 - kw) G = 0.0;
        else if (gk < s0 + kw) G = 0.5 * sin(0.5 * 3.141592653589793 * (gk - s0) / kw) + 0.5;
        else G = 1.0;
    }
    mu_a[e] = a.ph.ad_ceps * (sqrt(evol[e]) / (double)P) * G;
}

template <int P>
hipError_t art_visc_p(bool mixed, const StageArgs &a, const double *evol, const double *w, double *mu_a,
                      hipStream_t st) {
    constexpr int EPB = 256 / (4 * P);
    const int nb = (a.K + EPB - 1) / EPB;
    if (mixed) hipLaunchKernelGGL((k_art_visc<P, float>), dim3(nb), dim3(256), 0, st, a, evol, w, mu_a);
    else hipLaunchKernelGGL((k_art_visc<P, double>), dim3(nb), dim3(256), 0, st, a, evol, w, mu_a);
    return hipGetLastError();
}

#define HG_INSTANTIATE_STAGE(PV, PHYS) \
    template hipError_t stage_phys_p<PV, PHYS>(int, bool, const StageArgs &, hipStream_t);

#define HG_INSTANTIATE_ORDER(PV)                                                                              \
    template int stage_blocks_p<PV>(int, int);                                                                \
    template int tile_threads_p<PV>(int);                                                                     \
    template int tile_elems_p<PV>(int);                                                                       \
    template hipError_t gradient_p<PV>(int, bool, const StageArgs &, hipStream_t);                            \
    template hipError_t prolong_p<PV>(bool, const StageArgs &, hipStream_t);                                  \
    template hipError_t art_visc_p<PV>(bool, const StageArgs &, const double *, const double *, double *,     \
                                       hipStream_t);                                                          \
    template hipError_t dt_elem_p<PV>(const double *, int, double, double, double *, double *, hipStream_t);  \
    template int cons_blocks_p<PV>(int);                                                                      \
    template hipError_t cons_partial_p<PV>(const double *, const double *, const double *, int, int, double,  \
                                           double, double *, hipStream_t);

}  // namespace hg
