// Order dispatch of the DGSEM stage pipeline and the kernels that do not depend on the
// polynomial order (residual reduction, time-step scan, conservation-sum finalisation).
// The order-templated kernels live in hg_kernels.h and are instantiated per order by
// hg_inst.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "hg_device.h"

namespace hg {

namespace {

__device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

// max|QDot| per equation over the stage kernel's per-workgroup partials (4 doubles each):
// workgroup g takes the partials g*1024 + t, stepping by the grid, every lane keeping four
// partials' loads in flight, then wavefront shuffles and LDS; it writes its four maxima to
// out[4 g].  The maximum does not depend on the order (NaN wins every compare), so the result
// is exact whatever the grid.
__global__ __launch_bounds__(1024) void k_resid_final(const double *part, int nb, double *out) {
    __shared__ double s[16][4];
    const int t = threadIdx.x, stride = 1024 * gridDim.x;
    const double2 *p2 = (const double2 *)part;
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    int b = blockIdx.x * 1024 + t;
    for (; b + 3 * stride < nb; b += 4 * stride) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[2 * u] = p2[(int64_t)(b + u * stride) * 2];
            v[2 * u + 1] = p2[(int64_t)(b + u * stride) * 2 + 1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            m[0] = nan_max(m[0], v[2 * u].x); m[1] = nan_max(m[1], v[2 * u].y);
            m[2] = nan_max(m[2], v[2 * u + 1].x); m[3] = nan_max(m[3], v[2 * u + 1].y);
        }
    }
    for (; b < nb; b += stride) {
        const double2 a = p2[(int64_t)b * 2], c = p2[(int64_t)b * 2 + 1];
        m[0] = nan_max(m[0], a.x); m[1] = nan_max(m[1], a.y); m[2] = nan_max(m[2], c.x); m[3] = nan_max(m[3], c.y);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
        for (int off = 32; off > 0; off >>= 1) m[c] = nan_max(m[c], __shfl_xor(m[c], off, 64));
    if ((t & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) s[t >> 6][c] = m[c];
    }
    __syncthreads();
    if (t < 4) {
        double v = s[0][t];
        for (int w = 1; w < 16; ++w) v = nan_max(v, s[w][t]);
        out[blockIdx.x * 4 + t] = v;
    }
}

// TimeIntegrator_EstimateTimeStep (DGTimeIntegrator.f90:462-511), second half: the running
// maxima of |u| and a advance in element order starting from the values saved by the
// previous call, and dt is the minimum over elements of cfl*dx/(umax+amax)/(N+1) taken
// with the maxima as they stood at that element.  Exclusive prefix max over 1024 chunks.
__global__ __launch_bounds__(1024) void k_dt_scan(int K, const double *umax, const double *amax, const double *dx,
                                                  double coef, double N1, double dt_max, double *state, double *dt_out,
                                                  const double *n1e) {
    __shared__ double su[1024], sa[1024], sd[1024];
    const int t = threadIdx.x, T = blockDim.x;
    const int chunk = (K + T - 1) / T, b = t * chunk, end = min(K, b + chunk);
    double mu = 0.0, ma = 0.0;
    for (int e = b; e < end; ++e) { mu = fmax(mu, umax[e]); ma = fmax(ma, amax[e]); }
    su[t] = mu; sa[t] = ma;
    __syncthreads();
    if (t == 0) {
        double ru = state[0], ra = state[1];
        for (int x = 0; x < T; ++x) {
            const double cu = su[x], ca = sa[x];
            su[x] = ru; sa[x] = ra;
            ru = fmax(ru, cu); ra = fmax(ra, ca);
        }
        state[0] = ru; state[1] = ra;
    }
    __syncthreads();
    double ru = su[t], ra = sa[t], dt = dt_max;
    for (int e = b; e < end; ++e) {
        ru = fmax(ru, umax[e]); ra = fmax(ra, amax[e]);
        const double de = coef * dx[e] / (ru + ra) / (n1e ? n1e[e] : N1);
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

__global__ void k_cons_final(const double *partial, int nb, double *out) {
    if (threadIdx.x < 4) {
        double acc = 0.0;
        for (int b = 0; b < nb; ++b) acc += partial[b * 4 + threadIdx.x];
        out[threadIdx.x] = acc;
    }
}

// Surface monitors of one zone (QuadMesh_Integrals.incf:67-346): one workgroup; each lane sums a
// strided share of the zone's face nodes from the stored element-side Q traces (the energy slot
// shifted back by Eref in the FP32 form) and, for Navier-Stokes, the stored viscous face-flux
// traces, whose momentum components are the stress tensor (mu (f1, f2, g2) = tau_xx, tau_xy,
// tau_yy, ViscousFluxes.f90:416-461); then a fixed-order reduction.
template <typename TR>
__global__ __launch_bounds__(1024) void k_surface_integrals(const int *slots, int nslots, int P, const TR *tr,
                                                            const TR *vtr, const int4 *face, const double *egeo,
                                                            const double *w, double gm1, double Eref,
                                                            const double *mu_a, double mu, double *out) {
    __shared__ double s[16][9];
    const int t = threadIdx.x;
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int x = t; x < nslots * P; x += 1024) {
        const int slot = slots[x / P], k = x % P;
        const int4 fr = face[slot];
        const int gi = (fr.x & FF_CURVED) ? ((fr.x & FF_SELFREV) ? P - 1 - k : k) : 0;
        const int wi = (fr.x & FF_SELFREV) ? P - 1 - k : k;
        const double *geo = egeo + 3 * ((int64_t)fr.z + gi);
        const double nx = geo[0], ny = geo[1], dS = geo[2], wq = w[wi];
        const TR *q = tr + ((int64_t)slot * P + k) * 4;
        const double r = q[0], ru = q[1], rv = q[2], E = (double)q[3] + Eref;
        const double p = gm1 * (E - 0.5 * (ru * ru + rv * rv) / r);
        const double un = ru * nx + rv * ny;
        acc[0] += wq * dS;
        acc[1] += p * dS * wq;
        acc[2] += r * dS * wq;
        acc[3] += p * wq * nx * dS;
        acc[4] += p * wq * ny * dS;
        acc[5] += wq * un * dS;
        acc[6] += wq * un * dS / r;
        if (vtr) {
            const TR *v = vtr + ((int64_t)slot * P + k) * VT;
            // the traces carry mu + mu_a of the (boundary) face; the monitor's stress has mu
            const double sc = mu_a ? mu / (mu + mu_a[slot >> 2]) : 1.0;
            const double txx = v[0] * sc, txy = v[1] * sc, tyy = v[3] * sc;
            acc[7] += wq * (txx * nx + txy * ny) * dS;
            acc[8] += wq * (txy * nx + tyy * ny) * dS;
        }
    }
#pragma unroll
    for (int c = 0; c < 9; ++c)
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    if ((t & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 9; ++c) s[t >> 6][c] = acc[c];
    }
    __syncthreads();
    if (t < 9) {
        double v = s[0][t];
        for (int wv = 1; wv < 16; ++wv) v += s[wv][t];
        out[t] = v;
    }
}

// Probe_Update (IO/Monitors.f90:739-846), one lane per probe: B = l_xi^T A l_eta with
// (A l_eta)(i) = sum_j A(i,j) l_eta(j) (BilinearForm, MatrixOperations.f90:185-209) for each
// conservative variable of Q (or of QDot for rhot .. rhoet), then the tracked variable.
__global__ __launch_bounds__(64) void k_probes(const double *Q, const double *QDot, const int2 *ev, const double *l,
                                               int n, int P, double gm1, double gamma, double *out) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= n) return;
    const int2 x = ev[k];
    const double *lx = l + (int64_t)k * 2 * P, *ly = lx + P;
    const double *A = ((x.y >= 4 && x.y <= 7) ? QDot : Q) + (int64_t)x.x * P * P * 4;
    double v[4];
    for (int eq = 0; eq < 4; ++eq) {
        double b = 0.0;
        for (int i = 0; i < P; ++i) {
            double r = 0.0;
            for (int j = 0; j < P; ++j) r += A[(i + P * j) * 4 + eq] * ly[j];
            b += lx[i] * r;
        }
        v[eq] = b;
    }
    double o;
    switch (x.y) {
    case 8: o = v[1] / v[0]; break;
    case 9: o = v[2] / v[0]; break;
    case 10: o = gm1 * (v[3] - 0.5 * (v[1] * v[1] + v[2] * v[2]) / v[0]); break;
    case 11: o = sqrt(v[1] * v[1] + v[2] * v[2]) / v[0] / sqrt(gamma); break;
    default: o = v[x.y & 3]; break;
    }
    out[k] = o;
}

// Residual of each case of a batch: max over its tiles' partials (NaN propagates).
__global__ __launch_bounds__(256) void k_batch_resid(const double *part, const int2 *ranges, double *const *out) {
    __shared__ double s[256][4];
    const int t = threadIdx.x, c = blockIdx.x;
    const int2 r = ranges[c];
    double m[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = r.x + t; b < r.y; b += 256)
        for (int q = 0; q < 4; ++q) m[q] = nan_max(m[q], part[(int64_t)b * 4 + q]);
    for (int q = 0; q < 4; ++q) s[t][q] = m[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int q = 0; q < 4; ++q) s[t][q] = nan_max(s[t][q], s[t + w][q]);
        __syncthreads();
    }
    if (t < 4) out[c][t] = s[0][t];
}

}  // namespace

bool supported_order(int N) { return N >= 1 && N <= 12; }

#define HG_DISPATCH_P(P_, CALL)          \
    switch (P_) {                        \
    case 2: return CALL(2);              \
    case 3: return CALL(3);              \
    case 4: return CALL(4);              \
    case 5: return CALL(5);              \
    case 6: return CALL(6);              \
    case 7: return CALL(7);              \
    case 8: return CALL(8);              \
    case 9: return CALL(9);              \
    case 10: return CALL(10);            \
    case 11: return CALL(11);            \
    case 12: return CALL(12);            \
    case 13: return CALL(13);            \
    default: return hipErrorInvalidValue; \
    }

hipError_t launch_art_visc(int P, bool mixed, const StageArgs &a, const double *evol, const double *w, double *mu_a,
                           hipStream_t st) {
#define HG_CALL(PV) art_visc_p<PV>(mixed, a, evol, w, mu_a, st)
    HG_DISPATCH_P(P, HG_CALL)
#undef HG_CALL
}

hipError_t launch_prolong(int P, bool mixed, const StageArgs &a, hipStream_t st) {
#define HG_CALL(PV) prolong_p<PV>(mixed, a, st)
    HG_DISPATCH_P(P, HG_CALL)
#undef HG_CALL
}

template <int PHYS>
static hipError_t launch_stage_phys(int P, int tile, bool mixed, const StageArgs &a, hipStream_t st) {
#define HG_CALL(PV) stage_phys_p<PV, PHYS>(tile, mixed, a, st)
    HG_DISPATCH_P(P, HG_CALL)
#undef HG_CALL
}

hipError_t launch_stage(int P, int tile, bool mixed, const StageArgs &a, hipStream_t st) {
    if (!a.ph.ns) return launch_stage_phys<0>(P, tile, mixed, a, st);
    if (a.ph.viscous == 2) return launch_stage_phys<2>(P, tile, mixed, a, st);
    return launch_stage_phys<1>(P, tile, mixed, a, st);
}

hipError_t launch_gradient(int P, int tile, bool mixed, const StageArgs &a, hipStream_t st) {
#define HG_CALL(PV) gradient_p<PV>(tile, mixed, a, st)
    HG_DISPATCH_P(P, HG_CALL)
#undef HG_CALL
}

hipError_t launch_resid_final(const double *part, int nb, double *out, hipStream_t st, double *scratch) {
    // one workgroup (one lane per partial) when that covers them; beyond, a first pass of up to
    // kResidGroups workgroups into scratch, then one workgroup over their maxima (one workgroup
    // over 64 k partials, the 256 x 256 order-7 mesh, took 37 us of a 250 us stage)
    const int groups = std::min(kResidGroups, (nb + 1023) / 1024);
    if (groups <= 1 || !scratch) {
        hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(1024), 0, st, part, nb, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_resid_final, dim3(groups), dim3(1024), 0, st, part, nb, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(1024), 0, st, scratch, groups, out);
    return hipGetLastError();
}

int tile_threads(int P, int tile) {
#define HG_CALL(PV) tile_threads_p<PV>(tile)
    switch (P) {
    case 2: return HG_CALL(2);
    case 3: return HG_CALL(3);
    case 4: return HG_CALL(4);
    case 5: return HG_CALL(5);
    case 6: return HG_CALL(6);
    case 7: return HG_CALL(7);
    case 8: return HG_CALL(8);
    case 9: return HG_CALL(9);
    case 10: return HG_CALL(10);
    case 11: return HG_CALL(11);
    case 12: return HG_CALL(12);
    case 13: return HG_CALL(13);
    default: return 0;
    }
#undef HG_CALL
}

int tile_elems(int P, int tile) {
#define HG_CALL(PV) tile_elems_p<PV>(tile)
    switch (P) {
    case 2: return HG_CALL(2);
    case 3: return HG_CALL(3);
    case 4: return HG_CALL(4);
    case 5: return HG_CALL(5);
    case 6: return HG_CALL(6);
    case 7: return HG_CALL(7);
    case 8: return HG_CALL(8);
    case 9: return HG_CALL(9);
    case 10: return HG_CALL(10);
    case 11: return HG_CALL(11);
    case 12: return HG_CALL(12);
    case 13: return HG_CALL(13);
    default: return 0;
    }
#undef HG_CALL
}

int stage_blocks(int P, int tile, int K) {
#define HG_CALL(PV) stage_blocks_p<PV>(tile, K)
    switch (P) {
    case 2: return HG_CALL(2);
    case 3: return HG_CALL(3);
    case 4: return HG_CALL(4);
    case 5: return HG_CALL(5);
    case 6: return HG_CALL(6);
    case 7: return HG_CALL(7);
    case 8: return HG_CALL(8);
    case 9: return HG_CALL(9);
    case 10: return HG_CALL(10);
    case 11: return HG_CALL(11);
    case 12: return HG_CALL(12);
    case 13: return HG_CALL(13);
    default: return 0;
    }
#undef HG_CALL
}

hipError_t launch_dt_elem(int P, const double *Q, int K, double gamma, double gm1, double *umax, double *amax,
                          hipStream_t st) {
#define HG_CALL(PV) dt_elem_p<PV>(Q, K, gamma, gm1, umax, amax, st)
    HG_DISPATCH_P(P, HG_CALL)
#undef HG_CALL
}

hipError_t launch_dt_scan(int K, const double *umax, const double *amax, const double *dx, double coef, double N1,
                          double dt_max, double *state, double *dt_out, hipStream_t st, const double *n1e) {
    hipLaunchKernelGGL(k_dt_scan, dim3(1), dim3(1024), 0, st, K, umax, amax, dx, coef, N1, dt_max, state, dt_out,
                       n1e);
    return hipGetLastError();
}

static hipError_t cons_partial(int P, const double *Q, const double *jac, const double *w, int K, int mode,
                               double gamma, double gm1, double *partial,
                               int nbk, hipStream_t st) {
#define HG_CALL(PV) (cons_blocks_p<PV>(K) > nbk ? hipErrorInvalidValue : cons_partial_p<PV>(Q, jac, w, K, mode, gamma, gm1, partial, st))
    HG_DISPATCH_P(P, HG_CALL)
#undef HG_CALL
}

static int cons_blocks(int P, int K) {
#define HG_CALL(PV) cons_blocks_p<PV>(K)
    switch (P) {
    case 2: return HG_CALL(2);
    case 3: return HG_CALL(3);
    case 4: return HG_CALL(4);
    case 5: return HG_CALL(5);
    case 6: return HG_CALL(6);
    case 7: return HG_CALL(7);
    case 8: return HG_CALL(8);
    case 9: return HG_CALL(9);
    case 10: return HG_CALL(10);
    case 11: return HG_CALL(11);
    case 12: return HG_CALL(12);
    case 13: return HG_CALL(13);
    default: return 0;
    }
#undef HG_CALL
}

hipError_t launch_cons_sums(int P, const double *Q, const double *jac, const double *w, int K, double *partial,
                            int nbk, double *out, hipStream_t st) {
    hipError_t e = cons_partial(P, Q, jac, w, K, 0, 0.0, 0.0, partial, nbk, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, st, partial, cons_blocks(P, K), out);
    return hipGetLastError();
}

hipError_t launch_volume_integrals(int P, const double *Q, const double *jac, const double *w, int K, double gamma,
                                   double gm1, double *partial, int nbk, double *out, hipStream_t st) {
    hipError_t e = cons_partial(P, Q, jac, w, K, 1, gamma, gm1, partial, nbk, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cons_final, dim3(1), dim3(64), 0, st, partial, cons_blocks(P, K), out);
    return hipGetLastError();
}

hipError_t launch_surface_integrals(bool mixed, const int *slots, int nslots, int P, const void *tr, const void *vtr,
                                    const int4 *face, const double *egeo, const double *w, double gm1, double Eref,
                                    const double *mu_a, double mu, double *out, hipStream_t st) {
    if (mixed)
        hipLaunchKernelGGL((k_surface_integrals<float>), dim3(1), dim3(1024), 0, st, slots, nslots, P,
                           (const float *)tr, (const float *)vtr, face, egeo, w, gm1, Eref, mu_a, mu, out);
    else
        hipLaunchKernelGGL((k_surface_integrals<double>), dim3(1), dim3(1024), 0, st, slots, nslots, P,
                           (const double *)tr, (const double *)vtr, face, egeo, w, gm1, Eref, mu_a, mu, out);
    return hipGetLastError();
}

hipError_t launch_probes(const double *Q, const double *QDot, const int2 *ev, const double *l, int n, int P,
                         double gm1, double gamma, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_probes, dim3((n + 63) / 64), dim3(64), 0, st, Q, QDot, ev, l, n, P, gm1, gamma, out);
    return hipGetLastError();
}

}  // namespace hg

namespace hg {

hipError_t launch_batch_stage(int variant, int rs, int tile, const BatchArgs &b, int block, hipStream_t st) {
#define HG_RS(V)                                                    \
    switch (rs) {                                                   \
    case -1: return batch_stage_v##V##_rs0(tile, b, block, st);     \
    case 0: return batch_stage_v##V##_rs1(tile, b, block, st);      \
    case 1: return batch_stage_v##V##_rs2(tile, b, block, st);      \
    case 2: return batch_stage_v##V##_rs3(tile, b, block, st);      \
    case 3: return batch_stage_v##V##_rs4(tile, b, block, st);      \
    default: return hipErrorInvalidValue;                           \
    }
    switch (variant) {
    case 0: HG_RS(0)
    case 1: HG_RS(1)
    case 2: HG_RS(2)
    case 3: HG_RS(3)
    case 4: HG_RS(4)
    case 5: HG_RS(5)
    default: return hipErrorInvalidValue;
    }
#undef HG_RS
}

hipError_t launch_batch_gradient(int variant, int tile, const BatchArgs &b, int block, hipStream_t st) {
    switch (variant) {
    case 2: return batch_gradient_v2(tile, b, block, st);
    case 3: return batch_gradient_v3(tile, b, block, st);
    case 4: return batch_gradient_v4(tile, b, block, st);
    case 5: return batch_gradient_v5(tile, b, block, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_stage_p(int variant, int rs, int P, int tile, const BatchArgs &b, hipStream_t st) {
#define HG_RS(V)                                                        \
    switch (rs) {                                                       \
    case -1: return batch_stage_v##V##_rs0_p(P, tile, b, st);           \
    case 0: return batch_stage_v##V##_rs1_p(P, tile, b, st);            \
    case 1: return batch_stage_v##V##_rs2_p(P, tile, b, st);            \
    case 2: return batch_stage_v##V##_rs3_p(P, tile, b, st);            \
    case 3: return batch_stage_v##V##_rs4_p(P, tile, b, st);            \
    default: return hipErrorInvalidValue;                               \
    }
    switch (variant) {
    case 0: HG_RS(0)
    case 1: HG_RS(1)
    case 2: HG_RS(2)
    case 3: HG_RS(3)
    case 4: HG_RS(4)
    case 5: HG_RS(5)
    default: return hipErrorInvalidValue;
    }
#undef HG_RS
}

hipError_t launch_batch_gradient_p(int variant, int P, int tile, const BatchArgs &b, hipStream_t st) {
    switch (variant) {
    case 2: return batch_gradient_v2_p(P, tile, b, st);
    case 3: return batch_gradient_v3_p(P, tile, b, st);
    case 4: return batch_gradient_v4_p(P, tile, b, st);
    case 5: return batch_gradient_v5_p(P, tile, b, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_resid(const double *part, const int2 *ranges, double *const *out, int ncases, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_resid, dim3(ncases), dim3(256), 0, st, part, ranges, out);
    return hipGetLastError();
}

}  // namespace hg
