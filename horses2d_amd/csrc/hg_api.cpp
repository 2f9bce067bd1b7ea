// C ABI of the MI355X DGSEM library (declared in include/horses2d_gpu.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "hg_device.h"
#include "hg_handle.h"
#include "hg_mesh.h"
#include "horses2d_gpu.h"

namespace {

thread_local std::string g_err;

int set_err(const std::string &m) {
    g_err = m;
    return 1;
}

#define HG_CHECK(call)                                                                          \
    do {                                                                                        \
        hipError_t _e = (call);                                                                 \
        if (_e != hipSuccess) return set_err(std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)


}  // namespace


namespace {

template <typename T>
hipError_t upload(hg_handle *h, T **dst, const std::vector<T> &src) {
    hipError_t e = h->alloc(dst, src.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

hipError_t upload_real(hg_handle *h, void **dst, const std::vector<double> &src, bool as_float) {
    if (!as_float) {
        double *d;
        hipError_t e = upload(h, &d, src);
        *dst = d;
        return e;
    }
    std::vector<float> f(src.begin(), src.end());
    float *d;
    hipError_t e = upload(h, &d, f);
    *dst = d;
    return e;
}

hg::StageArgs base_args(hg_handle *h) {
    hg::StageArgs a{};
    a.K = h->K;
    a.nodes = h->nodes;
    a.Q = h->Q;
    a.G = h->G;
    a.QDot = h->QDot;
    a.dtr = h->dtr[h->ip_fused ? h->cur : 0];
    a.dtr_out = h->dtr[h->ip_fused ? h->cur : 0];
    a.metric = h->metric;
    a.metric_elem = h->metric_elem;
    a.face = h->face;
    a.forder = h->forder;
    a.egeo = h->egeo;
    a.einvh = h->einvh;
    a.zone = h->zone;
    a.fmet = h->fmet;
    a.hatD = h->hatD;
    a.D = h->D;
    a.lb = h->lb;
    a.lbw = h->lbw;
    a.dlbw = h->dlbw;
    a.resid_part = h->resid_part;
    a.ph = h->ph;
    a.mu_a = (h->ph.ad && !h->lap) ? h->mu_a : nullptr;
    a.lap = h->lap ? h->mu_a : nullptr;
    a.earea = h->earea;
    a.ip_epi = h->ip_fused ? 1 : 0;
    a.mustar = h->mustar;
    a.mflux = h->mflux;
    a.mdq = h->mdq;
    a.mgrad = h->mgrad;
    a.fstate = h->fstate;
    return a;
}

hipEvent_t take_event(hg_handle *h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Launch with optional event timing (kernel id 0: stage, 1: gradient).
template <typename F>
hipError_t timed(hg_handle *h, int which, F &&launch) {
    if (!h->prof) return launch();
    hipEvent_t a = take_event(h), b = take_event(h);
    if (!a || !b) return hipErrorOutOfMemory;
    hipEventRecord(a, h->stream);
    hipError_t e = launch();
    hipEventRecord(b, h->stream);
    h->ev_pending.push_back({which, {a, b}});
    return e;
}

// Chooses, for each kernel whose tile is not fixed yet, the faster of the two workgroup
// tiles by timing it on the current state.  The stage kernel is timed as it runs in a step (RK
// update, next traces) on a scratch copy of Q with cg = 0, G in the QDot buffer, and the next
// stage's trace buffers as outputs (the first real stage overwrites them).  Both tiles
// produce bit-identical results.
int autotune(hg_handle *h) {
    const bool need_g = h->ph.ns && h->tile[1] < 0, need_s = h->tile[0] < 0;
    if (!need_g && !need_s) return 0;
    h->qdot_valid = false;   // the timed stage launches use the QDot buffer as their G
    hipEvent_t ev[2];
    HG_CHECK(hipEventCreate(&ev[0]));
    HG_CHECK(hipEventCreate(&ev[1]));
    auto time_of = [&](auto launch, float *ms) -> hipError_t {
        hipError_t e = launch();
        if (e != hipSuccess) return e;
        hipEventRecord(ev[0], h->stream);
        for (int r = 0; r < 6; ++r)
            if ((e = launch()) != hipSuccess) return e;
        hipEventRecord(ev[1], h->stream);
        if ((e = hipEventSynchronize(ev[1])) != hipSuccess) return e;
        return hipEventElapsedTime(ms, ev[0], ev[1]);
    };
    hipError_t err = hipSuccess;
    hg::StageArgs g = base_args(h);
    g.tr_in = h->tr[h->cur];
    if (need_g) {
        float t[2] = {0.f, 0.f};
        for (int k = 0; k < 2 && err == hipSuccess; ++k)
            err = time_of([&] { return hg::launch_gradient(h->P, k, h->mixed, g, h->stream); }, &t[k]);
        if (err == hipSuccess) h->tile[1] = t[0] <= t[1] ? 0 : 1;
    }
    double *Qs = nullptr;
    if (need_s && err == hipSuccess) err = hipMalloc((void **)&Qs, (size_t)h->ndof * sizeof(double));
    if (need_s && err == hipSuccess)
        err = hipMemcpyAsync(Qs, h->Q, (size_t)h->ndof * sizeof(double), hipMemcpyDeviceToDevice, h->stream);
    if (need_s && err == hipSuccess) {
        hg::StageArgs a = base_args(h);
        a.Q = Qs;
        a.G = h->QDot;
        a.tr_in = h->tr[h->cur];
        a.tr_out = h->tr[h->cur ^ 1];
        a.dtr_out = h->dtr[h->ip_fused ? (h->cur ^ 1) : 0];
        a.am = 0.0;
        a.cq = 1.0;
        a.cg = 0.0;
        a.first_stage = 1;
        a.mode = 0;
        float t[2] = {0.f, 0.f};
        for (int k = 0; k < 2 && err == hipSuccess; ++k)
            err = time_of([&] { return hg::launch_stage(h->P, k, h->mixed, a, h->stream); }, &t[k]);
        if (err == hipSuccess) h->tile[0] = t[0] <= t[1] ? 0 : 1;
    }
    if (Qs) {
        hipStreamSynchronize(h->stream);
        hipFree(Qs);
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    HG_CHECK(err);
    return 0;
}

// Williamson coefficients of the handle's scheme: G = am G + cq QDot, Q += cg G, with cq and/or
// cg multiplied by dt (DGTimeIntegrator.f90:263-365).
struct StepCoefs {
    int nst;
    double am[5], cq[5], cg[5];
    int dt_cq, dt_cg;
};

StepCoefs step_coefs(int scheme) {
    StepCoefs r{};
    if (scheme == HG_SCHEME_RK3) {
        // TimeIntegrator_WilliamsonRK3 (DGTimeIntegrator.f90:282-317)
        const double am[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
        const double gm[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
        r.nst = 3;
        for (int s = 0; s < 3; ++s) { r.am[s] = am[s]; r.cq[s] = 1.0; r.cg[s] = gm[s]; }
        r.dt_cq = 0; r.dt_cg = 1;
    } else if (scheme == HG_SCHEME_RK5) {
        // TimeIntegrator_WilliamsonRK5 (DGTimeIntegrator.f90:319-365): G = am G + dt QDot, Q += gm G
        const double am[5] = {0.0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
        const double gm[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681};
        r.nst = 5;
        for (int s = 0; s < 5; ++s) { r.am[s] = am[s]; r.cq[s] = 1.0; r.cg[s] = gm[s]; }
        r.dt_cq = 1; r.dt_cg = 0;
    } else {
        // TimeIntegrator_ExplicitEuler (DGTimeIntegrator.f90:263-280)
        r.nst = 1;
        r.am[0] = 0.0; r.cq[0] = 1.0; r.cg[0] = 1.0;
        r.dt_cq = 0; r.dt_cg = 1;
    }
    return r;
}

// parts: 1 the gradient pass (with the indicator), 2 the stage kernel (with the residual and
// the trace-parity flip), 3 both.
int run_stage(hg_handle *h, double am, double cq, double cg, bool first, bool last, int mode,
              const double *dt_dev = nullptr, int dt_cq = 0, int dt_cg = 0, int parts = 3) {
    if (autotune(h)) return 1;
    if (parts & 1) {
        // artificial viscosity of the stage's state, before anything evaluates a viscous flux
        // (DGSpatial_QDotVolumeLoop computes it per element first, :255)
        if (h->ph.ad) {
            hg::StageArgs v = base_args(h);
            v.tr_in = h->tr[h->cur];
            HG_CHECK(hg::launch_art_visc(h->P, h->mixed, v, h->evol, h->w, h->mu_a, h->stream));
        }
        // BR1: the viscous face-flux traces of the current state, every stage; interior penalty:
        // only when the previous stage's epilogue did not write them (after an upload)
        if (h->ph.ns && (!h->ip_fused || !h->vtr_valid)) {
            hg::StageArgs g = base_args(h);
            g.tr_in = h->tr[h->cur];
            HG_CHECK(timed(h, 1, [&] { return hg::launch_gradient(h->P, h->tile[1], h->mixed, g, h->stream); }));
            if (h->ip_fused) h->vtr_valid = true;
        }
    }
    if (!(parts & 2)) return 0;
    hg::StageArgs a = base_args(h);
    a.tr_in = h->tr[h->cur];
    a.tr_out = h->tr[h->cur ^ 1];
    a.dtr_out = h->dtr[h->ip_fused ? (h->cur ^ 1) : 0];
    a.am = am;
    a.cq = cq;
    a.cg = cg;
    a.first_stage = first ? 1 : 0;
    a.mode = mode;
    a.do_resid = last ? 1 : 0;
    a.dt_dev = dt_dev;
    a.dt_cq = dt_cq;
    a.dt_cg = dt_cg;
    HG_CHECK(timed(h, 0, [&] { return hg::launch_stage(h->P, h->tile[0], h->mixed, a, h->stream); }));
    if (last)
        HG_CHECK(hg::launch_resid_final(h->resid_part, hg::stage_blocks(h->P, h->tile[0], h->K), h->resid, h->stream,
                                        h->resid_scratch));
    if (mode == 0) h->cur ^= 1;
    return 0;
}

void drop_graphs(hg_handle *h) {
    for (int p = 0; p < 2; ++p) {
        if (h->gexec[p]) hipGraphExecDestroy(h->gexec[p]);
        if (h->graph[p]) hipGraphDestroy(h->graph[p]);
        h->gexec[p] = nullptr;
        h->graph[p] = nullptr;
    }
}

// Captures one RK step for each trace parity (the kernels read dt from h->dt_dev).
int build_graphs(hg_handle *h, const StepCoefs &k) {
    const int cur0 = h->cur;
    for (int p = 0; p < 2; ++p) {
        h->cur = p;
        HG_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int bad = 0;
        for (int s = 0; s < k.nst && !bad; ++s)
            bad = run_stage(h, k.am[s], k.cq[s], k.cg[s], s == 0, s == k.nst - 1, 0, h->dt_dev, k.dt_cq, k.dt_cg);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        h->cur = cur0;
        if (bad) {
            if (g) hipGraphDestroy(g);
            return 1;
        }
        HG_CHECK(e);
        h->graph[p] = g;
        HG_CHECK(hipGraphInstantiate(&h->gexec[p], g, nullptr, nullptr, 0));
    }
    return 0;
}

// ---- meshes with several polynomial orders: one sub-handle per order, coupled by the mortar
// kernels (p-non-conforming edges, QuadElement_EdgeProcedures.incf:749-1206)

// The mortar kernels' arguments at the sub-handles' current trace parity.
hg::MortarArgs mortar_args(hg_handle *h) {
    hg::MortarArgs a{};
    a.edges = h->medges;
    a.nedges = h->nmortar;
    a.subs = h->sedges;
    a.nsubs = h->nsubedge;
    a.sgeo = h->sgeo;
    a.mats = h->mmats;
    a.ph = h->sub[0]->ph;
    a.mixed = h->sub[0]->mixed;
    a.Eref = h->sub[0]->mixed ? h->sub[0]->ph.Eref : 0.0;
    for (size_t g = 0; g < h->sub.size(); ++g) {
        hg_handle *s = h->sub[g];
        a.tr[g] = s->tr[s->cur];
        a.mdq[g] = s->mdq;
        a.mustar[g] = s->mustar;
        a.mflux[g] = s->mflux;
        a.mgrad[g] = s->mgrad;
    }
    return a;
}

// One RK stage (or the time derivative, mode 1) of every sub-handle: BR1 uStar of the mortar
// edges, each sub-handle's gradient pass (which also keeps its mortar faces' gradient traces),
// the mortar edges' fluxes, each sub-handle's stage kernel.
int composite_stage(hg_handle *h, double am, double cq, double cg, bool first, bool last, int mode,
                    const double *dt_dev = nullptr, int dt_cq = 0, int dt_cg = 0) {
    if (h->ph.ns) {
        if (h->ph.viscous == HG_VISCOUS_BR1) {
            HG_CHECK(hg::launch_mortar_ustar(mortar_args(h), h->stream));
            HG_CHECK(hg::launch_sub_ustar(mortar_args(h), h->stream));
        }
        for (auto *s : h->sub)
            if (run_stage(s, am, cq, cg, first, last, mode, dt_dev, dt_cq, dt_cg, 1)) return 1;
    }
    HG_CHECK(hg::launch_mortar_flux(mortar_args(h), h->stream));
    HG_CHECK(hg::launch_sub_flux(mortar_args(h), h->stream));
    for (auto *s : h->sub)
        if (run_stage(s, am, cq, cg, first, last, mode, dt_dev, dt_cq, dt_cg, 2)) return 1;
    if (mode == 0) h->cur ^= 1;
    return 0;
}

int build_composite_graphs(hg_handle *h, const StepCoefs &k) {
    const int cur0 = h->cur;
    for (int p = 0; p < 2; ++p) {
        h->cur = p;
        for (auto *s : h->sub) s->cur = p;
        HG_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        int bad = 0;
        for (int st = 0; st < k.nst && !bad; ++st)
            bad = composite_stage(h, k.am[st], k.cq[st], k.cg[st], st == 0, st == k.nst - 1, 0, h->dt_dev, k.dt_cq,
                                  k.dt_cg);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        h->cur = cur0;
        for (auto *s : h->sub) s->cur = cur0;
        if (bad) {
            if (g) hipGraphDestroy(g);
            return 1;
        }
        HG_CHECK(e);
        h->graph[p] = g;
        HG_CHECK(hipGraphInstantiate(&h->gexec[p], g, nullptr, nullptr, 0));
    }
    return 0;
}

int composite_rk_step(hg_handle *h, double dt) {
    const StepCoefs k = step_coefs(h->opt.scheme);
    for (auto *s : h->sub) {
        s->qdot_valid = false;
        if (autotune(s)) return 1;
        // interior penalty: the viscous traces (and the mortar faces' gradient traces) of the
        // current state before a captured step, whose stages expect the previous epilogue's
        if (hg::ensure_vtr(s)) return 1;
    }
    if (h->use_graph && !h->prof) {
        if (!h->dt_valid || dt != h->dt_last) {
            h->dt_last = dt;
            h->dt_valid = true;
            HG_CHECK(hipMemcpyAsync(h->dt_dev, &h->dt_last, sizeof(double), hipMemcpyHostToDevice, h->stream));
        }
        if (!h->gexec[0] && build_composite_graphs(h, k)) {
            drop_graphs(h);
            return 1;
        }
        HG_CHECK(hipGraphLaunch(h->gexec[h->cur], h->stream));
        if (k.nst & 1) {
            h->cur ^= 1;
            for (auto *s : h->sub) s->cur ^= 1;
        }
    } else {
        for (int st = 0; st < k.nst; ++st) {
            const double cq = k.dt_cq ? k.cq[st] * dt : k.cq[st];
            const double cg = k.dt_cg ? k.cg[st] * dt : k.cg[st];
            if (composite_stage(h, k.am[st], cq, cg, st == 0, st == k.nst - 1, 0)) return 1;
        }
    }
    for (auto *s : h->sub) hg::mark_last_evaluation(s, s->cur ^ 1);
    return 0;
}

}  // namespace

namespace hg {
StageArgs handle_args(hg_handle *h) { return base_args(h); }

void *vtr_current(hg_handle *h) { return h->dtr[h->ip_fused ? h->cur : 0]; }
void *vtr_next(hg_handle *h) { return h->dtr[h->ip_fused ? (h->cur ^ 1) : 0]; }

int ensure_vtr(hg_handle *h) {
    if (!h->ip_fused || h->vtr_valid) return 0;
    if (autotune(h)) return 1;
    hg::StageArgs g = base_args(h);
    g.tr_in = h->tr[h->cur];
    HG_CHECK(hg::launch_gradient(h->P, h->tile[1], h->mixed, g, h->stream));
    h->vtr_valid = true;
    return 0;
}
int handle_error(const std::string &m) { return set_err(m); }
void mark_last_evaluation(hg_handle *h, int tr) {
    h->mon_tr = tr;
    h->mon_vtr = h->ph.ns ? h->dtr[h->ip_fused ? tr : 0] : nullptr;
}
}  // namespace hg

namespace {
// Physics parameters, tables and device buffers of a handle over the elements `els` (global
// ids, ascending, all of one order) of the mesh m: loc maps a global element to its index in
// this handle (-1: another handle), mort a global edge to its index among the p-non-conforming
// edges (-1: conforming).  A handle whose stream is already set shares it.
int setup_handle(hg_handle *h, const hg::Mesh &m, const hg_physics *phys, const hg_options *opt,
                 const std::vector<int> &els, const std::vector<int> &loc, const std::vector<int> &mort) {
    h->K = (int)els.size();
    h->N = m.el[els[0]].N;
    if (!hg::supported_order(h->N)) return set_err("hg_create: polynomial order not compiled (1..12)");
    h->P = h->N + 1;
    const int P = h->P, PP = P * P, K = h->K;
    h->nodes = (int64_t)K * PP;
    h->ndof = 4 * h->nodes;
    h->mixed = opt->precision == HG_PRECISION_MIXED;
    h->opt = *opt;
    h->phys = *phys;
    // PhysicsNS.f90:405-481 (air)
    hg::PhysParams &ph = h->ph;
    ph.gamma = 1.4; ph.gm1 = 0.4; ph.cp = 3.5; ph.cv = 2.5; ph.lambda = -2.0 / 3.0;
    ph.invSGM = 1.0 / (std::sqrt(1.4) * phys->mach);
    ph.mach_sqrtg = phys->mach * std::sqrt(1.4);
    ph.mu = phys->reynolds == 0.0 ? 0.0 : 1.0 / phys->reynolds;
    ph.cpPr = 3.5 / phys->prandtl;
    ph.sigma0 = opt->sigma0;
    ph.pref = 0.0;
    ph.Eref = 0.0;
    ph.riemann = opt->riemann;
    ph.viscous = opt->navier_stokes ? opt->viscous : 0;
    ph.ns = opt->navier_stokes ? 1 : 0;
    if (opt->ip_method == HG_IP_SIPG) { ph.ip_eps = -1.0; ph.ip_gradflux = 1; }
    else if (opt->ip_method == HG_IP_NIPG) { ph.ip_eps = 1.0; ph.ip_gradflux = 1; }
    else { ph.ip_eps = 0.0; ph.ip_gradflux = 0; }   // IIPG (DGViscousMethods.f90:362-378)
    if (ph.viscous != 2) ph.ip_gradflux = 0;
    // the stage kernel compiled for the interior solver, unless a zone asks for another one
    ph.rs_kernel = opt->riemann;
    for (const auto &Z : m.zones) {
        if (Z.riemann >= 0 && Z.riemann != opt->riemann) ph.rs_kernel = -1;
        if (Z.type == HG_BC_EULERWALL || Z.type == HG_BC_DIRICHLET || Z.type == HG_BC_PRESSUREOUTLET ||
            Z.type == HG_BC_PRESSUREINLET || Z.type == HG_BC_NEWDIRICHLET)
            ph.xbc = 1;
    }
    if (opt->artificial_dissipation != HG_AD_NONE && opt->artificial_dissipation != HG_AD_PHYSICAL &&
        opt->artificial_dissipation != HG_AD_LAPLACIAN)
        return set_err("hg_create: unknown artificial dissipation type");
    // DGArtificialDissipation is a Navier-Stokes module
    ph.ad = (ph.ns && opt->artificial_dissipation != HG_AD_NONE) ? 1 : 0;
    h->lap = ph.ad && opt->artificial_dissipation == HG_AD_LAPLACIAN;
    ph.ad_ceps = opt->ad_intensity;
    if (ph.ad) ph.xbc = 1;
    // the run-time-solver form and the forms with those kinds exist for the wide tile only
    if (ph.rs_kernel < 0 || ph.xbc) h->tile[0] = 1;

    HG_CHECK(hipGetDevice(&h->device));
    if (hg::noalloc_setup(h->device)) return 1;
    if (!h->stream) HG_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    // ---- host-side tables
    const int E = (int)m.ed.size();
    std::vector<int> geo_off(E);
    std::vector<double> egeo;
    std::vector<double> einvh(E);
    for (int k = 0; k < E; ++k) {
        const hg::Edge &ed = m.ed[k];
        geo_off[k] = (int)(egeo.size() / 3);
        int np = ed.curved ? ed.N + 1 : 1;
        for (int i = 0; i < np; ++i) {
            egeo.push_back(ed.n[2 * i]);
            egeo.push_back(ed.n[2 * i + 1]);
            egeo.push_back(ed.dS[i]);
        }
        einvh[k] = ed.invh;
    }
    // a subdivided edge's small sides see their mortar's normal and half the edge's length (the
    // BR1 gradient lift of a straight edge, DGViscousMethods.f90:693-698)
    std::vector<int> half_off(E, -1);
    for (int k = 0; k < E; ++k) {
        const hg::Edge &ed = m.ed[k];
        if (ed.hang < 0) continue;
        half_off[k] = (int)(egeo.size() / 3);
        egeo.push_back(ed.n[0]);
        egeo.push_back(ed.n[1]);
        egeo.push_back(0.5 * ed.dS[0]);
    }
    std::vector<int4> face(4 * (size_t)K);
    std::vector<int> forder(K);
    std::vector<double> zone(8 * std::max<size_t>(m.zones.size(), 1), 0.0);
    for (size_t z = 0; z < m.zones.size(); ++z) {
        const hg::ZoneDesc &Z = m.zones[z];
        if (Z.type == HG_BC_VISCOUSWALL) { zone[8 * z] = Z.Tw; zone[8 * z + 1] = Z.vw[0]; zone[8 * z + 2] = Z.vw[1]; }
        else for (int c = 0; c < 4; ++c) zone[8 * z + c] = Z.q[c];   // the zone state
    }
    bool has_mortar = false;
    for (int e = 0; e < K; ++e) {
        const hg::Elem &el = m.el[els[e]];
        int ord[4] = {0, 1, 2, 3};
        std::sort(ord, ord + 4, [&](int a, int b) { return el.edges[a] < el.edges[b]; });
        forder[e] = ord[0] | (ord[1] << 2) | (ord[2] << 4) | (ord[3] << 6);
        for (int f = 0; f < 4; ++f) {
            const int k = el.edges[f];
            const hg::Edge &ed = m.ed[k];
            const int side = el.qpos[f];
            int flags = 0, y = 0;
            int goff = geo_off[k];
            if (!ed.bdry && mort[k] >= 0) {
                flags = hg::FK_MORTAR;   // evaluated by the mortar kernels
                y = mort[k];
                has_mortar = true;
                if (ed.hang >= 0 && side > 0) goff = half_off[k];
                // a curved mortar edge returns the BR1 lift terms per direction (bit 8 + f)
                if (ed.curved) {
                    forder[e] |= 1 << (8 + f);
                    flags |= hg::FF_MCURVED;
                }
            } else if (!ed.bdry) {
                flags = side == 0 ? hg::FK_INT_L : hg::FK_INT_R;
                int other = loc[ed.quads[1 - side]];
                y = other * 4 + (ed.loc[1 - side] - 1);
                if (ed.inverse) flags |= hg::FF_REV;
                if (side == 1 && ed.inverse) flags |= hg::FF_SELFREV;
            } else {
                const hg::ZoneDesc &Z = m.zones[ed.marker - 1];
                if (Z.type == HG_BC_PERIODIC) {
                    const hg::Edge &pe = m.ed[ed.partner];
                    flags = hg::FK_PERIODIC;
                    if (loc[pe.quads[0]] < 0)
                        return set_err("hg_create: periodic edges between elements of different orders");
                    y = loc[pe.quads[0]] * 4 + (pe.loc[0] - 1);
                    if (ed.inverse) flags |= hg::FF_REV;
                } else {
                    switch (Z.type) {
                    case HG_BC_FARFIELD: flags = hg::FK_FARFIELD; break;
                    case HG_BC_VISCOUSWALL: flags = hg::FK_WALL | (Z.isothermal ? hg::FF_ISOTHERMAL : 0); break;
                    case HG_BC_EULERWALL: flags = hg::FK_EULERWALL; break;
                    case HG_BC_DIRICHLET: flags = hg::FK_DIRICHLET; break;
                    case HG_BC_PRESSUREOUTLET:
                        flags = hg::FK_POUTLET | (Z.isothermal == HG_OUTFLOW_REFLECTIVE ? hg::FF_REFLECTIVE : 0);
                        break;
                    case HG_BC_NEWDIRICHLET:
                        flags = hg::FK_NEWDIRICHLET |
                                (Z.nd_mode == HG_NEWDIRICHLET_SPECIFY_TOTAL_PRESSURE ? hg::FF_TOTAL_INFLOW : 0) |
                                (Z.nd_outflow == HG_NEWDIRICHLET_TOTAL_PRESSURE ? hg::FF_TOTAL_OUTFLOW : 0);
                        break;
                    default: flags = hg::FK_PINLET; break;
                    }
                    y = ed.marker - 1;
                    if (!ph.ns && Z.type == HG_BC_VISCOUSWALL)
                        return set_err("hg_create: viscous walls need the Navier-Stokes form");
                }
                if (Z.riemann >= 0) flags |= (Z.riemann + 1) << hg::FF_RS_SHIFT;
            }
            // the mortar kernels hold a curved subdivided edge's geometry; its faces keep the
            // edge's first node's (read by no kernel: the BR1 form refuses such edges)
            if (ed.curved && ed.hang < 0) flags |= hg::FF_CURVED;
            face[4 * (size_t)e + f] = make_int4(flags, y, goff, k);
        }
    }
    // metric planes and per-face-node metric for the IP gradient face terms
    std::vector<double> metric(5 * (size_t)h->nodes), jac(h->nodes), dxe(K);
    std::vector<double> fmet;
    const hg::Spectral &sp = hg::spectral(h->N);
    if (ph.ip_gradflux) fmet.assign((size_t)K * 4 * P * 5, 0.0);
    for (int e = 0; e < K; ++e) {
        const hg::Elem &el = m.el[els[e]];
        for (int n = 0; n < PP; ++n) {
            size_t g = (size_t)e * PP + n;
            for (int c = 0; c < 4; ++c) metric[c * h->nodes + g] = el.Ja[c * PP + n];
            metric[4 * h->nodes + g] = 1.0 / el.jac[n];   // the kernels multiply by 1/jac
            jac[g] = el.jac[n];
        }
        double best = 1e300;
        for (int c = 0; c < 4; ++c) {
            int a = el.nodes[c], b = el.nodes[(c + 1) % 4];
            double ddx = m.pts[2 * a] - m.pts[2 * b], ddy = m.pts[2 * a + 1] - m.pts[2 * b + 1];
            best = std::fmin(best, std::sqrt(ddx * ddx + ddy * ddy));
        }
        dxe[e] = best;
        if (!fmet.empty())
            for (int f = 0; f < 4; ++f) {
                const double *lbv = (f == 0 || f == 3) ? sp.lb[0] : sp.lb[1];
                bool horiz = (f == 0 || f == 2);
                for (int k = 0; k < P; ++k) {
                    double v[5] = {0, 0, 0, 0, 0};
                    for (int l = 0; l < P; ++l) {
                        int n = horiz ? (k + P * l) : (l + P * k);
                        v[0] += el.Ja[0 * PP + n] * lbv[l];  // Ja(IX,IX) -> Jxi_x
                        v[1] += el.Ja[2 * PP + n] * lbv[l];  // Ja(IY,IX) -> Jxi_y
                        v[2] += el.Ja[1 * PP + n] * lbv[l];  // Ja(IX,IY) -> Jeta_x
                        v[3] += el.Ja[3 * PP + n] * lbv[l];  // Ja(IY,IY) -> Jeta_y
                        v[4] += el.jac[n] * lbv[l];
                    }
                    v[4] = 1.0 / v[4];
                    for (int c = 0; c < 5; ++c) fmet[(((size_t)e * 4 + f) * P + k) * 5 + c] = v[c];
                }
            }
    }
    std::vector<double> hat(PP), Dm(PP), lb(2 * P), lbw(2 * P), dlbw(2 * P), wv(P);
    for (int i = 0; i < P; ++i) {
        for (int l = 0; l < P; ++l) { hat[i * P + l] = sp.hatD[i][l]; Dm[i * P + l] = sp.D[i][l]; }
        for (int s = 0; s < 2; ++s) { lb[s * P + i] = sp.lb[s][i]; lbw[s * P + i] = sp.lbw[s][i]; dlbw[s * P + i] = sp.dlbw[s][i]; }
        wv[i] = sp.w[i];
    }
    // ---- device allocations
    HG_CHECK(h->alloc(&h->Q, h->ndof));
    HG_CHECK(h->alloc(&h->G, h->ndof));
    HG_CHECK(h->alloc(&h->QDot, h->ndof));
    const size_t tsz = (size_t)K * 4 * P * 4 * (h->mixed ? sizeof(float) : sizeof(double));
    HG_CHECK(h->alloc_raw(&h->tr[0], tsz));
    HG_CHECK(h->alloc_raw(&h->tr[1], tsz));
    if (ph.ns) {
        const size_t rs = h->mixed ? sizeof(float) : sizeof(double);
        // interior penalty forms the next stage's viscous traces in the stage epilogue, unless
        // they depend on the next stage's artificial viscosity (all elements' jumps first)
        // (the Laplacian type leaves mu alone: its viscous traces do not wait for the indicator)
        h->ip_fused = ph.viscous == HG_VISCOUS_IP && (!ph.ad || h->lap);
        for (int b = 0; b < (h->ip_fused ? 2 : 1); ++b)
            HG_CHECK(h->alloc_raw(&h->dtr[b], (size_t)K * 4 * P * hg::VT * rs));
    }
    // The FP32 form stores one metric record per element when every element's metric is
    // constant to FP32 resolution (straight-sided parallelograms): each node's value differs
    // from the element's first node by less than half an FP32 ulp of the element's largest
    // Ja component (1/jac: of its own magnitude), i.e. by less than the FP32 rounding the form
    // applies to the metric anyway.
    std::vector<double> metric_e;
    if (h->mixed) {
        const size_t nn = (size_t)h->nodes;
        bool constant = true;
        for (int e = 0; e < K && constant; ++e) {
            const size_t g0 = (size_t)e * PP;
            double scale = 0.0;
            for (int c = 0; c < 4; ++c)
                for (int n = 0; n < PP; ++n) scale = std::fmax(scale, std::fabs(metric[c * nn + g0 + n]));
            for (int c = 0; c < 5 && constant; ++c) {
                const double ref = metric[c * nn + g0];
                const double tol = std::ldexp(c < 4 ? scale : std::fabs(ref), -25);
                for (int n = 1; n < PP; ++n)
                    if (std::fabs(metric[c * nn + g0 + n] - ref) > tol) { constant = false; break; }
            }
        }
        if (constant) {
            metric_e.resize(5 * (size_t)K);
            for (int c = 0; c < 5; ++c)
                for (int e = 0; e < K; ++e) metric_e[(size_t)c * K + e] = metric[c * nn + (size_t)e * PP];
            h->metric_elem = 1;
        }
    }
    HG_CHECK(upload_real(h, &h->metric, h->metric_elem ? metric_e : metric, h->mixed));
    if (!fmet.empty()) HG_CHECK(upload_real(h, &h->fmet, fmet, h->mixed));
    HG_CHECK(upload(h, &h->face, face));
    HG_CHECK(upload(h, &h->forder, forder));
    HG_CHECK(upload(h, &h->egeo, egeo));
    HG_CHECK(upload(h, &h->einvh, einvh));
    {
        std::vector<double> earea(E);
        for (int k = 0; k < E; ++k) earea[k] = m.ed[k].area;
        HG_CHECK(upload(h, &h->earea, earea));
    }
    HG_CHECK(upload(h, &h->zone, zone));
    HG_CHECK(upload(h, &h->hatD, hat));
    HG_CHECK(upload(h, &h->D, Dm));
    HG_CHECK(upload(h, &h->lb, lb));
    HG_CHECK(upload(h, &h->lbw, lbw));
    HG_CHECK(upload(h, &h->dlbw, dlbw));
    HG_CHECK(upload(h, &h->w, wv));
    HG_CHECK(upload(h, &h->jac64, jac));
    if (ph.ad) {
        // element volumes (BilinearForm_F(jac, w, w), QuadMesh_Construct.incf:266)
        std::vector<double> vol(K);
        for (int e = 0; e < K; ++e) {
            double v = 0.0;
            for (int j = 0; j < P; ++j) {
                double acc = 0.0;
                for (int i = 0; i < P; ++i) acc = acc + wv[i] * jac[(size_t)e * PP + i + P * j];
                v = v + acc * wv[j];
            }
            vol[e] = v;
        }
        HG_CHECK(upload(h, &h->evol, vol));
        HG_CHECK(h->alloc(&h->mu_a, K));
        HG_CHECK(hipMemset(h->mu_a, 0, K * sizeof(double)));
    }
    HG_CHECK(upload(h, &h->dx, dxe));
    HG_CHECK(h->alloc(&h->umax, K));
    HG_CHECK(h->alloc(&h->amax, K));
    HG_CHECK(h->alloc(&h->dtstate, 2));
    HG_CHECK(h->alloc(&h->dtout, 1));
    HG_CHECK(hipMemset(h->dtstate, 0, 2 * sizeof(double)));
    h->cons_nb = (K + std::max(1, 256 / PP) - 1) / std::max(1, 256 / PP);
    HG_CHECK(h->alloc(&h->cons_partial, 4 * (size_t)h->cons_nb));
    HG_CHECK(h->alloc(&h->cons_out, 4));
    {
        // monitor surfaces: the boundary faces of every zone in ascending edge id
        std::vector<std::vector<int>> per(m.zones.size());
        for (size_t k = 0; k < m.ed.size(); ++k) {
            const hg::Edge &ed = m.ed[k];
            if (ed.bdry && ed.marker >= 1 && ed.marker <= (int)m.zones.size() && loc[ed.quads[0]] >= 0)
                per[ed.marker - 1].push_back(loc[ed.quads[0]] * 4 + (ed.loc[0] - 1));
        }
        std::vector<int> flat;
        h->bslot_off.assign(1, 0);
        for (const auto &v : per) {
            flat.insert(flat.end(), v.begin(), v.end());
            h->bslot_off.push_back((int)flat.size());
        }
        HG_CHECK(upload(h, &h->bslots, flat));
        HG_CHECK(h->alloc(&h->monitor_out, 12));
    }
    HG_CHECK(h->alloc(&h->dt_dev, 1));
    HG_CHECK(h->alloc(&h->resid, 4));
    HG_CHECK(h->alloc(&h->resid_part, 4 * (size_t)K));   // at least one element per workgroup
    HG_CHECK(h->alloc(&h->resid_scratch, 4 * (size_t)hg::kResidGroups));
    HG_CHECK(hipMemset(h->resid, 0, 4 * sizeof(double)));
    HG_CHECK(hipMemset(h->Q, 0, h->ndof * sizeof(double)));
    HG_CHECK(hipMemset(h->G, 0, h->ndof * sizeof(double)));
    HG_CHECK(hipMemset(h->QDot, 0, h->ndof * sizeof(double)));
    if (has_mortar) {
        // per face node: the projected flux (and for BR1 uStar and this side's gradient trace)
        const size_t fn = (size_t)K * 4 * P;
        HG_CHECK(h->alloc(&h->mflux, 4 * fn));
        HG_CHECK(hipMemset(h->mflux, 0, 4 * fn * sizeof(double)));
        if (ph.ns) {
            HG_CHECK(h->alloc(&h->mustar, 8 * fn));
            HG_CHECK(h->alloc(&h->mdq, 8 * fn));
            HG_CHECK(hipMemset(h->mustar, 0, 8 * fn * sizeof(double)));
            HG_CHECK(hipMemset(h->mdq, 0, 8 * fn * sizeof(double)));
            if (ph.ip_gradflux) {
                HG_CHECK(h->alloc(&h->mgrad, 8 * fn));
                HG_CHECK(hipMemset(h->mgrad, 0, 8 * fn * sizeof(double)));
            }
        }
    }
    HG_CHECK(hipDeviceSynchronize());
    return 0;
}


// A mesh with several polynomial orders (p-refinement zones, MeshFile.f90:143-180): one
// sub-handle per order over its elements, the p-non-conforming edges between them (edge order
// the higher of its two elements', QuadMesh_Construct.incf:104-118) in the mortar tables, and
// the whole mesh's time-step scan.  The FP64 and mixed forms (Euler, BR1 or interior penalty)
// without artificial dissipation; straight and curved mortar edges (curved ones under BR1 in the
// FP64 form only).  The mixed form's sub-handles share one reference pressure, so that the
// mortar kernels add one energy shift back to both sides' FP32 traces.
int create_composite(hg_handle *h, const hg_physics *phys, const hg_options *opt, const std::vector<int> &orders) {
    const hg::Mesh &m = h->mesh;
    if (opt->navier_stokes && opt->artificial_dissipation != HG_AD_NONE)
        return set_err("hg_create: a mesh with several polynomial orders or subdivided edges runs without artificial dissipation");
    if ((int)orders.size() > hg::kMaxSub) return set_err("hg_create: too many polynomial orders in one mesh");
    const int K = (int)m.el.size(), E = (int)m.ed.size();
    std::vector<int> group(K), loc(K);
    h->sub_els.assign(orders.size(), {});
    for (int e = 0; e < K; ++e) {
        group[e] = (int)(std::lower_bound(orders.begin(), orders.end(), m.el[e].N) - orders.begin());
        loc[e] = (int)h->sub_els[group[e]].size();
        h->sub_els[group[e]].push_back(e);
    }
    std::vector<int> mort(E, -1);
    std::vector<hg::MortarEdge> medges;
    std::vector<double> mats;
    std::vector<double> sgeo;   // per node (nx, ny, dS): curved p-non-conforming edges, subdivided edges' mortars
    for (int k = 0; k < E; ++k) {
        const hg::Edge &ed = m.ed[k];
        if (ed.bdry || ed.hang >= 0 || m.el[ed.quads[0]].N == m.el[ed.quads[1]].N) continue;
        hg::MortarEdge r{};
        for (int side = 0; side < 2; ++side) {
            const int q = ed.quads[side];
            r.sub[side] = group[q];
            r.slot[side] = loc[q] * 4 + (ed.loc[side] - 1);
            r.Ps[side] = m.el[q].P;
        }
        r.P = ed.N + 1;
        r.low = r.Ps[0] < r.Ps[1] ? 0 : 1;
        r.inverse = ed.inverse;
        r.nx = ed.n[0];
        r.ny = ed.n[1];
        r.dS = ed.dS[0];
        r.invh = ed.invh;
        r.geo = -1;
        if (ed.curved) {
            // the edge's normals and surface Jacobians at its N + 1 nodes (CurvedEdge_t)
            r.geo = (int)(sgeo.size() / 3);
            for (int i = 0; i <= ed.N; ++i) {
                sgeo.push_back(ed.n[2 * i]);
                sgeo.push_back(ed.n[2 * i + 1]);
                sgeo.push_back(ed.dS[i]);
            }
        }
        r.mat = (int)mats.size();
        const int Pl = r.Ps[r.low];
        mats.resize(mats.size() + 2 * (size_t)r.P * Pl);
        hg::mortar_matrices(Pl - 1, r.P - 1, mats.data() + r.mat, mats.data() + r.mat + (size_t)r.P * Pl);
        mort[k] = (int)medges.size();
        medges.push_back(r);
    }
    // subdivided edges (SubdividedEdge_LinkWithElements, QuadElement_EdgeProcedures.incf:201-355)
    std::vector<hg::SubMortarEdge> sedges;
    for (int k = 0; k < E; ++k) {
        const hg::Edge &ed = m.ed[k];
        if (ed.hang < 0) continue;
        const hg::SubEdge &d = *ed.sd;
        hg::SubMortarEdge r{};
        r.curved = ed.curved ? 1 : 0;
        r.invh = ed.invh;
        for (int side = 0; side < 3; ++side) {
            const int q = ed.quads[side];
            r.sub[side] = group[q];
            r.slot[side] = loc[q] * 4 + (ed.loc[side] - 1);
            r.Ps[side] = m.el[q].P;
        }
        r.PM[0] = d.NN + 1;
        r.PM[1] = d.NS + 1;
        r.mat = (int)mats.size();
        for (const std::vector<double> *t : {&d.TLN_F, &d.TLS_F, &d.TLN_B, &d.TLS_B, &d.TRN_F, &d.TRS_F, &d.TRN_B, &d.TRS_B})
            mats.insert(mats.end(), t->begin(), t->end());
        r.geo = (int)(sgeo.size() / 3);
        for (int h = 0; h < 2; ++h) {
            const std::vector<double> &nrm = h == 0 ? d.nN : d.nS, &dS = h == 0 ? d.dSN : d.dSS;
            for (int p = 0; p < r.PM[h]; ++p) {
                sgeo.push_back(nrm[2 * p]);
                sgeo.push_back(nrm[2 * p + 1]);
                sgeo.push_back(dS[p]);
            }
        }
        mort[k] = (int)(medges.size() + sedges.size());
        sedges.push_back(r);
    }
    HG_CHECK(hipGetDevice(&h->device));
    if (hg::noalloc_setup(h->device)) return 1;
    HG_CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (size_t g = 0; g < orders.size(); ++g) {
        hg_handle *s = new hg_handle();
        s->stream = h->stream;
        s->shared_stream = true;
        h->sub.push_back(s);
        std::vector<int> lg(K, -1);
        for (int e : h->sub_els[g]) lg[e] = loc[e];
        if (setup_handle(s, m, phys, opt, h->sub_els[g], lg, mort)) return 1;
    }
    if (opt->precision == HG_PRECISION_MIXED && opt->navier_stokes && opt->viscous == HG_VISCOUS_BR1) {
        // the y lift of a curved mortar face's uStar is added in the FP64 gradient only
        bool curved = false;
        for (const hg::MortarEdge &r : medges) curved = curved || r.geo >= 0;
        for (const hg::SubMortarEdge &r : sedges) curved = curved || r.curved;
        if (curved)
            return set_err("hg_create: curved mortar edges under BR1 run in the FP64 form only");
    }
    h->opt = *opt;
    h->phys = *phys;
    h->ph = h->sub[0]->ph;
    h->mixed = false;
    h->K = K;
    h->N = orders.back();
    h->P = h->N + 1;
    h->nodes = m.nodes_total;
    h->ndof = 4 * h->nodes;
    h->el_off.resize(K);
    for (int e = 0; e < K; ++e) h->el_off[e] = m.el[e].off;
    h->nmortar = (int)medges.size();
    h->nsubedge = (int)sedges.size();
    HG_CHECK(upload(h, &h->medges, medges));
    HG_CHECK(upload(h, &h->sedges, sedges));
    HG_CHECK(upload(h, &h->sgeo, sgeo));
    HG_CHECK(upload(h, &h->mmats, mats));
    for (size_t g = 0; g < orders.size(); ++g) HG_CHECK(upload(h, &h->sub_gmap[g], h->sub_els[g]));
    // the time-step scan over the whole mesh in element order: N + 1 and the shortest side per
    // element (as each sub-handle has them)
    std::vector<double> n1(K), dx(K);
    for (int e = 0; e < K; ++e) {
        const hg::Elem &el = m.el[e];
        n1[e] = (double)(el.N + 1);
        double best = 1e300;
        for (int c = 0; c < 4; ++c) {
            int a = el.nodes[c], b = el.nodes[(c + 1) % 4];
            double ddx = m.pts[2 * a] - m.pts[2 * b], ddy = m.pts[2 * a + 1] - m.pts[2 * b + 1];
            best = std::fmin(best, std::sqrt(ddx * ddx + ddy * ddy));
        }
        dx[e] = best;
    }
    HG_CHECK(upload(h, &h->n1e, n1));
    HG_CHECK(upload(h, &h->dx, dx));
    HG_CHECK(h->alloc(&h->umax, K));
    HG_CHECK(h->alloc(&h->amax, K));
    HG_CHECK(h->alloc(&h->dtstate, 2));
    HG_CHECK(h->alloc(&h->dtout, 1));
    HG_CHECK(h->alloc(&h->dt_dev, 1));
    HG_CHECK(hipMemset(h->dtstate, 0, 2 * sizeof(double)));
    HG_CHECK(hipDeviceSynchronize());
    return 0;
}

}  // namespace

extern "C" {

int hg_last_error(char *buf, int32_t n) {
    if (!buf || n <= 0) return 1;
    std::snprintf(buf, (size_t)n, "%s", g_err.c_str());
    return 0;
}

int hg_create(const hg_mesh_desc *mesh, const hg_physics *phys, const hg_options *opt, hg_handle **out) {
    if (!mesh || !phys || !opt || !out) return set_err("hg_create: null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err("hg_create: no HIP device available");
    if (opt->riemann < 0 || opt->riemann > 3) return set_err("hg_create: unknown Riemann solver");
    if (opt->navier_stokes && (opt->viscous < 1 || opt->viscous > 2))
        return set_err("hg_create: Navier-Stokes needs viscous = BR1 or IP");
    if (opt->precision != HG_PRECISION_FP64 && opt->precision != HG_PRECISION_MIXED)
        return set_err("hg_create: unknown precision");
    if (opt->scheme < 0 || opt->scheme > 2) return set_err("hg_create: unknown scheme");
    if (opt->navier_stokes && opt->viscous == HG_VISCOUS_IP &&
        (opt->ip_method < HG_IP_SIPG || opt->ip_method > HG_IP_IIPG))
        return set_err("hg_create: unknown interior penalty method (SIPG, NIPG or IIPG)");
    std::unique_ptr<hg_handle> h(new hg_handle());
    std::string err = hg::build_mesh(*mesh, h->mesh);
    if (!err.empty()) return set_err("hg_create: " + err);
    const hg::Mesh &m = h->mesh;
    std::vector<int> orders;
    for (const auto &e : m.el) orders.push_back(e.N);
    std::sort(orders.begin(), orders.end());
    orders.erase(std::unique(orders.begin(), orders.end()), orders.end());
    for (int N : orders)
        if (!hg::supported_order(N)) return set_err("hg_create: polynomial order not compiled (1..12)");
    // several orders or subdivided edges: sub-handles coupled by the mortar kernels
    bool subdivided = false;
    for (const auto &e : m.ed) subdivided = subdivided || e.hang >= 0;
    if (orders.size() > 1 || subdivided) {
        if (create_composite(h.get(), phys, opt, orders)) {
            const std::string msg = g_err;
            hg_destroy(h.release());   // the sub-handles made so far
            return set_err(msg);
        }
    } else {
        std::vector<int> all(m.el.size()), mort(m.ed.size(), -1);
        std::iota(all.begin(), all.end(), 0);
        if (setup_handle(h.get(), m, phys, opt, all, all, mort)) return 1;
    }
    *out = h.release();
    return 0;
}

int hg_destroy(hg_handle *h) {
    if (!h) return 0;
    if (h->stream) hipStreamSynchronize(h->stream);
    for (auto *s : h->sub) {
        s->stream = nullptr;   // shared with this handle
        hg_destroy(s);
    }
    h->sub.clear();
    drop_graphs(h);
    for (auto &b : h->bufs) hipFree(b.p);
    for (void *p : {(void *)h->probe_ev, (void *)h->probe_l, (void *)h->probe_out})
        if (p) hipFree(p);
    for (auto &e : h->ev_pool) hipEventDestroy(e);
    for (auto &p : h->ev_pending) { hipEventDestroy(p.second.first); hipEventDestroy(p.second.second); }
    if (h->stream && !h->shared_stream) hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

int64_t hg_ndof(const hg_handle *h) { return h ? h->ndof : -1; }

int hg_get_node_coords(hg_handle *h, double *X, int64_t nnodes) {
    if (!h || !X) return set_err("hg_get_node_coords: null argument");
    if (nnodes != h->nodes) return set_err("hg_get_node_coords: size mismatch");
    for (const auto &el : h->mesh.el)
        std::memcpy(X + 2 * el.off, el.X.data(), sizeof(double) * el.X.size());
    return 0;
}

int hg_get_connectivity(hg_handle *h, int32_t *edges, int32_t *elems, int32_t *partner) {
    if (!h) return set_err("hg_get_connectivity: null handle");
    const hg::Mesh &m = h->mesh;
    for (size_t k = 0; k < m.ed.size(); ++k) {
        const hg::Edge &ed = m.ed[k];
        if (edges) {
            int32_t *o = edges + 8 * k;
            o[0] = ed.nodes[0]; o[1] = ed.nodes[1]; o[2] = ed.quads[0]; o[3] = ed.quads[1];
            o[4] = ed.loc[0]; o[5] = ed.bdry ? 0 : ed.loc[1]; o[6] = ed.inverse; o[7] = ed.marker;
        }
        if (partner) partner[k] = ed.partner;
    }
    if (elems)
        for (size_t e = 0; e < m.el.size(); ++e)
            for (int c = 0; c < 4; ++c) {
                elems[12 * e + c] = m.el[e].edges[c];
                elems[12 * e + 4 + c] = m.el[e].dir[c];
                elems[12 * e + 8 + c] = m.el[e].qpos[c];
            }
    return 0;
}

int hg_num_edges(hg_handle *h) { return h ? (int)h->mesh.ed.size() : -1; }

int hg_build_connectivity(const hg_mesh_desc *mesh, int32_t *nedges, int32_t *edges, int32_t *elems,
                          int32_t *partner) {
    if (!mesh || !nedges) return set_err("hg_build_connectivity: null argument");
    hg::Mesh m;
    std::string err = hg::build_mesh(*mesh, m);
    if (!err.empty()) return set_err("hg_build_connectivity: " + err);
    *nedges = (int32_t)m.ed.size();
    if (!edges) return 0;
    for (size_t k = 0; k < m.ed.size(); ++k) {
        const hg::Edge &ed = m.ed[k];
        int32_t *o = edges + 8 * k;
        o[0] = ed.nodes[0]; o[1] = ed.nodes[1]; o[2] = ed.quads[0]; o[3] = ed.quads[1];
        o[4] = ed.loc[0]; o[5] = ed.bdry ? 0 : ed.loc[1]; o[6] = ed.inverse; o[7] = ed.marker;
        if (partner) partner[k] = ed.partner;
    }
    if (elems)
        for (size_t e = 0; e < m.el.size(); ++e)
            for (int c = 0; c < 4; ++c) {
                elems[12 * e + c] = m.el[e].edges[c];
                elems[12 * e + 4 + c] = m.el[e].dir[c];
                elems[12 * e + 8 + c] = m.el[e].qpos[c];
            }
    return 0;
}

int hg_build_geometry(const hg_mesh_desc *mesh, int64_t nnodes, double *X, double *Ja, double *jac) {
    if (!mesh) return set_err("hg_build_geometry: null argument");
    hg::Mesh m;
    std::string err = hg::build_mesh(*mesh, m);
    if (!err.empty()) return set_err("hg_build_geometry: " + err);
    if (nnodes != m.nodes_total) return set_err("hg_build_geometry: node count mismatch");
    for (const auto &el : m.el) {
        const int PP = el.P * el.P;
        for (int n = 0; n < PP; ++n) {
            const int64_t g = el.off + n;
            if (X) { X[2 * g] = el.X[2 * n]; X[2 * g + 1] = el.X[2 * n + 1]; }
            if (Ja) for (int c = 0; c < 4; ++c) Ja[4 * g + c] = el.Ja[c * PP + n];
            if (jac) jac[g] = el.jac[n];
        }
    }
    return 0;
}

// The mean nodal pressure of a state (0 when it is not finite).
static double mean_pressure(double gm1, const double *Q, int64_t nodes) {
    double sum = 0.0;
    for (int64_t n = 0; n < nodes; ++n) {
        const double *q = Q + 4 * n;
        sum += gm1 * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / q[0]);
    }
    const double pref = nodes > 0 ? sum / (double)nodes : 0.0;
    return std::isfinite(pref) ? pref : 0.0;
}

int hg_upload_state(hg_handle *h, const double *Q, int64_t ndof) {
    if (!h || !Q) return set_err("hg_upload_state: null argument");
    if (ndof != h->ndof) return set_err("hg_upload_state: ndof mismatch");
    if (!h->sub.empty()) {
        // each order's elements, from the reference layout, to its sub-handle (the mixed form's
        // reference pressure over the whole mesh, common to all of them)
        if (h->sub[0]->mixed) {
            const double pref = mean_pressure(h->sub[0]->ph.gm1, Q, h->nodes);
            // the captured mortar kernels carry the energy shift by value
            if (pref != h->sub[0]->ph.pref) drop_graphs(h);
            for (hg_handle *s : h->sub) {
                s->has_pref = true;
                s->pref_in = pref;
            }
        }
        for (size_t g = 0; g < h->sub.size(); ++g) {
            hg_handle *s = h->sub[g];
            const size_t n = 4 * (size_t)s->P * s->P;
            std::vector<double> q(s->ndof);
            for (size_t e = 0; e < h->sub_els[g].size(); ++e)
                std::memcpy(q.data() + e * n, Q + 4 * h->el_off[h->sub_els[g][e]], n * sizeof(double));
            if (hg_upload_state(s, q.data(), s->ndof)) return 1;
        }
        h->cur = 0;
        return 0;
    }
    const double pref_old = h->ph.pref;
    if (h->mixed) {
        // Reference pressure of the FP32 form: the mean nodal pressure of the uploaded state
        // (of the whole mesh, for a sub-handle of a mesh with several orders).
        h->ph.pref = h->has_pref ? h->pref_in : mean_pressure(h->ph.gm1, Q, h->nodes);
        h->ph.Eref = h->ph.pref / h->ph.gm1;
    }
    // the captured kernels carry the physics parameters (the FP32 form's reference pressure)
    // by value: recapture only if it changed
    if (h->ph.pref != pref_old) drop_graphs(h);
    HG_CHECK(hipMemcpyAsync(h->Q, Q, ndof * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hg::StageArgs a = base_args(h);
    h->cur = 0;
    a.tr_out = h->tr[0];
    HG_CHECK(hg::launch_prolong(h->P, h->mixed, a, h->stream));
    h->vtr_valid = false;
    h->qdot_valid = false;
    h->mon_tr = -1;
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

// The per-order arrays of a mesh with several orders into the reference layout.
static int composite_gather(hg_handle *h, bool qdot, double *out) {
    for (size_t g = 0; g < h->sub.size(); ++g) {
        hg_handle *s = h->sub[g];
        const size_t n = 4 * (size_t)s->P * s->P;
        std::vector<double> q(s->ndof);
        if ((qdot ? hg_download_qdot(s, q.data(), s->ndof) : hg_download_state(s, q.data(), s->ndof))) return 1;
        for (size_t e = 0; e < h->sub_els[g].size(); ++e)
            std::memcpy(out + 4 * h->el_off[h->sub_els[g][e]], q.data() + e * n, n * sizeof(double));
    }
    return 0;
}

int hg_download_state(hg_handle *h, double *Q, int64_t ndof) {
    if (!h || !Q) return set_err("hg_download_state: null argument");
    if (ndof != h->ndof) return set_err("hg_download_state: ndof mismatch");
    if (!h->sub.empty()) return composite_gather(h, false, Q);
    HG_CHECK(hipMemcpyAsync(Q, h->Q, ndof * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_download_qdot(hg_handle *h, double *QDot, int64_t ndof) {
    if (!h || !QDot) return set_err("hg_download_qdot: null argument");
    if (ndof != h->ndof) return set_err("hg_download_qdot: ndof mismatch");
    if (!h->sub.empty()) return composite_gather(h, true, QDot);
    HG_CHECK(hipMemcpyAsync(QDot, h->QDot, ndof * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_time_derivative(hg_handle *h, double t) {
    (void)t;  // no time-dependent boundary data on the GPU path
    if (!h) return set_err("hg_time_derivative: null handle");
    if (!h->sub.empty()) {
        if (composite_stage(h, 0.0, 1.0, 0.0, true, true, 1)) return 1;
        for (auto *s : h->sub) {
            s->qdot_valid = true;
            hg::mark_last_evaluation(s, s->cur);
        }
        return 0;
    }
    if (run_stage(h, 0.0, 1.0, 0.0, true, true, 1)) return 1;
    h->qdot_valid = true;
    hg::mark_last_evaluation(h, h->cur);
    return 0;
}

int hg_rk_step(hg_handle *h, double t, double dt) {
    (void)t;  // no time-dependent boundary data on the GPU path
    if (!h) return set_err("hg_rk_step: null handle");
    if (!h->sub.empty()) return composite_rk_step(h, dt);
    h->qdot_valid = false;
    const StepCoefs k = step_coefs(h->opt.scheme);
    if (h->use_graph && !h->prof) {
        // replay the captured step (per-kernel event timing uses the direct path below)
        if (autotune(h)) return 1;
        if (hg::ensure_vtr(h)) return 1;   // the captured step holds no gradient launch for IP
        if (!h->dt_valid || dt != h->dt_last) {
            h->dt_last = dt;
            h->dt_valid = true;
            HG_CHECK(hipMemcpyAsync(h->dt_dev, &h->dt_last, sizeof(double), hipMemcpyHostToDevice, h->stream));
        }
        if (!h->gexec[0] && build_graphs(h, k)) {
            drop_graphs(h);
            return 1;
        }
        HG_CHECK(hipGraphLaunch(h->gexec[h->cur], h->stream));
        if (k.nst & 1) h->cur ^= 1;
        hg::mark_last_evaluation(h, h->cur ^ 1);   // the last stage read the other buffer
        return 0;
    }
    for (int s = 0; s < k.nst; ++s) {
        const double cq = k.dt_cq ? k.cq[s] * dt : k.cq[s];
        const double cg = k.dt_cg ? k.cg[s] * dt : k.cg[s];
        if (run_stage(h, k.am[s], cq, cg, s == 0, s == k.nst - 1, 0)) return 1;
    }
    hg::mark_last_evaluation(h, h->cur ^ 1);
    return 0;
}

int hg_set_graph(hg_handle *h, int32_t on) {
    if (!h) return set_err("hg_set_graph: null handle");
    h->use_graph = on != 0;
    return 0;
}

int hg_residuals(hg_handle *h, double res[4]) {
    if (!h || !res) return set_err("hg_residuals: null argument");
    if (!h->sub.empty()) {
        // max over the orders (a NaN stays)
        for (int c = 0; c < 4; ++c) res[c] = 0.0;
        for (auto *s : h->sub) {
            double r[4];
            if (hg_residuals(s, r)) return 1;
            for (int c = 0; c < 4; ++c)
                if (r[c] > res[c] || std::isnan(r[c])) res[c] = std::isnan(res[c]) ? res[c] : r[c];
        }
        return 0;
    }
    HG_CHECK(hipMemcpyAsync(res, h->resid, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_estimate_dt(hg_handle *h, double dt_max, double cfl, double *dt) {
    if (!h || !dt) return set_err("hg_estimate_dt: null argument");
    const double coef = h->phys.mach * std::sqrt(h->ph.gamma) * cfl;
    if (!h->sub.empty()) {
        // per-element maxima of each order, into mesh order, then one scan over all elements
        for (size_t g = 0; g < h->sub.size(); ++g) {
            hg_handle *s = h->sub[g];
            HG_CHECK(hg::launch_dt_elem(s->P, s->Q, s->K, s->ph.gamma, s->ph.gm1, s->umax, s->amax, h->stream));
            HG_CHECK(hg::launch_scatter(s->umax, h->sub_gmap[g], s->K, h->umax, h->stream));
            HG_CHECK(hg::launch_scatter(s->amax, h->sub_gmap[g], s->K, h->amax, h->stream));
        }
        HG_CHECK(hg::launch_dt_scan(h->K, h->umax, h->amax, h->dx, coef, 0.0, dt_max, h->dtstate, h->dtout, h->stream,
                                    h->n1e));
    } else {
        HG_CHECK(hg::launch_dt_elem(h->P, h->Q, h->K, h->ph.gamma, h->ph.gm1, h->umax, h->amax, h->stream));
        HG_CHECK(hg::launch_dt_scan(h->K, h->umax, h->amax, h->dx, coef, (double)(h->N + 1), dt_max, h->dtstate,
                                    h->dtout, h->stream));
    }
    HG_CHECK(hipMemcpyAsync(dt, h->dtout, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_reset_dt_state(hg_handle *h) {
    if (!h) return set_err("hg_reset_dt_state: null handle");
    HG_CHECK(hipMemsetAsync(h->dtstate, 0, 2 * sizeof(double), h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_conservation_sums(hg_handle *h, double out[4]) {
    if (!h || !out) return set_err("hg_conservation_sums: null argument");
    if (!h->sub.empty()) {
        for (int c = 0; c < 4; ++c) out[c] = 0.0;
        for (auto *s : h->sub) {
            double v[4];
            if (hg_conservation_sums(s, v)) return 1;
            for (int c = 0; c < 4; ++c) out[c] += v[c];
        }
        return 0;
    }
    HG_CHECK(hg::launch_cons_sums(h->P, h->Q, h->jac64, h->w, h->K, h->cons_partial, h->cons_nb, h->cons_out, h->stream));
    HG_CHECK(hipMemcpyAsync(out, h->cons_out, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_volume_integrals(hg_handle *h, double out[2]) {
    if (!h || !out) return set_err("hg_volume_integrals: null argument");
    if (!h->sub.empty()) {
        double one = 0.0, sq = 0.0;
        for (auto *s : h->sub) {
            double v[2];
            if (hg_volume_integrals(s, v)) return 1;
            one += v[0];
            sq += v[1] * v[1];
        }
        out[0] = one;
        out[1] = std::sqrt(sq);
        return 0;
    }
    HG_CHECK(hg::launch_volume_integrals(h->P, h->Q, h->jac64, h->w, h->K, h->ph.gamma, h->ph.gm1, h->cons_partial,
                                         h->cons_nb, h->monitor_out, h->stream));
    double v[4];
    HG_CHECK(hipMemcpyAsync(v, h->monitor_out, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    out[0] = v[0];
    out[1] = std::sqrt(v[1]);
    return 0;
}

// ---- boundary states from problem-file functions
// The zone's boundary edges in ascending edge id (the node order of the boundary-state calls).
static std::vector<int> zone_edges(const hg::Mesh &m, int marker) {
    std::vector<int> r;
    for (size_t k = 0; k < m.ed.size(); ++k)
        if (m.ed[k].bdry && m.ed[k].marker == marker) r.push_back((int)k);
    return r;
}

int hg_boundary_node_count(hg_handle *h, int32_t marker, int64_t *n) {
    if (!h || !n) return set_err("hg_boundary_node_count: null argument");
    if (marker < 1 || marker > (int)h->mesh.zones.size()) return set_err("hg_boundary_node_count: marker out of range");
    int64_t c = 0;
    for (int k : zone_edges(h->mesh, marker)) c += h->mesh.ed[k].N + 1;
    *n = c;
    return 0;
}

int hg_boundary_nodes(hg_handle *h, int32_t marker, int64_t n, double *x) {
    int64_t c = 0;
    if (hg_boundary_node_count(h, marker, &c)) return 1;
    if (!x || n != c) return set_err("hg_boundary_nodes: null argument or node count mismatch");
    int64_t o = 0;
    for (int k : zone_edges(h->mesh, marker)) {
        const hg::Edge &ed = h->mesh.ed[k];
        for (int i = 0; i <= ed.N; ++i, ++o) { x[2 * o] = ed.X[2 * i]; x[2 * o + 1] = ed.X[2 * i + 1]; }
    }
    return 0;
}

int hg_set_boundary_state(hg_handle *h, int32_t marker, int64_t n, const double *q) {
    int64_t c = 0;
    if (hg_boundary_node_count(h, marker, &c)) return 1;
    if (q && n != c) return set_err("hg_set_boundary_state: node count mismatch");
    const hg::Mesh &m = h->mesh;
    const int type = m.zones[marker - 1].type;
    if (type != HG_BC_FARFIELD && type != HG_BC_DIRICHLET && type != HG_BC_NEWDIRICHLET &&
        type != HG_BC_PRESSUREOUTLET && type != HG_BC_PRESSUREINLET)
        return set_err("hg_set_boundary_state: the zone's kind has no boundary state");
    if (q)
        for (int64_t i = 0; i < 4 * n; ++i)
            if (!std::isfinite(q[i])) return set_err("hg_set_boundary_state: non-finite state");
    // the handle (or sub-handle) and local element of every mesh element
    const int K = (int)m.el.size();
    std::vector<hg_handle *> owner(K, h);
    std::vector<int> loc(K);
    if (h->sub.empty()) {
        std::iota(loc.begin(), loc.end(), 0);
    } else {
        for (size_t g = 0; g < h->sub.size(); ++g)
            for (size_t e = 0; e < h->sub_els[g].size(); ++e) {
                owner[h->sub_els[g][e]] = h->sub[g];
                loc[h->sub_els[g][e]] = (int)e;
            }
    }
    std::vector<hg_handle *> hs = h->sub.empty() ? std::vector<hg_handle *>{h} : h->sub;
    for (hg_handle *s : hs) HG_CHECK(hipStreamSynchronize(s->stream));
    // per target handle: its face records and (when states are set) its per-face-node states
    std::vector<std::vector<int4>> faces(hs.size());
    std::vector<std::vector<double>> fst(hs.size());
    for (size_t g = 0; g < hs.size(); ++g) {
        hg_handle *s = hs[g];
        faces[g].resize(4 * (size_t)s->K);
        HG_CHECK(hipMemcpy(faces[g].data(), s->face, faces[g].size() * sizeof(int4), hipMemcpyDeviceToHost));
        const size_t fn = (size_t)s->K * 4 * s->P * 4;
        fst[g].assign(fn, 0.0);
        if (s->fstate) HG_CHECK(hipMemcpy(fst[g].data(), s->fstate, fn * sizeof(double), hipMemcpyDeviceToHost));
    }
    int64_t o = 0;
    for (int k : zone_edges(m, marker)) {
        const hg::Edge &ed = m.ed[k];
        const int e = ed.quads[0], f = ed.loc[0] - 1;
        hg_handle *s = owner[e];
        const size_t g = h->sub.empty() ? 0 : (size_t)(std::find(h->sub.begin(), h->sub.end(), s) - h->sub.begin());
        int4 &fr = faces[g][4 * (size_t)loc[e] + f];
        if (!q) {
            fr.x &= ~hg::FF_FSTATE;
            continue;
        }
        fr.x |= hg::FF_FSTATE;
        // a boundary face is its edge's LEFT side: face node i is edge node i
        for (int i = 0; i <= ed.N; ++i, ++o)
            for (int v = 0; v < 4; ++v) fst[g][(((size_t)loc[e] * 4 + f) * s->P + i) * 4 + v] = q[4 * o + v];
    }
    for (size_t g = 0; g < hs.size(); ++g) {
        hg_handle *s = hs[g];
        HG_CHECK(hipMemcpy(s->face, faces[g].data(), faces[g].size() * sizeof(int4), hipMemcpyHostToDevice));
        if (q) {
            if (!s->fstate) HG_CHECK(s->alloc(&s->fstate, fst[g].size()));
            HG_CHECK(hipMemcpy(s->fstate, fst[g].data(), fst[g].size() * sizeof(double), hipMemcpyHostToDevice));
        }
        drop_graphs(s);   // the captured kernels hold the state pointer by value
        s->qdot_valid = false;
        if (s->ip_fused) s->vtr_valid = false;   // the epilogue's viscous traces used the old states
    }
    drop_graphs(h);
    h->qdot_valid = false;
    return 0;
}

int hg_surface_integrals(hg_handle *h, int32_t marker, double out[9]) {
    if (!h || !out) return set_err("hg_surface_integrals: null argument");
    if (!h->sub.empty()) {
        for (int c = 0; c < 9; ++c) out[c] = 0.0;
        for (auto *s : h->sub) {
            double v[9];
            if (hg_surface_integrals(s, marker, v)) return 1;
            for (int c = 0; c < 9; ++c) out[c] += v[c];
        }
        return 0;
    }
    const int nz = (int)h->bslot_off.size() - 1;
    if (marker < 1 || marker > nz) return set_err("hg_surface_integrals: marker out of range");
    const int first = h->bslot_off[marker - 1], n = h->bslot_off[marker] - first;
    const void *vtr = nullptr, *tr = h->tr[h->cur];
    if (h->mon_tr >= 0) {
        // the edge storage of the last derivative evaluation
        tr = h->tr[h->mon_tr];
        vtr = h->mon_vtr;
    } else if (h->ph.ns) {
        // no evaluation since the upload: the viscous face-flux traces of the uploaded state
        if (autotune(h)) return 1;
        if (h->ip_fused) {
            if (hg::ensure_vtr(h)) return 1;
        } else {
            hg::StageArgs g = base_args(h);
            g.tr_in = h->tr[h->cur];
            if (h->ph.ad && !h->lap)
                HG_CHECK(hg::launch_art_visc(h->P, h->mixed, g, h->evol, h->w, h->mu_a, h->stream));
            HG_CHECK(hg::launch_gradient(h->P, h->tile[1], h->mixed, g, h->stream));
        }
        vtr = hg::vtr_current(h);
        hg::mark_last_evaluation(h, h->cur);
    } else {
        hg::mark_last_evaluation(h, h->cur);
    }
    if (n > 0)
        HG_CHECK(hg::launch_surface_integrals(h->mixed, h->bslots + first, n, h->P, tr, vtr, h->face,
                                              h->egeo, h->w, h->ph.gm1, h->mixed ? h->ph.Eref : 0.0,
                                              (h->ph.ad && !h->lap) ? h->mu_a : nullptr, h->ph.mu, h->monitor_out,
                                              h->stream));
    else
        HG_CHECK(hipMemsetAsync(h->monitor_out, 0, 9 * sizeof(double), h->stream));
    HG_CHECK(hipMemcpyAsync(out, h->monitor_out, 9 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_probes_set(hg_handle *h, int32_t n, const int32_t *elem, const double *xi, const double *eta,
                  const int32_t *var) {
    if (!h) return set_err("hg_probes_set: null handle");
    if (n < 0) return set_err("hg_probes_set: negative probe count");
    if (n > 0 && (!elem || !xi || !eta || !var)) return set_err("hg_probes_set: null argument");
    if (!h->sub.empty()) {
        // each probe to the sub-handle of its element's order, with that handle's element index
        std::vector<int> group(h->K, -1), loc(h->K, -1);
        for (size_t g = 0; g < h->sub.size(); ++g)
            for (size_t e = 0; e < h->sub_els[g].size(); ++e) {
                group[h->sub_els[g][e]] = (int)g;
                loc[h->sub_els[g][e]] = (int)e;
            }
        for (int k = 0; k < n; ++k)
            if (elem[k] < 0 || elem[k] >= h->K) return set_err("hg_probes_set: element id out of range");
        h->sub_probes.assign(h->sub.size(), {});
        for (int k = 0; k < n; ++k) h->sub_probes[group[elem[k]]].push_back(k);
        for (size_t g = 0; g < h->sub.size(); ++g) {
            const auto &ix = h->sub_probes[g];
            std::vector<int32_t> el(ix.size()), vr(ix.size());
            std::vector<double> x(ix.size()), y(ix.size());
            for (size_t j = 0; j < ix.size(); ++j) {
                el[j] = loc[elem[ix[j]]]; vr[j] = var[ix[j]]; x[j] = xi[ix[j]]; y[j] = eta[ix[j]];
            }
            if (hg_probes_set(h->sub[g], (int32_t)ix.size(), el.data(), x.data(), y.data(), vr.data())) return 1;
        }
        h->nprobes = n;
        return 0;
    }
    const int P = h->P;
    std::vector<int2> ev(n);
    std::vector<double> l((size_t)n * 2 * P);
    const hg::Spectral &sp = hg::spectral(h->N);
    for (int k = 0; k < n; ++k) {
        if (elem[k] < 0 || elem[k] >= h->K) return set_err("hg_probes_set: element id out of range");
        if (var[k] < HG_PROBE_RHO || var[k] > HG_PROBE_MACH) return set_err("hg_probes_set: unknown probe variable");
        if (!std::isfinite(xi[k]) || !std::isfinite(eta[k])) return set_err("hg_probes_set: non-finite coordinate");
        ev[k] = make_int2(elem[k], var[k]);
        hg::lagrange_at(sp, xi[k], l.data() + (size_t)k * 2 * P);
        hg::lagrange_at(sp, eta[k], l.data() + (size_t)k * 2 * P + P);
    }
    HG_CHECK(hipStreamSynchronize(h->stream));   // earlier updates may still read the old set
    h->nprobes = 0;                               // no set until the new one is on the device
    h->probe_var.clear();
    if (n > h->probe_cap) {
        for (void **p : {(void **)&h->probe_ev, (void **)&h->probe_l, (void **)&h->probe_out})
            if (*p) {
                HG_CHECK(hipFree(*p));
                *p = nullptr;
            }
        h->probe_cap = 0;
        HG_CHECK(hipMalloc((void **)&h->probe_ev, sizeof(int2) * n));
        HG_CHECK(hipMalloc((void **)&h->probe_l, sizeof(double) * 2 * P * n));
        HG_CHECK(hipMalloc((void **)&h->probe_out, sizeof(double) * n));
        h->probe_cap = n;
    }
    if (n > 0) {
        HG_CHECK(hipMemcpy(h->probe_ev, ev.data(), sizeof(int2) * n, hipMemcpyHostToDevice));
        HG_CHECK(hipMemcpy(h->probe_l, l.data(), sizeof(double) * l.size(), hipMemcpyHostToDevice));
    }
    h->nprobes = n;
    h->probe_var.assign(var, var + n);
    return 0;
}

int hg_probes_update(hg_handle *h, double *out) {
    if (!h) return set_err("hg_probes_update: null handle");
    if (h->nprobes == 0) return 0;
    if (!out) return set_err("hg_probes_update: null argument");
    if (!h->sub.empty()) {
        for (size_t g = 0; g < h->sub.size(); ++g) {
            const auto &ix = h->sub_probes[g];
            if (ix.empty()) continue;
            std::vector<double> v(ix.size());
            if (hg_probes_update(h->sub[g], v.data())) return 1;
            for (size_t j = 0; j < ix.size(); ++j) out[ix[j]] = v[j];
        }
        return 0;
    }
    for (int v : h->probe_var)
        if (v >= HG_PROBE_RHOT && v <= HG_PROBE_RHOET && !h->qdot_valid)
            return set_err("hg_probes_update: rhot/rhout/rhovt/rhoet probes need hg_time_derivative after the "
                           "last step or upload");
    HG_CHECK(hg::launch_probes(h->Q, h->QDot, h->probe_ev, h->probe_l, h->nprobes, h->P, h->ph.gm1, h->ph.gamma,
                               h->probe_out, h->stream));
    HG_CHECK(hipMemcpyAsync(out, h->probe_out, sizeof(double) * h->nprobes, hipMemcpyDeviceToHost, h->stream));
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_synchronize(hg_handle *h) {
    if (!h) return set_err("hg_synchronize: null handle");
    HG_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int hg_set_tiles(hg_handle *h, int32_t stage_tile, int32_t gradient_tile) {
    if (!h) return set_err("hg_set_tiles: null handle");
    if (stage_tile < -1 || stage_tile > 1 || gradient_tile < -1 || gradient_tile > 1)
        return set_err("hg_set_tiles: tile must be -1 (autotune), 0 (narrow) or 1 (wide)");
    // only the wide tile has the run-time solver form and the forms with the newer boundary kinds
    if ((h->ph.rs_kernel < 0 || h->ph.xbc) && stage_tile != 1) stage_tile = 1;
    h->tile[0] = stage_tile;
    h->tile[1] = gradient_tile;
    drop_graphs(h);
    for (auto *s : h->sub)
        if (hg_set_tiles(s, stage_tile, gradient_tile)) return 1;
    return 0;
}

int hg_get_tiles(hg_handle *h, int32_t *stage_tile, int32_t *gradient_tile) {
    if (!h) return set_err("hg_get_tiles: null handle");
    if (!h->sub.empty()) return hg_get_tiles(h->sub.back(), stage_tile, gradient_tile);
    if (stage_tile) *stage_tile = h->tile[0];
    if (gradient_tile) *gradient_tile = h->tile[1];
    return 0;
}

int hg_profile_enable(hg_handle *h, int32_t on) {
    if (!h) return set_err("hg_profile_enable: null handle");
    for (auto *s : h->sub)
        if (hg_profile_enable(s, on)) return 1;
    h->prof = on != 0;
    h->prof_n[0] = h->prof_n[1] = 0;
    h->prof_ms[0] = h->prof_ms[1] = 0.0;
    return 0;
}

int hg_profile_read(hg_handle *h, int32_t which, int64_t *launches, double *ms) {
    if (!h || which < 0 || which > 1) return set_err("hg_profile_read: bad argument");
    if (!h->sub.empty()) {
        // the launches of every order together
        int64_t n = 0;
        double t = 0.0;
        for (auto *s : h->sub) {
            int64_t ns = 0;
            double ts = 0.0;
            if (hg_profile_read(s, which, &ns, &ts)) return 1;
            n += ns;
            t += ts;
        }
        if (launches) *launches = n;
        if (ms) *ms = t;
        return 0;
    }
    HG_CHECK(hipStreamSynchronize(h->stream));
    for (auto &p : h->ev_pending) {
        float t = 0.0f;
        HG_CHECK(hipEventElapsedTime(&t, p.second.first, p.second.second));
        h->prof_n[p.first] += 1;
        h->prof_ms[p.first] += t;
        h->ev_pool.push_back(p.second.first);
        h->ev_pool.push_back(p.second.second);
    }
    h->ev_pending.clear();
    if (launches) *launches = h->prof_n[which];
    if (ms) *ms = h->prof_ms[which];
    return 0;
}

}  // extern "C"
