// Batched form: many independent cases advanced together (host side).
//
// Every RK stage of every case is one entry of a per-case slot table (StageArgs with the
// case's buffers, the stage's RK coefficients and the trace buffers of the current parity).
// A work list holds one entry per element tile (case, first element, order, residual slot);
// one launch per stage evaluates all of them, and the step (all stages, the gradient passes
// and the per-case residual reduction) is captured once per trace parity into a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "hg_device.h"
#include "hg_handle.h"
#include "horses2d_gpu.h"

namespace {

#define HB_CHECK(call)                                                                                     \
    do {                                                                                                   \
        hipError_t _e = (call);                                                                            \
        if (_e != hipSuccess) return hg::handle_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Williamson coefficients (DGTimeIntegrator.f90:263-365), as in hg_rk_step.
struct Scheme {
    int nst;
    double am[5], cq[5], cg[5];
    int dt_cq, dt_cg;
};

Scheme scheme_of(int s) {
    Scheme r{};
    if (s == HG_SCHEME_RK3) {
        r.nst = 3;
        const double am[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0}, gm[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
        for (int i = 0; i < 3; ++i) { r.am[i] = am[i]; r.cq[i] = 1.0; r.cg[i] = gm[i]; }
        r.dt_cq = 0; r.dt_cg = 1;
    } else if (s == HG_SCHEME_RK5) {
        r.nst = 5;
        const double am[5] = {0.0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
        const double gm[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488, 0.1530572479681};
        for (int i = 0; i < 5; ++i) { r.am[i] = am[i]; r.cq[i] = 1.0; r.cg[i] = gm[i]; }
        r.dt_cq = 1; r.dt_cg = 0;
    } else {
        r.nst = 1;
        r.am[0] = 0.0; r.cq[0] = 1.0; r.cg[0] = 1.0;
        r.dt_cq = 0; r.dt_cg = 1;
    }
    return r;
}

}  // namespace

struct hg_batch {
    std::vector<hg_handle *> cases;
    hg_batch_options opt{};
    Scheme sch{};
    int variant = 0, nslots = 0;
    int rs = -1;                           // Riemann solver compiled into the stage kernel (-1: per face)
    hipStream_t stream = nullptr;
    std::vector<void *> bufs;
    hg::StageArgs *slots = nullptr;        // [case][nslots]
    std::vector<hg::StageArgs> hslots;     // host copy: the indicator launches take their args
    int4 *items[2] = {nullptr, nullptr};   // per tile
    int nitems[2] = {0, 0};
    int block[2] = {0, 0};
    int *sched[2][3] = {};                 // per tile, per order
    // One launch per polynomial order P (points per direction; schedule order 3).  Candidates:
    // per tile and P, the case-major positions in items[tile] of that order's work items.
    // Chosen: a tile per order (ptile), one work list mixing the two tiles' items with unique
    // residual slots (items_po, ranges_po, sched_po), and side streams on which the orders'
    // launches of one stage run concurrently between a fork and a join event.
    int *psched[2][12] = {};
    int pnitems[2][12] = {};
    std::vector<int> porders;              // the orders present, costliest first
    int ptile[12] = {};
    int4 *items_po = nullptr;
    int2 *ranges_po = nullptr;
    int *sched_po[12] = {};
    int nitems_po[12] = {};
    size_t max_items = 0;                  // capacity of resid_part, in work items
    hipStream_t side[4] = {};
    int nside = 0;
    hipEvent_t fork = nullptr, join[4] = {};
    int2 *ranges[2] = {nullptr, nullptr};  // per tile: per case residual partial range
    double **resid_out = nullptr;          // per case: the handle's residual slots
    double *dt = nullptr;                  // per case
    double *resid_part = nullptr;
    int tile = -1, order = -1;
    int parity = 0;
    // what the slot table was built from, per case: the trace parity (h->cur) at parity 0 and
    // the FP32 form's reference pressure (carried by value in every slot)
    std::vector<int> cur0;
    std::vector<double> pref0;
    std::vector<const double *> fst0;   // the boundary-function state buffer (hg_set_boundary_state)
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    hipGraph_t graph[2] = {nullptr, nullptr};
    // kernel timing (hg_batch_profile_enable): HIP events around every stage (0), gradient (1)
    // and residual (2) launch, on the batch stream; the step is then launched kernel by kernel
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    int64_t prof_n[3] = {0, 0, 0};
    double prof_ms[3] = {0.0, 0.0, 0.0};

    template <typename T>
    hipError_t upload(T **dst, const std::vector<T> &src) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<size_t>(src.size(), 1) * sizeof(T));
        if (e != hipSuccess) return e;
        bufs.push_back(p);
        *dst = (T *)p;
        return src.empty() ? hipSuccess : hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
    }
};

namespace {

constexpr int kPerOrder = 3;   // schedule order: one launch per polynomial order
constexpr int kPersist = 4;    // schedule order: persistent workgroups (one launch per stage)

hg::BatchArgs args_of(const hg_batch *b, int tile, int order, int slot) {
    hg::BatchArgs a{};
    a.slots = b->slots;
    a.nslots = b->nslots;
    a.slot = slot;
    a.items = b->items[tile];
    a.sched = b->sched[tile][order == kPerOrder ? 0 : order];
    a.nitems = b->nitems[tile];
    a.dt = b->dt;
    a.dt_cq = b->sch.dt_cq;
    a.dt_cg = b->sch.dt_cg;
    return a;
}

// The persistent form over one tile's list: each order's items by their positions.
hg::BatchArgs args_persist(const hg_batch *b, int tile, int slot) {
    hg::BatchArgs a = args_of(b, tile, 0, slot);
    a.persist = 1;
    for (int P = 0; P < 12; ++P) {
        a.pn[P] = b->pnitems[tile][P];
        a.psched[P] = b->psched[tile][P];
    }
    return a;
}

// A candidate of the autotuner: order P's items of one tile's list.
hg::BatchArgs args_of_order(const hg_batch *b, int tile, int P, int slot) {
    hg::BatchArgs a = args_of(b, tile, 0, slot);
    a.sched = b->psched[tile][P];
    a.nitems = b->pnitems[tile][P];
    return a;
}

// The chosen per-order schedule: order P's items of the mixed-tile list.
hg::BatchArgs args_po(const hg_batch *b, int P, int slot) {
    hg::BatchArgs a = args_of(b, 0, 0, slot);
    a.items = b->items_po;
    a.sched = b->sched_po[P];
    a.nitems = b->nitems_po[P];
    return a;
}

// One launch per order, each on a side stream: the side streams wait for what the batch stream
// holds (fork), the batch stream for every side stream (join).  Captured into the step's graph,
// the orders' kernels become parallel branches.
template <typename F>
hipError_t per_order(hg_batch *b, F &&launch) {
    hipError_t e = hipEventRecord(b->fork, b->stream);
    for (int k = 0; k < b->nside && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->side[k], b->fork, 0);
    for (size_t i = 0; i < b->porders.size() && e == hipSuccess; ++i)
        e = launch(b->porders[i], b->side[i % b->nside]);
    for (int k = 0; k < b->nside && e == hipSuccess; ++k) {
        e = hipEventRecord(b->join[k], b->side[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join[k], 0);
    }
    return e;
}

// The gradient pass and the stage pass of one slot: one launch each over the whole work list,
// or one per polynomial order (all gradient launches before the stage launches, so that each
// pass is timed as one interval).
hipError_t gradient_pass(hg_batch *b, int tile, int order, int slot) {
    if (order == kPersist)
        return hg::launch_batch_gradient(b->variant, tile, args_persist(b, tile, slot), b->block[tile], b->stream);
    if (order != kPerOrder)
        return hg::launch_batch_gradient(b->variant, tile, args_of(b, tile, order, slot), b->block[tile], b->stream);
    return per_order(b, [&](int P, hipStream_t s) {
        return hg::launch_batch_gradient_p(b->variant, P, b->ptile[P], args_po(b, P, slot), s);
    });
}

hipError_t stage_pass(hg_batch *b, int tile, int order, int slot) {
    if (order == kPersist)
        return hg::launch_batch_stage(b->variant, b->rs, tile, args_persist(b, tile, slot), b->block[tile], b->stream);
    if (order != kPerOrder)
        return hg::launch_batch_stage(b->variant, b->rs, tile, args_of(b, tile, order, slot), b->block[tile], b->stream);
    return per_order(b, [&](int P, hipStream_t s) {
        return hg::launch_batch_stage_p(b->variant, b->rs, P, b->ptile[P], args_po(b, P, slot), s);
    });
}

// The per-order work list for the chosen tiles: each case's element tiles in the tile of its
// order, residual slots numbered over the whole list (case-major), per order the positions of
// its items.  Replaces an earlier list (the autotuner runs once per batch).
int build_per_order_list(hg_batch *b) {
    std::vector<int4> items;
    std::vector<int2> ranges;
    std::vector<int> per[12];
    for (int c = 0; c < (int)b->cases.size(); ++c) {
        const hg_handle *h = b->cases[c];
        const int epb = hg::tile_elems(h->P, b->ptile[h->P]);
        const int first = (int)items.size();
        for (int e0 = 0; e0 < h->K; e0 += epb) {
            per[h->P].push_back((int)items.size());
            items.push_back(make_int4(c, e0, h->P, (int)items.size()));
        }
        ranges.push_back(make_int2(first, (int)items.size()));
    }
    if (items.size() > b->max_items) return hg::handle_error("hg_batch: per-order work list exceeds the residual buffer");
    HB_CHECK(b->upload(&b->items_po, items));
    HB_CHECK(b->upload(&b->ranges_po, ranges));
    for (int P = 0; P < 12; ++P) {
        b->nitems_po[P] = (int)per[P].size();
        b->sched_po[P] = nullptr;
        if (!per[P].empty()) HB_CHECK(b->upload(&b->sched_po[P], per[P]));
    }
    return 0;
}

// One RK step of all cases for the given trace parity (launches only).  Interior penalty: no
// gradient launch, each stage's epilogue forms the next stage's viscous face-flux traces.
// The artificial-viscosity indicator of every case from its slot's current traces: one launch
// per case, before the gradient pass that reads mu_a (DGSpatialDiscretizationMethods.f90:255).
hipError_t art_visc_all(hg_batch *b, int slot) {
    for (size_t c = 0; c < b->cases.size(); ++c) {
        hg_handle *h = b->cases[c];
        if (!h->ph.ad) continue;
        const hg::StageArgs &v = b->hslots[c * b->nslots + slot];
        hipError_t e = hg::launch_art_visc(h->P, h->mixed, v, h->evol, h->w, h->mu_a, b->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipEvent_t take_event(hg_batch *b) {
    if (!b->ev_pool.empty()) {
        hipEvent_t e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// A launch, between two recorded events when kernel timing is on (kernel id `which`).
template <typename F>
hipError_t timed(hg_batch *b, int which, F &&launch) {
    if (!b->prof) return launch();
    hipEvent_t e0 = take_event(b), e1 = take_event(b);
    if (!e0 || !e1) return hipErrorOutOfMemory;
    hipError_t e = hipEventRecord(e0, b->stream);
    if (e == hipSuccess) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
    b->ev_pending.push_back({which, {e0, e1}});
    return e;
}

hipError_t enqueue_step(hg_batch *b, int parity) {
    const bool ns = b->cases[0]->ph.ns != 0 && !b->cases[0]->ip_fused;
    const int t = b->tile;
    for (int s = 0; s < b->sch.nst; ++s) {
        const int slot = parity * b->sch.nst + s;
        hipError_t e;
        if ((e = art_visc_all(b, slot)) != hipSuccess) return e;
        if (ns && (e = timed(b, 1, [&] { return gradient_pass(b, t, b->order, slot); })) != hipSuccess) return e;
        if ((e = timed(b, 0, [&] { return stage_pass(b, t, b->order, slot); })) != hipSuccess) return e;
    }
    return timed(b, 2, [&] {
        const int2 *ranges = b->order == kPerOrder ? b->ranges_po : b->ranges[t];
        return hg::launch_batch_resid(b->resid_part, ranges, b->resid_out, (int)b->cases.size(), b->stream);
    });
}

// Times derivative-only evaluations (plus the gradient pass) of every (tile, order)
// candidate not fixed by the options and keeps the fastest.  The per-order candidate first
// picks each order's tile from its own launches, then is timed as it runs (concurrent orders).
int autotune(hg_batch *b) {
    // the gradient pass runs as in enqueue_step (interior penalty: the stage epilogue forms it)
    const bool ns = b->cases[0]->ph.ns != 0 && !b->cases[0]->ip_fused;
    const int dslot = 2 * b->sch.nst;
    hipEvent_t ev[2];
    HB_CHECK(hipEventCreate(&ev[0]));
    HB_CHECK(hipEventCreate(&ev[1]));
    // ms of three runs of a launch sequence on the batch stream, after one untimed run
    auto time_of = [&](auto &&run, float *ms) -> hipError_t {
        hipError_t e = run();
        if (e == hipSuccess) e = hipEventRecord(ev[0], b->stream);
        for (int r = 0; r < 3 && e == hipSuccess; ++r) e = run();
        if (e == hipSuccess) e = hipEventRecord(ev[1], b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(ms, ev[0], ev[1]);
        return e;
    };
    float best = 1e30f, ms = 0.f;
    int bt = 1, bo = 0;
    for (int t = 0; t < 2; ++t) {
        if (b->opt.tile >= 0 && t != b->opt.tile) continue;
        for (int o = 0; o < kPerOrder; ++o) {
            if (b->opt.order >= 0 && o != b->opt.order) continue;
            HB_CHECK(time_of([&]() -> hipError_t {
                hipError_t e = art_visc_all(b, dslot);
                if (e == hipSuccess && ns) e = gradient_pass(b, t, o, dslot);
                if (e == hipSuccess) e = stage_pass(b, t, o, dslot);
                return e;
            }, &ms));
            if (ms < best) { best = ms; bt = t; bo = o; }
        }
    }
    // the persistent form, both tiles
    for (int t = 0; t < 2; ++t) {
        if (b->opt.tile >= 0 && t != b->opt.tile) continue;
        if (b->opt.order >= 0 && b->opt.order != kPersist) continue;
        HB_CHECK(time_of([&]() -> hipError_t {
            hipError_t e = art_visc_all(b, dslot);
            if (e == hipSuccess && ns) e = gradient_pass(b, t, kPersist, dslot);
            if (e == hipSuccess) e = stage_pass(b, t, kPersist, dslot);
            return e;
        }, &ms));
        if (ms < best) { best = ms; bt = t; bo = kPersist; }
    }
    if (b->opt.order < 0 || b->opt.order == kPerOrder) {
        for (int P : b->porders) {
            float pbest = 1e30f;
            for (int t = 0; t < 2; ++t) {
                if (b->opt.tile >= 0 && t != b->opt.tile) continue;
                const hg::BatchArgs a = args_of_order(b, t, P, dslot);
                HB_CHECK(time_of([&]() -> hipError_t {
                    hipError_t e = ns ? hg::launch_batch_gradient_p(b->variant, P, t, a, b->stream) : hipSuccess;
                    if (e == hipSuccess) e = hg::launch_batch_stage_p(b->variant, b->rs, P, t, a, b->stream);
                    return e;
                }, &ms));
                if (ms < pbest) { pbest = ms; b->ptile[P] = t; }
            }
        }
        if (build_per_order_list(b)) return 1;
        HB_CHECK(time_of([&]() -> hipError_t {
            hipError_t e = art_visc_all(b, dslot);
            if (e == hipSuccess && ns) e = gradient_pass(b, 0, kPerOrder, dslot);
            if (e == hipSuccess) e = stage_pass(b, 0, kPerOrder, dslot);
            return e;
        }, &ms));
        if (ms < best) {
            best = ms;
            bo = kPerOrder;
            // reported tile: the one holding most of the work items
            int n[2] = {0, 0};
            for (int P : b->porders) n[b->ptile[P]] += b->nitems_po[P];
            bt = n[1] >= n[0] ? 1 : 0;
        }
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    b->tile = bt;
    b->order = bo;
    return 0;
}

// The slot table from every case's current state: parity p, stage s -> p*nst + s, the
// derivative-only slot 2*nst.  Records each case's trace parity and reference pressure.
std::vector<hg::StageArgs> slot_table(hg_batch *b) {
    std::vector<hg::StageArgs> slots(b->cases.size() * b->nslots);
    b->cur0.assign(b->cases.size(), 0);
    b->pref0.assign(b->cases.size(), 0.0);
    b->fst0.assign(b->cases.size(), nullptr);
    for (int c = 0; c < (int)b->cases.size(); ++c) {
        hg_handle *h = b->cases[c];
        b->cur0[c] = h->cur;
        b->pref0[c] = h->ph.pref;
        b->fst0[c] = h->fstate;
        for (int p = 0; p < 2; ++p)
            for (int s = 0; s < b->sch.nst; ++s) {
                hg::StageArgs a = hg::handle_args(h);
                const int in = h->cur ^ p ^ (s & 1);
                a.tr_in = h->tr[in];
                a.tr_out = h->tr[in ^ 1];
                a.dtr = h->dtr[h->ip_fused ? in : 0];
                a.dtr_out = h->dtr[h->ip_fused ? (in ^ 1) : 0];
                a.am = b->sch.am[s];
                a.cq = b->sch.cq[s];
                a.cg = b->sch.cg[s];
                a.first_stage = s == 0 ? 1 : 0;
                a.mode = 0;
                a.do_resid = s == b->sch.nst - 1 ? 1 : 0;
                a.resid_part = b->resid_part;
                slots[(size_t)c * b->nslots + p * b->sch.nst + s] = a;
            }
        hg::StageArgs d = hg::handle_args(h);   // dtr, dtr_out: the current traces
        d.tr_in = h->tr[h->cur];
        d.tr_out = h->tr[h->cur ^ 1];
        d.cq = 1.0;
        d.first_stage = 1;
        d.mode = 1;
        d.do_resid = 0;
        d.resid_part = b->resid_part;
        slots[(size_t)c * b->nslots + 2 * b->sch.nst] = d;
    }
    return slots;
}

void drop_graphs(hg_batch *b) {
    for (int p = 0; p < 2; ++p) {
        if (b->exec[p]) hipGraphExecDestroy(b->exec[p]);
        if (b->graph[p]) hipGraphDestroy(b->graph[p]);
        b->exec[p] = nullptr;
        b->graph[p] = nullptr;
    }
}

int build_graphs(hg_batch *b) {
    for (int p = 0; p < 2; ++p) {
        HB_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue_step(b, p);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->stream, &g);
        HB_CHECK(e);
        HB_CHECK(e2);
        b->graph[p] = g;
        HB_CHECK(hipGraphInstantiate(&b->exec[p], g, nullptr, nullptr, 0));
    }
    return 0;
}

}  // namespace

extern "C" {

int hg_batch_create(hg_handle *const *cases, int32_t ncases, const hg_batch_options *opt, hg_batch **out) {
    if (!cases || ncases <= 0 || !out) return hg::handle_error("hg_batch_create: bad argument");
    *out = nullptr;
    hg_batch_options o{-1, -1, 1};
    if (opt) o = *opt;
    if (o.tile < -1 || o.tile > 1 || o.order < -1 || o.order > kPersist)
        return hg::handle_error("hg_batch_create: tile must be -1..1 and order -1..3");
    hg_handle *h0 = cases[0];
    for (int c = 0; c < ncases; ++c) {
        hg_handle *h = cases[c];
        if (!h) return hg::handle_error("hg_batch_create: null case handle");
        if (!h->sub.empty()) return hg::handle_error("hg_batch_create: a case with several polynomial orders");
        if (h->P > 11) return hg::handle_error("hg_batch_create: the batched kernels take polynomial orders 1..10");
        // artificial dissipation all or none: it decides whether the gradient pass runs every stage
        if (h->ph.ns != h0->ph.ns || h->ph.viscous != h0->ph.viscous || h->mixed != h0->mixed ||
            h->opt.scheme != h0->opt.scheme || h->ph.ad != h0->ph.ad || h->lap != h0->lap)
            return hg::handle_error("hg_batch_create: the cases of a batch must share equations, viscous method, "
                                    "precision, time scheme and artificial dissipation");
        for (int k = 0; k < c; ++k)
            if (cases[k] == h) return hg::handle_error("hg_batch_create: a case appears twice");
    }
    std::unique_ptr<hg_batch> b(new hg_batch());
    b->cases.assign(cases, cases + ncases);
    b->opt = o;
    b->sch = scheme_of(h0->opt.scheme);
    b->variant = 2 * (h0->ph.ns ? (h0->ph.viscous == 2 ? 2 : 1) : 0) + (h0->mixed ? 1 : 0);
    b->nslots = 2 * b->sch.nst + 1;
    b->rs = h0->ph.rs_kernel;
    for (int c = 0; c < ncases; ++c) {
        if (cases[c]->ph.rs_kernel != b->rs) b->rs = -1;
        if (cases[c]->ph.xbc) b->rs = -1;   // the batched kernels compile those kinds into that form only
    }
    HB_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    for (auto *h : b->cases)
        if (hg::ensure_vtr(h)) return 1;
    for (auto *h : b->cases) HB_CHECK(hipStreamSynchronize(h->stream));

    // work lists per tile: (case, first element, P, residual slot); residual ranges per case
    size_t max_items = 0;
    std::vector<int4> items[2];
    std::vector<int2> ranges[2];
    for (int t = 0; t < 2; ++t) {
        int blk = 64;
        for (int c = 0; c < ncases; ++c) {
            const hg_handle *h = cases[c];
            const int epb = hg::tile_elems(h->P, t);
            if (epb <= 0) return hg::handle_error("hg_batch_create: polynomial order not compiled");
            blk = std::max(blk, hg::tile_threads(h->P, t));
            const int first = (int)items[t].size();
            for (int e0 = 0; e0 < h->K; e0 += epb) items[t].push_back(make_int4(c, e0, h->P, (int)items[t].size()));
            ranges[t].push_back(make_int2(first, (int)items[t].size()));
        }
        b->nitems[t] = (int)items[t].size();
        b->block[t] = blk;
        max_items = std::max(max_items, items[t].size());
        HB_CHECK(b->upload(&b->items[t], items[t]));
        HB_CHECK(b->upload(&b->ranges[t], ranges[t]));
        // orders: case-major; costliest tiles first (work ~ elements * P^3); cases interleaved
        const int n = b->nitems[t];
        std::vector<int> s0(n), s1(n), s2;
        std::iota(s0.begin(), s0.end(), 0);
        s1 = s0;
        auto cost = [&](int i) {
            const int4 v = items[t][i];
            const hg_handle *h = cases[v.x];
            const int ne = std::min(hg::tile_elems(h->P, t), h->K - v.y);
            return (double)ne * h->P * h->P * h->P;
        };
        std::stable_sort(s1.begin(), s1.end(), [&](int a, int c) { return cost(a) > cost(c); });
        s2.reserve(n);
        std::vector<int> pos(ncases);
        for (int c = 0; c < ncases; ++c) pos[c] = ranges[t][c].x;
        for (bool more = true; more;) {
            more = false;
            for (int c = 0; c < ncases; ++c)
                if (pos[c] < ranges[t][c].y) { s2.push_back(pos[c]++); more = true; }
        }
        HB_CHECK(b->upload(&b->sched[t][0], s0));
        HB_CHECK(b->upload(&b->sched[t][1], s1));
        HB_CHECK(b->upload(&b->sched[t][2], s2));
        // per order: the case-major list restricted to that order
        std::vector<int> per[12];
        for (int i : s0) per[items[t][i].z].push_back(i);
        for (int P = 0; P < 12; ++P) {
            b->pnitems[t][P] = (int)per[P].size();
            if (!per[P].empty()) HB_CHECK(b->upload(&b->psched[t][P], per[P]));
        }
    }
    for (int P = 11; P >= 0; --P)
        if (b->pnitems[0][P] > 0) b->porders.push_back(P);
    b->max_items = max_items;
    b->nside = std::min<int>(4, (int)b->porders.size());
    for (int k = 0; k < b->nside; ++k) {
        HB_CHECK(hipStreamCreateWithFlags(&b->side[k], hipStreamNonBlocking));
        HB_CHECK(hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming));
    }
    HB_CHECK(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    std::vector<double> zeros(4 * std::max<size_t>(max_items, 1), 0.0);
    HB_CHECK(b->upload(&b->resid_part, zeros));
    std::vector<double> dt0(ncases, 0.0);
    HB_CHECK(b->upload(&b->dt, dt0));
    std::vector<double *> rout(ncases);
    for (int c = 0; c < ncases; ++c) rout[c] = cases[c]->resid;
    HB_CHECK(b->upload(&b->resid_out, rout));

    std::vector<hg::StageArgs> slots = slot_table(b.get());
    HB_CHECK(b->upload(&b->slots, slots));
    b->hslots = slots;
    HB_CHECK(hipDeviceSynchronize());
    *out = b.release();
    return 0;
}

int hg_batch_rk_step(hg_batch *b, const double *t, const double *dt) {
    (void)t;  // no time-dependent boundary data on the GPU path
    if (!b || !dt) return hg::handle_error("hg_batch_rk_step: bad argument");
    // A case stepped or uploaded on its own since the last batch step has another trace parity,
    // possibly another reference pressure (FP32 form) and stale viscous traces (interior
    // penalty): rebuild the slot table from the cases' current state.  The captured kernels
    // read it from device memory, so the graphs stay valid, except the indicator launches of
    // artificial dissipation, which take their slot by value: those graphs are captured again.
    bool stale = false;
    for (size_t c = 0; c < b->cases.size(); ++c) {
        hg_handle *h = b->cases[c];
        if (h->cur != (b->cur0[c] ^ b->parity) || h->ph.pref != b->pref0[c] || h->fstate != b->fst0[c]) stale = true;
        if (hg::ensure_vtr(h)) return 1;
    }
    for (auto *h : b->cases) HB_CHECK(hipStreamSynchronize(h->stream));
    if (stale) {
        HB_CHECK(hipStreamSynchronize(b->stream));
        std::vector<hg::StageArgs> slots = slot_table(b);
        HB_CHECK(hipMemcpy(b->slots, slots.data(), slots.size() * sizeof(hg::StageArgs), hipMemcpyHostToDevice));
        b->hslots = slots;
        if (b->cases[0]->ph.ad) drop_graphs(b);
        b->parity = 0;
    }
    HB_CHECK(hipMemcpyAsync(b->dt, dt, b->cases.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
    if (b->tile < 0) {
        if (b->parity != 0) return hg::handle_error("hg_batch_rk_step: internal parity error");
        if (autotune(b)) return 1;
    }
    const bool graph = b->opt.use_graph && !b->prof;
    if (graph && !b->exec[0] && build_graphs(b)) {
        drop_graphs(b);
        return 1;
    }
    if (graph) HB_CHECK(hipGraphLaunch(b->exec[b->parity], b->stream));
    else HB_CHECK(enqueue_step(b, b->parity));
    b->parity ^= 1;
    for (auto *h : b->cases) {
        h->cur ^= 1;   // every scheme has an odd number of stages
        h->qdot_valid = false;
        hg::mark_last_evaluation(h, h->cur ^ 1);
    }
    return 0;
}

int hg_batch_synchronize(hg_batch *b) {
    if (!b) return hg::handle_error("hg_batch_synchronize: null batch");
    HB_CHECK(hipStreamSynchronize(b->stream));
    return 0;
}

int hg_batch_schedule(hg_batch *b, int32_t *tile, int32_t *order) {
    if (!b) return hg::handle_error("hg_batch_schedule: null batch");
    if (tile) *tile = b->tile;
    if (order) *order = b->order;
    return 0;
}

int hg_batch_profile_enable(hg_batch *b, int32_t on) {
    if (!b) return hg::handle_error("hg_batch_profile_enable: null batch");
    HB_CHECK(hipStreamSynchronize(b->stream));
    for (auto &p : b->ev_pending) {
        b->ev_pool.push_back(p.second.first);
        b->ev_pool.push_back(p.second.second);
    }
    b->ev_pending.clear();
    b->prof = on != 0;
    for (int k = 0; k < 3; ++k) {
        b->prof_n[k] = 0;
        b->prof_ms[k] = 0.0;
    }
    return 0;
}

int hg_batch_profile_read(hg_batch *b, int32_t which, int64_t *launches, double *ms) {
    if (!b || which < 0 || which > 2) return hg::handle_error("hg_batch_profile_read: bad argument");
    HB_CHECK(hipStreamSynchronize(b->stream));
    for (auto &p : b->ev_pending) {
        float t = 0.0f;
        HB_CHECK(hipEventElapsedTime(&t, p.second.first, p.second.second));
        b->prof_n[p.first] += 1;
        b->prof_ms[p.first] += t;
        b->ev_pool.push_back(p.second.first);
        b->ev_pool.push_back(p.second.second);
    }
    b->ev_pending.clear();
    if (launches) *launches = b->prof_n[which];
    if (ms) *ms = b->prof_ms[which];
    return 0;
}

int hg_batch_destroy(hg_batch *b) {
    if (!b) return 0;
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (auto &e : b->ev_pool) hipEventDestroy(e);
    for (auto &p : b->ev_pending) {
        hipEventDestroy(p.second.first);
        hipEventDestroy(p.second.second);
    }
    for (void *p : b->bufs) hipFree(p);
    for (int k = 0; k < 4; ++k) {
        if (b->side[k]) hipStreamDestroy(b->side[k]);
        if (b->join[k]) hipEventDestroy(b->join[k]);
    }
    if (b->fork) hipEventDestroy(b->fork);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

}  // extern "C"
