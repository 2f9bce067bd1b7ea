// Internal definition of the library handle (shared by the C ABI and the batched form).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "hg_device.h"
#include "hg_mesh.h"
#include "horses2d_gpu.h"

namespace hg {
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};
}  // namespace hg

struct hg_handle {
    hg::Mesh mesh;
    int K = 0, N = 0, P = 0;
    int64_t nodes = 0, ndof = 0;
    bool mixed = false;
    bool has_pref = false;   // the mixed form's reference pressure is given (pref_in), not the
    double pref_in = 0.0;    // uploaded state's mean (sub-handles of a mesh with several orders)
    int device = -1;   // the HIP device current at hg_create (the handle's stream and buffers)
    hg_options opt{};
    hg_physics phys{};
    hg::PhysParams ph{};
    hipStream_t stream = nullptr;
    std::vector<hg::DevBuf> bufs;
    // device arrays
    double *Q = nullptr, *G = nullptr, *QDot = nullptr;
    void *tr[2] = {nullptr, nullptr};
    // viscous face-flux traces: one buffer (BR1: k_gradient writes it before each stage), or a
    // ping-pong pair indexed like tr (interior penalty: each stage's epilogue writes the next
    // stage's; k_gradient runs only when the current one is stale, after an upload)
    void *dtr[2] = {nullptr, nullptr};
    bool ip_fused = false;
    bool vtr_valid = false;
    void *metric = nullptr, *fmet = nullptr;
    int metric_elem = 0;   // metric stored per element (FP32 form, element-constant metric)
    int4 *face = nullptr;
    int *forder = nullptr;
    double *egeo = nullptr, *einvh = nullptr, *zone = nullptr;
    double *hatD = nullptr, *D = nullptr, *lb = nullptr, *lbw = nullptr, *dlbw = nullptr, *w = nullptr;
    double *jac64 = nullptr;
    double *dx = nullptr, *umax = nullptr, *amax = nullptr, *dtstate = nullptr, *dtout = nullptr;
    double *cons_partial = nullptr, *cons_out = nullptr;
    int cons_nb = 0;
    // boundary face slots (element*4 + face) of every zone, zones in marker order, each zone's
    // edges in ascending edge id (the reference's zone edge lists); monitor output
    int *bslots = nullptr;
    std::vector<int> bslot_off;   // zone z (marker z+1): [bslot_off[z], bslot_off[z+1])
    double *monitor_out = nullptr;
    // probes: (element, variable) and the interpolants l(xi), l(eta) (2P per probe), owned here
    // (freed on replacement and in hg_destroy, not through bufs)
    int nprobes = 0, probe_cap = 0;
    std::vector<int> probe_var;
    int2 *probe_ev = nullptr;
    double *probe_l = nullptr, *probe_out = nullptr;
    bool qdot_valid = false;   // QDot holds hg_time_derivative of the current state
    // the edge storage the surface monitors read: the Q trace buffer and viscous face-flux
    // traces of the last derivative evaluation (-1: none since the last upload)
    int mon_tr = -1;
    void *mon_vtr = nullptr;
    // artificial dissipation: per-element viscosity (written before each stage) and volume;
    // lap: the Laplacian type (the viscosity enters its own fluxes, not mu); per-edge Area
    double *mu_a = nullptr, *evol = nullptr, *earea = nullptr;
    bool lap = false;
    double *resid = nullptr, *resid_part = nullptr;  // max|QDot| of the last stage; per-workgroup partials
    double *resid_scratch = nullptr;                 // first-pass maxima of a many-workgroup reduction
    // p-non-conforming faces of this handle: per face node the mortar kernels' projected flux,
    // BR1 uStar, and this side's gradient trace (null when every edge is conforming)
    double *mflux = nullptr, *mustar = nullptr, *mdq = nullptr;
    double *mgrad = nullptr;   // interior penalty: the projected gradient fluxes per mortar face node
    // boundary-state functions: per face node the state (boundary faces of zones given one by
    // hg_set_boundary_state; null until the first such call)
    double *fstate = nullptr;
    std::vector<int> fzone;    // zones given per-node states
    // A mesh with several polynomial orders: one sub-handle per order (its elements in mesh
    // order, sharing this handle's stream) coupled by the mortar kernels; this handle keeps the
    // whole mesh, the time-step scan over all elements in mesh order and the mortar tables.
    std::vector<hg_handle *> sub;
    std::vector<std::vector<int>> sub_els;   // per sub-handle: its elements' mesh ids
    std::vector<int64_t> el_off;             // per mesh element: node offset in the reference layout
    int *sub_gmap[hg::kMaxSub] = {};         // per sub-handle (device): local element -> mesh element
    hg::MortarEdge *medges = nullptr;
    double *mmats = nullptr;
    int nmortar = 0;
    hg::SubMortarEdge *sedges = nullptr;     // subdivided edges and their mortars' geometry
    double *sgeo = nullptr;
    int nsubedge = 0;
    double *n1e = nullptr;                   // per mesh element: N + 1 (time-step scan)
    std::vector<std::vector<int>> sub_probes;   // per sub-handle: the probe indices it holds
    bool shared_stream = false;              // a sub-handle: the stream belongs to its parent
    int cur = 0;
    int tile[2] = {-1, -1};  // workgroup tile of the stage and gradient kernels, -1 = not chosen yet
    // one RK step captured as a HIP graph per trace parity; the step size is read from dt_dev
    bool use_graph = true;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    double *dt_dev = nullptr;
    double dt_last = 0.0;
    bool dt_valid = false;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    int64_t prof_n[2] = {0, 0};
    double prof_ms[2] = {0.0, 0.0};

    template <typename T>
    hipError_t alloc(T **p, size_t count) {
        void *raw = nullptr;
        hipError_t e = hipMalloc(&raw, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        bufs.push_back({raw, count * sizeof(T)});
        *p = (T *)raw;
        return hipSuccess;
    }
    hipError_t alloc_raw(void **p, size_t bytes) {
        hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 1));
        if (e != hipSuccess) return e;
        bufs.push_back({*p, bytes});
        return hipSuccess;
    }
};

namespace hg {
// StageArgs with every per-handle pointer and the physics filled in (per-stage fields zero).
StageArgs handle_args(hg_handle *h);
// Viscous face-flux traces of the current state: the buffer, and (interior penalty) making
// sure it is up to date by running k_gradient once if the state was uploaded since.
void *vtr_current(hg_handle *h);
void *vtr_next(hg_handle *h);
int ensure_vtr(hg_handle *h);
// Records which trace buffers hold the last derivative evaluation's edge storage (the Q trace
// buffer `tr` and, for Navier-Stokes, its viscous face-flux traces), for the surface monitors.
void mark_last_evaluation(hg_handle *h, int tr);
// Records the message returned by hg_last_error and returns 1.
int handle_error(const std::string &m);
// Locates the allocation-free entry points' scratch array on a device (once per device).
int noalloc_setup(int device);
}  // namespace hg
