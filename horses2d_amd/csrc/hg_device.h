// Device-side data model and kernel launchers of the DGSEM stage pipeline.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace hg {

// Face-record flag bits (per element face, element-local orientation).
enum FaceKind : int {
    FK_INT_L = 0,     // this element is the LEFT (lowest id) side of an interior edge
    FK_INT_R = 1,     // RIGHT side: Riemann arguments swapped, flux negated
    FK_PERIODIC = 2,  // periodic boundary edge: partner trace is the outer state
    FK_FARFIELD = 3,  // characteristic far field (Farfield.incf:138-376)
    FK_WALL = 4,      // viscous wall (ViscousWall.incf:158-254)
    FK_EULERWALL = 5, // Euler wall, prescribed flux (EulerWall.incf:96-159)
    FK_DIRICHLET = 6, // Dirichlet state (Dirichlet.incf:106-337)
    FK_POUTLET = 7,   // pressure outlet (PressureOutlet.incf:151-450)
    FK_PINLET = 8,    // pressure inlet, prescribed flux (PressureInlet.incf:134-410)
    FK_NEWDIRICHLET = 9, // characteristic Dirichlet state (newDirichlet.incf:146-433)
    FK_MORTAR = 10,   // p-non-conforming interior edge: the flux (and BR1 uStar) of this side
                      // come projected from the mortar kernels (Edge_ProjectFluxesType3-6)
};
constexpr int FF_CURVED = 1 << 4;    // per-node normals and dS
constexpr int FF_REV = 1 << 5;       // neighbor trace index runs backwards
constexpr int FF_SELFREV = 1 << 6;   // edge geometry index runs backwards (RIGHT side of an inverse edge)
constexpr int FF_ISOTHERMAL = 1 << 7;  // viscous wall: isothermal
constexpr int FF_REFLECTIVE = 1 << 7;  // pressure outlet: reflective outflow (same bit, other kind)
constexpr int FF_TOTAL_INFLOW = 1 << 7; // newDirichlet: subsonic inflow at the total pressure (same bit)
constexpr int FF_TOTAL_OUTFLOW = 1 << 12; // newDirichlet: subsonic outflow at the total pressure
constexpr int FF_RS_SHIFT = 8;       // (zone Riemann solver + 1), 0 = interior method
constexpr int FF_FSTATE = 1 << 13;   // boundary face whose zone state comes per node from a
                                     // boundary-state function (StageArgs::fstate)
constexpr int FF_MCURVED = 1 << 14;  // mortar face of a curved edge: BR1 uStar dS n per direction

// Viscous face-flux traces: per element face node, mu times this side's viscous flux from its
// own Q and gradient traces, (f1, f2, f3, g2, g3) and one pad (f0 = g0 = 0, g1 = f2).
constexpr int VT = 6;

struct PhysParams {
    double gamma, gm1, cp, cv, lambda, invSGM, mu, cpPr, sigma0, ip_eps, mach_sqrtg;
    // Reference pressure of the FP32 form (pref, Eref = pref/(gamma-1)); zero in the FP64 form.
    double pref, Eref;
    int riemann, viscous, ip_gradflux, ns;
    int rs_kernel;  // Riemann solver compiled into the stage kernel (0..3), -1: chosen per face at run time
    int xbc;        // a zone is an Euler wall, Dirichlet, pressure outlet or inlet, or artificial
                    // dissipation is on: the kernels compiled with those (stage: the run-time-
                    // solver form) are launched
    int ad;         // artificial dissipation (jumps-based indicator): "Physical" or "Laplacian" type
    double ad_ceps; // its intensity
};

// Everything a stage or gradient kernel reads or writes.  Uniform P per handle.
struct StageArgs {
    int K;                      // elements
    int64_t nodes;              // K*P*P
    const double *Q;            // state (FP64), updated in place by the stage kernel
    double *G;                  // low-storage RK register (FP64)
    double *QDot;               // derivative output (derivative-only mode) or nullptr
    const void *tr_in;          // Q traces [(e*4+f)*P + k][4] (FP64 or FP32)
    void *tr_out;               // next-stage Q traces
    const void *dtr;            // viscous face-flux traces [(e*4+f)*P + k][VT] (FP64 or FP32)
    void *dtr_out;
    const void *metric;         // 5 planes: Ja11, Ja12, Ja21, Ja22, 1/jac, per node or per element
    int metric_elem;            // 1: one metric record per element (element-constant metric)
    const int4 *face;           // per element face record
    const int *forder;          // per element: faces sorted by global edge id (2 bits each)
    const double *egeo;         // edge geometry triplets (nx, ny, dS)
    const double *einvh;        // per edge 1/h for the IP penalty
    const double *zone;         // per zone 8 doubles
    const void *fmet;           // IP: per face node (Jxi_x, Jxi_y, Jeta_x, Jeta_y, jac)
    const double *hatD, *D, *lb, *lbw, *dlbw;  // spectral operators (P*P, P*P, 2P, 2P, 2P)
    double am, cq, cg;          // G = am*G + cq*QDot ; Q = Q + cg*G
    const double *dt_dev;       // if set, cq (dt_cq) and/or cg (dt_cg) are multiplied by *dt_dev
    int dt_cq, dt_cg;
    int first_stage;            // G is not read (am = 0)
    int mode;                   // 0: RK update + next traces, 1: derivative only
    double *resid_part;         // per workgroup max|QDot| per equation (4 doubles per workgroup)
    int do_resid;
    PhysParams ph;
    const double *mu_a;         // artificial viscosity per element (null: none)
    const double *lap;          // Laplacian dissipation: the indicator's viscosity per element (null: none)
    const double *earea;        // per edge: its Area (the Laplacian face flux)
    int ip_epi;                 // interior penalty: the stage epilogue forms the next viscous traces
    // p-non-conforming faces (FK_MORTAR), per element face node [(e*4+f)*P + k] (FP64):
    const double *mustar;       // BR1 uStar of this side, projected and signed [4]
    const double *mflux;        // (Fi - Fv) dS of this side, projected and signed [4]
    double *mdq;                // this side's gradient trace [8], written by the gradient kernel
    const double *mgrad;        // interior penalty: this side's gradient flux, projected, times dS [8]
    // boundary-state functions: per boundary face node [(e*4+f)*P + k][4], the state the ghost
    // state starts from in place of the zone's (faces flagged FF_FSTATE)
    const double *fstate;
};

// p-non-conforming edges of a handle whose elements have several orders: the elements of each
// order form a sub-handle; an edge between two of them is evaluated on its own nodes (the
// higher order) by the mortar kernels.
constexpr int kMaxSub = 12;
struct MortarEdge {
    int sub[2];      // sub-handle of the LEFT and RIGHT element
    int slot[2];     // their face slots (local element * 4 + face)
    int Ps[2];       // their points per direction
    int P;           // the edge's points per direction (the higher)
    int low;         // the side of lower order (0 LEFT, 1 RIGHT)
    int inverse;     // the RIGHT element runs the edge backwards
    int mat;         // offset of T_forward [P][Plow] and then T_backward [Plow][P] in the matrix pool
    int geo;         // curved edge: (nx, ny, dS) per edge node in the geometry pool; -1: straight
    double nx, ny, dS;   // straight edge: normal (LEFT to RIGHT) and length
    double invh;     // interior penalty: the edge's 1/h
};
// A subdivided edge (one element's edge meeting two elements, MeshFile.f90:364-490): its big
// (LEFT), NORTH and SOUTH elements and the two mortars between them.
struct SubMortarEdge {
    int sub[3];      // sub-handle of the LEFT, NORTH and SOUTH elements
    int slot[3];     // their face slots (local element * 4 + face)
    int Ps[3];       // their points per direction
    int PM[2];       // points of the NORTH and SOUTH mortars
    int mat;         // T_LN_F, T_LS_F, T_LN_B, T_LS_B, T_RN_F, T_RS_F, T_RN_B, T_RS_B in the pool, row-major
    int geo;         // (nx, ny, dS) per mortar node in the geometry pool: NORTH then SOUTH
    int curved;      // curved edge: the BR1 uStar is scaled per mortar node before its return
    int pad;
    double invh;     // interior penalty: the edge's 1/h
};
struct MortarArgs {
    const MortarEdge *edges;
    int nedges;
    const SubMortarEdge *subs;
    int nsubs;
    const double *sgeo;
    const double *mats;
    const void *tr[kMaxSub];     // Q traces of the current stage, per sub-handle (FP64, or FP32
                                 // in the mixed form with the energy slot less Eref)
    int mixed;                   // the traces are FP32
    double Eref;                 // the mixed form's energy shift, common to all sub-handles
    const double *mdq[kMaxSub];  // gradient traces at mortar faces, per sub-handle
    double *mustar[kMaxSub];     // per mortar face node [8]: straight edges uStar [0..3]; curved
                                 // edges uStar dS nx [0..3] and uStar dS ny [4..7]
    double *mflux[kMaxSub];
    double *mgrad[kMaxSub];      // interior penalty gradient fluxes per mortar face node [8]
    PhysParams ph;
};
// BR1 uStar of every mortar edge (from the sub-handles' current Q traces), to each side.
hipError_t launch_mortar_ustar(const MortarArgs &m, hipStream_t s);
// The numerical flux (Fi - Fv) dS of every mortar edge on its nodes, to each side.
hipError_t launch_mortar_flux(const MortarArgs &m, hipStream_t s);
// The same two steps for the subdivided edges (two mortars each).
hipError_t launch_sub_ustar(const MortarArgs &m, hipStream_t s);
hipError_t launch_sub_flux(const MortarArgs &m, hipStream_t s);
// dst[map[i]] = src[i], i < n (a sub-handle's per-element values into the whole mesh's order).
hipError_t launch_scatter(const double *src, const int *map, int n, double *dst, hipStream_t s);

// The batched form: one launch per RK stage over a work list of element tiles of many cases.
struct BatchArgs {
    const StageArgs *slots;  // per case, nslots StageArgs (parity x stage, then derivative-only)
    int nslots, slot;        // slot of this launch
    const int4 *items;       // (case, first element, points per direction P, residual slot)
    const int *sched;        // processing order of the items
    int nitems;
    const double *dt;        // per case time step
    int dt_cq, dt_cg;        // multiply the slot's cq / cg by the case's dt
    // persistent form (persist = 1): as many workgroups as the GPU holds at once, a multiple
    // of eight per polynomial order; the workgroups of order P loop over its items
    // (psched[P][0..pn[P]), positions in `items`), each XCD over a contiguous eighth of them
    int persist;
    int pn[12];
    const int *psched[12];
};

// variant = 2 * physics (0 Euler, 1 Navier-Stokes BR1, 2 Navier-Stokes IP) + (1 if mixed)
hipError_t launch_batch_stage(int variant, int rs, int tile, const BatchArgs &b, int block, hipStream_t s);
hipError_t launch_batch_gradient(int variant, int tile, const BatchArgs &b, int block, hipStream_t s);
// One order per launch: the work items (b.items, b.sched, b.nitems) of order P only.
hipError_t launch_batch_stage_p(int variant, int rs, int P, int tile, const BatchArgs &b, hipStream_t s);
hipError_t launch_batch_gradient_p(int variant, int P, int tile, const BatchArgs &b, hipStream_t s);
// out[c][0..4) = max over the residual partials [ranges[c].x, ranges[c].y) (one workgroup per case)
hipError_t launch_batch_resid(const double *part, const int2 *ranges, double *const *out, int ncases, hipStream_t s);
int tile_threads(int P, int tile);  // workgroup size of a tile
int tile_elems(int P, int tile);    // elements per tile

// Per-variant launchers (one object file per variant and Riemann solver, from
// hg_batch_inst.hip): batch_stage_v<V>_rs<RS + 1>, batch_gradient_v<V>.
#define HG_DECLARE_BATCH_STAGE(V, RS1) \
    hipError_t batch_stage_v##V##_rs##RS1(int tile, const BatchArgs &b, int block, hipStream_t s); \
    hipError_t batch_stage_v##V##_rs##RS1##_p(int P, int tile, const BatchArgs &b, hipStream_t s);
#define HG_DECLARE_BATCH_VARIANT(V)                                                            \
    HG_DECLARE_BATCH_STAGE(V, 0)                                                               \
    HG_DECLARE_BATCH_STAGE(V, 1)                                                               \
    HG_DECLARE_BATCH_STAGE(V, 2)                                                               \
    HG_DECLARE_BATCH_STAGE(V, 3)                                                               \
    HG_DECLARE_BATCH_STAGE(V, 4)                                                               \
    hipError_t batch_gradient_v##V(int tile, const BatchArgs &b, int block, hipStream_t s);               \
    hipError_t batch_gradient_v##V##_p(int P, int tile, const BatchArgs &b, hipStream_t s);
HG_DECLARE_BATCH_VARIANT(0)
HG_DECLARE_BATCH_VARIANT(1)
HG_DECLARE_BATCH_VARIANT(2)
HG_DECLARE_BATCH_VARIANT(3)
HG_DECLARE_BATCH_VARIANT(4)
HG_DECLARE_BATCH_VARIANT(5)
#undef HG_DECLARE_BATCH_VARIANT
#undef HG_DECLARE_BATCH_STAGE

// Launchers (return hipError_t of the launch).
hipError_t launch_prolong(int P, bool mixed, const StageArgs &a, hipStream_t s);
// tile: 0 = narrow workgroup (the widest element tile that fits one wavefront),
//       1 = wide workgroup (up to four wavefronts).  Results are bit-identical.
hipError_t launch_stage(int P, int tile, bool mixed, const StageArgs &a, hipStream_t s);
hipError_t launch_gradient(int P, int tile, bool mixed, const StageArgs &a, hipStream_t s);
// Workgroups of the stage kernel (= residual partials it writes) for a tile choice.
int stage_blocks(int P, int tile, int K);
// Reduces nb residual partials to out[4] (NaN propagates); with scratch (4 * kResidGroups
// doubles) many partials are reduced by several workgroups first.
constexpr int kResidGroups = 256;
hipError_t launch_resid_final(const double *part, int nb, double *out, hipStream_t s, double *scratch = nullptr);

// Time-step estimate: per element maxima, then an ordered prefix-max scan.
hipError_t launch_dt_elem(int P, const double *Q, int K, double gamma, double gm1, double *umax, double *amax,
                          hipStream_t s);
// N1: N + 1 of every element, or per element n1e[e] when n1e is set (several orders).
hipError_t launch_dt_scan(int K, const double *umax, const double *amax, const double *dx, double coef, double N1,
                          double dt_max, double *state /* [umax, amax] */, double *dt_out, hipStream_t s,
                          const double *n1e = nullptr);
// Conservation sums (deterministic two-pass).
// Volume monitors [One, dSnorm^2, 0, 0] (QuadMesh_Integrals.incf:22-65) into out[4].
hipError_t launch_volume_integrals(int P, const double *Q, const double *jac, const double *w, int K, double gamma,
                                   double gm1, double *partial, int nbk, double *out, hipStream_t s);
// Surface monitors of one zone (QuadMesh_Integrals.incf:67-346) over its boundary face slots
// (element*4 + face): out[9] = Surface, pressure, density, pressure force (2), mass-flow, flow,
// viscous force (2); vtr null for Euler.  Eref: the FP32 form's energy-slot shift.
hipError_t launch_surface_integrals(bool mixed, const int *slots, int nslots, int P, const void *tr, const void *vtr,
                                    const int4 *face, const double *egeo, const double *w, double gm1, double Eref,
                                    const double *mu_a, double mu, double *out, hipStream_t s);
// Jumps-based artificial viscosity per element from the stage's Q traces (and boundary states).
hipError_t launch_art_visc(int P, bool mixed, const StageArgs &a, const double *evol, const double *w, double *mu_a,
                           hipStream_t st);
// Probe values (Probe_Update): n probes, per probe an int2 (element, variable) and 2P
// interpolant values; variables 4..7 read QDot, the others Q.
hipError_t launch_probes(const double *Q, const double *QDot, const int2 *ev, const double *l, int n, int P,
                         double gm1, double gamma, double *out, hipStream_t st);
hipError_t launch_cons_sums(int P, const double *Q, const double *jac, const double *w, int K, double *partial,
                            int nblocks, double *out, hipStream_t s);

bool supported_order(int N);

// Per-order launchers, explicitly instantiated once per P (one object file per order).
template <int P, int PHYS> hipError_t stage_phys_p(int tile, bool mixed, const StageArgs &a, hipStream_t s);
template <int P> int stage_blocks_p(int tile, int K);
template <int P> int tile_threads_p(int tile);
template <int P> int tile_elems_p(int tile);
template <int P> hipError_t gradient_p(int tile, bool mixed, const StageArgs &a, hipStream_t s);
template <int P> hipError_t prolong_p(bool mixed, const StageArgs &a, hipStream_t s);
template <int P>
hipError_t art_visc_p(bool mixed, const StageArgs &a, const double *evol, const double *w, double *mu_a, hipStream_t s);
template <int P> hipError_t dt_elem_p(const double *Q, int K, double gamma, double gm1, double *umax, double *amax,
                                      hipStream_t s);
template <int P> int cons_blocks_p(int K);
template <int P> hipError_t cons_partial_p(const double *Q, const double *jac, const double *w, int K, int mode,
                                           double gamma, double gm1, double *partial, hipStream_t s);

}  // namespace hg
