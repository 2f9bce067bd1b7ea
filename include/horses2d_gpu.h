/*
 * horses2d_gpu.h -- C ABI of the MI355X DGSEM residual + explicit Runge-Kutta library.
 *
 * Drop-in boundary for HORSES2D's hot path.  Each entry point names the
 * reference interface it replaces (paths relative to the HORSES2D source tree).
 * All functions return 0 on success and a nonzero code on failure; the message
 * is available through hg_last_error().  No function exits or aborts.
 *
 * Data layout contract (same as the reference's flat Storage%Q,
 * Generic/Storage.f90:54-73, QuadElement_QuadProcedures.incf:22-40):
 *   element e occupies Q[4*off_e .. 4*off_e + 4*P_e^2), off_e = sum_{k<e} P_k^2,
 *   Q(eq,i,j) = Q[4*off_e + eq + 4*(i + P_e*j)], eq fastest, then xi, then eta.
 * All state buffers passed in or out are host memory in that layout (FP64).
 *
 * Threading: a handle is used from one host thread; it owns one HIP stream.
 */
#ifndef HORSES2D_GPU_H
#define HORSES2D_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hg_handle hg_handle;
typedef struct hg_batch hg_batch;

/* Boundary condition kinds (Generic/Defines.h:103-111 numbering; "Type" of a zone block,
 * BoundaryConditions/DGBoundaryConditions.f90:244-357). */
#define HG_BC_PERIODIC 1
#define HG_BC_DIRICHLET 2
#define HG_BC_EULERWALL 3
#define HG_BC_VISCOUSWALL 4
#define HG_BC_FARFIELD 5
#define HG_BC_PRESSUREOUTLET 6
#define HG_BC_PRESSUREINLET 7
#define HG_BC_NEWDIRICHLET 9   /* 8 (Riemann) is not supported: its outer state is undefined */

/* Pressure outlet "Outflow type" (BoundaryConditions/PressureOutlet.incf:66-72). */
#define HG_OUTFLOW_PARTIALLY_NON_REFLECTIVE 0
#define HG_OUTFLOW_REFLECTIVE 1

/* newDirichlet "Mode" and "Outflow pressure" (BoundaryConditions/DGBoundaryConditions.f90:47-54,
 * newDirichlet.incf:79-105; any other value selects the reference's default, the first). */
#define HG_NEWDIRICHLET_SPECIFY_SPEED 1
#define HG_NEWDIRICHLET_SPECIFY_TOTAL_PRESSURE 2
#define HG_NEWDIRICHLET_STATIC_PRESSURE 1
#define HG_NEWDIRICHLET_TOTAL_PRESSURE 2

/* Riemann solvers: Roe, HLL, exact (Physics/RiemannSolvers.f90:33-399) and the
 * Lax-Friedrichs flux, which the reference does not have. */
#define HG_RIEMANN_ROE 0
#define HG_RIEMANN_HLL 1
#define HG_RIEMANN_EXACT 2
#define HG_RIEMANN_LF 3

#define HG_VISCOUS_NONE 0
#define HG_VISCOUS_BR1 1
#define HG_VISCOUS_IP 2

#define HG_IP_SIPG 0
#define HG_IP_NIPG 1
#define HG_IP_IIPG 2

#define HG_PRECISION_FP64 0
#define HG_PRECISION_MIXED 1 /* FP64 state/RK/reductions, FP32 fluxes and contractions */

/* Artificial dissipation (Approximation/DGArtificialDissipation.f90), with the "Jumps-based"
 * indicator: none, the "Physical" type (the element and edge viscosities mu_a added to mu in
 * every viscous flux, :252-275) or the "Laplacian" type (a volume flux mu times the element's
 * interior gradient and a face flux -mu Area (QL - QR), :207-249).  The "Residuals-based"
 * indicator (which the reference leaves without a viscosity function) is not provided. */
#define HG_AD_NONE 0
#define HG_AD_PHYSICAL 1
#define HG_AD_LAPLACIAN 2

#define HG_SCHEME_EULER 0
#define HG_SCHEME_RK3 1
#define HG_SCHEME_RK5 2

/*
 * Mesh description: the content of a .HiOMesh file plus the boundary-zone
 * blocks of the .HiOCase file.  Mirrors MeshFile_t (Mesh/MeshFile.f90:31-64)
 * after ReadMesh (:77-279): coordinates already divided by the Reynolds
 * length, ids 0-based.  The library rebuilds connectivity with the
 * reference's numbering (edges in first-appearance order over
 * (element, face 1..4), LEFT = lowest element id) and the geometry
 * (QuadMesh_Construct.incf:22-305), so a Fortran caller passes the arrays it
 * read from the mesh file.
 */
typedef struct {
    int32_t nnodes;
    const double *points;     /* [nnodes][2] */
    int32_t nelements;
    const int32_t *quads;     /* [K][4] counter-clockwise node ids */
    const int32_t *order;     /* [K] polynomial order (uniform per mesh on the GPU path) */
    int32_t nbedges;
    const int32_t *bedges;    /* [nb][2] */
    const int32_t *bmarker;   /* [nb] 1-based marker */
    int32_t ncurved;          /* number of curvilinear edges (0 if none) */
    int32_t np1;              /* points per curvilinear edge */
    const int32_t *cedges;    /* [nc][2] */
    const double *cxy;        /* [nc][np1][2] at Chebyshev-Gauss-Lobatto points */
    int32_t nmarkers;
    const int32_t *bc_type;   /* [nmarkers] HG_BC_* */
    const int32_t *bc_int;    /* [nmarkers][2] periodic: (direction 1=x|2=y, partner marker);
                                 viscous wall: (isothermal 0|1, 0);
                                 pressure outlet: (HG_OUTFLOW_*, 0);
                                 newDirichlet: (mode, outflow pressure); others (0, 0) */
    const double *bc_real;    /* [nmarkers][4] farfield, Dirichlet, pressure outlet, pressure
                                 inlet, newDirichlet: the zone state q (dimensionless, as their _Construct
                                 forms it; "Use function" states are not supported);
                                 viscous wall: (Tw, vx, vy, 0); Euler wall: unused */
    const int32_t *bc_riemann;/* [nmarkers] per-zone Riemann solver or -1 (may be NULL) */
} hg_mesh_desc;

/* PhysicsNS.f90:405-481 inputs (air: gamma = 1.4). */
typedef struct {
    double mach;
    double reynolds;  /* 0 => inviscid */
    double prandtl;
} hg_physics;

typedef struct {
    int32_t navier_stokes; /* 0: HORSES2D.Euler, 1: HORSES2D.NS (Solver/Makefile:163,180-186) */
    int32_t riemann;       /* HG_RIEMANN_* ("Inviscid Riemann solver", Setup.f90:280-281) */
    int32_t viscous;       /* HG_VISCOUS_* ("Viscous discretization", Setup.f90:295-296) */
    int32_t ip_method;     /* HG_IP_* */
    double sigma0;         /* "Jumps penalty parameter" */
    int32_t precision;     /* HG_PRECISION_* */
    int32_t scheme;        /* HG_SCHEME_* ("Integration scheme", Setup.f90:354-355) */
    int32_t artificial_dissipation; /* HG_AD_* ("Artificial dissipation (0/1)", Setup.f90:333-349);
                                       Navier-Stokes only, as DGArtificialDissipation.f90 */
    double ad_intensity;   /* "Artificial dissipation intensity" (Ceps) */
} hg_options;

/* Builds connectivity/geometry on the host, uploads it, allocates device state.
 * Replaces the device-side part of DGSEM_construct (Approximation/DGSEM_class.f90:104-185)
 * and DGSpatial_Initialization (DGSpatialDiscretizationMethods.f90:67-86). */
int hg_create(const hg_mesh_desc *mesh, const hg_physics *phys, const hg_options *opt, hg_handle **out);

/* Frees device memory and the stream owned by the handle. */
int hg_destroy(hg_handle *h);

/* Number of scalar unknowns, size(Storage%Q) (DGSEM_class.f90:339-340). */
int64_t hg_ndof(const hg_handle *h);

/* Node coordinates X[node][2] (node = off_e + i + P j), the element mapping of
 * QuadElement_QuadMappings.incf:325-334; used by the host to evaluate the
 * libproblemfile initial condition (QuadMesh.f90:192-213). */
int hg_get_node_coords(hg_handle *h, double *X, int64_t nnodes);

/* Connectivity as built (for bit-exact comparison with the Fortran mesh):
 *   edges[k][8] = {node0, node1, quadLEFT, quadRIGHT, locLEFT, locRIGHT, inverse, marker}
 *   elems[e][12] = {edges[4], edgesDirection[4], quadPosition[4]}
 *   partner[k] = periodic partner edge or -1.
 * (MeshFile.f90:287-618, QuadElement_EdgeProcedures.incf:30-199,357-424,
 *  QuadMesh_Zones.incf:22-131). */
int hg_get_connectivity(hg_handle *h, int32_t *edges, int32_t *elems, int32_t *partner);

/* Host-only mesh build (no device needed): connectivity tables as in
 * hg_get_connectivity.  Call with edges == NULL to query *nedges first.
 * Lets the caller check the numbering against its own mesh (MeshFile.f90:287-618). */
int hg_build_connectivity(const hg_mesh_desc *mesh, int32_t *nedges, int32_t *edges, int32_t *elems,
                          int32_t *partner);

/* Host-only geometry of the built mesh: node coordinates X[node][2], metric
 * Ja[node][4] = (Ja11, Ja12, Ja21, Ja22) and jac[node]
 * (QuadElement_QuadMappings.incf:22-367).  Arrays sized by the total node count. */
int hg_build_geometry(const hg_mesh_desc *mesh, int64_t nnodes, double *X, double *Ja, double *jac);

/* Number of edges of the built mesh (size of the hg_get_connectivity edge table). */
int hg_num_edges(hg_handle *h);

/* Host <-> device state (Storage%Q layout).  Upload replaces the state after the
 * initial condition or a restart (DGSEM_class.f90:201-205, 263-291). */
int hg_upload_state(hg_handle *h, const double *Q, int64_t ndof);
int hg_download_state(hg_handle *h, double *Q, int64_t ndof);
int hg_download_qdot(hg_handle *h, double *QDot, int64_t ndof);

/* QDot = residual of the current state: DGSpatial_computeTimeDerivative
 * (DGSpatialDiscretizationMethods.f90:94-116). */
int hg_time_derivative(hg_handle *h, double t);

/* One time step of the configured scheme (Williamson RK3 by default):
 * the TimeScheme procedure pointer (DGTimeIntegrator.f90:71-83) bound to
 * TimeIntegrator_WilliamsonRK3 (:282-317), RK5 (:319-365) or ExplicitEuler (:263-280). */
int hg_rk_step(hg_handle *h, double t, double dt);

/* max|QDot| per equation of the last stage (Mesh_ComputeResiduals,
 * Mesh/QuadMesh_Auxiliars.incf:22-38), as used by Monitors%UpdateValues. */
int hg_residuals(hg_handle *h, double res[4]);

/* TimeIntegrator_EstimateTimeStep (DGTimeIntegrator.f90:462-511), including the
 * running maxima umax/amax that the reference keeps across calls. */
int hg_estimate_dt(hg_handle *h, double dt_max, double cfl, double *dt);

/* Forgets the running maxima kept by hg_estimate_dt, as a fresh process would
 * (the reference's SAVEd umax/amax start at zero in every run). */
int hg_reset_dt_state(hg_handle *h);

/* FP64 conservation sums  sum_e sum_ij w_i w_j J Q  (no reference counterpart;
 * a monitor for the mixed-precision form). */
int hg_conservation_sums(hg_handle *h, double out[4]);

/* hg_conservation_sums without allocating or synchronising: the same sums, by the same kernels
 * in the same order, written to out_dev[4] (device memory) on the handle's stream.  No
 * counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  The per-workgroup partial sums
 * live in a library-owned device array of 2 MiB per GPU (kept until the process ends), which
 * limits the mesh to 65536 workgroups of the sum kernel (the 512 x 512 order-7 grid); larger
 * meshes get an error.  Calls that overlap on one GPU share that array and race.  Read out_dev
 * after hg_synchronize (or an event on the handle's stream). */
int hg_conservation_sums_noalloc(hg_handle *h, double *out_dev);

/* Volume monitors of the current state, QuadMesh_Integrals.incf:22-65
 * (Compute_volumeIntegral): out[0] "One" = sum_e w^T J w (the domain area), out[1]
 * "dSnorm" = sqrt(sum_e w^T ((p / rho^gamma - 1)^2 J) w).  FP64 in both forms. */
int hg_volume_integrals(hg_handle *h, double out[2]);

/* Surface monitors over the edges of boundary zone `marker` (1-based, the HiOMesh marker),
 * QuadMesh_Integrals.incf:67-346, from the element-side edge storage (storage(1)) as the last
 * derivative evaluation left it: after hg_rk_step (or hg_batch_rk_step) the traces and gradients
 * of the last RK stage's input state, not of the updated state, as the reference's monitors read
 * them after TimeStep; after hg_time_derivative those of the current state; after
 * hg_upload_state they are formed from the uploaded state.  out[0] "Surface" sum w dS, out[1] "pressure" sum w p dS, out[2] "density"
 * sum w rho dS, out[3..4] "pressure" force sum w p n dS, out[5] "mass-flow" sum w (rho u.n) dS,
 * out[6] "flow" sum w (rho u.n)/rho dS, out[7..8] "viscous" force sum w (tau.n) dS (Navier-Stokes
 * only, tau as ComputeViscousTensor, ViscousFluxes.f90:416-461; zero for Euler).  Replaces the
 * Monitors.f90 surface-monitor calls into Compute_*SurfaceIntegral. */
int hg_surface_integrals(hg_handle *h, int32_t marker, double out[9]);

/* Probe variables, in the order of Probe_Update's cases (IO/Monitors.f90:762-836). */
#define HG_PROBE_RHO 0
#define HG_PROBE_RHOU 1
#define HG_PROBE_RHOV 2
#define HG_PROBE_RHOE 3
#define HG_PROBE_RHOT 4
#define HG_PROBE_RHOUT 5
#define HG_PROBE_RHOVT 6
#define HG_PROBE_RHOET 7
#define HG_PROBE_U 8
#define HG_PROBE_V 9
#define HG_PROBE_P 10
#define HG_PROBE_MACH 11

/* Boundary states from problem-file functions ("Use function = BC1..BC4" or "Initial condition"
 * in a zone block; AssociateBoundaryConditionFunction, BoundaryConditionFunctionDrivers.incf:22-63).
 * The reference evaluates q = BoundaryState(edge % x(:, iXi), time) at every node of the zone's
 * edges in each _UpdateSolution (far field Farfield.incf:179-180, Dirichlet Dirichlet.incf:
 * 182-183, newDirichlet newDirichlet.incf:195-196, pressure outlet PressureOutlet.incf:187-188 and
 * inlet PressureInlet.incf:160-161) in place of the zone state.  Here the caller evaluates the
 * function at the zone's nodes and hands the values over; the kernels read them wherever the
 * zone's state is read.  The nodes of zone `marker` (1-based) are those of its edges in ascending
 * edge id, each edge's N + 1 nodes in the edge's own order (the element side's face nodes).
 * hg_boundary_node_count gives their number, hg_boundary_nodes their coordinates x[2 i + d]
 * (edge % x), hg_set_boundary_state the dimensionless states q[4 i + eq] (q = NULL: back to the
 * zone state).  The values hold until the next call: a function of time is re-evaluated by the
 * caller between steps (the reference takes each RK stage's time; the GPU path keeps one value
 * per step).  Zones of other kinds are refused. */
int hg_boundary_node_count(hg_handle *h, int32_t marker, int64_t *n);
int hg_boundary_nodes(hg_handle *h, int32_t marker, int64_t n, double *x);
int hg_set_boundary_state(hg_handle *h, int32_t marker, int64_t n, const double *q);

/* Probes, Probe_Initialization (IO/Monitors.f90:647-737) after findElementWithCoords: n probes,
 * probe k in element elem[k] (0-based) at reference coordinates (xi[k], eta[k]) (the element's
 * [0,1]^2), tracking variable var[k] (HG_PROBE_*).  The library evaluates the Lagrange
 * interpolants l_j(xi), l_j(eta) once (barycentric form, as NodesAndWeightsClass.f90:334-342)
 * and keeps them on the device.  Replaces any previous set; n = 0 removes it. */
int hg_probes_set(hg_handle *h, int32_t n, const int32_t *elem, const double *xi, const double *eta,
                  const int32_t *var);

/* Probe_Update (IO/Monitors.f90:739-846): out[k] = the dimensionless value of probe k,
 * l_xi^T A l_eta per conservative variable (BilinearForm, MatrixOperations.f90:185-209) and
 * u, v, p, Mach from those; the caller applies Probe_Update's refValues factor.  A = the current
 * state for rho .. rhoe, u, v, p, Mach; for rhot .. rhoet, the QDot of the last
 * hg_time_derivative, which must come after the last step or upload (a Runge-Kutta step keeps
 * no QDot; the reference's QDot at that point is its last stage's). */
int hg_probes_update(hg_handle *h, double *out);

/* Blocks until the handle's stream is idle. */
int hg_synchronize(hg_handle *h);

/* Copies the last error message (NUL-terminated, truncated to n bytes). */
int hg_last_error(char *buf, int32_t n);

/* Kernel-level timing of the handle's stage kernels over the last hg_rk_step
 * calls recorded with HIP events on the handle's stream: returns the number
 * of recorded launches and their total milliseconds for kernel id `which`
 * (0: stage kernel, 1: gradient kernel). Recording is enabled by
 * hg_profile_enable(h, 1).  No reference counterpart (measurement only). */
int hg_profile_enable(hg_handle *h, int32_t on);
int hg_profile_read(hg_handle *h, int32_t which, int64_t *launches, double *ms);

/* Workgroup tile of the stage and gradient kernels: -1 = choose the faster at the
 * first evaluation (timed derivative-only launches), 0 = narrow (elements that fit
 * one wavefront), 1 = wide (up to four wavefronts).  Results are bit-identical for
 * every tile.  No reference counterpart (performance control only). */
int hg_set_tiles(hg_handle *h, int32_t stage_tile, int32_t gradient_tile);

/* hg_rk_step replays one captured HIP graph per step (on = 1, the default) or launches the
 * kernels one by one (on = 0; also used while hg_profile_enable is on).  Results are
 * bit-identical.  No reference counterpart (performance control only). */
int hg_set_graph(hg_handle *h, int32_t on);
int hg_get_tiles(hg_handle *h, int32_t *stage_tile, int32_t *gradient_tile);

/* ------------------------------------------------------------------------ */
/* Batched form: many independent cases (own mesh, polynomial order, Riemann solver and
 * dt) advanced together, one launch per RK stage (plus one per gradient pass for
 * Navier-Stokes) over a work list of element tiles of all cases, the whole step captured
 * in a HIP graph.  The workgroup tile and the order of the work list are chosen by a
 * timing autotuner at the first step.  All cases of a batch share the equations
 * (Euler / Navier-Stokes BR1 / Navier-Stokes IP), the precision, the time scheme and
 * whether artificial dissipation is on (its indicator runs once per case and stage).
 * No reference counterpart: the reference runs one case per process
 * (TimeIntegrator_t%dt per case, DGTimeIntegrator.f90:159-255). */
typedef struct {
    int32_t tile;      /* -1: autotune, 0: narrow, 1: wide workgroup tile */
    int32_t order;     /* -1: autotune, 0: case-major, 1: costliest tiles first, 2: cases interleaved,
                          3: one launch per polynomial order (case-major within it),
                          4: persistent (resident workgroups looping over each order's tiles) */
    int32_t use_graph; /* 1: replay the captured step as a HIP graph, 0: launch kernel by kernel */
} hg_batch_options;

/* The batch keeps pointers to the handles (not copies); they must outlive it.  Between
 * hg_batch_rk_step and hg_batch_synchronize the cases must not be used directly.  A case
 * stepped (hg_rk_step) or given a new state (hg_upload_state) on its own after
 * hg_batch_synchronize is taken as it is at the next hg_batch_rk_step. */
int hg_batch_create(hg_handle *const *cases, int32_t ncases, const hg_batch_options *opt, hg_batch **out);
/* One RK step of every case: t[c], dt[c] per case (hg_rk_step semantics per case). */
int hg_batch_rk_step(hg_batch *b, const double *t, const double *dt);
int hg_batch_synchronize(hg_batch *b);
/* The autotuned (or fixed) tile and order; -1 before the first step. */
int hg_batch_schedule(hg_batch *b, int32_t *tile, int32_t *order);
/* Kernel-level timing of the batched step: with on = 1, hg_batch_rk_step launches the step
 * kernel by kernel (no graph replay; same results) with HIP events around every launch on the
 * batch's stream.  hg_batch_profile_read returns the recorded launches and their total
 * milliseconds for kernel `which` (0: batched stage kernel, 1: batched gradient kernel,
 * 2: per-case residual reduction) since the last hg_batch_profile_enable.  No reference
 * counterpart (measurement only). */
int hg_batch_profile_enable(hg_batch *b, int32_t on);
int hg_batch_profile_read(hg_batch *b, int32_t which, int64_t *launches, double *ms);
int hg_batch_destroy(hg_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* HORSES2D_GPU_H */
