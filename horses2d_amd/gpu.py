"""Host-side mirror of HORSES2D's solver objects over the C ABI of ``libhorses2d_gpu.so``.

``DGSEM`` plays the role of ``DGSEM_t`` (Approximation/DGSEM_class.f90:64-78): it builds
the mesh from a HiOMesh description, applies the initial condition on the host (the
``libproblemfile`` hook stays on the host) and owns the device state.  ``integrate``
restates ``TimeIntegrator_Integrate`` (Approximation/DGTimeIntegrator.f90:159-255):
final-time clamp, one ``TimeScheme`` call per iteration, residual monitor, and the
time-step re-estimate every 100 iterations.

The HIP library is required: nothing here falls back to a CPU path.  A missing or
unloadable library raises ``RuntimeError`` at first use.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import cases as C

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhorses2d_gpu.so")

PRECISION = {"fp64": 0, "mixed": 1}


class _MeshDesc(ctypes.Structure):
    _fields_ = [
        ("nnodes", ctypes.c_int32), ("points", ctypes.c_void_p),
        ("nelements", ctypes.c_int32), ("quads", ctypes.c_void_p), ("order", ctypes.c_void_p),
        ("nbedges", ctypes.c_int32), ("bedges", ctypes.c_void_p), ("bmarker", ctypes.c_void_p),
        ("ncurved", ctypes.c_int32), ("np1", ctypes.c_int32), ("cedges", ctypes.c_void_p), ("cxy", ctypes.c_void_p),
        ("nmarkers", ctypes.c_int32),
        ("bc_type", ctypes.c_void_p), ("bc_int", ctypes.c_void_p), ("bc_real", ctypes.c_void_p),
        ("bc_riemann", ctypes.c_void_p),
    ]


class _Physics(ctypes.Structure):
    _fields_ = [("mach", ctypes.c_double), ("reynolds", ctypes.c_double), ("prandtl", ctypes.c_double)]


class _Options(ctypes.Structure):
    _fields_ = [("navier_stokes", ctypes.c_int32), ("riemann", ctypes.c_int32), ("viscous", ctypes.c_int32),
                ("ip_method", ctypes.c_int32), ("sigma0", ctypes.c_double), ("precision", ctypes.c_int32),
                ("scheme", ctypes.c_int32), ("artificial_dissipation", ctypes.c_int32),
                ("ad_intensity", ctypes.c_double)]


class _BatchOptions(ctypes.Structure):
    _fields_ = [("tile", ctypes.c_int32), ("order", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


# every symbol declared in include/horses2d_gpu.h
EXPORTS = [
    "hg_create", "hg_destroy", "hg_ndof", "hg_get_node_coords", "hg_get_connectivity", "hg_num_edges",
    "hg_build_connectivity", "hg_build_geometry",
    "hg_upload_state", "hg_download_state", "hg_download_qdot", "hg_time_derivative", "hg_rk_step",
    "hg_residuals", "hg_estimate_dt", "hg_reset_dt_state", "hg_conservation_sums", "hg_conservation_sums_noalloc",
    "hg_volume_integrals",
    "hg_surface_integrals", "hg_boundary_node_count", "hg_boundary_nodes", "hg_set_boundary_state",
    "hg_probes_set", "hg_probes_update", "hg_synchronize",
    "hg_last_error", "hg_profile_enable", "hg_profile_read", "hg_set_tiles", "hg_get_tiles", "hg_set_graph", "hg_batch_create", "hg_batch_rk_step",
    "hg_batch_synchronize", "hg_batch_schedule", "hg_batch_profile_enable", "hg_batch_profile_read",
    "hg_batch_destroy",
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads the HIP library and declares the C signatures (raises if missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library {path} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(path)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    sig = {
        "hg_create": ([vp, vp, vp, vp], ctypes.c_int),
        "hg_destroy": ([vp], ctypes.c_int),
        "hg_ndof": ([vp], i64),
        "hg_get_node_coords": ([vp, vp, i64], ctypes.c_int),
        "hg_get_connectivity": ([vp, vp, vp, vp], ctypes.c_int),
        "hg_num_edges": ([vp], ctypes.c_int),
        "hg_build_connectivity": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "hg_build_geometry": ([vp, i64, vp, vp, vp], ctypes.c_int),
        "hg_upload_state": ([vp, vp, i64], ctypes.c_int),
        "hg_download_state": ([vp, vp, i64], ctypes.c_int),
        "hg_download_qdot": ([vp, vp, i64], ctypes.c_int),
        "hg_time_derivative": ([vp, dbl], ctypes.c_int),
        "hg_rk_step": ([vp, dbl, dbl], ctypes.c_int),
        "hg_residuals": ([vp, vp], ctypes.c_int),
        "hg_estimate_dt": ([vp, dbl, dbl, vp], ctypes.c_int),
        "hg_reset_dt_state": ([vp], ctypes.c_int),
        "hg_conservation_sums": ([vp, vp], ctypes.c_int),
        "hg_conservation_sums_noalloc": ([vp, vp], ctypes.c_int),
        "hg_volume_integrals": ([vp, vp], ctypes.c_int),
        "hg_surface_integrals": ([vp, i32, vp], ctypes.c_int),
        "hg_boundary_node_count": ([vp, i32, vp], ctypes.c_int),
        "hg_boundary_nodes": ([vp, i32, i64, vp], ctypes.c_int),
        "hg_set_boundary_state": ([vp, i32, i64, vp], ctypes.c_int),
        "hg_probes_set": ([vp, i32, vp, vp, vp, vp], ctypes.c_int),
        "hg_probes_update": ([vp, vp], ctypes.c_int),
        "hg_synchronize": ([vp], ctypes.c_int),
        "hg_last_error": ([vp, i32], ctypes.c_int),
        "hg_profile_enable": ([vp, i32], ctypes.c_int),
        "hg_profile_read": ([vp, i32, vp, vp], ctypes.c_int),
        "hg_set_tiles": ([vp, i32, i32], ctypes.c_int),
        "hg_get_tiles": ([vp, vp, vp], ctypes.c_int),
        "hg_set_graph": ([vp, i32], ctypes.c_int),
        "hg_batch_create": ([vp, i32, vp, vp], ctypes.c_int),
        "hg_batch_rk_step": ([vp, vp, vp], ctypes.c_int),
        "hg_batch_synchronize": ([vp], ctypes.c_int),
        "hg_batch_schedule": ([vp, vp, vp], ctypes.c_int),
        "hg_batch_profile_enable": ([vp, i32], ctypes.c_int),
        "hg_batch_profile_read": ([vp, i32, vp, vp], ctypes.c_int),
        "hg_batch_destroy": ([vp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _almost_equal(a: float, b: float) -> bool:
    """AlmostEqual (Generic/Utilities.f90:37-63): within two machine epsilons, relative."""
    eps = 2.0 * np.finfo(np.float64).eps
    if a == 0.0 or b == 0.0:
        return abs(a - b) <= eps
    return abs(b - a) <= eps * max(abs(a), abs(b))


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load_library().hg_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {last_error()}")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def mesh_desc(case: C.Case, keep: list) -> _MeshDesc:
    """Fills the ``hg_mesh_desc`` of a case; ``keep`` holds the arrays alive."""
    m = case.mesh

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    pts = arr(m.points, np.float64)
    quads = arr(m.quads, np.int32)
    order = arr(case.N, np.int32)
    be = arr(m.bedges, np.int32)
    bm = arr(m.bmarker, np.int32)
    md = _MeshDesc()
    md.nnodes = pts.shape[0]; md.points = _ptr(pts)
    md.nelements = quads.shape[0]; md.quads = _ptr(quads); md.order = _ptr(order)
    md.nbedges = be.shape[0]; md.bedges = _ptr(be); md.bmarker = _ptr(bm)
    if m.cedges is not None and len(m.cedges):
        ce = arr(m.cedges, np.int32); cx = arr(m.cxy, np.float64)
        md.ncurved = ce.shape[0]; md.np1 = cx.shape[1]; md.cedges = _ptr(ce); md.cxy = _ptr(cx)
    zones = m.zones
    md.nmarkers = len(zones)
    md.bc_type = _ptr(arr([z.kind for z in zones], np.int32))
    md.bc_int = _ptr(arr([z.ints() for z in zones], np.int32))
    md.bc_real = _ptr(arr([z.reals() for z in zones], np.float64))
    md.bc_riemann = _ptr(arr([z.riemann for z in zones], np.int32))
    return md


def build_connectivity(case: C.Case):
    """Host-only connectivity of the product's mesh builder (no GPU needed)."""
    L = load_library()
    keep = []
    md = mesh_desc(case, keep)
    ne = ctypes.c_int32()
    _check(L.hg_build_connectivity(ctypes.byref(md), ctypes.byref(ne), None, None, None), "hg_build_connectivity")
    E = ne.value
    edges = np.zeros((E, 8), np.int32)
    elems = np.zeros((case.mesh.K, 12), np.int32)
    partner = np.zeros(E, np.int32)
    _check(L.hg_build_connectivity(ctypes.byref(md), ctypes.byref(ne), _ptr(edges), _ptr(elems), _ptr(partner)),
           "hg_build_connectivity")
    return edges, elems, partner


def build_geometry(case: C.Case):
    """Host-only node coordinates, metric terms and Jacobian of the product's builder."""
    L = load_library()
    keep = []
    md = mesh_desc(case, keep)
    nodes = case.ndof // 4
    X = np.zeros((nodes, 2)); Ja = np.zeros((nodes, 4)); jac = np.zeros(nodes)
    _check(L.hg_build_geometry(ctypes.byref(md), nodes, _ptr(X), _ptr(Ja), _ptr(jac)), "hg_build_geometry")
    return X, Ja, jac


class DGSEM:
    """One case on the GPU (the HORSES2D ``DGSEM_t`` object, device side)."""

    def __init__(self, case: C.Case, precision: str = "fp64", apply_ic: bool = True):
        L = load_library()
        keep = []
        md = mesh_desc(case, keep)

        ph = _Physics(case.Mach, case.Re, case.Pr)
        op = _Options(int(case.navier_stokes), C.RIEMANN[case.riemann],
                      C.VISCOUS[case.viscous] if case.navier_stokes else 0, C.IP_METHOD[case.ip_method],
                      case.sigma0, PRECISION[precision], C.SCHEME[case.scheme],
                      C.ARTIFICIAL_DISSIPATION[case.artificial_dissipation], case.ad_intensity)
        h = ctypes.c_void_p()
        _check(L.hg_create(ctypes.byref(md), ctypes.byref(ph), ctypes.byref(op), ctypes.byref(h)), "hg_create")
        self._h = h
        self._L = L
        self.case = case
        self.precision = precision
        self.ndof = int(L.hg_ndof(h))
        self.nodes = self.ndof // 4
        self.t = 0.0
        self.iter = 0
        if apply_ic and case.ic is not None:
            self.upload(case.initial_state(self.coords()))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.hg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- data movement
    def coords(self) -> np.ndarray:
        X = np.zeros((self.nodes, 2))
        _check(self._L.hg_get_node_coords(self._h, _ptr(X), self.nodes), "hg_get_node_coords")
        return X

    def connectivity(self):
        E = int(self._L.hg_num_edges(self._h))
        edges = np.zeros((E, 8), np.int32)
        elems = np.zeros((self.case.mesh.K, 12), np.int32)
        partner = np.zeros(E, np.int32)
        _check(self._L.hg_get_connectivity(self._h, _ptr(edges), _ptr(elems), _ptr(partner)), "hg_get_connectivity")
        return edges, elems, partner

    def upload(self, Q: np.ndarray):
        Q = np.ascontiguousarray(Q, np.float64)
        _check(self._L.hg_upload_state(self._h, _ptr(Q), Q.size), "hg_upload_state")

    def state(self) -> np.ndarray:
        Q = np.zeros(self.ndof)
        _check(self._L.hg_download_state(self._h, _ptr(Q), self.ndof), "hg_download_state")
        return Q

    def qdot(self) -> np.ndarray:
        Q = np.zeros(self.ndof)
        _check(self._L.hg_download_qdot(self._h, _ptr(Q), self.ndof), "hg_download_qdot")
        return Q

    # ---- hot path
    def time_derivative(self, t: float = 0.0) -> np.ndarray:
        _check(self._L.hg_time_derivative(self._h, t), "hg_time_derivative")
        return self.qdot()

    def rk_step(self, t: float, dt: float):
        _check(self._L.hg_rk_step(self._h, t, dt), "hg_rk_step")

    def residuals(self) -> np.ndarray:
        r = np.zeros(4)
        _check(self._L.hg_residuals(self._h, _ptr(r)), "hg_residuals")
        return r

    def estimate_dt(self, dt_max: float, cfl: float) -> float:
        out = np.zeros(1)
        _check(self._L.hg_estimate_dt(self._h, dt_max, cfl, _ptr(out)), "hg_estimate_dt")
        return float(out[0])

    def reset_dt_state(self):
        _check(self._L.hg_reset_dt_state(self._h), "hg_reset_dt_state")

    def conservation_sums(self) -> np.ndarray:
        r = np.zeros(4)
        _check(self._L.hg_conservation_sums(self._h, _ptr(r)), "hg_conservation_sums")
        return r

    def volume_integrals(self) -> np.ndarray:
        """[One (area), dSnorm] of the current state (QuadMesh_Integrals.incf:22-65)."""
        r = np.zeros(2)
        _check(self._L.hg_volume_integrals(self._h, _ptr(r)), "hg_volume_integrals")
        return r

    def surface_integrals(self, marker: int) -> np.ndarray:
        """[Surface, pressure, density, pressure force x, y, mass-flow, flow, viscous force x, y]
        over the edges of zone ``marker`` (1-based; QuadMesh_Integrals.incf:67-346)."""
        r = np.zeros(9)
        _check(self._L.hg_surface_integrals(self._h, int(marker), _ptr(r)), "hg_surface_integrals")
        return r

    def boundary_nodes(self, marker: int) -> np.ndarray:
        """Coordinates [n, 2] of zone ``marker``'s edge nodes (ascending edge id, each edge's nodes
        in its own order): where the caller evaluates a boundary-state function."""
        n = ctypes.c_int64(0)
        _check(self._L.hg_boundary_node_count(self._h, int(marker), ctypes.byref(n)), "hg_boundary_node_count")
        x = np.zeros((n.value, 2))
        _check(self._L.hg_boundary_nodes(self._h, int(marker), n.value, _ptr(x)), "hg_boundary_nodes")
        return x

    def set_boundary_state(self, marker: int, q):
        """A "Use function" zone's dimensionless states [n, 4] at those nodes (None: back to the
        zone state), as hg_set_boundary_state."""
        if q is None:
            _check(self._L.hg_set_boundary_state(self._h, int(marker), 0, None), "hg_set_boundary_state")
            return
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 4)
        _check(self._L.hg_set_boundary_state(self._h, int(marker), q.shape[0], _ptr(q)), "hg_set_boundary_state")

    PROBE_VARIABLES = ("rho", "rhou", "rhov", "rhoe", "rhot", "rhout", "rhovt", "rhoet", "u", "v", "p", "Mach")

    def probes_set(self, elem, xi, eta, variables):
        """Probes in elements ``elem`` (0-based) at reference coordinates (xi, eta) tracking
        ``variables`` (names of Probe_Update, IO/Monitors.f90:762-836, or HG_PROBE_* codes)."""
        e = np.ascontiguousarray(elem, dtype=np.int32)
        x = np.ascontiguousarray(xi, dtype=np.float64)
        y = np.ascontiguousarray(eta, dtype=np.float64)
        v = np.ascontiguousarray([self.PROBE_VARIABLES.index(a) if isinstance(a, str) else int(a)
                                  for a in variables], dtype=np.int32)
        if not (len(e) == len(x) == len(y) == len(v)):
            raise ValueError("probes_set: arrays of different lengths")
        _check(self._L.hg_probes_set(self._h, len(e), _ptr(e), _ptr(x), _ptr(y), _ptr(v)), "hg_probes_set")
        self._nprobes = len(e)

    def probes_update(self) -> np.ndarray:
        """Dimensionless probe values of the current state (Probe_Update)."""
        r = np.zeros(getattr(self, "_nprobes", 0))
        _check(self._L.hg_probes_update(self._h, _ptr(r) if len(r) else None), "hg_probes_update")
        return r

    def synchronize(self):
        _check(self._L.hg_synchronize(self._h), "hg_synchronize")

    def profile(self, on: bool = True):
        _check(self._L.hg_profile_enable(self._h, 1 if on else 0), "hg_profile_enable")

    def profile_read(self, which: int = 0):
        n = ctypes.c_int64(); ms = ctypes.c_double()
        _check(self._L.hg_profile_read(self._h, which, ctypes.byref(n), ctypes.byref(ms)), "hg_profile_read")
        return int(n.value), float(ms.value)

    def set_tiles(self, stage: int = -1, gradient: int = -1):
        """Workgroup tile per kernel: -1 autotune at the next evaluation, 0 narrow, 1 wide."""
        _check(self._L.hg_set_tiles(self._h, stage, gradient), "hg_set_tiles")

    def set_graph(self, on: bool = True):
        """Replay each RK step as a captured HIP graph (default) or launch kernel by kernel."""
        _check(self._L.hg_set_graph(self._h, 1 if on else 0), "hg_set_graph")

    def tiles(self):
        s = ctypes.c_int32(); g = ctypes.c_int32()
        _check(self._L.hg_get_tiles(self._h, ctypes.byref(s), ctypes.byref(g)), "hg_get_tiles")
        return int(s.value), int(g.value)

    @property
    def handle(self):
        return self._h

    def integrate(self, steps: Optional[int] = None, dt_max: Optional[float] = None,
                  t_end: Optional[float] = None, monitor: bool = False, mode: Optional[str] = None,
                  residual_target: Optional[float] = None):
        """TimeIntegrator_Integrate (DGTimeIntegrator.f90:159-255) after NewTimeIntegrator's time-step
        estimate (:86-157).  Transient: the last step is clamped to end at t_end and the loop
        stops there.  Steady: after each step the residuals (max |QDot| of the last stage,
        Monitors%UpdateValues) are read and the loop stops once their maximum is below the
        residual target.  The time step is re-estimated every 100 iterations.  Defaults come
        from the case (steps, dt cap, mode, target, t_end); returns the residual history
        (every iteration when ``monitor`` or steady, else empty); ``self.iter`` is the last
        iteration, as the reference's Integrator % iter."""
        case = self.case
        steps = case.steps if steps is None else steps
        dt_max = case.dt_max if dt_max is None else dt_max
        t_end = case.t_end if t_end is None else t_end
        steady = (case.mode if mode is None else mode) == "steady"
        target = case.residual_target if residual_target is None else residual_target
        dt = self.estimate_dt(dt_max, case.cfl)
        history = []
        first = self.iter + 1
        for it in range(first, first + steps):
            self.iter = it
            if not steady and self.t + dt >= t_end:
                dt = t_end - self.t
            self.rk_step(self.t, dt)
            self.t = self.t + dt
            if not steady and _almost_equal(self.t, t_end):
                if monitor:
                    history.append(self.residuals())
                break
            res = self.residuals() if (monitor or steady) else None
            if res is not None:
                history.append(res)
            if it % 100 == 0:
                dt = self.estimate_dt(dt_max, case.cfl)
            if steady and np.max(np.abs(res)) < target:
                break
        self.dt = dt
        return np.array(history)


class Batch:
    """Many DGSEM handles advanced together (hg_batch_*): one launch per RK stage over the element
    tiles of all cases, the step replayed as a HIP graph; each case keeps its own dt."""

    def __init__(self, members, tile: int = -1, order: int = -1, use_graph: bool = True):
        self._L = load_library()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.handle for m in self.members])
        opt = _BatchOptions(tile, order, 1 if use_graph else 0)
        self._b = ctypes.c_void_p()
        _check(self._L.hg_batch_create(arr, len(self.members), ctypes.byref(opt), ctypes.byref(self._b)),
               "hg_batch_create")

    def rk_step(self, t, dt):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(t, np.float64), (len(self.members),)))
        dt = np.ascontiguousarray(np.asarray(dt, np.float64))
        assert dt.shape == (len(self.members),)
        _check(self._L.hg_batch_rk_step(self._b, _ptr(t), _ptr(dt)), "hg_batch_rk_step")

    def synchronize(self):
        _check(self._L.hg_batch_synchronize(self._b), "hg_batch_synchronize")

    def schedule(self):
        t = ctypes.c_int32(); o = ctypes.c_int32()
        _check(self._L.hg_batch_schedule(self._b, ctypes.byref(t), ctypes.byref(o)), "hg_batch_schedule")
        return int(t.value), int(o.value)

    def profile(self, on: bool = True):
        """Kernel timing: launch the step kernel by kernel with HIP events around each launch."""
        _check(self._L.hg_batch_profile_enable(self._b, 1 if on else 0), "hg_batch_profile_enable")

    def profile_read(self, which: int = 0):
        """(launches, total ms) of kernel ``which``: 0 stage, 1 gradient, 2 residual reduction."""
        n = ctypes.c_int64(); ms = ctypes.c_double()
        _check(self._L.hg_batch_profile_read(self._b, which, ctypes.byref(n), ctypes.byref(ms)),
               "hg_batch_profile_read")
        return int(n.value), float(ms.value)

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            self._L.hg_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
