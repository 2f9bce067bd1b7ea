"""ctypes wrapper of the CPU oracle (``oracle/horses_oracle.c``).

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only to check results
or as the timed CPU baseline.  The product package ``horses2d_amd`` never imports it.

The oracle restates, in plain C and in the reference's operation order, the HORSES2D
functions on the DGSEM residual + explicit time-integration path (file:line citations are
on every C function).  What pins each part (see also tests/test_oracle_*.py):

* Gauss nodes/weights, D, hatD, lb, lbw, dlb, dlbw -- numpy ``leggauss`` and closed forms
  (exactness on polynomials, hatD(i,l) = w_l D(l,i) / w_i, lbw = lb / w, dlb = D^T lb).
* Point fluxes -- consistency F*(q,q,n) = F(q).n, rotational invariance, Roe/HLL upwinding
  for supersonic states, exact solver on Toro's textbook tests (known p*).
* Whole residual -- free-stream preservation on straight and curved meshes, discrete
  conservation on periodic meshes, spectral convergence of QDot to the analytic
  -div F of the isentropic vortex, BR1 gradient exact on polynomials.
* RK3 -- the stability polynomial 1 + z + z^2/2 + z^3/6 on y' = lambda y.
* Connectivity -- a brute-force O(E^2) restatement of MeshFile.f90; compared bit-exactly
  with the product's O(E) builder.
* Mortars -- p-non-conforming edges (test_oracle_mortar.py: exact interpolation, T_BKW T_FWD = I,
  free stream, conservation) and subdivided edges (test_oracle_subdivided.py: the same on each
  half, the hanging nodes moved back onto the big edge); end to end, the reference's
  CylinderRe100 (orders 2-8) and VortexTransport (34 subdivided edges, 10 curved) goldens.
* Whole runs -- the reference's Ramp, CylinderRe100 and VortexTransport Finalize values
  (tests/test_benchmark_goldens.py).
* Lax-Friedrichs flux and the far-field viscous treatment -- ABSENT from the reference;
  parity unpinned (consistency / antisymmetry / dissipation-sign checks only).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libhorses_oracle.so")
SRC = os.path.join(HERE, "horses_oracle.c")


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (no FMA contraction, OpenMP over elements)."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        cmd = ["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-o", LIB, SRC, "-lm"]
        subprocess.check_call(cmd)
    return LIB


class _MeshIn(ctypes.Structure):
    _fields_ = [
        ("nnodes", ctypes.c_int), ("points", ctypes.c_void_p),
        ("K", ctypes.c_int), ("quads", ctypes.c_void_p), ("order", ctypes.c_void_p),
        ("nbedges", ctypes.c_int), ("bedges", ctypes.c_void_p), ("bmarker", ctypes.c_void_p),
        ("ncurved", ctypes.c_int), ("np1", ctypes.c_int), ("cedges", ctypes.c_void_p), ("cxy", ctypes.c_void_p),
        ("nmarkers", ctypes.c_int),
        ("bc_type", ctypes.c_void_p), ("bc_int", ctypes.c_void_p), ("bc_real", ctypes.c_void_p),
        ("bc_rs", ctypes.c_void_p),
    ]


class _Options(ctypes.Structure):
    _fields_ = [("Mach", ctypes.c_double), ("Re", ctypes.c_double), ("Pr", ctypes.c_double),
                ("ns", ctypes.c_int), ("riemann", ctypes.c_int), ("viscous", ctypes.c_int),
                ("ip_method", ctypes.c_int), ("sigma0", ctypes.c_double),
                ("ad", ctypes.c_int), ("ad_ceps", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(build())
        _lib.or_create.restype = ctypes.c_void_p
        _lib.or_create.argtypes = [ctypes.POINTER(_MeshIn), ctypes.POINTER(_Options)]
        _lib.or_destroy.argtypes = [ctypes.c_void_p]
        _lib.or_ndof.restype = ctypes.c_int64
        _lib.or_ndof.argtypes = [ctypes.c_void_p]
        _lib.or_num_edges.argtypes = [ctypes.c_void_p]
        _lib.or_last_error.restype = ctypes.c_char_p
        for name in ("or_set_state", "or_get_state", "or_get_qdot", "or_get_grad", "or_get_coords",
                     "or_residuals", "or_conservation_sums", "or_volume_integrals"):
            getattr(_lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_surface_integrals.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        _lib.or_boundary_states.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5
        _lib.or_probes.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5
        _lib.or_boundary_nodes.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
        _lib.or_boundary_nodes.restype = ctypes.c_int64
        _lib.or_set_boundary_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p]
        _lib.or_get_metrics.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_get_connectivity.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 3
        _lib.or_get_edge_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 4
        _lib.or_time_derivative.argtypes = [ctypes.c_void_p, ctypes.c_double]
        _lib.or_edge_mortar.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 5
        _lib.or_edge_subdivided.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _lib.or_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        _lib.or_estimate_dt.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        _lib.or_reset_dt_state.argtypes = [ctypes.c_void_p]
        _lib.or_spectral.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 8
        _lib.or_riemann.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 4
        _lib.or_inviscid_flux.argtypes = [ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 3
        _lib.or_viscous_flux.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 4
        _lib.or_set_threads.argtypes = [ctypes.c_int]
        _lib.or_set_threads.restype = ctypes.c_int
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def set_threads(n: int) -> int:
    """Threads of the oracle's OpenMP loops (n <= 0 leaves them); returns the count in effect."""
    return int(lib().or_set_threads(int(n)))


def spectral(N: int):
    P = N + 1
    out = dict(xi=np.zeros(P), w=np.zeros(P), D=np.zeros((P, P)), hatD=np.zeros((P, P)),
               lb=np.zeros((2, P)), lbw=np.zeros((2, P)), dlb=np.zeros((2, P)), dlbw=np.zeros((2, P)))
    rc = lib().or_spectral(N, *[_ptr(out[k]) for k in ("xi", "w", "D", "hatD", "lb", "lbw", "dlb", "dlbw")])
    assert rc == 0
    return out


def riemann(solver: int, Mach: float, qL, qR, nrm):
    qL = np.ascontiguousarray(qL, np.float64).reshape(-1, 4)
    qR = np.ascontiguousarray(qR, np.float64).reshape(-1, 4)
    nrm = np.ascontiguousarray(nrm, np.float64).reshape(-1, 2)
    F = np.zeros_like(qL)
    lib().or_riemann(solver, Mach, qL.shape[0], _ptr(qL), _ptr(qR), _ptr(nrm), _ptr(F))
    return F


def inviscid_flux(Mach: float, q):
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 4)
    f = np.zeros_like(q); g = np.zeros_like(q)
    lib().or_inviscid_flux(Mach, q.shape[0], _ptr(q), _ptr(f), _ptr(g))
    return f, g


def viscous_flux(Re: float, Pr: float, q, dq):
    q = np.ascontiguousarray(q, np.float64).reshape(-1, 4)
    dq = np.ascontiguousarray(dq, np.float64).reshape(-1, 8)
    f = np.zeros_like(q); g = np.zeros_like(q)
    lib().or_viscous_flux(Re, Pr, q.shape[0], _ptr(q), _ptr(dq), _ptr(f), _ptr(g))
    return f, g


class Oracle:
    """One HORSES2D-like mesh + state on the CPU."""

    def __init__(self, case):
        from horses2d_amd import cases as C
        m = case.mesh
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        pts = arr(m.points, np.float64)
        quads = arr(m.quads, np.int32)
        order = arr(case.N, np.int32)
        be = arr(m.bedges, np.int32)
        bm = arr(m.bmarker, np.int32)
        zones = m.zones
        bct = arr([z.kind for z in zones], np.int32)
        bci = arr([z.ints() for z in zones], np.int32)
        bcr = arr([z.reals() for z in zones], np.float64)
        bcrs = arr([z.riemann for z in zones], np.int32)
        mi = _MeshIn()
        mi.nnodes = pts.shape[0]; mi.points = _ptr(pts)
        mi.K = quads.shape[0]; mi.quads = _ptr(quads); mi.order = _ptr(order)
        mi.nbedges = be.shape[0]; mi.bedges = _ptr(be); mi.bmarker = _ptr(bm)
        if m.cedges is not None and len(m.cedges):
            ce = arr(m.cedges, np.int32); cx = arr(m.cxy, np.float64)
            mi.ncurved = ce.shape[0]; mi.np1 = cx.shape[1]; mi.cedges = _ptr(ce); mi.cxy = _ptr(cx)
        mi.nmarkers = len(zones)
        mi.bc_type = _ptr(bct); mi.bc_int = _ptr(bci); mi.bc_real = _ptr(bcr); mi.bc_rs = _ptr(bcrs)
        op = _Options(case.Mach, case.Re, case.Pr, int(case.navier_stokes), C.RIEMANN[case.riemann],
                      C.VISCOUS[case.viscous], C.IP_METHOD[case.ip_method], case.sigma0,
                      C.ARTIFICIAL_DISSIPATION[case.artificial_dissipation], case.ad_intensity)
        h = lib().or_create(ctypes.byref(mi), ctypes.byref(op))
        if not h:
            raise RuntimeError("oracle: " + lib().or_last_error().decode())
        self.h = h
        self.case = case
        self.ndof = int(lib().or_ndof(h))
        self.nodes = self.ndof // 4
        self.E = int(lib().or_num_edges(h))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def coords(self):
        X = np.zeros((self.nodes, 2))
        lib().or_get_coords(self.h, _ptr(X))
        return X

    def metrics(self):
        Ja = np.zeros((self.nodes, 4)); jac = np.zeros(self.nodes)
        lib().or_get_metrics(self.h, _ptr(Ja), _ptr(jac))
        return Ja, jac

    def connectivity(self):
        edges = np.zeros((self.E, 8), np.int32)
        elem = np.zeros((self.case.mesh.K, 12), np.int32)
        partner = np.zeros(self.E, np.int32)
        lib().or_get_connectivity(self.h, _ptr(edges), _ptr(elem), _ptr(partner))
        return edges, elem, partner

    def edge_geometry(self, maxP: int = 16):
        nx = np.zeros((self.E, maxP)); ny = np.zeros((self.E, maxP)); dS = np.zeros((self.E, maxP))
        invh = np.zeros(self.E)
        lib().or_get_edge_geometry(self.h, maxP, _ptr(nx), _ptr(ny), _ptr(dS), _ptr(invh))
        return nx, ny, dS, invh

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, np.float64)
        assert Q.size == self.ndof
        lib().or_set_state(self.h, _ptr(Q))

    def state(self):
        Q = np.zeros(self.ndof); lib().or_get_state(self.h, _ptr(Q)); return Q

    def qdot(self):
        Q = np.zeros(self.ndof); lib().or_get_qdot(self.h, _ptr(Q)); return Q

    def grad(self):
        Q = np.zeros(2 * self.ndof); lib().or_get_grad(self.h, _ptr(Q)); return Q

    def edge_mortar(self, k: int):
        """(low side or -1, N, NLow, T_forward (N+1, NLow+1), T_backward (NLow+1, N+1)) of edge k."""
        low, N, NL = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        Tf = np.zeros(17 * 17); Tb = np.zeros(17 * 17)
        rc = lib().or_edge_mortar(self.h, k, ctypes.byref(low), ctypes.byref(N), ctypes.byref(NL), _ptr(Tf), _ptr(Tb))
        assert rc == 0
        n, nl = N.value + 1, NL.value + 1
        return low.value, N.value, NL.value, Tf[:n * nl].reshape(n, nl), Tb[:n * nl].reshape(nl, n)

    def edge_subdivided(self, k: int):
        """A subdivided edge's linking and its eight mortar matrices (None for other edges): the
        orders of the big element and the mortars, the small elements' directions (1 forward, -1
        backward), the elements LEFT / NORTH / SOUTH and their edge locations."""
        info = np.zeros(13, np.int32)
        mats = np.zeros(8 * 17 * 17)
        rc = lib().or_edge_subdivided(self.h, k, _ptr(info), _ptr(mats))
        if rc == 1:
            return None
        assert rc == 0
        NL, NN, NS, NRN, NRS = (int(v) + 1 for v in info[:5])
        shapes = {"TLN_F": (NN, NL), "TLS_F": (NS, NL), "TLN_B": (NL, NN), "TLS_B": (NL, NS),
                  "TRN_F": (NN, NRN), "TRS_F": (NS, NRS), "TRN_B": (NRN, NN), "TRS_B": (NRS, NS)}
        out = {"N_left": NL - 1, "N_north": NN - 1, "N_south": NS - 1, "dir_north": int(info[5]),
               "dir_south": int(info[6]), "elements": info[7:10].tolist(), "locations": info[10:13].tolist()}
        for t, (name, (r, c)) in enumerate(shapes.items()):
            out[name] = mats[t * 289: t * 289 + r * c].reshape(r, c).copy()
        return out

    def time_derivative(self, t: float = 0.0):
        lib().or_time_derivative(self.h, t)
        return self.qdot()

    def step(self, t: float, dt: float, scheme: str = "rk3"):
        from horses2d_amd.cases import SCHEME
        rc = lib().or_step(self.h, SCHEME[scheme], t, dt)
        assert rc == 0

    def residuals(self):
        r = np.zeros(4); lib().or_residuals(self.h, _ptr(r)); return r

    def estimate_dt(self, dt_max: float, cfl: float) -> float:
        out = np.zeros(1)
        lib().or_estimate_dt(self.h, dt_max, cfl, _ptr(out))
        return float(out[0])

    def conservation_sums(self):
        r = np.zeros(4); lib().or_conservation_sums(self.h, _ptr(r)); return r

    def volume_integrals(self):
        """[One, dSnorm] of the current state (QuadMesh_Integrals.incf:22-65)."""
        r = np.zeros(2); lib().or_volume_integrals(self.h, _ptr(r)); return r

    def surface_integrals(self, marker: int):
        """[Surface, pressure, density, pressure force x, y, mass-flow, flow, viscous force x, y]
        over the edges of zone ``marker`` (1-based; QuadMesh_Integrals.incf:67-346)."""
        r = np.zeros(9)
        if lib().or_surface_integrals(self.h, int(marker), _ptr(r)) != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return r

    def boundary_nodes(self, marker: int):
        """Coordinates [n, 2] of zone ``marker``'s edge nodes (ascending edge id, each edge's nodes
        in its own order): where a boundary-state function is evaluated."""
        n = lib().or_boundary_nodes(self.h, int(marker), 0, None)
        if n < 0:
            raise RuntimeError(lib().or_last_error().decode())
        x = np.zeros((int(n), 2))
        if lib().or_boundary_nodes(self.h, int(marker), int(n), _ptr(x)) != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return x

    def set_boundary_state(self, marker: int, q):
        """Boundary-state function values [n, 4] (dimensionless) at those nodes, in place of the
        zone state; None drops them."""
        if q is None:
            if lib().or_set_boundary_state(self.h, int(marker), -1, None) != 0:
                raise RuntimeError(lib().or_last_error().decode())
            return
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 4)
        if lib().or_set_boundary_state(self.h, int(marker), q.shape[0], _ptr(q)) != 0:
            raise RuntimeError(lib().or_last_error().decode())

    def boundary_states(self, marker: int):
        """(trace, normal, uB, uSB, viscous type) per edge node of zone ``marker`` from the
        last time derivative or stage; the viscous types are 1 Dirichlet, 2 Neumann, 4 adiabatic."""
        n = lib().or_boundary_states(self.h, int(marker), None, None, None, None, None)
        if n < 0:
            raise RuntimeError(lib().or_last_error().decode())
        qt, nrm, ub, usb = np.zeros((n, 4)), np.zeros((n, 2)), np.zeros((n, 4)), np.zeros((n, 4))
        vbc = np.zeros(n, np.int32)
        lib().or_boundary_states(self.h, int(marker), _ptr(qt), _ptr(nrm), _ptr(ub), _ptr(usb), _ptr(vbc))
        return qt, nrm, ub, usb, vbc

    PROBE_VARIABLES = ("rho", "rhou", "rhov", "rhoe", "rhot", "rhout", "rhovt", "rhoet", "u", "v", "p", "Mach")

    def probes(self, elem, xi, eta, variables):
        """Probe_Update (IO/Monitors.f90:739-846) at reference coordinates (xi, eta) of elements
        ``elem`` (0-based); QDot variables read the last time derivative or stage."""
        e = np.ascontiguousarray(elem, dtype=np.int32)
        x = np.ascontiguousarray(xi, dtype=np.float64)
        y = np.ascontiguousarray(eta, dtype=np.float64)
        v = np.ascontiguousarray([self.PROBE_VARIABLES.index(a) if isinstance(a, str) else int(a)
                                  for a in variables], dtype=np.int32)
        r = np.zeros(len(e))
        if lib().or_probes(self.h, len(e), _ptr(e), _ptr(x), _ptr(y), _ptr(v), _ptr(r)) != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return r

    def integrate(self, steps: int, t0: float = 0.0, dt_max: float = 1.0e30, t_end: float = 1.0e30,
                  scheme: str = "rk3", mode: str = "transient", residual_target: float = 0.0):
        """TimeIntegrator_Integrate (DGTimeIntegrator.f90:159-255): returns (t, dts, iterations,
        residual history).  Transient: the last step ends at t_end; steady: stop after the first
        iteration whose max residual (of its last stage) is below the target."""
        t = t0
        dt = self.estimate_dt(dt_max, self.case.cfl)
        dts, hist = [], []
        steady = mode == "steady"
        eps = 2.0 * np.finfo(np.float64).eps
        it = 0
        for it in range(1, steps + 1):
            if not steady and t + dt >= t_end:
                dt = t_end - t
            self.step(t, dt, scheme)
            dts.append(dt)
            t = t + dt
            res = self.residuals()
            hist.append(res)
            if not steady and abs(t_end - t) <= eps * max(abs(t), abs(t_end)):
                break
            if it % 100 == 0:
                dt = self.estimate_dt(dt_max, self.case.cfl)
            if steady and np.max(np.abs(res)) < residual_target:
                break
        return t, np.array(dts), it, np.array(hist)
