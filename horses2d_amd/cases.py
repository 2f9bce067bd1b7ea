"""Case descriptions: HiOMesh-equivalent meshes, boundary zones and initial conditions.

This is the host-side stand-in for what HORSES2D reads from a ``.HiOMesh`` file
(Solver/src/Mesh/MeshFile.f90:77-279; on-disk schema written by
Utils/MeshPreprocessor/Functions/MeshFile.m:478-537), the boundary-zone blocks of a
``.HiOCase`` file (BoundaryConditions/DGBoundaryConditions.f90:244-357) and the
``UserDefinedInitialCondition`` hook of ``libproblemfile`` (IO/ProblemFile.f90:24-199).

Conventions follow the reference meshes shipped in TestCases/Cases/VortexTransport/MESH:
nodes are numbered with y fastest, elements are numbered with y fastest, and
``points_of_quads`` lists the four corners counter-clockwise starting bottom-left.
Coordinates are already divided by the Reynolds length (MeshFile.f90:136).

All state vectors are dimensionless: rho/rho_ref, rho*u/(rho_ref*a_ref), rho*e/p_ref with
a_ref = sqrt(p_ref/rho_ref) (VariableConversion.f90:32-42, PhysicsNS.f90:424-437).
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional

import numpy as np

GAMMA = 1.4
GM1 = 0.4

# Boundary-condition codes of Generic/Defines.h:103-111
BC_PERIODIC = 1
BC_DIRICHLET = 2
BC_EULERWALL = 3
BC_VISCOUSWALL = 4
BC_FARFIELD = 5
BC_PRESSUREOUTLET = 6
BC_PRESSUREINLET = 7
BC_NEWDIRICHLET = 9
# pressure outlet "Outflow type" (PressureOutlet.incf:66-72)
OUTFLOW = {"partially non reflective": 0, "reflective": 1}
# newDirichlet "Mode" and "Outflow pressure" (DGBoundaryConditions.f90:47-54)
NEWDIRICHLET_MODE = {"specify speed": 1, "specify total pressure": 2}
NEWDIRICHLET_OUTFLOW = {"static": 1, "total": 2}

RIEMANN = {"roe": 0, "hll": 1, "exact": 2, "lf": 3}
VISCOUS = {"none": 0, "br1": 1, "ip": 2}
IP_METHOD = {"sipg": 0, "nipg": 1, "iipg": 2}
SCHEME = {"euler": 0, "rk3": 1, "rk5": 2}
# "Artificial dissipation (0/1)" with "type" Physical or Laplacian and "indicator" Jumps-based
# (Setup.f90:333-349, DGArtificialDissipation.f90:100-137)
ARTIFICIAL_DISSIPATION = {"none": 0, "physical": 1, "laplacian": 2}


@dataclasses.dataclass
class Zone:
    """One ``# define zone`` block."""

    kind: int
    direction: int = 0          # periodic: 1 = x-translation, 2 = y-translation
    partner: int = 0            # periodic: connected marker (1-based)
    isothermal: int = 0         # viscous wall
    q: tuple = (0.0, 0.0, 0.0, 0.0)   # far field, (new)Dirichlet, pressure outlet/inlet state (dimensionless)
    Tw: float = 1.0
    wall_velocity: tuple = (0.0, 0.0)
    riemann: int = -1           # per-zone Riemann solver, -1 = interior method
    outflow: int = 0            # pressure outlet: OUTFLOW[...]
    mode: int = 1               # newDirichlet: NEWDIRICHLET_MODE[...]
    outflow_pressure: int = 1   # newDirichlet: NEWDIRICHLET_OUTFLOW[...]
    function: str = ""          # "Use function": the problem-file function giving the state per node

    def ints(self):
        if self.kind == BC_PERIODIC:
            return (self.direction, self.partner)
        if self.kind == BC_PRESSUREOUTLET:
            return (self.outflow, 0)
        if self.kind == BC_NEWDIRICHLET:
            return (self.mode, self.outflow_pressure)
        return (self.isothermal, 0)

    def reals(self):
        if self.kind == BC_VISCOUSWALL:
            return (float(self.Tw), float(self.wall_velocity[0]), float(self.wall_velocity[1]), 0.0)
        return tuple(float(v) for v in self.q)


@dataclasses.dataclass
class HiOMesh:
    points: np.ndarray          # (nnodes, 2) float64
    quads: np.ndarray           # (K, 4) int32, 0-based
    bedges: np.ndarray          # (nb, 2) int32, 0-based
    bmarker: np.ndarray         # (nb,) int32, 1-based marker
    zones: List[Zone]
    cedges: Optional[np.ndarray] = None   # (nc, 2) int32
    cxy: Optional[np.ndarray] = None      # (nc, np1, 2) float64 at Chebyshev-Gauss-Lobatto points

    @property
    def K(self):
        return int(self.quads.shape[0])


@dataclasses.dataclass
class Case:
    name: str
    mesh: HiOMesh
    N: np.ndarray                        # (K,) polynomial order per element
    Mach: float
    Re: float = 0.0
    Pr: float = 0.72
    navier_stokes: bool = False
    riemann: str = "roe"
    viscous: str = "none"
    ip_method: str = "sipg"
    sigma0: float = 1.0
    cfl: float = 0.4
    dt_max: float = 1.0e30
    steps: int = 1
    scheme: str = "rk3"
    ic: Optional[Callable] = None        # ic(x, y) -> (..., 4) dimensionless state
    artificial_dissipation: str = "none"  # ARTIFICIAL_DISSIPATION key
    ad_intensity: float = 1.0            # "Artificial dissipation intensity" (Ceps)
    mode: str = "transient"              # "Integration mode": transient (to t_end) or steady
    residual_target: float = 0.0         # steady: stop when max |residual| < this
    t_end: float = 1.0e30                # transient: "Simulation time"
    t0: float = 0.0                      # "Initial time"
    p_ref: float = 101325.0              # "Reference pressure" (Pa)
    T_ref: float = 273.15                # "Reference Temperature" (K)

    def initial_state(self, coords: np.ndarray) -> np.ndarray:
        """Evaluate the IC at node coordinates ``coords`` (nnodes, 2) -> flat Q (4*nnodes)."""
        q = np.asarray(self.ic(coords[:, 0], coords[:, 1]), dtype=np.float64)
        return np.ascontiguousarray(q.reshape(-1))

    @property
    def ndof(self):
        P = self.N.astype(np.int64) + 1
        return int(4 * np.sum(P * P))


# ---------------------------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------------------------
def _grid_quads(nx: int, ny: int) -> np.ndarray:
    """Quads of a structured (nx x ny) grid, y fastest, nodes (nx+1)*(ny+1) with y fastest."""
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    ix = ix.ravel()
    iy = iy.ravel()
    n0 = ix * (ny + 1) + iy
    return np.stack([n0, n0 + (ny + 1), n0 + (ny + 1) + 1, n0 + 1], axis=1).astype(np.int32)


def periodic_rectangle(nx: int, ny: int, x0: float, x1: float, y0: float, y1: float) -> HiOMesh:
    """Uniform quads on [x0,x1]x[y0,y1] with four periodic zones.

    Markers as in TestCases/Cases/VortexTransport: 1 bottom, 2 top (periodic in y,
    ``Direction = y``), 3 left, 4 right (periodic in x, ``Direction = x``)."""
    xs = x0 + (x1 - x0) * np.arange(nx + 1) / nx
    ys = y0 + (y1 - y0) * np.arange(ny + 1) / ny
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    points = np.stack([X.ravel(), Y.ravel()], axis=1)
    quads = _grid_quads(nx, ny)
    nid = lambda i, j: i * (ny + 1) + j
    bottom = [(nid(i, 0), nid(i + 1, 0)) for i in range(nx)]
    top = [(nid(i + 1, ny), nid(i, ny)) for i in range(nx)]
    left = [(nid(0, j + 1), nid(0, j)) for j in range(ny)]
    right = [(nid(nx, j), nid(nx, j + 1)) for j in range(ny)]
    bedges = np.array(bottom + top + left + right, dtype=np.int32)
    bmarker = np.array([1] * nx + [2] * nx + [3] * ny + [4] * ny, dtype=np.int32)
    zones = [Zone(BC_PERIODIC, direction=2, partner=2), Zone(BC_PERIODIC, direction=2, partner=1),
             Zone(BC_PERIODIC, direction=1, partner=4), Zone(BC_PERIODIC, direction=1, partner=3)]
    return HiOMesh(points, quads, bedges, bmarker, zones)


def channel(nx: int, ny: int, L: float, H: float, zones: List[Zone], skew: float = 0.0) -> HiOMesh:
    """Quads on [0,L]x[0,H] (interior nodes shifted by ``skew`` times a smooth bump, so the
    elements are general quadrilaterals) with four zones: markers 1 bottom, 2 right, 3 top,
    4 left, in that order in ``zones``."""
    xs = L * np.arange(nx + 1) / nx
    ys = H * np.arange(ny + 1) / ny
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    bump = np.sin(np.pi * X / L) * np.sin(np.pi * Y / H)
    X = X + skew * (L / nx) * bump
    Y = Y + skew * (H / ny) * bump * np.cos(np.pi * X / L)
    points = np.stack([X.ravel(), Y.ravel()], axis=1)
    quads = _grid_quads(nx, ny)
    nid = lambda i, j: i * (ny + 1) + j
    bottom = [(nid(i, 0), nid(i + 1, 0)) for i in range(nx)]
    right = [(nid(nx, j), nid(nx, j + 1)) for j in range(ny)]
    top = [(nid(i + 1, ny), nid(i, ny)) for i in range(nx)]
    left = [(nid(0, j + 1), nid(0, j)) for j in range(ny)]
    bedges = np.array(bottom + right + top + left, dtype=np.int32)
    bmarker = np.array([1] * nx + [2] * ny + [3] * nx + [4] * ny, dtype=np.int32)
    return HiOMesh(points, quads, bedges, bmarker, list(zones))


def refined_channel(nx: int, ny: int, L: float, H: float, zones: List[Zone], refine, skew: float = 0.0) -> HiOMesh:
    """The channel of ``channel`` with the cells (i, j) in ``refine`` split into 2 x 2 quads: where
    a split cell meets an unsplit one, the unsplit element's edge meets two elements and their
    common node hangs on it (the subdivided edges of MeshFile.f90:364-490).  The smooth skew map
    moves every node, the hanging ones too, so the reference's builder has to move them back
    onto the big edge (QuadMesh_Construct.incf:168-173).  Markers 1 bottom, 2 right, 3 top,
    4 left."""
    refine = set((int(i), int(j)) for i, j in refine)
    ids = {}

    def node(a, b):          # lattice point (a, b) of the 2nx x 2ny half-cell lattice
        if (a, b) not in ids:
            ids[(a, b)] = len(ids)
        return ids[(a, b)]

    quads = []
    for i in range(nx):
        for j in range(ny):
            a, b = 2 * i, 2 * j
            if (i, j) in refine:
                for di in (0, 1):
                    for dj in (0, 1):
                        quads.append((node(a + di, b + dj), node(a + di + 1, b + dj), node(a + di + 1, b + dj + 1),
                                      node(a + di, b + dj + 1)))
            else:
                quads.append((node(a, b), node(a + 2, b), node(a + 2, b + 2), node(a, b + 2)))

    def side(cells, first, step):
        out = []
        for c, (a, b) in zip(cells, first):
            da, db = step
            if c in refine:
                out += [(node(a, b), node(a + da, b + db)), (node(a + da, b + db), node(a + 2 * da, b + 2 * db))]
            else:
                out.append((node(a, b), node(a + 2 * da, b + 2 * db)))
        return out
    bottom = side([(i, 0) for i in range(nx)], [(2 * i, 0) for i in range(nx)], (1, 0))
    right = side([(nx - 1, j) for j in range(ny)], [(2 * nx, 2 * j) for j in range(ny)], (0, 1))
    top = side([(i, ny - 1) for i in range(nx)], [(2 * i + 2, 2 * ny) for i in range(nx)], (-1, 0))
    left = side([(0, j) for j in range(ny)], [(0, 2 * j + 2) for j in range(ny)], (0, -1))
    lat = np.zeros((len(ids), 2))
    for (a, b), k in ids.items():
        lat[k] = (a, b)
    X = L * lat[:, 0] / (2 * nx)
    Y = H * lat[:, 1] / (2 * ny)
    bump = np.sin(np.pi * X / L) * np.sin(np.pi * Y / H)
    X = X + skew * (L / nx) * bump
    Y = Y + skew * (H / ny) * bump * np.cos(np.pi * X / L)
    bedges = np.array(bottom + right + top + left, dtype=np.int32)
    bmarker = np.array([1] * len(bottom) + [2] * len(right) + [3] * len(top) + [4] * len(left), dtype=np.int32)
    return HiOMesh(np.stack([X, Y], axis=1), np.array(quads, dtype=np.int32), bedges, bmarker, list(zones))


def annulus(nr: int, nt: int, r0: float, r1: float, q_inf, wall_velocity=(0.0, 0.0),
            stretch: bool = True, np1: int = 3) -> HiOMesh:
    """O-grid annulus r in [r0, r1]: every azimuthal edge is a curved (quadratic) arc.

    Marker 1: inner viscous wall (adiabatic); marker 2: outer far field.  The element
    index runs radially fastest.  Curved edges are written like ``x/y_curvilinear_edges``
    of MeshFile.f90:186-232 at ``np1`` Chebyshev-Gauss-Lobatto points."""
    if stretch:
        rs = r0 * (r1 / r0) ** (np.arange(nr + 1) / nr)
    else:
        rs = r0 + (r1 - r0) * np.arange(nr + 1) / nr
    th = 2.0 * np.pi * np.arange(nt) / nt
    # node (it, ir) -> it*(nr+1) + ir
    R, T = np.meshgrid(rs, th, indexing="xy")      # shape (nt, nr+1)
    points = np.stack([(R * np.cos(T)).ravel(), (R * np.sin(T)).ravel()], axis=1)
    nid = lambda it, ir: (it % nt) * (nr + 1) + ir
    quads = []
    for it in range(nt):
        for ir in range(nr):
            # counter-clockwise: (r_i, th_j) -> (r_i+1, th_j) -> (r_i+1, th_j+1) -> (r_i, th_j+1)
            quads.append((nid(it, ir), nid(it, ir + 1), nid(it + 1, ir + 1), nid(it + 1, ir)))
    quads = np.array(quads, dtype=np.int32)
    # curved arcs: between (it, ir) and (it+1, ir) for every ir
    cgl = 0.5 - 0.5 * np.cos(np.pi * np.arange(np1) / (np1 - 1))
    cedges, cxy = [], []
    for ir in range(nr + 1):
        for it in range(nt):
            a0 = th[it]
            cedges.append((nid(it, ir), nid(it + 1, ir)))
            ang = a0 + (2.0 * np.pi / nt) * cgl
            cxy.append(np.stack([rs[ir] * np.cos(ang), rs[ir] * np.sin(ang)], axis=1))
    bedges, bmarker = [], []
    for it in range(nt):
        bedges.append((nid(it, 0), nid(it + 1, 0)))
        bmarker.append(1)
    for it in range(nt):
        bedges.append((nid(it + 1, nr), nid(it, nr)))
        bmarker.append(2)
    zones = [Zone(BC_VISCOUSWALL, isothermal=0, wall_velocity=wall_velocity),
             Zone(BC_FARFIELD, q=tuple(q_inf))]
    return HiOMesh(points, quads, np.array(bedges, dtype=np.int32), np.array(bmarker, dtype=np.int32),
                   zones, np.array(cedges, dtype=np.int32), np.array(cxy, dtype=np.float64))


# ---------------------------------------------------------------------------------------------
# Initial conditions (dimensionless; the UserDefinedInitialCondition of each case)
# ---------------------------------------------------------------------------------------------
def freestream_state(Mach: float, angle: float = 0.0, rho: float = 1.0, p: float = 1.0):
    V = Mach * np.sqrt(GAMMA)        # |u| / a_ref
    u, v = V * np.cos(angle), V * np.sin(angle)
    return np.array([rho, rho * u, rho * v, p / GM1 + 0.5 * rho * (u * u + v * v)])


def _pack(rho, u, v, p):
    return np.stack([rho, rho * u, rho * v, p / GM1 + 0.5 * rho * (u * u + v * v)], axis=-1)


def isentropic_vortex(Mach: float, beta: float = 5.0, xc: float = 0.0, yc: float = 0.0):
    """Shu's isentropic vortex of strength ``beta`` in a freestream of Mach ``Mach`` along +x.

    In the reference's dimensionless variables velocities are scaled by a_ref, so the
    freestream speed is Mach*sqrt(gamma) and the vortex is an exact travelling solution
    moving one length unit per unit of dimensionless time."""
    def ic(x, y):
        dx, dy = x - xc, y - yc
        r2 = dx * dx + dy * dy
        f = np.exp(0.5 * (1.0 - r2))
        u = Mach * np.sqrt(GAMMA) - beta / (2.0 * np.pi) * dy * f
        v = beta / (2.0 * np.pi) * dx * f
        T = 1.0 - GM1 * beta * beta / (8.0 * GAMMA * np.pi * np.pi) * f * f
        rho = T ** (1.0 / GM1)
        return _pack(rho, u, v, rho * T)
    return ic


def gaussian_vortex(Mach: float, radius: float = 0.1, beta: float = 0.05, xc: float = 0.0, yc: float = 0.0):
    """The vortex of the reference's VortexTransport benchmark
    (TestCases/Benchmark/Simulations/VortexTransport/SETUP/ProblemFile.f90:36-76), in the
    dimensionless variables its initial-condition driver stores (rho / rho_ref, rho u / (rho_ref
    a_ref), rho e / p_ref): velocity V (1 - beta dy / R e^(-r2/2), beta dx / R e^(-r2/2)) with
    r2 = (dx^2 + dy^2) / R^2, temperature T_ref (1 - gamma M^2 beta^2 / (2 cp) e^(-r2)) with the
    dimensionless cp = gamma / (gamma - 1), isentropic density (T / T_ref)^(1 / (gamma - 1))."""
    V = Mach * np.sqrt(GAMMA)
    cp = GAMMA / GM1

    def ic(x, y):
        dx, dy = x - xc, y - yc
        r2 = (dx * dx + dy * dy) / (radius * radius)
        u = V * (1.0 - beta * dy / radius * np.exp(-0.5 * r2))
        v = V * (beta * dx / radius * np.exp(-0.5 * r2))
        T = 1.0 - GAMMA * Mach * Mach * beta * beta / (2.0 * cp) * np.exp(-r2)
        rho = T ** (1.0 / GM1)
        return _pack(rho, u, v, rho * T)
    return ic


def taylor_green(Mach: float, seed: int = 0, amp: float = 1.0e-3):
    """2D Taylor-Green vortex on [0, 2pi]^2 with seeded +-amp*V0 uniform velocity noise."""
    V0 = Mach * np.sqrt(GAMMA)

    def ic(x, y):
        rng = np.random.Generator(np.random.PCG64(seed))
        du = rng.uniform(-amp, amp, size=x.shape) * V0
        dv = rng.uniform(-amp, amp, size=x.shape) * V0
        u = V0 * np.sin(x) * np.cos(y) + du
        v = -V0 * np.cos(x) * np.sin(y) + dv
        p = 1.0 + 0.25 * V0 * V0 * (np.cos(2.0 * x) + np.cos(2.0 * y))
        rho = np.ones_like(x) * p
        return _pack(rho, u, v, p)
    return ic


def shear_layer(Mach: float, delta: float = 1.0 / 28.0, eps: float = 1.0e-2, seed: int = 0):
    """Double (Kelvin-Helmholtz) shear layer on the periodic unit square."""
    U = Mach * np.sqrt(GAMMA)

    def ic(x, y):
        rng = np.random.Generator(np.random.PCG64(seed))
        u = np.where(y <= 0.5, U * np.tanh((y - 0.25) / delta), U * np.tanh((0.75 - y) / delta))
        v = U * (eps * np.sin(2.0 * np.pi * x) + 0.1 * eps * rng.uniform(-1.0, 1.0, size=x.shape))
        rho = np.ones_like(x)
        p = np.ones_like(x)
        return _pack(rho, u, v, p)
    return ic


def couette_flow(Mach: float, Pr: float, p_ref: float = 101325.0, T_ref: float = 273.15, perturb: bool = False):
    """The Couette profile of the reference's Couette benchmark
    (TestCases/Benchmark/Simulations/Couette/SETUP/ProblemFile.f90: BoundaryConditionFunction1
    :194-218, UserDefinedInitialCondition :32-61 with ``perturb``: rho u times 1.005 where
    x^2 + y^2 < 0.005), evaluated like the problem file in dimensional variables of air
    (R = 287.15, gamma = 1.4, PhysicsNS.f90:97-98) with the reference values of the
    dimensionless-tau form (PhysicsNS.f90:423-437: rho_ref = p_ref / (R T_ref), a_ref =
    sqrt(p_ref / rho_ref), V = sqrt(gamma) Mach a_ref) and made dimensionless by
    getDimensionlessVariables (VariableConversion.f90:32-43): T = T_ref - V^2 Pr / (8 cp)
    (y + 1)(y - 1), rho = p_ref / (R T), rho u = rho V (y + 1) / 2, rho v = 0,
    rho e = cv rho T + (rho u)^2 / (2 rho)."""
    R, gamma = 287.15, 1.4
    cp, cv = R * 3.5, R * 2.5
    rho_ref = p_ref / (R * T_ref)
    a_ref = np.sqrt(p_ref / rho_ref)
    V = np.sqrt(gamma) * Mach * a_ref

    def ic(x, y):
        T = T_ref - (V * V) * Pr / (8.0 * cp) * (y + 1.0) * (y - 1.0)
        rho = p_ref / (R * T)
        rhou = rho * V * 0.5 * (y + 1.0)
        rhoe = cv * rho * T + 0.5 * rhou * rhou / rho
        if perturb:
            rhou = np.where(x * x + y * y < 0.005, 1.005 * rhou, rhou)
        return np.stack([rho / rho_ref, rhou / (rho_ref * a_ref), np.zeros_like(rho), rhoe / p_ref], axis=-1)
    return ic


def uniform(q):
    q = np.asarray(q, dtype=np.float64)

    def ic(x, y):
        return np.broadcast_to(q, x.shape + (4,)).copy()
    return ic


# ---------------------------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------------------------
def euler_vortex_case(n: int = 16, N: int = 4, steps: int = 200, beta: float = 5.0, Mach: float = 0.5,
                      riemann: str = "roe", xc: float = 0.0, yc: float = 0.0, L: float = 5.0) -> Case:
    """Plain FP64 Euler: isentropic vortex on a periodic [-L,L]^2 square, Roe, RK3, CFL 0.4."""
    mesh = periodic_rectangle(n, n, -L, L, -L, L)
    return Case(f"euler_vortex_{n}x{n}_N{N}", mesh, np.full(mesh.K, N, np.int32), Mach, 0.0,
                riemann=riemann, cfl=0.4 if N <= 4 else 0.1, steps=steps,
                ic=isentropic_vortex(Mach, beta, xc, yc))


def stable_cfl(N: int) -> float:
    """CFL used with the reference's time-step estimate (DGTimeIntegrator.f90:499-500) when a
    configuration does not state one: the reference default 0.1 (Setup.f90:359-360).

    The estimate divides by (N+1) while the spectral radius of the Gauss-node DG operator
    grows like (N+1)^2; measured with the oracle, RK3 diverges at CFL 0.4 for N >= 5 on the
    vortex and at CFL 0.2 for N = 6 on the Taylor-Green vortex, and is stable at 0.1."""
    return 0.4 if N <= 3 else 0.1


def ns_tgv_case(n: int = 128, N: int = 6, steps: int = 500, Re: float = 1600.0, Mach: float = 0.1,
                seed: int = 0, riemann: str = "lf", viscous: str = "br1") -> Case:
    """FP64 Navier-Stokes Taylor-Green vortex, BR1, Lax-Friedrichs, RK3."""
    mesh = periodic_rectangle(n, n, 0.0, 2.0 * np.pi, 0.0, 2.0 * np.pi)
    return Case(f"ns_tgv_{n}x{n}_N{N}", mesh, np.full(mesh.K, N, np.int32), Mach, Re, 0.72, True,
                riemann=riemann, viscous=viscous, cfl=stable_cfl(N), steps=steps, ic=taylor_green(Mach, seed))


def ns_kh_case(n: int = 256, N: int = 7, steps: int = 200, Re: float = 1.0e4, Mach: float = 0.1,
               seed: int = 0) -> Case:
    """Kelvin-Helmholtz double shear layer, interior penalty (SIPG, sigma0 = 1), Roe."""
    mesh = periodic_rectangle(n, n, 0.0, 1.0, 0.0, 1.0)
    return Case(f"ns_kh_{n}x{n}_N{N}", mesh, np.full(mesh.K, N, np.int32), Mach, Re, 0.72, True,
                riemann="roe", viscous="ip", ip_method="sipg", sigma0=1.0, cfl=stable_cfl(N), steps=steps,
                ic=shear_layer(Mach, seed=seed))


def viscous_dt_cap(mesh: HiOMesh, N: int, Re: float, c: float = 0.5) -> float:
    """Upper bound on dt for the explicit viscous terms, c h_min^2 Re / (N+1)^4 with h_min the
    shortest element side.  The reference's estimate (DGTimeIntegrator.f90:462-511) has no
    viscous term; its case-file time step (Setup%dt) caps dt, and this value is used for it.
    On the 512 x 512 cylinder grid (h_min = 3.6e-3) RK3 is stable at 5e-7 and diverges at 2e-6
    from the impulsive start; the cap gives 3.2e-7."""
    p = mesh.points
    q = mesh.quads
    h = np.inf
    for c0 in range(4):
        d = p[q[:, c0]] - p[q[:, (c0 + 1) % 4]]
        h = min(h, float(np.min(np.hypot(d[:, 0], d[:, 1]))))
    return c * h * h * Re / float(N + 1) ** 4


def ns_cylinder_case(nr: int = 512, nt: int = 512, N: int = 7, steps: int = 2000, Re: float = 200.0,
                     Mach: float = 0.2) -> Case:
    """Flow past a cylinder (radius 0.5) on an O-grid to R=20; BR1, Roe, far field + adiabatic wall."""
    qinf = freestream_state(Mach)
    mesh = annulus(nr, nt, 0.5, 20.0, qinf)
    return Case(f"ns_cylinder_{nr}x{nt}_N{N}", mesh, np.full(mesh.K, N, np.int32), Mach, Re, 0.72, True,
                riemann="roe", viscous="br1", cfl=stable_cfl(N), dt_max=viscous_dt_cap(mesh, N, Re), steps=steps,
                ic=uniform(qinf))


def ensemble_cases(count: int = 64, seed_base: int = 1234, nmin: int = 32, nmax: int = 128,
                   Nmin: int = 3, Nmax: int = 8, steps: int = 10) -> List[Case]:
    """Batched ensemble of isentropic-vortex Euler cases with seeded sizes, orders and vortices."""
    out = []
    for s in range(count):
        rng = np.random.Generator(np.random.PCG64(seed_base + s))
        n = int(rng.integers(nmin, nmax + 1))
        N = int(rng.integers(Nmin, Nmax + 1))
        beta = float(rng.uniform(3.0, 6.0))
        xc, yc = (float(v) for v in rng.uniform(-2.0, 2.0, size=2))
        c = euler_vortex_case(n, N, steps, beta=beta, xc=xc, yc=yc)
        c.name = f"ensemble_{s}_{n}x{n}_N{N}"
        # no CFL stated for the ensemble: the reference default (0.4 stays only where N <= 3;
        # at N = 4 on the finer ensemble meshes CFL 0.4 amplifies rounding-level perturbations
        # ~1e4-fold within 20 steps before they decay)
        c.cfl = stable_cfl(N)
        out.append(c)
    return out
