"""HORSES2D file formats on the host: .HiOMesh meshes, .HiOCase parameters and zone blocks, and
.HiORst restart files, so that a reference case runs on this library from its own files.

* ``read_hiomesh`` -- MeshFile_t after ReadMesh (Mesh/MeshFile.f90:77-279): the NetCDF classic
  variables points, points_of_quads, points_of_bdryedges, bdrymarker_of_edges, marker names and
  the curvilinear edges, ids made 0-based, coordinates divided by the Reynolds length.
* ``read_value`` / ``read_value_in_region`` -- readValue / readValueInRegion
  (IO/Paramfile.f90:60-357): the first line (or the first line inside the region) whose text
  before its first '=' or ':' equals the label once spaces are removed; '!' starts a comment.
* ``case_from_hiocase`` -- Setup_Initialization (Generic/Setup.f90:175-430) for the keys this
  library uses, the zone blocks as BoundaryConditions_construct and each condition's
  _Construct form them (BoundaryConditions/*.incf), and the "p-Refinement" block.
* ``read_hiorst`` / ``write_hiorst`` -- DGSEM_loadRestartFile (Approximation/DGSEM_class.f90:
  264-285) and TimeIntegrator_Autosave (DGTimeIntegrator.f90:413-454): t, iter and the flat Q.

The initial condition of a case is not in these files (it is problem-file code or a restart):
``case_from_hiocase`` leaves ``Case.ic`` to the caller unless the case asks for "Uniform" or
"Restart".
"""
from __future__ import annotations

import math
import os
import re
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import cases as C

COMMENT = "!"


# ---------------------------------------------------------------------------------------------
# .HiOMesh
# ---------------------------------------------------------------------------------------------
def read_hiomesh(path: str, reynolds_length: float = 1.0) -> Dict[str, np.ndarray]:
    """Arrays of a .HiOMesh file: points [n,2], quads [K,4], bedges [nb,2], bmarker [nb]
    (1-based markers), marker_names, cedges [nc,2] / cxy [nc,np1,2] when the mesh has
    curvilinear edges, and pref_zones [K] (p-refinement zone per element, 0 = default order)
    when it has p-refinement zones.  Ids are 0-based; coordinates are divided by ``reynolds_length``."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables
        out = {
            "points": np.array(v["points"][:], dtype=np.float64) / reynolds_length,
            "quads": np.array(v["points_of_quads"][:], dtype=np.int32) - 1,
            "bedges": np.array(v["points_of_bdryedges"][:], dtype=np.int32) - 1,
            "bmarker": np.array(v["bdrymarker_of_edges"][:], dtype=np.int32),
        }
        nm = int(nc.dimensions["no_of_markers"])
        out["marker_names"] = [v[f"marker{m}"][:].tobytes().decode(errors="replace").strip(" \x00")
                               for m in range(1, nm + 1) if f"marker{m}" in v]
        if "pRefinement_zones" in v:
            out["pref_zones"] = np.array(v["pRefinement_zones"][:], dtype=np.int32)
        if "curvilinear_edges" in v:
            out["cedges"] = np.array(v["curvilinear_edges"][:], dtype=np.int32) - 1
            x = np.array(v["x_curvilinear_edges"][:], dtype=np.float64)
            y = np.array(v["y_curvilinear_edges"][:], dtype=np.float64)
            out["cxy"] = np.stack([x, y], axis=-1) / reynolds_length
    return out


# ---------------------------------------------------------------------------------------------
# .HiOCase
# ---------------------------------------------------------------------------------------------
def _squash(s: str) -> str:
    """getSquashedLine (Paramfile.f90:359-396): the text before '!' without spaces."""
    p = s.find(COMMENT)
    if p >= 0:
        s = s[:p]
    return s.strip().replace(" ", "")


def _value_of(line: str, label: str) -> Optional[str]:
    p = line.find(COMMENT)
    if p >= 0:
        line = line[:p]
    line = line.rstrip()
    if label not in line:
        return None
    eq = [i for i in (line.find("="), line.find(":")) if i >= 0]
    if not eq:
        return None
    pos = max(eq)
    if _squash(line[:pos]) != _squash(label):
        return None
    return line[pos + 1:].strip()


def read_value(path: str, label: str) -> Optional[str]:
    """readValue: the value of ``label`` anywhere in the file, or None."""
    with open(path) as f:
        for line in f:
            v = _value_of(line.rstrip("\n"), label)
            if v is not None:
                return v
    return None


def read_value_in_region(path: str, label: str, in_label: str, out_label: str = "# end") -> Optional[str]:
    """readValueInRegion: the value of ``label`` between the lines ``in_label`` and
    ``out_label`` (compared without spaces), or None."""
    inside = False
    a, b = _squash(in_label), _squash(out_label)
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            sq = _squash(line)
            if sq == a:
                inside = True
                continue
            if sq == b:
                inside = False
                continue
            if inside:
                v = _value_of(line, label)
                if v is not None:
                    return v
    return None


def _real(s: Optional[str]) -> Optional[float]:
    if s is None or not s.strip():
        return None
    return float(s.split()[0].replace("d", "e").replace("D", "e"))


def _int(s: Optional[str]) -> Optional[int]:
    if s is None or not s.strip():
        return None
    return int(s.split()[0])


def _array(s: str) -> List[float]:
    """getArrayFromString (Generic/Utilities.f90:173-240): the numbers between '[' and ']'."""
    a, b = s.find("["), s.find("]")
    body = s[a + 1:b] if a >= 0 and b > a else s
    return [float(t.replace("d", "e")) for t in re.split(r"[,\s]+", body.strip()) if t]


RIEMANN_NAMES = {"Roe": "roe", "HLL": "hll", "Exact": "exact", "Lax-Friedrichs": "lf"}
SCHEME_NAMES = {"Explicit-Euler": "euler", "Williamson RK3": "rk3", "Williamson RK5": "rk5"}
ZONE_TYPES = {"Periodic": C.BC_PERIODIC, "Dirichlet": C.BC_DIRICHLET, "Euler wall": C.BC_EULERWALL,
              "Viscous wall": C.BC_VISCOUSWALL, "Farfield": C.BC_FARFIELD,
              "Pressure outlet": C.BC_PRESSUREOUTLET, "Pressure inlet": C.BC_PRESSUREINLET,
              "newDirichlet": C.BC_NEWDIRICHLET}


def zone_state(path: str, in_label: str, Mach: float, T_ref: float, p_ref: float) -> Tuple[float, ...]:
    """The dimensionless zone state the far field, Dirichlet and pressure conditions form in
    their _Construct (e.g. Farfield.incf:31-101): pressure, temperature, Mach number and angle
    of attack (defaults: the reference pressure and temperature, the case's Mach number, 0),
    V = Mach sqrt(gamma R T), rho = p / (R T), then getDimensionlessVariables
    (Physics/VariableConversion.f90:32-43) with rho_ref = p_ref / (R T_ref), a_ref = sqrt(R T_ref)
    and p_ref (PhysicsNS.f90:424-437): only the ratios p / p_ref and T / T_ref remain."""
    p = _real(read_value_in_region(path, "pressure", in_label))
    T = _real(read_value_in_region(path, "Temperature", in_label))
    M = _real(read_value_in_region(path, "Mach", in_label))
    aoa = _real(read_value_in_region(path, "Angle of attack", in_label))
    p = p_ref if p is None else p
    T = T_ref if T is None else T
    M = Mach if M is None else M
    aoa = 0.0 if aoa is None else aoa * math.pi / 180.0
    rho = (p / p_ref) * (T_ref / T)
    V = M * math.sqrt(C.GAMMA * T / T_ref)
    pd = p / p_ref
    return (rho, rho * V * math.cos(aoa), rho * V * math.sin(aoa), pd / C.GM1 + 0.5 * rho * V * V)


def read_zone(path: str, marker: int, Mach: float, T_ref: float, p_ref: float) -> C.Zone:
    """One ``# define zone <marker>`` block (BoundaryConditions_construct,
    DGBoundaryConditions.f90:244-357, and the _Construct of its type)."""
    lab = f"# define zone {marker}"
    t = read_value_in_region(path, "Type", lab)
    if t is None:
        raise ValueError(f"zone {marker}: no Type")
    if t not in ZONE_TYPES:
        raise ValueError(f"zone {marker}: boundary condition '{t}' is not supported")
    kind = ZONE_TYPES[t]
    rs = read_value_in_region(path, "Riemann solver", lab)
    riemann = C.RIEMANN[RIEMANN_NAMES[rs]] if rs in RIEMANN_NAMES else -1
    z = C.Zone(kind, riemann=riemann)
    if kind == C.BC_PERIODIC:
        # PeriodicBC_Construct (Periodic.incf:30-60)
        d = read_value_in_region(path, "Direction", lab)
        if d not in ("x", "y"):
            raise ValueError(f"zone {marker}: Direction must be 'x' or 'y'")
        z.direction = 1 if d == "x" else 2
        z.partner = _int(read_value_in_region(path, "Marker", lab))
        if z.partner is None:
            raise ValueError(f"zone {marker}: periodic zone without Marker")
    elif kind == C.BC_VISCOUSWALL:
        # ViscousWall_Construct (ViscousWall.incf:29-96)
        wt = read_value_in_region(path, "Wall type", lab)
        if wt not in ("Isothermal", "Adiabatic"):
            raise ValueError(f"zone {marker}: Wall type must be Isothermal or Adiabatic")
        z.isothermal = 1 if wt == "Isothermal" else 0
        Tw = _real(read_value_in_region(path, "Wall temperature", lab))
        z.Tw = 1.0 if Tw is None else Tw / T_ref
        Mw = _real(read_value_in_region(path, "Mach number", lab))
        v = 0.0 if Mw is None else math.sqrt(C.GAMMA) * Mw
        vd = read_value_in_region(path, "Velocity direction", lab)
        d = [1.0, 0.0] if not vd else _array(vd)[:2]
        z.wall_velocity = (v * d[0], v * d[1])
    elif kind in (C.BC_FARFIELD, C.BC_DIRICHLET, C.BC_PRESSUREOUTLET, C.BC_PRESSUREINLET, C.BC_NEWDIRICHLET):
        z.q = zone_state(path, lab, Mach, T_ref, p_ref)
        # "Use function" (AssociateBoundaryConditionFunction, BoundaryConditionFunctionDrivers.incf:
        # 22-63): the state comes node by node from a problem-file function, which the caller
        # evaluates at the zone's boundary nodes (DGSEM.set_boundary_state); q stays the default
        fn = read_value_in_region(path, "Use function", lab)
        if fn:
            if fn not in ("Initial condition", "BC1", "BC2", "BC3", "BC4"):
                raise ValueError(f"zone {marker}: unknown boundary-state function '{fn}'")
            z.function = fn
        if kind == C.BC_PRESSUREOUTLET:
            # PressureOutletBC_Construct (PressureOutlet.incf:66-72): default partially non reflective
            of = read_value_in_region(path, "Outflow type", lab)
            z.outflow = C.OUTFLOW["reflective"] if of == "Reflective" else C.OUTFLOW["partially non reflective"]
        if kind == C.BC_NEWDIRICHLET:
            # newDirichletBC_Construct (newDirichlet.incf:79-105): defaults Specify speed, Static
            md = read_value_in_region(path, "Mode", lab)
            z.mode = C.NEWDIRICHLET_MODE["specify total pressure" if md == "Specify total pressure" else "specify speed"]
            op = read_value_in_region(path, "Outflow pressure", lab)
            z.outflow_pressure = C.NEWDIRICHLET_OUTFLOW["total" if op == "Total" else "static"]
    return z


def read_p_refinement(path: str, pref_zones: Optional[np.ndarray], K: int, N: int) -> np.ndarray:
    """Per-element orders (MeshFile.f90:142-168): the mesh's p-refinement zone of each element
    (0: the default order N) mapped through the case file's ``# define p-Refinement`` block,
    whose ``zone: order`` lines are read like readValueInRegion (a zone without a line keeps N)."""
    if pref_zones is None:
        return np.full(K, N, np.int32)
    nz = int(pref_zones.max()) if len(pref_zones) else 0
    order = [N]
    for z in range(1, nz + 1):
        o = _int(read_value_in_region(path, str(z), "# define p-Refinement"))
        order.append(N if o is None else o)
    return np.asarray(order, dtype=np.int32)[pref_zones]


def case_from_hiocase(case_path: str, mesh_path: Optional[str] = None, ic=None, steps: Optional[int] = None,
                      navier_stokes: Optional[bool] = None):
    """A ``cases.Case`` from a .HiOCase file (and its mesh, by default the file's "Mesh file"
    relative to the case file's directory, then its parent).  Keys: Reference pressure /
    Temperature, Reynolds length / number, Prandtl number, Mach number, Default polynomial
    order, Inviscid Riemann solver (default Roe), Viscous discretization (default BR1),
    Interior penalty method, Jumps penalty parameter, Integration scheme (default Williamson
    RK3), CFL Number (default 0.1), Time step (default 0.01, the dt cap), Number of iterations;
    zone blocks as ``read_zone``; "Boundary file: _this_" (or empty) means the case file."""
    rv = lambda k: read_value(case_path, k)
    base = os.path.dirname(os.path.abspath(case_path))
    if mesh_path is None:
        mf = rv("Mesh file")
        if not mf:
            raise ValueError("no Mesh file in the case file")
        for root in (base, os.path.dirname(base)):
            cand = os.path.normpath(os.path.join(root, mf))
            if os.path.exists(cand):
                mesh_path = cand
                break
        if mesh_path is None:
            raise FileNotFoundError(mf)
    bf = rv("Boundary file")
    bdry = case_path if (not bf or bf == "_this_") else os.path.normpath(os.path.join(base, bf))
    L = _real(rv("Reynolds length")) or 1.0
    p_ref = _real(rv("Reference pressure")) or 101325.0
    T_ref = _real(rv("Reference Temperature")) or 273.15
    Re = _real(rv("Reynolds number")) or 0.0
    Pr = _real(rv("Prandtl number")) or 0.72
    Mach = _real(rv("Mach number"))
    if Mach is None:
        raise ValueError("no Mach number in the case file")
    N = _int(rv("Default polynomial order"))
    if N is None:
        raise ValueError("no Default polynomial order in the case file")
    rs = rv("Inviscid Riemann solver") or "Roe"
    if rs not in RIEMANN_NAMES:
        raise ValueError(f"Inviscid Riemann solver '{rs}' is not supported")
    visc = rv("Viscous discretization") or "BR1"
    visc = "IP" if visc == "Interior-penalty" else visc
    ipm = (rv("Interior penalty method") or "SIPG").lower()
    sigma0 = _real(rv("Jumps penalty parameter"))
    scheme = rv("Integration scheme") or "Williamson RK3"
    if scheme not in SCHEME_NAMES:
        raise ValueError(f"Integration scheme '{scheme}' is not supported")
    # artificial dissipation (Setup.f90:333-349; DGArtificialDissipation.f90:100-137): type
    # "Physical" (default) or "Laplacian" with the "Jumps-based" indicator (default); another
    # type name is the base class, which adds nothing; "Residuals-based" leaves the viscosity
    # function unassociated in the reference and is not provided
    ad_on = _int(rv("Artificial dissipation (0/1)")) == 1
    ad_type = rv("Artificial dissipation type") or "Physical"
    ad_ind = rv("Artificial dissipation indicator") or "Jumps-based"
    ad_ceps = _real(rv("Artificial dissipation intensity"))
    ad_kind = "none"
    if ad_on and ad_type in ("Physical", "Laplacian"):
        if ad_ind != "Jumps-based":
            raise ValueError(f"artificial dissipation indicator '{ad_ind}' is not supported")
        ad_kind = ad_type.lower()
    cfl = _real(rv("CFL Number"))
    dt = _real(rv("Time step"))
    iters = _int(rv("Number of iterations"))
    # TimeIntegrator_NewTimeIntegrator (DGTimeIntegrator.f90:86-157): "Steady" stops at the
    # residual target, "Transient" at the simulation time
    mode = (rv("Integration mode") or "Transient").strip()
    if mode not in ("Steady", "Transient"):
        raise ValueError(f"Integration mode '{mode}' is not supported")
    target = _real(rv("Residual target"))
    t_end = _real(rv("Simulation time"))
    t0 = _real(rv("Initial time"))

    m = read_hiomesh(mesh_path, L)
    nmark = int(m["bmarker"].max()) if len(m["bmarker"]) else 0
    zones = [read_zone(bdry, z, Mach, T_ref, p_ref) for z in range(1, nmark + 1)]
    mesh = C.HiOMesh(m["points"], m["quads"], m["bedges"], m["bmarker"], zones,
                     m.get("cedges"), m.get("cxy"))
    orders = read_p_refinement(case_path, m.get("pref_zones"), len(m["quads"]), N)
    # the binary decides the equations (HORSES2D.NS / HORSES2D.Euler, Solver/Makefile:163-186);
    # by default Navier-Stokes when the case has a viscosity or artificial dissipation
    ns = (Re > 0.0 or ad_kind != "none") if navier_stokes is None else bool(navier_stokes)
    case = C.Case(os.path.splitext(os.path.basename(case_path))[0], mesh, orders, Mach, Re, Pr, ns,
                  riemann=RIEMANN_NAMES[rs], viscous=(visc.lower() if ns else "none"),
                  ip_method=ipm if ipm in C.IP_METHOD else "sipg",
                  sigma0=1.0 if sigma0 is None else sigma0, cfl=0.1 if cfl is None else cfl,
                  dt_max=0.01 if dt is None else dt, steps=steps or iters or 1,
                  scheme=SCHEME_NAMES[scheme], ic=ic,
                  artificial_dissipation=ad_kind if ns else "none",
                  ad_intensity=1.0 if ad_ceps is None else ad_ceps, mode=mode.lower(),
                  residual_target=0.0 if (mode == "Transient" or target is None) else target,
                  t_end=1.0e30 if (mode == "Steady" or t_end is None) else t_end,
                  t0=0.0 if t0 is None else t0)
    case.p_ref, case.T_ref = p_ref, T_ref
    ic_name = rv("Initial condition") or ""
    if case.ic is None and ic_name == "Uniform":
        case.ic = C.uniform(C.freestream_state(Mach))
    elif case.ic is None and ic_name == "Restart":
        # DGSEM_SetInitialCondition -> loadRestartFile (DGSEM_class.f90:186-200, 264-285)
        rf = rv("Restart file")
        if not rf:
            raise ValueError("Initial condition Restart without a Restart file")
        for root in (base, os.path.dirname(base)):
            cand = os.path.normpath(os.path.join(root, rf))
            if os.path.exists(cand):
                Q, t0, it0 = read_hiorst(cand)
                case.restart = (t0, it0)
                case.ic = lambda x, y, Q=Q: Q.reshape(-1, 4)
                break
        else:
            raise FileNotFoundError(rf)
    return case


# ---------------------------------------------------------------------------------------------
# .HiORst
# ---------------------------------------------------------------------------------------------
def write_hiorst(path: str, Q: np.ndarray, t: float, iteration: int, N: int, no_of_elements: int) -> None:
    """TimeIntegrator_Autosave (DGTimeIntegrator.f90:413-454): NetCDF classic with dimensions
    one, NDOF, N, NCONS, no_of_elements and variables t(one), Q(NDOF), iter(one)."""
    from scipy.io import netcdf_file
    Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(-1)
    with netcdf_file(path, "w") as nc:
        nc.createDimension("one", 1)
        nc.createDimension("NDOF", Q.size)
        nc.createDimension("N", int(N))
        nc.createDimension("NCONS", 4)
        nc.createDimension("no_of_elements", int(no_of_elements))
        nc.createVariable("t", "d", ("one",))[:] = [float(t)]
        nc.createVariable("Q", "d", ("NDOF",))[:] = Q
        nc.createVariable("iter", "i", ("one",))[:] = [int(iteration)]


def read_hiorst(path: str) -> Tuple[np.ndarray, float, int]:
    """DGSEM_loadRestartFile (DGSEM_class.f90:264-285): (Q, t, iter)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables
        return (np.array(v["Q"][:], dtype=np.float64), float(v["t"][0]), int(v["iter"][0]))
