"""Host file formats (horses2d_amd/hio.py): .HiOMesh, .HiOCase and .HiORst, against the
reference's parsing rules (IO/Paramfile.f90:60-396), its zone constructors
(BoundaryConditions/*.incf) and nondimensionalisation (Physics/PhysicsNS.f90:424-437), and the
mesh fixtures extracted from the reference's own test cases."""
import math
import os

import numpy as np
import pytest

from horses2d_amd import cases as C
from horses2d_amd import hio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _write_hiomesh(path, d, pref_zones=None):
    """A NetCDF classic .HiOMesh with the variables ReadMesh reads (MeshFile.f90:99-232)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "w") as nc:
        names = [str(n) for n in d["marker_names"]]
        nc.createDimension("no_of_nodes", len(d["points"]))
        nc.createDimension("no_of_elements", len(d["quads"]))
        nc.createDimension("no_of_bdryedges", len(d["bedges"]))
        nc.createDimension("no_of_markers", len(names))
        nc.createDimension("two", 2)
        nc.createDimension("four", 4)
        nc.createDimension("STR_LEN", 32)
        nc.createVariable("points", "d", ("no_of_nodes", "two"))[:] = d["points"]
        nc.createVariable("points_of_quads", "i", ("no_of_elements", "four"))[:] = d["quads"] + 1
        nc.createVariable("points_of_bdryedges", "i", ("no_of_bdryedges", "two"))[:] = d["bedges"] + 1
        nc.createVariable("bdrymarker_of_edges", "i", ("no_of_bdryedges",))[:] = d["bmarker"]
        for m, name in enumerate(names, 1):
            nc.createVariable(f"marker{m}", "c", ("STR_LEN",))[:] = np.array(list(name.ljust(32)), dtype="S1")
        if pref_zones is not None:
            nc.createDimension("no_of_pRefinementzones", int(pref_zones.max()))
            nc.createVariable("pRefinement_zones", "i", ("no_of_elements",))[:] = pref_zones
        if "cedges" in d:
            nc.createDimension("no_of_curvilinearedges", len(d["cedges"]))
            nc.createDimension("np1", d["cxy"].shape[1])
            nc.createVariable("curvilinear_edges", "i", ("no_of_curvilinearedges", "two"))[:] = d["cedges"] + 1
            nc.createVariable("x_curvilinear_edges", "d", ("no_of_curvilinearedges", "np1"))[:] = d["cxy"][..., 0]
            nc.createVariable("y_curvilinear_edges", "d", ("no_of_curvilinearedges", "np1"))[:] = d["cxy"][..., 1]


@pytest.mark.parametrize("name", ["vortex_transport_16x16", "cylinder_very_coarse"])
def test_hiomesh_round_trip_of_the_reference_meshes(tmp_path, name):
    d = dict(np.load(os.path.join(GOLDEN, f"mesh_{name}.npz")))
    path = str(tmp_path / "m.HiOMesh")
    _write_hiomesh(path, d)
    m = hio.read_hiomesh(path, reynolds_length=0.5)
    np.testing.assert_array_equal(m["points"], d["points"] / 0.5)
    for k in ("quads", "bedges", "bmarker"):
        np.testing.assert_array_equal(m[k], d[k])
    assert m["marker_names"] == [str(n) for n in d["marker_names"]]
    if "cedges" in d:
        np.testing.assert_array_equal(m["cedges"], d["cedges"])
        np.testing.assert_array_equal(m["cxy"], d["cxy"] / 0.5)


def test_read_value_follows_the_reference_rules(tmp_path):
    """readValue: comment '!' removed first; the line must contain the label as written, and
    the text before the first '=' or ':' (the later of the two) must equal it once spaces are
    removed; the first such line wins; regions are delimited by lines equal to the in/out
    labels without spaces."""
    f = tmp_path / "c.HiOCase"
    f.write_text("\n".join([
        "! Mach number : 9.0",                      # a comment
        "   Reference Mach number : 3.0",           # another label that contains the label
        "   Mach   number   :  0.75",               # the label is not a substring: skipped
        "   Mach number   :  0.25   ! trailing",    # spaces before ':' squashed, comment cut
        "   Mach number : 0.5",                     # later duplicate: ignored
        "Time step = 1.0d-3",
        "#  define   zone 2",
        "   Type = Farfield",
        "   Mach = 0.4",
        "# end",
        "   Type = Euler wall",                     # outside any region
    ]) + "\n")
    assert hio.read_value(str(f), "Mach number") == "0.25"
    assert hio._real(hio.read_value(str(f), "Time step")) == pytest.approx(1e-3)
    assert hio.read_value(str(f), "Missing") is None
    assert hio.read_value_in_region(str(f), "Type", "# define zone 2") == "Farfield"
    assert hio.read_value_in_region(str(f), "Mach", "# define zone 2") == "0.4"
    assert hio.read_value_in_region(str(f), "Type", "# define zone 3") is None


CASE = """!-------------------------------------: Geometry
                             Mesh file: ./MESH/box.HiOMesh
                         Boundary file: _this_
!-------------------------------------: Physics
                Reference pressure    : 101325.0
                Reference Temperature : 300.0
                Reynolds length       : 2.0
                Reynolds number       : 500.0
                Prandtl number        : 0.7
                Mach number           : 0.2
         Default polynomial order    : 5
                   Initial condition : Uniform
             Inviscid Riemann solver : HLL
               Viscous discretization: IP
              Interior penalty method: NIPG
              Jumps penalty parameter: 2.5
               Integration scheme   : Williamson RK5
               CFL Number           : 0.3
               Time step            : 1.0e-3
               Number of iterations : 77

# define p-Refinement
   1: 7
# end

# define zone 1
   Name = Wall
   Type = Viscous wall
   Wall type = Isothermal
   Wall temperature = 450.0
   Mach number = 0.1
   Velocity direction = [0.0, -1.0]
# end

# define zone 2
   Type = Pressure outlet
   pressure = 202650.0
   Outflow type = Reflective
# end

# define zone 3
   Type = Euler wall
   Riemann solver = Exact
# end

# define zone 4
   Type = Farfield
   pressure = 202650.0
   Temperature = 450.0
   Mach = 0.5
   Angle of attack = 30.0
# end
"""


def _box_files(tmp_path, pref=False):
    d = dict(np.load(os.path.join(GOLDEN, "mesh_vortex_transport_16x16.npz")))
    (tmp_path / "MESH").mkdir()
    (tmp_path / "SETUP").mkdir()
    K = len(d["quads"])
    pz = (np.arange(K) % 2).astype(np.int32) if pref else None
    _write_hiomesh(str(tmp_path / "MESH" / "box.HiOMesh"), d, pz)
    case = tmp_path / "SETUP" / "box.HiOCase"
    case.write_text(CASE)
    return d, str(case)


def test_case_from_hiocase_keys_zones_and_nondimensional_states(tmp_path):
    d, path = _box_files(tmp_path)
    c = hio.case_from_hiocase(path)
    assert (c.Mach, c.Re, c.Pr) == (0.2, 500.0, 0.7)
    assert c.navier_stokes and c.viscous == "ip" and c.ip_method == "nipg" and c.sigma0 == 2.5
    assert (c.riemann, c.scheme, c.cfl, c.dt_max, c.steps) == ("hll", "rk5", 0.3, 1e-3, 77)
    assert np.all(c.N == 5)
    np.testing.assert_array_equal(c.mesh.points, d["points"] / 2.0)       # Reynolds length
    z1, z2, z3, z4 = c.mesh.zones
    # viscous wall: Tw / T_ref, wall speed sqrt(gamma) M along the given direction
    assert (z1.kind, z1.isothermal) == (C.BC_VISCOUSWALL, 1)
    assert z1.Tw == pytest.approx(1.5)
    np.testing.assert_allclose(z1.wall_velocity, [0.0, -0.1 * math.sqrt(1.4)], atol=1e-15)
    # pressure outlet at twice p_ref, T_ref, the case Mach: rho = 2, |u| = M sqrt(gamma)
    assert (z2.kind, z2.outflow) == (C.BC_PRESSUREOUTLET, 1)
    np.testing.assert_allclose(z2.q, C.freestream_state(0.2, 0.0, rho=2.0, p=2.0), rtol=1e-15)
    assert (z3.kind, z3.riemann) == (C.BC_EULERWALL, C.RIEMANN["exact"])
    # far field at 2 p_ref and 1.5 T_ref: rho = 2 / 1.5, |u| = M sqrt(gamma 1.5), 30 degrees
    rho, V, a = 2.0 / 1.5, 0.5 * math.sqrt(1.4 * 1.5), math.pi / 6
    np.testing.assert_allclose(z4.q, [rho, rho * V * math.cos(a), rho * V * math.sin(a),
                                      2.0 / 0.4 + 0.5 * rho * V * V], rtol=1e-14)
    X = np.zeros((len(c.mesh.quads) * 36, 2))
    np.testing.assert_allclose(c.initial_state(X).reshape(-1, 4)[0], C.freestream_state(0.2))


def test_newdirichlet_zone_keys(tmp_path):
    """newDirichletBC_Construct (newDirichlet.incf:22-125): the zone state as the far field
    forms it; Mode and Outflow pressure take their exact strings, anything else the defaults
    Specify speed and Static."""
    _, path = _box_files(tmp_path)
    with open(path) as f:
        text = f.read()
    far = "   Type = Farfield\n"
    for extra, mode, outp in (("", 1, 1),
                              ("   Mode = Specify total pressure\n   Outflow pressure = Total\n", 2, 2),
                              ("   Mode = Specify Speed\n   Outflow pressure = Static\n", 1, 1),
                              ("   Outflow pressure = Total\n", 1, 2)):
        with open(path, "w") as f:
            f.write(text.replace(far, "   Type = newDirichlet\n" + extra))
        z4 = hio.case_from_hiocase(path).mesh.zones[3]
        assert (z4.kind, z4.mode, z4.outflow_pressure) == (C.BC_NEWDIRICHLET, mode, outp)
        assert z4.ints() == (mode, outp)
        rho, V, a = 2.0 / 1.5, 0.5 * math.sqrt(1.4 * 1.5), math.pi / 6
        np.testing.assert_allclose(z4.q, [rho, rho * V * math.cos(a), rho * V * math.sin(a),
                                          2.0 / 0.4 + 0.5 * rho * V * V], rtol=1e-14)


def test_artificial_dissipation_keys(tmp_path):
    _, path = _box_files(tmp_path)
    with open(path) as f:
        text = f.read()
    assert hio.case_from_hiocase(path).artificial_dissipation == "none"
    extra = "\n   Artificial dissipation (0/1) : 1\n   Artificial dissipation intensity : 0.3\n"
    with open(path, "w") as f:
        f.write(text.replace("Number of iterations : 77", "Number of iterations : 77" + extra))
    c = hio.case_from_hiocase(path)
    assert (c.artificial_dissipation, c.ad_intensity) == ("physical", 0.3)
    lap = "   Artificial dissipation type : Laplacian\n"
    with open(path, "w") as f:
        f.write(text.replace("Number of iterations : 77", "Number of iterations : 77" + extra + lap))
    assert hio.case_from_hiocase(path).artificial_dissipation == "laplacian"
    # the dissipation module is Navier-Stokes only (DGArtificialDissipation.f90:22): the Euler
    # binary runs without it
    assert hio.case_from_hiocase(path, navier_stokes=False).artificial_dissipation == "none"
    bad = "   Artificial dissipation indicator : Residuals-based\n"
    with open(path, "w") as f:
        f.write(text.replace("Number of iterations : 77", "Number of iterations : 77" + extra + bad))
    with pytest.raises(ValueError):
        hio.case_from_hiocase(path)


def test_p_refinement_zones_map_to_orders(tmp_path):
    _, path = _box_files(tmp_path, pref=True)
    c = hio.case_from_hiocase(path)
    K = len(c.N)
    np.testing.assert_array_equal(c.N, np.where(np.arange(K) % 2 == 1, 7, 5))


def test_unsupported_zones_are_refused(tmp_path):
    _, path = _box_files(tmp_path)
    with open(path) as f:
        text = f.read()
    for bad in ("Type = Riemann", "Type = Wall"):
        with open(path, "w") as f:
            f.write(text.replace("Type = Euler wall", bad))
        with pytest.raises(ValueError):
            hio.case_from_hiocase(path)
    with open(path, "w") as f:
        f.write(text.replace("   Angle of attack = 30.0", "   Use function = MyState"))
    with pytest.raises(ValueError):
        hio.case_from_hiocase(path)


def test_use_function_zones_record_the_function(tmp_path):
    """"Use function = BC1" (AssociateBoundaryConditionFunction, BoundaryConditionFunctionDrivers.
    incf:22-63): the zone keeps its default state and names the function its caller evaluates."""
    _, path = _box_files(tmp_path)
    with open(path) as f:
        text = f.read()
    with open(path, "w") as f:
        f.write(text.replace("   Angle of attack = 30.0", "   Use function = BC1"))
    c = hio.case_from_hiocase(path)
    assert [z.function for z in c.mesh.zones if z.function] == ["BC1"]


def test_hiorst_round_trip_and_restart_initial_condition(tmp_path):
    from oracle.oracle import Oracle
    d, path = _box_files(tmp_path)
    case = C.euler_vortex_case(4, 3)
    o = Oracle(case)
    o.set_state(case.initial_state(o.coords()))
    dt = o.estimate_dt(case.dt_max, case.cfl)
    for it in range(3):
        o.step(it * dt, dt, case.scheme)
    Q = o.state()
    rst = str(tmp_path / "RESULTS.HiORst")
    hio.write_hiorst(rst, Q, 3 * dt, 3, 3, case.mesh.K)
    Q2, t2, it2 = hio.read_hiorst(rst)
    np.testing.assert_array_equal(Q2, Q)
    assert (t2, it2) == (3 * dt, 3)
    # a case whose initial condition is the restart file
    with open(path) as f:
        text = f.read()
    (tmp_path / "RESULTS").mkdir()
    hio.write_hiorst(str(tmp_path / "RESULTS" / "box.HiORst"), np.arange(4.0 * 256 * 36), 0.5, 10, 5, 256)
    with open(path, "w") as f:
        f.write(text.replace("Initial condition : Uniform", "Initial condition : Restart\n"
                             "   Restart file : ./RESULTS/box.HiORst"))
    c = hio.case_from_hiocase(path)
    assert c.restart == (0.5, 10)
    np.testing.assert_array_equal(c.initial_state(np.zeros((256 * 36, 2))), np.arange(4.0 * 256 * 36))


@pytest.mark.gpu
def test_case_read_from_files_runs_on_the_device_like_the_oracle(tmp_path):
    """The case above (isothermal moving wall, reflective pressure outlet, Euler wall with the
    exact solver, far field; NIPG, HLL, RK5) from its files, through the C ABI: QDot and three
    steps within 1e-12 of the oracle."""
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    _, path = _box_files(tmp_path)
    with open(path) as f:       # the zone states of the parsing test, made mild enough to step
        text = f.read()
    for a, b in (("pressure = 202650.0", "pressure = 106000.0"), ("Temperature = 450.0", "Temperature = 310.0"),
                 ("Wall temperature = 450.0", "Wall temperature = 320.0"),
                 ("Reynolds length       : 2.0", "Reynolds length       : 0.1"),   # the unit square
                 ("CFL Number           : 0.3", "CFL Number           : 0.1")):     # stable at N = 5
        text = text.replace(a, b)
    with open(path, "w") as f:
        f.write(text)
    case = hio.case_from_hiocase(path)
    dg = gpu.DGSEM(case, "fp64")
    orc = Oracle(case)
    Q0 = case.initial_state(dg.coords())
    dg.upload(Q0)
    orc.set_state(Q0)
    qo = orc.time_derivative(0.0)
    assert np.max(np.abs(dg.time_derivative(0.0) - qo)) <= 1e-12 * np.max(np.abs(qo))
    dt = orc.estimate_dt(case.dt_max, case.cfl)
    for it in range(3):
        orc.step(it * dt, dt, case.scheme)
        dg.rk_step(it * dt, dt)
    s = orc.state()
    assert np.all(np.isfinite(s))
    assert np.max(np.abs(dg.state() - s)) <= 1e-12 * np.max(np.abs(s))
