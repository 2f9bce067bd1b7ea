"""End-to-end pins against the reference's own benchmark goldens: the numbers each
TestCases/Benchmark/Simulations/*/SETUP/ProblemFile.f90 ``Finalize`` checks after a full run.

Ramp (Ramp/SETUP/ProblemFile.f90:79-89), reproduced by the oracle and the GPU path: supersonic flow (Mach 2) over a ramp, 180 straight
quads of order 3, Euler walls and newDirichlet inflow/outflow, the Navier-Stokes equations at
Re = 0 with interior penalty (IIPG), the Laplacian artificial dissipation with the jumps-based
indicator (intensity 0.1), Williamson RK5, CFL 0.1, dt cap 2, steady mode with residual target
1e-8.  Finalize requires the run to stop at iteration 9884 with the four residuals (max |QDot|
of the last stage) within 1e-11 of its stored values.

CylinderRe100 (CylinderRe100/SETUP/ProblemFile.f90:79-99), reproduced by the oracle and the GPU
path: Mach 0.1, Re 100 past a cylinder, 640 quads in four p-refinement zones of orders 2, 4, 6
and 8 (p-non-conforming edges between them), BR1, HLL inside and the exact solver on every
zone, Dirichlet far boundaries, partially non-reflective pressure outlet, adiabatic wall,
Williamson RK5, CFL 0.1, dt cap 0.01, steady mode for 250 iterations.  Finalize requires the
four residuals within 1e-11 and the lift and drag coefficients of the wall within 1e-12; the
surface monitors read the edge storage of the last stage's derivative evaluation.

VortexTransport (VortexTransport/SETUP/ProblemFile.f90:107-150), reproduced by the oracle and
the GPU path: a
vortex carried through a periodic square at Mach 0.6 (Euler: the case has Re = 0 and no
artificial dissipation), 144 quads of orders 5, 7 and 12 with 34 subdivided edges (10 of them
curved: hanging nodes, two mortars per edge), Roe (the case file's "Inviscid Riemann Flux"
label is not the one the solver reads, so its default applies), Williamson RK5, CFL 0.4, dt cap
0.01, transient to t = 2.  Finalize requires the run to end at iteration 3237 with the max-norm
errors against the initial vortex within 1e-12 of its stored values.

The cases come from committed fixtures (tests/golden/benchmark_*.json: the case file's
parameters as horses2d_amd.hio reads them; mesh_*.npz: the HiOMesh arrays), written by
tests/golden/make_golden.py from the reference tree.  Uniform initial condition
(InitialConditions.f90:154-170), except VortexTransport's problem-file vortex
(horses2d_amd.cases.gaussian_vortex).
"""
import json
import os

import numpy as np
import pytest

from horses2d_amd import cases as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def benchmark_case(name):
    with open(os.path.join(GOLDEN, f"benchmark_{name}.json")) as f:
        d = json.load(f)
    p = d["case"]
    m = np.load(os.path.join(GOLDEN, d["mesh"]), allow_pickle=False)
    zones = [C.Zone(z["kind"], direction=z["direction"], partner=z["partner"], isothermal=z["isothermal"],
                    q=tuple(z["q"]), Tw=z["Tw"], wall_velocity=tuple(z["wall_velocity"]), riemann=z["riemann"],
                    outflow=z["outflow"], mode=z["mode"], outflow_pressure=z["outflow_pressure"],
                    function=z.get("function", "")) for z in p["zones"]]
    mesh = C.HiOMesh(m["points"], m["quads"], m["bedges"], m["bmarker"], zones,
                     m["cedges"] if "cedges" in m else None, m["cxy"] if "cxy" in m else None)
    orders = np.array(p["orders"], np.int32) if p.get("orders") else np.full(mesh.K, p["N"][0], np.int32)
    # the reference's benchmark build solves the Navier-Stokes equations; at Re = 0 without artificial
    # dissipation every viscous term is a product with mu = 0, the Euler form's computation
    ns = p["Re"] > 0.0 or p["artificial_dissipation"] != "none"
    p_ref, T_ref = p.get("p_ref", 101325.0), p.get("T_ref", 273.15)
    if name == "vortex_transport":
        ic = C.gaussian_vortex(p["Mach"])
    elif name == "couette":
        ic = C.couette_flow(p["Mach"], p["Pr"], p_ref, T_ref, perturb=True)
    else:
        ic = C.uniform(C.freestream_state(p["Mach"]))
    case = C.Case(name, mesh, orders, p["Mach"], p["Re"], p["Pr"], ns,
                  riemann=p["riemann"], viscous=p["viscous"], ip_method=p["ip_method"], sigma0=p["sigma0"],
                  cfl=p["cfl"], dt_max=p["dt_max"], steps=p["steps"], scheme=p["scheme"],
                  ic=ic, artificial_dissipation=p["artificial_dissipation"],
                  ad_intensity=p["ad_intensity"], mode=p["mode"], residual_target=p["residual_target"],
                  t_end=p["t_end"], t0=p["t0"], p_ref=p_ref, T_ref=T_ref)
    return case, d["finalize"]


# The problem-file boundary-state functions of the benchmarks ("Use function = BCn" zones):
# Couette's BC1 is its Couette profile (Couette/SETUP/ProblemFile.f90:194-218).
BOUNDARY_FUNCTIONS = {
    ("couette", "BC1"): lambda case: C.couette_flow(case.Mach, case.Pr, case.p_ref, case.T_ref),
}


def apply_boundary_functions(solver, case):
    """Every "Use function" zone's state, evaluated at its boundary nodes (ascending edge id,
    each edge's nodes in its own order), given to the solver (oracle or GPU handle)."""
    for marker, z in enumerate(case.mesh.zones, start=1):
        if z.function:
            fn = BOUNDARY_FUNCTIONS[(case.name, z.function)](case)
            x = solver.boundary_nodes(marker)
            solver.set_boundary_state(marker, fn(x[:, 0], x[:, 1]))


def test_ramp_fixture_matches_the_case_file():
    """The committed parameters are what the case reader makes of the reference's own files."""
    path = os.path.join(REFERENCE, "TestCases/Benchmark/Simulations/Ramp/SETUP/Ramp.HiOCase")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    from horses2d_amd import hio
    c = hio.case_from_hiocase(path, navier_stokes=True)
    f, _ = benchmark_case("ramp")
    for k in ("Mach", "Re", "viscous", "ip_method", "artificial_dissipation", "ad_intensity", "scheme", "mode",
              "cfl", "dt_max", "residual_target", "steps"):
        assert getattr(c, k) == getattr(f, k), k
    np.testing.assert_array_equal(c.mesh.quads, f.mesh.quads)
    np.testing.assert_array_equal(c.mesh.points, f.mesh.points)
    assert [(z.kind, z.q) for z in c.mesh.zones] == [(z.kind, tuple(z.q)) for z in f.mesh.zones]


def test_oracle_reproduces_the_ramp_golden():
    """The oracle run to the steady stop: iteration 9884 and the residuals within Finalize's 1e-11."""
    from oracle.oracle import Oracle
    case, gold = benchmark_case("ramp")
    orc = Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    _, _, it, hist = orc.integrate(case.steps, dt_max=case.dt_max, scheme=case.scheme, mode=case.mode,
                                   residual_target=case.residual_target)
    print("oracle: iteration", it, "residuals", hist[-1], "golden", gold["residuals"])
    assert it == gold["final_iteration"]
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]


@pytest.mark.gpu
def test_gpu_reproduces_the_ramp_golden():
    """The FP64 GPU path (hg_rk_step replays of the RK5 step, hg_residuals after each, the host
    loop of gpu.DGSEM.integrate) to the steady stop: iteration 9884, residuals within 1e-11."""
    from horses2d_amd import gpu
    case, gold = benchmark_case("ramp")
    dg = gpu.DGSEM(case, "fp64")
    hist = dg.integrate()
    print("gpu: iteration", dg.iter, "residuals", hist[-1], "golden", gold["residuals"])
    assert dg.iter == gold["final_iteration"]
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_ramp_steps_match_the_oracle(precision):
    """The first 60 RK5 steps (the shock forming, the dissipation switching on) against the
    oracle: FP64 within 1e-12 (max-norm), the mixed form within 1e-6 (relative L2)."""
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    case, _ = benchmark_case("ramp")
    dg = gpu.DGSEM(case, precision)
    orc = Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    dt = orc.estimate_dt(case.dt_max, case.cfl)
    assert dg.estimate_dt(case.dt_max, case.cfl) == pytest.approx(dt, rel=1e-14)
    for it in range(60):
        orc.step(0.0, dt, case.scheme)
        dg.rk_step(0.0, dt)
    a, b = dg.state(), orc.state()
    if precision == "fp64":
        e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        assert e <= 1e-12, e
        np.testing.assert_allclose(dg.residuals(), orc.residuals(), rtol=1e-10, atol=1e-10 * np.max(orc.residuals()))
    else:
        a = a.reshape(-1, 4); b = b.reshape(-1, 4)
        err = np.sqrt(np.sum((a - b) ** 2, axis=0) / np.sum(b ** 2, axis=0))
        print("mixed vs oracle after 60 steps", err)
        assert np.all(err <= 1e-6), err


def lift_drag(surface, Mach, reference_surface=1.0):
    """The lift and drag monitors (IO/Monitors.f90:1130-1138, dynamic pressure :1013-1033) from
    hg_surface_integrals' sums: (pressure force - sqrt(gamma) M viscous force) over the dynamic
    pressure 1/2 rho V^2 S, which in the reference's dimensionless variables is gamma M^2 S / 2."""
    F = surface[3:5] - surface[7:9] * np.sqrt(1.4) * Mach
    q = 0.5 * 1.4 * Mach * Mach * reference_surface
    return F[1] / q, F[0] / q


@pytest.mark.gpu
def test_gpu_reproduces_the_oracle_on_cylinder_mach2():
    """CylinderMach2 (CylinderMach2/SETUP/ProblemFile.f90:79-95): Mach 2 past a cylinder, 640
    quads of order 5 with curved (order-10 data) wall edges, Euler walls, Dirichlet inflow,
    pressure outlet, Navier-Stokes at Re = 0 with IIPG and the Laplacian dissipation (intensity
    0.4), RK5, CFL 0.05, dt cap 0.01, transient to t = 10.

    The oracle's run of this case (tests/golden/oracle_benchmark_runs.py; its end in
    oracle_benchmark_runs.json) ends with the lift coefficient of Finalize within 9e-13 of the
    stored value, the drag within 3e-8 and the residuals within 6e-6 relative, but after 16079
    iterations instead of 15599: its time steps are 3% shorter on average.  The step follows
    the largest speeds met so far (the reference's saved maxima, which grow from u + a = 3.4 to
    7.8 during the impulsive start), so a start-up that peaks 3% higher than the reference's
    would explain it; where it departs from the reference's has not been found.  Finalize's own
    bars (iteration, residuals within 1e-9, coefficients within 1e-12) are therefore not met by
    the oracle; this test holds the FP64 GPU path to the oracle's run: the same final iteration,
    residuals within 1e-9 and coefficients within 1e-11 after 16079 steps."""
    from horses2d_amd import gpu
    case, gold = benchmark_case("cylinder_mach2")
    with open(os.path.join(GOLDEN, "oracle_benchmark_runs.json")) as f:
        ref = json.load(f)["cylinder_mach2"]
    dg = gpu.DGSEM(case, "fp64")
    hist = dg.integrate(monitor=True)
    cl, cd = lift_drag(dg.surface_integrals(ref["marker"]), case.Mach)
    print("gpu: iteration", dg.iter, "t", dg.t, "residuals", hist[-1], "cl", cl, "cd", cd)
    print("oracle:", ref, "reference Finalize:", gold)
    assert dg.iter == ref["final_iteration"] and dg.t == ref["t"]
    assert np.max(np.abs(hist[-1] - np.array(ref["residuals"]))) <= 1e-9
    assert abs(cl - ref["cl"]) <= 1e-11 and abs(cd - ref["cd"]) <= 1e-11


def test_cylinder_re100_fixture_matches_the_case_file():
    path = os.path.join(REFERENCE, "TestCases/Benchmark/Simulations/CylinderRe100/SETUP/Cylinder.HiOCase")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    from horses2d_amd import hio
    c = hio.case_from_hiocase(path, navier_stokes=True)
    f, _ = benchmark_case("cylinder_re100")
    for k in ("Mach", "Re", "viscous", "riemann", "scheme", "mode", "cfl", "dt_max", "steps"):
        assert getattr(c, k) == getattr(f, k), k
    np.testing.assert_array_equal(c.N, f.N)
    assert sorted(set(int(n) for n in f.N)) == [2, 4, 6, 8]


def test_oracle_reproduces_the_cylinder_re100_golden():
    """250 RK5 iterations over the four orders: residuals within Finalize's 1e-11, lift and drag
    of the wall (marker 5) within its 1e-12."""
    from oracle.oracle import Oracle
    case, gold = benchmark_case("cylinder_re100")
    orc = Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    _, _, it, hist = orc.integrate(case.steps, dt_max=case.dt_max, scheme=case.scheme, mode=case.mode,
                                   residual_target=case.residual_target)
    cl, cd = lift_drag(orc.surface_integrals(5), case.Mach)
    print("oracle: iteration", it, "residuals", hist[-1], "cl", cl, "cd", cd, "golden", gold)
    assert it == case.steps
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]
    assert abs(cl - gold["cl"]) <= 1e-12 and abs(cd - gold["cd"]) <= 1e-12


@pytest.mark.gpu
def test_gpu_reproduces_the_cylinder_re100_golden():
    """The FP64 GPU path (one sub-handle per order, the mortar kernels between them, the RK5 step
    replayed as a graph) through the same 250 iterations: Finalize's bars."""
    from horses2d_amd import gpu
    case, gold = benchmark_case("cylinder_re100")
    dg = gpu.DGSEM(case, "fp64")
    hist = dg.integrate(monitor=True)
    cl, cd = lift_drag(dg.surface_integrals(5), case.Mach)
    print("gpu: iteration", dg.iter, "residuals", hist[-1], "cl", cl, "cd", cd, "golden", gold)
    assert dg.iter == case.steps
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]
    assert abs(cl - gold["cl"]) <= 1e-12 and abs(cd - gold["cd"]) <= 1e-12


def vortex_errors(state, case, coords):
    """Finalize's errors (VortexTransport/SETUP/ProblemFile.f90:107-131): the max over the nodes of
    |initial vortex - Q| per variable."""
    return np.max(np.abs(case.initial_state(coords).reshape(-1, 4) - state.reshape(-1, 4)), axis=0)


def test_vortex_transport_fixture_matches_the_case_file():
    path = os.path.join(REFERENCE, "TestCases/Benchmark/Simulations/VortexTransport/SETUP/VortexTransport.HiOCase")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    from horses2d_amd import hio
    c = hio.case_from_hiocase(path, navier_stokes=True)
    f, gold = benchmark_case("vortex_transport")
    for k in ("Mach", "Re", "riemann", "scheme", "mode", "cfl", "dt_max", "t_end"):
        assert getattr(c, k) == getattr(f, k), k
    np.testing.assert_array_equal(c.N, f.N)
    assert not f.navier_stokes and gold["final_iteration"] == 3237


def test_oracle_reproduces_the_vortex_transport_golden():
    """3237 RK5 steps to t = 2 across the subdivided and p-non-conforming edges: the final
    iteration exactly, the errors within Finalize's 1e-12."""
    from oracle.oracle import Oracle
    case, gold = benchmark_case("vortex_transport")
    orc = Oracle(case)
    X = orc.coords()
    orc.set_state(case.initial_state(X))
    t, _, it, _ = orc.integrate(case.steps, dt_max=case.dt_max, t_end=case.t_end, scheme=case.scheme, mode=case.mode)
    err = vortex_errors(orc.state(), case, X)
    print("oracle: iteration", it, "t", t, "errors", err, "golden", gold)
    assert it == gold["final_iteration"] and t == case.t_end
    assert np.max(np.abs(err - np.array(gold["errors"]))) <= gold["residual_tolerance"]


@pytest.mark.gpu
def test_gpu_reproduces_the_vortex_transport_golden():
    """The FP64 GPU path (one sub-handle per order, the p-non-conforming and subdivided mortar
    kernels between them, RK5 replayed as a graph) through the same 3237 steps: Finalize's bars."""
    from horses2d_amd import gpu
    case, gold = benchmark_case("vortex_transport")
    dg = gpu.DGSEM(case, "fp64")
    X = dg.coords()
    dg.integrate()
    err = vortex_errors(dg.state(), case, X)
    print("gpu: iteration", dg.iter, "t", dg.t, "errors", err, "golden", gold)
    assert dg.iter == gold["final_iteration"] and dg.t == case.t_end
    assert np.max(np.abs(err - np.array(gold["errors"]))) <= gold["residual_tolerance"]


def test_cylinder_re1000_fixture_matches_the_case_file():
    path = os.path.join(REFERENCE, "TestCases/Benchmark/Simulations/CylinderRe1000/SETUP/CylinderRe1000.HiOCase")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    from horses2d_amd import hio
    c = hio.case_from_hiocase(path, navier_stokes=True)
    f, gold = benchmark_case("cylinder_re1000")
    for k in ("Mach", "Re", "viscous", "ip_method", "sigma0", "riemann", "scheme", "mode", "cfl", "dt_max", "t_end"):
        assert getattr(c, k) == getattr(f, k), k
    np.testing.assert_array_equal(c.N, f.N)
    assert sorted(set(int(n) for n in f.N)) == [4, 6, 8, 10] and gold["final_iteration"] == 512


def test_oracle_reproduces_the_cylinder_re1000_golden():
    """CylinderRe1000 (CylinderRe1000/SETUP/ProblemFile.f90:79-99): interior penalty (SIPG, jumps
    penalty 10) across the p-non-conforming and subdivided edges of 1048 quads of orders 4, 6, 8
    and 10 (curved around the cylinder), HLL inside and the exact solver on every zone, RK5 to
    t = 0.013: iteration 512, residuals within Finalize's 1e-10, lift and drag within 1e-12."""
    from oracle.oracle import Oracle
    case, gold = benchmark_case("cylinder_re1000")
    orc = Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    t, _, it, hist = orc.integrate(case.steps, dt_max=case.dt_max, t_end=case.t_end, scheme=case.scheme,
                                   mode=case.mode)
    cl, cd = lift_drag(orc.surface_integrals(5), case.Mach)
    print("oracle: iteration", it, "t", t, "residuals", hist[-1], "cl", cl, "cd", cd, "golden", gold)
    assert it == gold["final_iteration"]
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]
    assert abs(cl - gold["cl"]) <= 1e-12 and abs(cd - gold["cd"]) <= 1e-12


@pytest.mark.gpu
def test_gpu_reproduces_the_cylinder_re1000_golden():
    """The FP64 GPU path (one sub-handle per order, the IP mortar kernels with their gradient
    fluxes, RK5 replayed as a graph) through the same 512 iterations: Finalize's bars."""
    from horses2d_amd import gpu
    case, gold = benchmark_case("cylinder_re1000")
    dg = gpu.DGSEM(case, "fp64")
    hist = dg.integrate(monitor=True)
    cl, cd = lift_drag(dg.surface_integrals(5), case.Mach)
    print("gpu: iteration", dg.iter, "t", dg.t, "residuals", hist[-1], "cl", cl, "cd", cd, "golden", gold)
    assert dg.iter == gold["final_iteration"]
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]
    assert abs(cl - gold["cl"]) <= 1e-12 and abs(cd - gold["cd"]) <= 1e-12


def test_couette_fixture_matches_the_case_file():
    path = os.path.join(REFERENCE, "TestCases/Benchmark/Simulations/Couette/SETUP/Couette.HiOCase")
    if not os.path.exists(path):
        pytest.skip("reference tree not present")
    from horses2d_amd import hio
    c = hio.case_from_hiocase(path, navier_stokes=True)
    f, gold = benchmark_case("couette")
    for k in ("Mach", "Re", "Pr", "viscous", "riemann", "scheme", "mode", "cfl", "dt_max", "residual_target",
              "p_ref", "T_ref"):
        assert getattr(c, k) == getattr(f, k), k
    np.testing.assert_array_equal(c.N, f.N)
    assert [z.function for z in c.mesh.zones] == ["", "", "BC1", ""]
    assert gold["final_iteration"] == 19585


def test_oracle_run_reproduces_the_couette_golden():
    """Couette (Couette/SETUP/ProblemFile.f90:90-92,156): BR1 and Roe on 144 quads of orders 5, 7
    and 10 with curved p-non-conforming edges, isothermal walls (the top one moving at Mach
    0.8), a Dirichlet inflow whose state is the problem file's BoundaryConditionFunction1 at each
    node, a pressure outlet, RK5 in steady mode to residual 1e-4 from the perturbed Couette
    profile.  The oracle's whole run (19585 iterations, about two minutes on 6 cores) is
    recorded by tests/golden/oracle_benchmark_runs.py; Finalize requires iteration 19585 and the
    residuals within 1e-12."""
    import json
    with open(os.path.join(GOLDEN, "oracle_benchmark_runs.json")) as f:
        run = json.load(f)["couette"]
    _, gold = benchmark_case("couette")
    assert run["final_iteration"] == gold["final_iteration"]
    assert np.max(np.abs(np.array(run["residuals"]) - np.array(gold["residuals"]))) <= gold["residual_tolerance"]


def test_oracle_starts_the_couette_run_like_its_record():
    """The first 20 iterations of the recorded run, from the committed fixtures: the same time
    step and the same residuals at iteration 20 (so the record belongs to this case and these
    boundary states)."""
    import json
    from oracle.oracle import Oracle
    with open(os.path.join(GOLDEN, "oracle_benchmark_runs.json")) as f:
        run = json.load(f)["couette"]
    case, _ = benchmark_case("couette")
    orc = Oracle(case)
    apply_boundary_functions(orc, case)
    orc.set_state(case.initial_state(orc.coords()))
    _, dts, it, hist = orc.integrate(20, dt_max=case.dt_max, scheme=case.scheme, mode=case.mode,
                                     residual_target=case.residual_target)
    assert it == 20 and dts[0] == run["first_dt"]
    np.testing.assert_allclose(hist[-1], run["residuals_at_20"], rtol=1e-13)


@pytest.mark.gpu
def test_gpu_reproduces_the_couette_golden():
    """The FP64 GPU path with the inflow states handed over per node (hg_set_boundary_state), the
    curved mortar kernels (BR1 uStar dS n per direction), RK5 steady to residual 1e-4:
    Finalize's bars (iteration 19585, residuals within 1e-12)."""
    from horses2d_amd import gpu
    case, gold = benchmark_case("couette")
    dg = gpu.DGSEM(case, "fp64")
    apply_boundary_functions(dg, case)
    dg.upload(case.initial_state(dg.coords()))
    hist = dg.integrate()
    print("gpu: iteration", dg.iter, "residuals", hist[-1], "golden", gold)
    assert dg.iter == gold["final_iteration"]
    assert np.max(np.abs(hist[-1] - np.array(gold["residuals"]))) <= gold["residual_tolerance"]
