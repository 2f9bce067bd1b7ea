"""Zones whose state comes from a problem-file function ("Use function = BC1..BC4",
AssociateBoundaryConditionFunction, BoundaryConditionFunctionDrivers.incf:22-63): the reference
evaluates q = BoundaryState(edge % x(:, iXi), time) at each node of the zone's edges in place of
the zone state (Dirichlet.incf:182-183, Farfield.incf:179-180, newDirichlet.incf:195-196,
PressureOutlet.incf:187-188, PressureInlet.incf:160-161).  The caller evaluates the function at
the nodes the library lists (hg_boundary_nodes) and hands the states over
(hg_set_boundary_state); the FP64 GPU path is held to the oracle given the same states, at the
parity bar of the other cases (1e-12 relative, max-norm), the mixed form to the FP64 one."""
import numpy as np
import pytest

from horses2d_amd import cases as C
from test_gpu_parity import TOL, pair, rel, run_both, _channel, _flow_ic

pytestmark = pytest.mark.gpu


def profile_state(Mach):
    """A state that varies along the boundary: a sheared, heated inflow (no closed form needed)."""
    V = Mach * np.sqrt(1.4)

    def fn(x, y):
        u = V * (1.0 + 0.3 * np.sin(2.0 * y + 0.5 * x))
        v = V * 0.1 * np.cos(1.5 * y)
        rho = 1.0 + 0.05 * np.cos(3.0 * y)
        p = 1.0 + 0.04 * np.sin(y - x)
        return C._pack(rho, u, v, p)
    return fn


def give_states(solvers, marker, fn):
    """The same node states to every solver (oracle and GPU handles list the same nodes)."""
    xs = [s.boundary_nodes(marker) for s in solvers]
    for x in xs[1:]:
        np.testing.assert_array_equal(x, xs[0])
    q = fn(xs[0][:, 0], xs[0][:, 1])
    for s in solvers:
        s.set_boundary_state(marker, q)
    return q


def channel_case(kind, viscous, amp=1.5):
    """The skewed channel with a `kind` inlet (marker 4) and a pressure outlet (marker 2)."""
    return _channel(kind, C.BC_PRESSUREOUTLET, 0.3, amp, viscous)


def mortar_case(viscous):
    """The rotated channel of several orders (p-non-conforming edges) with a Dirichlet inlet."""
    from test_gpu_mortar import br1_case
    c = br1_case(2, "hll")
    c.viscous, c.sigma0, c.cfl = viscous, 1.0, 0.02
    return c


CASES = {
    "euler_dirichlet": lambda: (channel_case(C.BC_DIRICHLET, "none"), 4),
    "br1_dirichlet": lambda: (channel_case(C.BC_DIRICHLET, "br1"), 4),
    "sipg_dirichlet": lambda: (channel_case(C.BC_DIRICHLET, "ip"), 4),
    "euler_pressure_inlet": lambda: (channel_case(C.BC_PRESSUREINLET, "none"), 4),
    "br1_newdirichlet": lambda: (channel_case(C.BC_NEWDIRICHLET, "br1"), 4),
    "euler_pressure_outlet": lambda: (channel_case(C.BC_DIRICHLET, "none"), 2),
    "br1_farfield_annulus": lambda: (C.ns_cylinder_case(4, 16, N=3, Re=200.0), 2),
    "br1_mortar_dirichlet": lambda: (mortar_case("br1"), 4),
    "sipg_mortar_dirichlet": lambda: (mortar_case("ip"), 3),
}


def test_boundary_nodes_are_the_oracles():
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    for name, make in CASES.items():
        case, marker = make()
        dg, orc = gpu.DGSEM(case, "fp64"), Oracle(case)
        x = dg.boundary_nodes(marker)
        np.testing.assert_array_equal(x, orc.boundary_nodes(marker))
        assert x.shape[0] > 0, name


@pytest.mark.parametrize("name", sorted(CASES))
def test_time_derivative_and_steps_match_the_oracle(name):
    case, marker = CASES[name]()
    dg, orc = pair(case)
    give_states([orc, dg], marker, profile_state(case.Mach))
    qd_o, qd_g = orc.time_derivative(), dg.time_derivative()
    e0 = rel(qd_g, qd_o)
    run_both(dg, orc, case, 10)
    e1 = rel(dg.state(), orc.state())
    print(name, "QDot rel", e0, "state rel after 10 steps", e1)
    assert e0 <= TOL and e1 <= TOL


def test_the_states_enter_the_flux():
    """A state different from the zone's changes QDot at the zone's faces (the nodes are
    used), and dropping it gives back the zone state's QDot bit for bit."""
    case, marker = CASES["br1_dirichlet"]()
    dg, _ = pair(case)
    q0 = dg.time_derivative()
    give_states([dg], marker, profile_state(case.Mach))
    q1 = dg.time_derivative()
    assert rel(q1, q0) > 1e-6
    dg.set_boundary_state(marker, None)
    np.testing.assert_array_equal(dg.time_derivative(), q0)


def test_zone_state_given_per_node_is_bit_identical():
    """The zone's own state handed over node by node gives the zone state's results bit for bit
    (graph replay of a few steps included)."""
    from horses2d_amd import gpu
    case, marker = CASES["sipg_dirichlet"]()
    a = gpu.DGSEM(case, "fp64")
    b = gpu.DGSEM(case, "fp64")
    Q0 = case.initial_state(a.coords())
    a.upload(Q0)
    b.upload(Q0)
    zq = np.array(case.mesh.zones[marker - 1].q)
    n = b.boundary_nodes(marker).shape[0]
    b.set_boundary_state(marker, np.tile(zq, (n, 1)))
    dt = a.estimate_dt(case.dt_max, case.cfl)
    b.estimate_dt(case.dt_max, case.cfl)
    for _ in range(3):
        a.rk_step(0.0, dt)
        b.rk_step(0.0, dt)
    np.testing.assert_array_equal(a.state(), b.state())


def test_mixed_form_follows_the_fp64_form():
    from horses2d_amd import gpu
    case, marker = CASES["br1_dirichlet"]()
    f64 = gpu.DGSEM(case, "fp64")
    mix = gpu.DGSEM(case, "mixed")
    Q0 = case.initial_state(f64.coords())
    fn = profile_state(case.Mach)
    for s in (f64, mix):
        s.upload(Q0)
        give_states([s], marker, fn)
    dt = f64.estimate_dt(case.dt_max, case.cfl)
    for _ in range(20):
        f64.rk_step(0.0, dt)
        mix.rk_step(0.0, dt)
    err = rel(mix.state(), f64.state())
    print("mixed vs FP64 after 20 steps", err)
    assert err <= 1e-6


def test_kinds_without_a_state_are_refused():
    from horses2d_amd import gpu
    case, _ = CASES["br1_farfield_annulus"]()
    dg = gpu.DGSEM(case, "fp64")
    n = dg.boundary_nodes(1).shape[0]   # the viscous wall
    with pytest.raises(RuntimeError, match="no boundary state"):
        dg.set_boundary_state(1, np.ones((n, 4)))
    with pytest.raises(RuntimeError, match="node count"):
        dg.set_boundary_state(2, np.ones((3, 4)))
