"""Meshes whose elements have several polynomial orders (p-non-conforming edges, mortars): the
FP64 GPU path against the oracle, through the C ABI, at the parity bar of the conforming cases
(1e-12 relative, max-norm).  The reference evaluates such an edge on its own nodes (the higher
order), interpolates the low side's traces up with T_forward and returns the flux with
T_backward (QuadElement_EdgeProcedures.incf:749-1206); the edge Types 3-6 (low side LEFT or
RIGHT, the RIGHT side running the edge forwards or backwards) all occur in these meshes, since
the element orders are drawn per element and the node lists are rotated per element."""
import numpy as np
import pytest

from horses2d_amd import cases as C
from test_gpu_parity import TOL, pair, rel, run_both, _flow_ic

pytestmark = pytest.mark.gpu


def rotated_channel(nx, ny, zones, skew, seed):
    """The channel mesh with each element's node list rotated by a seeded amount: the elements
    keep their shape, their local faces and edge directions change, so edges run forwards and
    backwards on their RIGHT side."""
    mesh = C.channel(nx, ny, 3.0, 1.0, zones, skew=skew)
    rng = np.random.default_rng(seed)
    q = mesh.quads.copy()
    for e in range(q.shape[0]):
        q[e] = np.roll(q[e], int(rng.integers(0, 4)))
    mesh.quads = q
    return mesh


def mortar_types(case):
    """The p-non-conforming edges of a case as the oracle links them: (low side, inverse) pairs."""
    from oracle.oracle import Oracle
    orc = Oracle(case)
    edges, _, _ = orc.connectivity()
    seen = set()
    for k in range(edges.shape[0]):
        low, N, NL, _, _ = orc.edge_mortar(k)
        if low >= 0:
            seen.add((low, int(edges[k, 6])))
    return seen


def euler_case(seed=1):
    q = C.freestream_state(0.4, 0.3)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    mesh = rotated_channel(6, 4, zones, 0.3, seed)
    N = np.random.default_rng(seed).integers(2, 7, size=mesh.K).astype(np.int32)
    return C.Case("mortar_euler", mesh, N, 0.4, 0.0, 0.72, False, riemann="roe", cfl=0.1,
                  ic=C.isentropic_vortex(0.4, beta=4.0, xc=1.4, yc=0.5))


def br1_case(seed=2, riemann="hll"):
    """BR1 on the rotated channel: adiabatic viscous wall below, pressure outlet right, Dirichlet
    above and left (the zone kinds of the reference's CylinderRe100 benchmark)."""
    q = C.freestream_state(0.3)
    zones = [C.Zone(C.BC_VISCOUSWALL), C.Zone(C.BC_PRESSUREOUTLET, q=tuple(q)),
             C.Zone(C.BC_DIRICHLET, q=tuple(q), riemann=C.RIEMANN["exact"]),
             C.Zone(C.BC_DIRICHLET, q=tuple(q), riemann=C.RIEMANN["exact"])]
    mesh = rotated_channel(6, 4, zones, 0.3, seed)
    N = np.random.default_rng(seed).integers(2, 7, size=mesh.K).astype(np.int32)
    return C.Case("mortar_br1", mesh, N, 0.3, 100.0, 0.72, True, riemann=riemann, viscous="br1", cfl=0.1,
                  ic=_flow_ic(0.3, 0.4))


def ip_case(seed, method, riemann="hll"):
    """Interior penalty across orders (ComputeRiemannSolver_InteriorEdge with the IP penalty at
    the edge's order and, for SIPG and NIPG, the gradient fluxes returned by
    ProjectGradientFluxes, QuadElement_EdgeProcedures.incf:1208-1339) on the rotated channel
    (CFL 0.02: the penalty's stiffness, which the time-step estimate does not see)."""
    c = br1_case(seed, riemann)
    c.viscous, c.ip_method, c.sigma0, c.cfl = "ip", method, 1.0, 0.02
    return c


def annulus_case(viscous="br1"):
    """The curved annulus (far field, adiabatic wall), BR1, the orders changing between
    azimuthal sectors: the p-non-conforming edges are the straight radial ones."""
    c = C.ns_cylinder_case(4, 16, N=3, Re=200.0)
    nr = 4
    c.N = np.array([2 + (e // nr) % 4 for e in range(c.mesh.K)], np.int32)
    c.viscous = viscous
    return c


def curved_annulus_case(viscous="br1"):
    """The curved annulus with the orders changing from ring to ring (4, 5, 6, 4 outwards): the
    p-non-conforming edges are the circumferential arcs (ComputeRiemannSolver_InteriorCurvedEdge;
    BR1 returns uStar dS n per node and direction, BaseClass_ComputeSolutionRiemann_CurvedInterior).
    An interior edge's curve takes the order NLow / 2 (QuadMesh_Construct.incf:104-118), so
    orders of at least 4 keep the arcs curved."""
    c = C.ns_cylinder_case(4, 16, N=3, Re=200.0)
    c.N = np.array([4 + (e % 4) % 3 for e in range(c.mesh.K)], np.int32)   # elements run radially fastest
    c.viscous, c.sigma0 = viscous, 1.0
    c.cfl, c.dt_max = C.stable_cfl(6), C.viscous_dt_cap(c.mesh, 6, 200.0)
    return c


CASES = {
    "euler_seed1": lambda: euler_case(1),
    "euler_seed5": lambda: euler_case(5),
    "br1_hll": lambda: br1_case(2, "hll"),
    "br1_roe": lambda: br1_case(7, "roe"),
    "br1_annulus": annulus_case,
    "sipg_hll": lambda: ip_case(3, "sipg"),
    "nipg_roe": lambda: ip_case(8, "nipg", "roe"),
    "iipg_hll": lambda: ip_case(5, "iipg"),
    "sipg_annulus": lambda: annulus_case("ip"),
    "br1_curved_annulus": lambda: curved_annulus_case("br1"),
    "sipg_curved_annulus": lambda: curved_annulus_case("ip"),
    "euler_curved_annulus": lambda: _euler(curved_annulus_case("br1")),
}


def _euler(c):
    c.navier_stokes, c.Re, c.viscous = False, 0.0, "none"
    c.mesh.zones[0] = C.Zone(C.BC_EULERWALL)
    return c


def test_the_cases_hold_every_edge_type():
    """Types 3-6: low side LEFT or RIGHT, the RIGHT side forwards or backwards."""
    seen = set()
    for make in CASES.values():
        seen |= mortar_types(make())
    assert seen == {(0, 0), (0, 1), (1, 0), (1, 1)}, seen


@pytest.mark.parametrize("name", sorted(CASES))
def test_qdot_matches_the_oracle(name):
    case = CASES[name]()
    dg, orc = pair(case)
    q_o = orc.time_derivative()
    q_g = dg.time_derivative()
    err = rel(q_g, q_o)
    print(name, "QDot rel", err)
    assert err <= TOL


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("scheme", ["rk3", "rk5"])
def test_steps_match_the_oracle(name, scheme):
    case = CASES[name]()
    case.scheme = scheme
    dg, orc = pair(case)
    run_both(dg, orc, case, 10)
    err = rel(dg.state(), orc.state())
    print(name, scheme, "state rel after 10 steps", err, "residuals", dg.residuals(), orc.residuals())
    assert err <= TOL
    np.testing.assert_allclose(dg.residuals(), orc.residuals(), rtol=1e-10, atol=1e-12 * np.max(orc.residuals()))


def test_monitors_match_the_oracle():
    """Conservation sums, volume and surface integrals and probes gathered over the orders."""
    case = br1_case(2)
    dg, orc = pair(case)
    run_both(dg, orc, case, 3)
    np.testing.assert_allclose(dg.conservation_sums(), orc.conservation_sums(), rtol=1e-12)
    np.testing.assert_allclose(dg.volume_integrals(), orc.volume_integrals(), rtol=1e-12)
    for marker in (1, 2, 3, 4):
        s_g, s_o = dg.surface_integrals(marker), orc.surface_integrals(marker)
        assert np.max(np.abs(s_g - s_o)) <= 1e-12 * max(np.max(np.abs(s_o)), 1.0), marker
    rng = np.random.default_rng(3)
    n = 40
    elem = rng.integers(0, case.mesh.K, n).astype(np.int32)
    xi, eta = rng.uniform(0.05, 0.95, n), rng.uniform(0.05, 0.95, n)
    var = rng.integers(0, 4, n).astype(np.int32)
    dg.probes_set(elem, xi, eta, var)
    np.testing.assert_allclose(dg.probes_update(), orc.probes(elem, xi, eta, var), rtol=1e-12, atol=1e-14)


def test_graph_and_direct_steps_agree_bit_for_bit():
    from horses2d_amd import gpu
    case = br1_case(2)
    a = gpu.DGSEM(case, "fp64")
    b = gpu.DGSEM(case, "fp64")
    b.set_graph(False)
    dt = a.estimate_dt(case.dt_max, case.cfl)
    b.estimate_dt(case.dt_max, case.cfl)
    for _ in range(4):
        a.rk_step(0.0, dt)
        b.rk_step(0.0, dt)
    np.testing.assert_array_equal(a.state(), b.state())
    np.testing.assert_array_equal(a.residuals(), b.residuals())


def test_curved_cases_have_curved_mortars():
    """The curved annulus cases put their p-non-conforming edges on the curved rings."""
    from oracle.oracle import Oracle
    orc = Oracle(curved_annulus_case())
    edges, _, _ = orc.connectivity()
    nx, _, _, _ = orc.edge_geometry(16)
    curved = []
    for k in range(edges.shape[0]):
        low, N, _, _, _ = orc.edge_mortar(k)
        if low >= 0 and np.ptp(nx[k, :N + 1]) > 1e-9:   # the normal turns along the edge
            curved.append(k)
    assert len(curved) >= 16


def test_unsupported_forms_are_refused():
    from horses2d_amd import gpu
    c = br1_case(2)
    c.artificial_dissipation = "physical"
    with pytest.raises(RuntimeError, match="without artificial dissipation"):
        gpu.DGSEM(c, "fp64")
    c = C.euler_vortex_case(4, 3)
    c.N = np.array([2, 3] * (c.mesh.K // 2), np.int32)
    with pytest.raises(RuntimeError, match="periodic edges between elements of different orders"):
        gpu.DGSEM(c, "fp64")
