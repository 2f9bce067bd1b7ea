"""Meshes with subdivided edges (one element's edge meets two elements whose common node hangs on
it, MeshFile.f90:364-490): the FP64 GPU path against the oracle, through the C ABI, at the parity
bar of the conforming cases (1e-12 relative, max-norm; for QDot, twice the oracle's own
sensitivity to last-place state changes where that is larger).  The reference solves such an edge's
Riemann problem on two mortars from the three sides' traces and returns it with the mortar
matrices (DGSpatialDiscretizationMethods.f90:761-1140, DGViscousMethods.f90:618-780); the
library does it in the subdivided-mortar kernels between the per-order sub-handles.  The small
elements run their halves forwards and backwards (node lists rotated per element), and the
hanging nodes, displaced by the skew map, are moved back onto the big edge by both builders."""
import numpy as np
import pytest

from horses2d_amd import cases as C
from test_gpu_parity import TOL, pair, rel, run_both, _flow_ic
from test_oracle_subdivided import refined_case

pytestmark = pytest.mark.gpu


def euler_case(seed, lo, hi):
    q = C.freestream_state(0.4, 0.3)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    c = refined_case(zones, seed, lo=lo, hi=hi)
    c.ic = C.isentropic_vortex(0.4, beta=4.0, xc=1.4, yc=0.5)
    c.cfl = 0.1
    return c


def br1_case(seed, riemann="hll"):
    """BR1 on the refined channel: adiabatic viscous wall below, pressure outlet right, Dirichlet
    above and left."""
    q = C.freestream_state(0.3)
    zones = [C.Zone(C.BC_VISCOUSWALL), C.Zone(C.BC_PRESSUREOUTLET, q=tuple(q)),
             C.Zone(C.BC_DIRICHLET, q=tuple(q), riemann=C.RIEMANN["exact"]),
             C.Zone(C.BC_DIRICHLET, q=tuple(q), riemann=C.RIEMANN["exact"])]
    c = refined_case(zones, seed, ns=True, lo=2, hi=8)
    c.riemann = riemann
    c.ic = _flow_ic(0.3, 0.4)
    c.cfl = 0.1
    return c


def vortex_transport_case():
    """The reference's VortexTransport benchmark case: 144 quads of orders 5, 7 and 12, 34
    subdivided edges (10 curved), periodic boundaries, Euler."""
    from test_benchmark_goldens import benchmark_case
    case, _ = benchmark_case("vortex_transport")
    return case


def ip_case(seed, method):
    """Interior penalty on the refined channel: the penalty at each mortar's order and the
    gradient fluxes returned with T_BKW to all three sides (:906-940)."""
    c = br1_case(seed, "hll")
    c.viscous, c.ip_method, c.sigma0, c.cfl = "ip", method, 1.0, 0.02
    return c


def vortex_transport_ns(viscous):
    """The VortexTransport mesh (curved subdivided edges) with the Navier-Stokes equations at
    Re = 100: BR1 returns uStar dS n per mortar node and direction
    (BaseClass_ComputeSolutionRiemann_CurvedSubdivided, DGViscousMethods.f90:700-780)."""
    c = vortex_transport_case()
    c.navier_stokes, c.viscous, c.Re, c.sigma0, c.cfl = True, viscous, 100.0, 1.0, 0.02
    return c


CASES = {
    "euler_one_order": lambda: euler_case(0, 4, 5),
    "euler_orders": lambda: euler_case(1, 2, 8),
    "euler_orders_rot": lambda: euler_case(3, 1, 9),
    "br1_hll": lambda: br1_case(2, "hll"),
    "br1_roe": lambda: br1_case(5, "roe"),
    "vortex_transport": vortex_transport_case,
    "sipg_hll": lambda: ip_case(2, "sipg"),
    "nipg_hll": lambda: ip_case(6, "nipg"),
    "vortex_transport_br1": lambda: vortex_transport_ns("br1"),
    "vortex_transport_sipg": lambda: vortex_transport_ns("ip"),
}


def test_the_cases_hold_every_orientation():
    from oracle.oracle import Oracle
    seen = set()
    for name in ("euler_orders", "euler_orders_rot", "br1_hll", "br1_roe"):
        orc = Oracle(CASES[name]())
        for k in range(orc.E):
            s = orc.edge_subdivided(k)
            if s is not None:
                seen.add((s["dir_north"], s["dir_south"]))
    assert seen == {(1, 1), (1, -1), (-1, 1), (-1, -1)}, seen


def rounding_sensitivity(case):
    """How far the oracle's own QDot moves (relative, max-norm) when every state value moves by
    one unit in the last place: the weak vortex of VortexTransport in its Mach 0.6 stream makes
    QDot a small difference of large fluxes (1e-11 here, 1e-15 on the channels)."""
    from oracle.oracle import Oracle
    orc = Oracle(case)
    Q0 = case.initial_state(orc.coords())
    orc.set_state(Q0)
    q0 = orc.time_derivative().copy()
    noise = np.random.default_rng(0).choice([-1.0, 1.0], size=Q0.shape)
    orc.set_state(Q0 * (1.0 + np.finfo(np.float64).eps * noise))
    return rel(orc.time_derivative(), q0)


@pytest.mark.parametrize("name", sorted(CASES))
def test_qdot_matches_the_oracle(name):
    """Bar: 1e-12, or twice the oracle's own sensitivity to last-place changes of the state where
    the field is worse conditioned than that."""
    case = CASES[name]()
    dg, orc = pair(case)
    err = rel(dg.time_derivative(), orc.time_derivative())
    bar = max(TOL, 2.0 * rounding_sensitivity(case))
    print(name, "QDot rel", err, "bar", bar)
    assert err <= bar


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
    # the residuals are max |QDot| of the last stage: QDot moves by the oracle's own rounding
    # sensitivity per last-place change of the state, times how many last places the states differ
    rtol = max(1e-10, 2.0 * rounding_sensitivity(case) * max(1.0, err / np.finfo(np.float64).eps))
    np.testing.assert_allclose(dg.residuals(), orc.residuals(), rtol=rtol, atol=1e-12 * np.max(orc.residuals()))


def test_monitors_match_the_oracle():
    case = br1_case(2)
    dg, orc = pair(case)
    run_both(dg, orc, case, 3)
    np.testing.assert_allclose(dg.conservation_sums(), orc.conservation_sums(), rtol=1e-12)
    np.testing.assert_allclose(dg.volume_integrals(), orc.volume_integrals(), rtol=1e-12)
    for marker in (1, 2, 3, 4):
        s_g, s_o = dg.surface_integrals(marker), orc.surface_integrals(marker)
        assert np.max(np.abs(s_g - s_o)) <= 1e-12 * max(np.max(np.abs(s_o)), 1.0), marker


def test_graph_and_direct_steps_agree_bit_for_bit():
    from horses2d_amd import gpu
    case = br1_case(5, "roe")
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


def test_unsupported_forms_are_refused():
    from horses2d_amd import gpu
    c = br1_case(2)
    c.artificial_dissipation = "physical"
    with pytest.raises(RuntimeError, match="without artificial dissipation"):
        gpu.DGSEM(c, "fp64")
