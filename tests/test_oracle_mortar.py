"""Pins of the oracle's p-non-conforming edges (mortars), which meshes with several polynomial
orders need (QuadElement_EdgeProcedures.incf:30-199, 749-1206; QuadMesh_Construct.incf:104-118):
what the mathematics fixes, independent of the restatement.

* T_forward (low-order nodes -> edge nodes) interpolates every polynomial of degree <= NLow
  exactly; T_backward = Minv_low T_forward^T M_edge is the L2 projection back, so
  T_backward T_forward = I and T_backward conserves the edge integral (sum w_low (Tb F) =
  sum w_edge F): a flux returned to the low side loses nothing.
* A uniform flow stays uniform (QDot at rounding) on skewed meshes of random orders whose
  element node lists are rotated (every edge Type 3-6 occurs), Euler and BR1.
* In a closed box of Euler walls the discrete mass and energy are conserved with mortar edges
  between orders.
The whole path is pinned end to end by the reference's goldens (tests/test_benchmark_goldens.py):
CylinderRe100 (BR1, orders 2, 4, 6 and 8), CylinderRe1000 (interior penalty with the gradient
fluxes of ProjectGradientFluxes, orders 4 to 10) and Couette (BR1 across curved edges between
orders 5, 7 and 10, a function-given inflow).
"""
import numpy as np
import pytest

from horses2d_amd import cases as C


def rotated_channel(nx, ny, zones, skew, seed):
    mesh = C.channel(nx, ny, 3.0, 1.0, zones, skew=skew)
    rng = np.random.default_rng(seed)
    q = mesh.quads.copy()
    for e in range(q.shape[0]):
        q[e] = np.roll(q[e], int(rng.integers(0, 4)))
    mesh.quads = q
    return mesh


def gauss01(N):
    x, w = np.polynomial.legendre.leggauss(N + 1)
    return 0.5 + 0.5 * x, 0.5 * w


def mortar_edges(orc):
    out = []
    for k in range(orc.E):
        low, N, NL, Tf, Tb = orc.edge_mortar(k)
        if low >= 0:
            out.append((N, NL, Tf, Tb))
    return out


def some_mesh(seed=4):
    q = C.freestream_state(0.4, 0.3)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    mesh = rotated_channel(6, 4, zones, 0.3, seed)
    N = np.random.default_rng(seed).integers(1, 9, size=mesh.K).astype(np.int32)
    return C.Case("m", mesh, N, 0.4, 0.0, 0.72, False, riemann="roe", ic=C.uniform(q))


def test_projection_matrices():
    from oracle.oracle import Oracle
    orc = Oracle(some_mesh())
    pairs = set()
    for N, NL, Tf, Tb in mortar_edges(orc):
        pairs.add((NL, N))
        xl, wl = gauss01(NL)
        xh, wh = gauss01(N)
        for d in range(NL + 1):   # exact interpolation of degree <= NLow
            np.testing.assert_allclose(Tf @ xl ** d, xh ** d, atol=1e-13)
        np.testing.assert_allclose(Tb @ Tf, np.eye(NL + 1), atol=1e-13)
        F = np.cos(3.0 * xh) + xh ** (N + 1)   # any edge values
        assert abs(wl @ (Tb @ F) - wh @ F) <= 1e-14
    assert len(pairs) >= 10


@pytest.mark.parametrize("viscous", ["none", "br1", "sipg", "nipg"])
@pytest.mark.parametrize("seed", [4, 9])
def test_free_stream_is_preserved_across_orders(viscous, seed):
    from oracle.oracle import Oracle
    q = C.freestream_state(0.3, 0.2)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    mesh = rotated_channel(6, 4, zones, 0.3, seed)
    N = np.random.default_rng(seed).integers(1, 9, size=mesh.K).astype(np.int32)
    ns = viscous != "none"
    c = C.Case("fs", mesh, N, 0.3, 100.0 if ns else 0.0, 0.72, ns, riemann="roe",
               viscous={"none": "none", "br1": "br1"}.get(viscous, "ip"),
               ip_method=viscous if viscous in ("sipg", "nipg") else "sipg", sigma0=10.0, ic=C.uniform(q))
    orc = Oracle(c)
    assert mortar_edges(orc)
    orc.set_state(c.initial_state(orc.coords()))
    # the penalty sigma0 mu N^2 / h (about 60 here) multiplies the rounding of the trace jumps
    assert np.max(np.abs(orc.time_derivative())) <= (1e-10 if viscous in ("sipg", "nipg") else 1e-11)


@pytest.mark.parametrize("viscous", ["none", "br1", "ip"])
def test_free_stream_is_preserved_across_curved_orders(viscous):
    """The curved annulus with the orders changing from ring to ring (curved p-non-conforming
    edges: n and dS per node, BR1 uStar dS n projected per direction), far field all around."""
    from oracle.oracle import Oracle
    c = C.ns_cylinder_case(4, 16, N=3, Re=200.0)
    q = c.mesh.zones[1].q
    c.mesh.zones[0] = C.Zone(C.BC_FARFIELD, q=q)
    c.N = np.array([4 + (e % 4) % 3 for e in range(c.mesh.K)], np.int32)   # arcs keep curving: NLow >= 4
    c.ic = C.uniform(np.array(q))
    if viscous == "none":
        c.navier_stokes, c.Re, c.viscous = False, 0.0, "none"
    else:
        c.viscous, c.sigma0 = viscous, 10.0
    orc = Oracle(c)
    assert mortar_edges(orc)
    orc.set_state(c.initial_state(orc.coords()))
    assert np.max(np.abs(orc.time_derivative())) <= 1e-10


def test_closed_box_conserves_mass_and_energy_across_orders():
    """Euler walls all around (prescribed pressure flux, no mass or energy through the walls):
    sum_e w^T J QDot vanishes for mass and energy, with mortar edges between random orders."""
    from oracle.oracle import Oracle
    zones = [C.Zone(C.BC_EULERWALL) for _ in range(4)]
    mesh = rotated_channel(5, 4, zones, 0.3, 11)
    N = np.random.default_rng(11).integers(1, 8, size=mesh.K).astype(np.int32)

    def ic(x, y):
        return C._pack(1.0 + 0.1 * np.sin(2 * x + y), 0.2 * np.cos(x - 2 * y), 0.1 * np.sin(x * y),
                       1.0 + 0.05 * np.cos(3 * x))
    c = C.Case("box", mesh, N, 0.3, 0.0, 0.72, False, riemann="roe", ic=ic)
    orc = Oracle(c)
    orc.set_state(c.initial_state(orc.coords()))
    qd = orc.time_derivative().reshape(-1, 4)
    _, jac = orc.metrics()
    w = []
    for n in N:
        _, wq = gauss01(int(n))
        w.append(np.outer(wq, wq).T.reshape(-1))
    wJ = np.concatenate(w) * jac
    scale = np.max(np.abs(qd), axis=0)
    rate = wJ @ qd
    assert abs(rate[0]) <= 1e-13 * scale[0] and abs(rate[3]) <= 1e-13 * scale[3], rate
