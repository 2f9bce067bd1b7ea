"""Pins of the oracle's subdivided (h-non-conforming) edges: one element's edge meets two elements
whose common node hangs on it (MeshFile.f90:364-490 merges such edges; the reference solves the
Riemann problem on two mortars, QuadElement_EdgeProcedures.incf:201-355, 530-748,
DGSpatialDiscretizationMethods.f90:761-1140, DGViscousMethods.f90:618-780).  What the
mathematics fixes, independent of the restatement:

* T_FWD from the big element interpolates every polynomial of its degree exactly on each half;
  T_BKW = Minv T_FWD^T M is the L2 projection back, so the two halves' returns add up to the
  identity on the big side (T_LN_BKW T_LN_FWD + T_LS_BKW T_LS_FWD = 2 I: each half carries
  half the edge's quadrature) and a small side's T_BKW T_FWD = I.
* A uniform flow stays uniform on skewed, refined meshes of random orders with rotated node
  lists (the small elements run their halves forwards and backwards), Euler and BR1: the
  hanging nodes, displaced by the skew map, have to be moved back onto the big edge for this.
* In a closed box of Euler walls, mass and energy are conserved through the mortars.
The whole path, curved subdivided edges included, is pinned end to end by the reference's
VortexTransport golden (tests/test_benchmark_goldens.py).
"""
import numpy as np
import pytest

from horses2d_amd import cases as C

REFINE = [(1, 1), (2, 1), (3, 2), (0, 3), (4, 0)]


def cell_orders(nx, ny, refine, rng, lo=1, hi=9):
    """One random order per cell of the coarse grid, the four parts of a split cell sharing it
    (the element order of refined_channel)."""
    out = []
    for i in range(nx):
        for j in range(ny):
            n = int(rng.integers(lo, hi))
            out += [n] * (4 if (i, j) in refine else 1)
    return np.array(out, np.int32)


def refined_case(zones, seed, ns=False, ic=None, skew=0.3, lo=1, hi=9):
    mesh = C.refined_channel(5, 4, 3.0, 1.0, zones, REFINE, skew=skew)
    rng = np.random.default_rng(seed)
    N = cell_orders(5, 4, set(REFINE), rng, lo, hi)
    q = mesh.quads.copy()
    for e in range(q.shape[0]):
        q[e] = np.roll(q[e], int(rng.integers(0, 4)))
    mesh.quads = q
    q0 = C.freestream_state(0.3, 0.2)
    return C.Case("refined", mesh, N, 0.3, 100.0 if ns else 0.0, 0.72, ns, riemann="roe",
                  viscous="br1" if ns else "none", ic=ic or C.uniform(q0))


def gauss01(N):
    x, w = np.polynomial.legendre.leggauss(N + 1)
    return 0.5 + 0.5 * x, 0.5 * w


def test_the_mesh_has_subdivided_edges_in_every_orientation():
    from oracle.oracle import Oracle
    q = C.freestream_state(0.3, 0.2)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    seen = set()
    for seed in range(4):
        orc = Oracle(refined_case(zones, seed))
        for k in range(orc.E):
            sub = orc.edge_subdivided(k)
            if sub is not None:
                seen.add((sub["dir_north"], sub["dir_south"]))
    assert seen == {(1, 1), (1, -1), (-1, 1), (-1, -1)}, seen


def test_mortar_matrices():
    from oracle.oracle import Oracle
    q = C.freestream_state(0.3, 0.2)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    n = 0
    for seed in range(3):
        orc = Oracle(refined_case(zones, seed))
        for k in range(orc.E):
            s = orc.edge_subdivided(k)
            if s is None:
                continue
            n += 1
            NL = s["N_left"]
            xl, _ = gauss01(NL)
            xn, _ = gauss01(s["N_north"])
            xs, _ = gauss01(s["N_south"])
            for d in range(NL + 1):   # the big element's polynomials on each half
                np.testing.assert_allclose(s["TLN_F"] @ xl ** d, (0.5 + 0.5 * xn) ** d, atol=1e-12)
                np.testing.assert_allclose(s["TLS_F"] @ xl ** d, (0.5 * xs) ** d, atol=1e-12)
            np.testing.assert_allclose(s["TLN_B"] @ s["TLN_F"] + s["TLS_B"] @ s["TLS_F"], 2.0 * np.eye(NL + 1),
                                       atol=1e-12)
            for side in ("N", "S"):
                F, B = s[f"TR{side}_F"], s[f"TR{side}_B"]
                np.testing.assert_allclose(B @ F, np.eye(F.shape[1]), atol=1e-12)
    assert n >= 20


@pytest.mark.parametrize("viscous", ["none", "br1", "ip"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_free_stream_is_preserved(viscous, seed):
    from oracle.oracle import Oracle
    q = C.freestream_state(0.3, 0.2)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    case = refined_case(zones, seed, viscous != "none")
    if viscous == "ip":
        case.viscous, case.ip_method, case.sigma0 = "ip", "sipg", 10.0
    orc = Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    # the penalty sigma0 mu N^2 / h (over 100 on the split cells) multiplies the rounding of the jumps
    assert np.max(np.abs(orc.time_derivative())) <= (1e-10 if viscous == "ip" else 1e-11)


def test_closed_box_conserves_mass_and_energy():
    from oracle.oracle import Oracle

    def ic(x, y):
        return C._pack(1.0 + 0.1 * np.sin(2 * x + y), 0.2 * np.cos(x - 2 * y), 0.1 * np.sin(x * y),
                       1.0 + 0.05 * np.cos(3 * x))
    zones = [C.Zone(C.BC_EULERWALL) for _ in range(4)]
    for seed in range(2):
        case = refined_case(zones, seed, ic=ic, hi=8)
        orc = Oracle(case)
        orc.set_state(case.initial_state(orc.coords()))
        qd = orc.time_derivative().reshape(-1, 4)
        _, jac = orc.metrics()
        w = []
        for n in case.N:
            _, wq = gauss01(int(n))
            w.append(np.outer(wq, wq).T.reshape(-1))
        rate = (np.concatenate(w) * jac) @ qd
        scale = np.max(np.abs(qd), axis=0)
        assert abs(rate[0]) <= 1e-13 * scale[0] and abs(rate[3]) <= 1e-13 * scale[3], rate


def test_unsupported_forms_are_refused():
    from oracle.oracle import Oracle
    q = C.freestream_state(0.3, 0.2)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    c = refined_case(zones, 0, ns=True, lo=3, hi=4)   # one order: no p-non-conforming edges
    c.artificial_dissipation = "physical"
    with pytest.raises(RuntimeError, match="subdivided edges: no artificial dissipation"):
        Oracle(c)
    c = refined_case(zones, 0)
    c.N = c.N.copy()
    # the four parts of split cell (1, 1), orders by row: its left edge meets two of them
    order = []
    for i in range(5):
        for j in range(4):
            order += [(i, j)] * (4 if (i, j) in set(REFINE) else 1)
    e = order.index((1, 1))
    c.N[e:e + 4] = [3, 4, 3, 4]
    with pytest.raises(RuntimeError, match="two small elements of one order"):
        Oracle(c)
