"""Pins of the oracle's boundary states from problem-file functions ("Use function",
BoundaryConditionFunctionDrivers.incf:22-63: q = BoundaryState(edge % x(:, iXi), time) at every
node of the zone's edges, in place of the zone state).  What fixes them independent of the
restatement:

* the zone's own state handed over node by node changes no bit of QDot;
* a uniform flow stays uniform when the function gives that flow's state at every node;
* the nodes are the edges' own nodes (edge % x): Gauss points along each boundary edge;
* a state that varies along the zone changes QDot only in elements on that zone.
End to end, the reference's Couette golden (a Dirichlet inflow given by its BC1 function,
tests/test_benchmark_goldens.py) pins the path.
"""
import numpy as np
import pytest

from horses2d_amd import cases as C


def channel(kind=C.BC_DIRICHLET, ns=True):
    q = C.freestream_state(0.3)
    zones = [C.Zone(C.BC_EULERWALL), C.Zone(C.BC_PRESSUREOUTLET, q=tuple(q)), C.Zone(C.BC_EULERWALL),
             C.Zone(kind, q=tuple(q))]
    mesh = C.channel(6, 3, 3.0, 1.0, zones, skew=0.3)
    V = 0.3 * np.sqrt(1.4)

    def ic(x, y):
        return C._pack(1.0 + 0.05 * np.cos(x + 2 * y), V * (1.0 + 0.2 * np.sin(2.1 * y)), 0.1 * V * np.sin(x),
                       1.0 + 0.05 * np.sin(x - y))
    return C.Case("channel", mesh, np.full(mesh.K, 4, np.int32), 0.3, 100.0 if ns else 0.0, 0.72, ns,
                  riemann="roe", viscous="br1" if ns else "none", ic=ic)


@pytest.mark.parametrize("kind", [C.BC_DIRICHLET, C.BC_NEWDIRICHLET, C.BC_PRESSUREINLET, C.BC_FARFIELD])
def test_the_zone_state_per_node_changes_nothing(kind):
    from oracle.oracle import Oracle
    c = channel(kind)
    a, b = Oracle(c), Oracle(c)
    Q0 = c.initial_state(a.coords())
    a.set_state(Q0)
    b.set_state(Q0)
    n = b.boundary_nodes(4).shape[0]
    b.set_boundary_state(4, np.tile(np.array(c.mesh.zones[3].q), (n, 1)))
    np.testing.assert_array_equal(a.time_derivative(), b.time_derivative())


def test_uniform_flow_given_by_the_function_stays_uniform():
    from oracle.oracle import Oracle
    q = C.freestream_state(0.3, 0.2)
    zones = [C.Zone(C.BC_DIRICHLET, q=(1.0, 0.0, 0.0, 2.5)) for _ in range(4)]   # zone states wrong on purpose
    mesh = C.channel(5, 4, 3.0, 1.0, zones, skew=0.3)
    c = C.Case("fs", mesh, np.full(mesh.K, 5, np.int32), 0.3, 100.0, 0.72, True, riemann="roe", viscous="br1",
               ic=C.uniform(q))
    orc = Oracle(c)
    orc.set_state(c.initial_state(orc.coords()))
    for m in range(1, 5):
        n = orc.boundary_nodes(m).shape[0]
        orc.set_boundary_state(m, np.tile(q, (n, 1)))
    assert np.max(np.abs(orc.time_derivative())) <= 1e-12


def test_the_nodes_are_gauss_points_of_the_edges():
    from oracle.oracle import Oracle
    c = channel()
    orc = Oracle(c)
    x = orc.boundary_nodes(4)   # the inlet x = skew map of x = 0, 3 edges of N = 4
    assert x.shape == (15, 2)
    xi = 0.5 + 0.5 * np.polynomial.legendre.leggauss(5)[0]
    for e in range(3):
        seg = x[5 * e:5 * e + 5]
        a, b = seg[0] - (seg[-1] - seg[0]) * xi[0] / (xi[-1] - xi[0]), None
        d = (seg[-1] - seg[0]) / (xi[-1] - xi[0])
        np.testing.assert_allclose(seg, a + np.outer(xi, d), atol=1e-14)   # equally parametrised along a line


def test_a_varying_state_acts_on_the_zone_elements_only():
    from oracle.oracle import Oracle
    c = channel()
    a, b = Oracle(c), Oracle(c)
    Q0 = c.initial_state(a.coords())
    a.set_state(Q0)
    b.set_state(Q0)
    x = b.boundary_nodes(4)
    q = np.tile(np.array(c.mesh.zones[3].q), (x.shape[0], 1))
    q[:, 1] *= 1.0 + 0.1 * np.sin(3.0 * x[:, 1])
    b.set_boundary_state(4, q)
    d = np.abs(a.time_derivative() - b.time_derivative()).reshape(c.mesh.K, -1).max(axis=1)
    changed = set(np.nonzero(d > 0)[0])
    edges, _, _ = a.connectivity()
    inlet = {int(edges[k, 2]) for k in range(edges.shape[0]) if edges[k, 7] == 4}
    # BR1: the inlet elements and, through the gradient, their neighbours
    assert inlet <= changed and len(changed) < c.mesh.K
    with pytest.raises(RuntimeError, match="no boundary state"):
        b.set_boundary_state(1, np.ones((b.boundary_nodes(1).shape[0], 4)))
