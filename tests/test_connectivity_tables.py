"""Literal connectivity tables, derived by hand from the reference's mesh construction, for two
small meshes; the oracle's brute-force builder and the product's O(E) builder must both give
them exactly.

Rules followed by hand (Mesh/MeshFile.f90:287-362 computeFaces, :498-565 computeElementOfEdges,
:754-783 searchSimpleEdgeInElements; Mesh/QuadElement_EdgeProcedures.incf:30-199,357-424
Edge_LinkWithElements / BdryEdge_LinkWithElements; Mesh/QuadElement_Auxiliars.incf:27-116
searchEdge / compareEdges; Mesh/QuadMesh_Zones.incf:22-181 Zone_LinkPeriodicZones /
checkIfAreInverted):

* edges are numbered in first-appearance order over (element, face 1..4), face f of an element
  being its corner nodes (f, f+1) (face 4: (4, 1)); an edge keeps the node order it was first
  added with;
* LEFT is the lower element id holding both nodes, RIGHT the other; a face runs FORWARD when
  its nodes are in counter-clockwise order on faces 1-2 (bottom, right) and in clockwise order
  on faces 3-4 (top, left); the LEFT (or only) element always links FORWARD, the edge's nodes
  being swapped where its own order runs BACKWARD for that element; the RIGHT element's
  direction decides ``inverse``;
* periodic zones pair their edges by equal coordinates across the translation, in edge order,
  and a pair is inverted when the two (linked) edges' tangents point opposite ways.

Row layout (the C ABI's hg_get_connectivity): edges[k] = (node0, node1, LEFT element, RIGHT
element or -1, LEFT face, RIGHT face or 0, inverse, marker or 0); elems[e] = (edge of faces
1..4, direction of faces 1..4 (1 FORWARD, -1 BACKWARD), side of faces 1..4 (0 LEFT or
boundary, 1 RIGHT)); partner[k] = periodic partner edge or -1.  Ids are 0-based.
"""
import numpy as np
import pytest

from horses2d_amd import cases as C


def _case(points, quads, bedges, bmarker, zones, q, N=2):
    mesh = C.HiOMesh(np.asarray(points, float), np.asarray(quads, np.int32), np.asarray(bedges, np.int32),
                     np.asarray(bmarker, np.int32), zones)
    return C.Case("hand", mesh, np.full(len(quads), N, np.int32), 0.3, ic=C.uniform(q))


def mesh_2x2_rotated():
    """Unit squares on [0,2]^2, nodes y-fastest (node = 3 i + j), the fourth element's corners
    listed from its top-right corner, so that both of its interior edges link BACKWARD."""
    pts = [(i, j) for i in range(3) for j in range(3)]
    quads = [[0, 3, 4, 1], [1, 4, 5, 2], [3, 6, 7, 4], [8, 5, 4, 7]]
    bedges = [(0, 3), (3, 6), (6, 7), (7, 8), (8, 5), (5, 2), (2, 1), (1, 0)]
    bmarker = [1, 1, 2, 2, 3, 3, 4, 4]
    zones = [C.Zone(C.BC_EULERWALL) for _ in range(4)]
    return _case(pts, quads, bedges, bmarker, zones, C.freestream_state(0.0))     # at rest in a closed box


EDGES_2X2 = [
    (0, 3, 0, -1, 1, 0, 0, 1),
    (3, 4, 0, 2, 2, 4, 0, 0),
    (1, 4, 0, 1, 3, 1, 0, 0),   # added as (4, 1), swapped: the LEFT element's top runs 1 -> 4
    (0, 1, 0, -1, 4, 0, 0, 4),
    (4, 5, 1, 3, 2, 2, 1, 0),   # the rotated element's face 2 is (5, 4): BACKWARD
    (2, 5, 1, -1, 3, 0, 0, 3),
    (1, 2, 1, -1, 4, 0, 0, 4),
    (3, 6, 2, -1, 1, 0, 0, 1),
    (6, 7, 2, -1, 2, 0, 0, 2),
    (4, 7, 2, 3, 3, 3, 1, 0),   # top of both elements, same node order: BACKWARD for the RIGHT
    (8, 5, 3, -1, 1, 0, 0, 3),
    (8, 7, 3, -1, 4, 0, 0, 2),
]
ELEMS_2X2 = [
    (0, 1, 2, 3, 1, 1, 1, 1, 0, 0, 0, 0),
    (2, 4, 5, 6, 1, 1, 1, 1, 1, 0, 0, 0),
    (7, 8, 9, 1, 1, 1, 1, 1, 0, 0, 0, 1),
    (10, 4, 9, 11, 1, -1, -1, 1, 0, 1, 1, 0),
]
PARTNER_2X2 = [-1] * 12


def mesh_3x2_periodic_rotated():
    """Unit squares on [0,3]x[0,2], periodic in x and y (markers 1 bottom, 2 top, 3 left,
    4 right, as the periodic squares of the benchmark cases), the last element's corners listed
    from its bottom-right corner, so that its geometric top edge is its face 2."""
    pts = [(i, j) for i in range(4) for j in range(3)]
    quads = [[0, 3, 4, 1], [1, 4, 5, 2], [3, 6, 7, 4], [4, 7, 8, 5], [6, 9, 10, 7], [10, 11, 8, 7]]
    bedges = [(0, 3), (3, 6), (6, 9), (5, 2), (8, 5), (11, 8), (1, 0), (2, 1), (9, 10), (10, 11)]
    bmarker = [1, 1, 1, 2, 2, 2, 3, 3, 4, 4]
    zones = [C.Zone(C.BC_PERIODIC, direction=2, partner=2), C.Zone(C.BC_PERIODIC, direction=2, partner=1),
             C.Zone(C.BC_PERIODIC, direction=1, partner=4), C.Zone(C.BC_PERIODIC, direction=1, partner=3)]
    return _case(pts, quads, bedges, bmarker, zones, C.freestream_state(0.3, 0.4))


EDGES_3X2 = [
    (0, 3, 0, -1, 1, 0, 0, 1),
    (3, 4, 0, 2, 2, 4, 0, 0),
    (1, 4, 0, 1, 3, 1, 0, 0),
    (0, 1, 0, -1, 4, 0, 0, 3),
    (4, 5, 1, 3, 2, 4, 0, 0),
    (2, 5, 1, -1, 3, 0, 0, 2),
    (1, 2, 1, -1, 4, 0, 0, 3),
    (3, 6, 2, -1, 1, 0, 0, 1),
    (6, 7, 2, 4, 2, 4, 0, 0),
    (4, 7, 2, 3, 3, 1, 0, 0),
    (7, 8, 3, 5, 2, 3, 0, 0),   # the rotated element's face 3 is (8, 7), a top face: FORWARD
    (5, 8, 3, -1, 3, 0, 0, 2),
    (6, 9, 4, -1, 1, 0, 1, 1),  # periodic with edge 16, whose tangent points along -x
    (9, 10, 4, -1, 2, 0, 0, 4),
    (7, 10, 4, 5, 3, 4, 1, 0),  # the rotated element's face 4 is (7, 10): BACKWARD
    (10, 11, 5, -1, 1, 0, 0, 4),
    (11, 8, 5, -1, 2, 0, 1, 2),
]
ELEMS_3X2 = [
    (0, 1, 2, 3, 1, 1, 1, 1, 0, 0, 0, 0),
    (2, 4, 5, 6, 1, 1, 1, 1, 1, 0, 0, 0),
    (7, 8, 9, 1, 1, 1, 1, 1, 0, 0, 0, 1),
    (9, 10, 11, 4, 1, 1, 1, 1, 1, 0, 0, 1),
    (12, 13, 14, 8, 1, 1, 1, 1, 0, 0, 0, 1),
    (15, 16, 10, 14, 1, 1, 1, -1, 0, 0, 1, 1),
]
PARTNER_3X2 = [5, -1, -1, 13, -1, 0, 15, 11, -1, -1, -1, 7, 16, 3, -1, 6, 12]

TABLES = {
    "2x2_rotated": (mesh_2x2_rotated, EDGES_2X2, ELEMS_2X2, PARTNER_2X2),
    "3x2_periodic_rotated": (mesh_3x2_periodic_rotated, EDGES_3X2, ELEMS_3X2, PARTNER_3X2),
}


def _check(tables, edges, elems, partner):
    np.testing.assert_array_equal(edges, np.asarray(tables[1], np.int32))
    np.testing.assert_array_equal(elems, np.asarray(tables[2], np.int32))
    np.testing.assert_array_equal(partner, np.asarray(tables[3], np.int32))


@pytest.mark.parametrize("name", list(TABLES))
def test_oracle_builds_the_hand_derived_tables(name):
    from oracle.oracle import Oracle
    t = TABLES[name]
    _check(t, *Oracle(t[0]()).connectivity())


@pytest.mark.parametrize("name", list(TABLES))
def test_product_builder_gives_the_hand_derived_tables(name, gpu_lib):
    from horses2d_amd import gpu
    t = TABLES[name]
    _check(t, *gpu.build_connectivity(t[0]()))


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(TABLES))
def test_device_handle_holds_the_hand_derived_tables(name):
    """The same tables from a device handle (hg_create, hg_get_connectivity), and its residual
    on a uniform flow: the edge records the kernels use send a free stream through unchanged."""
    from horses2d_amd import gpu
    t = TABLES[name]
    case = t[0]()
    dg = gpu.DGSEM(case, "fp64")
    _check(t, *dg.connectivity())
    assert np.max(np.abs(dg.time_derivative(0.0))) < 1e-12
