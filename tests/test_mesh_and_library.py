"""CPU tests of the product's host side: the library loads and exports the C ABI, and its
O(E) mesh builder reproduces the oracle's brute-force restatement of the reference
numbering (MeshFile.f90:287-618, QuadElement_EdgeProcedures.incf:30-424,
QuadMesh_Zones.incf:22-131) bit-exactly, on generated meshes and on the reference's own
HiOMesh files (fixtures under tests/golden)."""
import os
import re

import numpy as np
import pytest

from horses2d_amd import cases as C
from horses2d_amd import gpu
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    text = open(os.path.join(ROOT, "include", "horses2d_gpu.h")).read()
    return sorted(set(re.findall(r"\b(hg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(gpu_lib):
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(gpu_lib, s), s
    assert sorted(gpu.EXPORTS) == syms


def test_product_refuses_to_run_without_the_library(tmp_path, monkeypatch):
    monkeypatch.setattr(gpu, "_lib", None)
    with pytest.raises(RuntimeError):
        gpu.load_library(str(tmp_path / "missing.so"))


def reference_mesh_case(name, N=3):
    d = np.load(os.path.join(GOLDEN, f"mesh_{name}.npz"), allow_pickle=False)
    nm = len(d["marker_names"])
    q = C.freestream_state(0.3, 0.1)
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(nm)]
    mesh = C.HiOMesh(d["points"], d["quads"], d["bedges"], d["bmarker"], zones,
                     d["cedges"] if "cedges" in d else None, d["cxy"] if "cxy" in d else None)
    return C.Case(f"ref_{name}", mesh, np.full(mesh.K, N, np.int32), 0.3, ic=C.uniform(q))


def periodic_reference_vortex_mesh():
    d = np.load(os.path.join(GOLDEN, "mesh_vortex_transport_16x16.npz"), allow_pickle=False)
    # TestCases/Cases/VortexTransport/SETUP/VortexTransport.HiOCase:58-84 zones
    zones = [C.Zone(C.BC_PERIODIC, direction=2, partner=2), C.Zone(C.BC_PERIODIC, direction=2, partner=1),
             C.Zone(C.BC_PERIODIC, direction=1, partner=4), C.Zone(C.BC_PERIODIC, direction=1, partner=3)]
    mesh = C.HiOMesh(d["points"] / 0.1, d["quads"], d["bedges"], d["bmarker"], zones)   # Reynolds length 0.1
    return C.Case("ref_vortex_transport", mesh, np.full(mesh.K, 4, np.int32), 0.5,
                  ic=C.isentropic_vortex(0.5, 1.0, 0.5, 0.5))


def mixed_order_channel():
    """A skewed channel of random orders with rotated element node lists (p-non-conforming
    edges: the elements' mappings evaluate each interior edge at their own nodes)."""
    q = C.freestream_state(0.3, 0.1)
    mesh = C.channel(5, 4, 3.0, 1.0, [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)], skew=0.3)
    rng = np.random.default_rng(8)
    for e in range(mesh.K):
        mesh.quads[e] = np.roll(mesh.quads[e], int(rng.integers(0, 4)))
    return C.Case("mixed_orders", mesh, rng.integers(1, 9, size=mesh.K).astype(np.int32), 0.3, ic=C.uniform(q))


def reference_cylinder_re100():
    """The reference's CylinderRe100 mesh with its four p-refinement zones (orders 2, 4, 6, 8)."""
    d = np.load(os.path.join(GOLDEN, "mesh_cylinder_re100.npz"), allow_pickle=False)
    case = reference_mesh_case("cylinder_re100")
    orders = {1: 2, 2: 8, 3: 6, 4: 4}   # CylinderRe100/SETUP/Cylinder.HiOCase p-Refinement block
    case.N = np.array([orders[int(z)] for z in d["pref_zones"]], np.int32)
    return case


CASES = {
    "periodic_square": lambda: C.euler_vortex_case(7, 3),
    "mixed_orders_channel": mixed_order_channel,
    "ref_cylinder_re100": reference_cylinder_re100,
    "tgv": lambda: C.ns_tgv_case(5, 4),
    "annulus": lambda: C.ns_cylinder_case(3, 10, N=4),
    "ref_vortex_transport": periodic_reference_vortex_mesh,
    "ref_cylinder_very_coarse": lambda: reference_mesh_case("cylinder_very_coarse"),
    "ref_kelvin_helmholtz": lambda: reference_mesh_case("kelvin_helmholtz"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_connectivity_and_geometry_match_the_oracle(name, gpu_lib):
    case = CASES[name]()
    orc = O.Oracle(case)
    e1, el1, p1 = orc.connectivity()
    e2, el2, p2 = gpu.build_connectivity(case)
    np.testing.assert_array_equal(e1, e2)
    np.testing.assert_array_equal(el1, el2)
    np.testing.assert_array_equal(p1, p2)
    X1 = orc.coords(); Ja1, jac1 = orc.metrics()
    X2, Ja2, jac2 = gpu.build_geometry(case)
    np.testing.assert_array_equal(X1, X2)
    np.testing.assert_array_equal(Ja1, Ja2)
    np.testing.assert_array_equal(jac1, jac2)


@pytest.mark.parametrize("name", list(CASES))
def test_mesh_invariants(name, gpu_lib):
    """CheckMesh invariants (Generic/Checks.f90:109-306): positive Jacobian, each element
    owns four distinct edges, LEFT is the lower element id, every interior edge is seen
    from both sides."""
    case = CASES[name]()
    edges, elems, partner = gpu.build_connectivity(case)
    _, _, jac = gpu.build_geometry(case)
    assert np.all(jac > 0)
    for e in range(case.mesh.K):
        assert len(set(elems[e, :4])) == 4
    interior = edges[:, 3] >= 0
    assert np.all(edges[interior, 2] < edges[interior, 3])
    counts = np.bincount(elems[:, :4].ravel(), minlength=edges.shape[0])
    np.testing.assert_array_equal(counts, np.where(interior, 2, 1))
    per = partner >= 0
    np.testing.assert_array_equal(partner[partner[per]], np.nonzero(per)[0])


def test_reference_vortex_mesh_has_the_generated_topology(gpu_lib):
    """The generator follows the reference mesh convention: the shipped 16x16 periodic mesh and
    the generated one have identical connectivity tables."""
    ref = periodic_reference_vortex_mesh()
    gen = C.euler_vortex_case(16, 4)
    a = gpu.build_connectivity(ref)
    b = gpu.build_connectivity(gen)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_free_stream_on_reference_meshes():
    """Uniform far-field state on the reference's straight-sided KH mesh: QDot vanishes.
    The shipped cylinder mesh's curved-edge end points sit ~1e-6 away from its corner nodes
    (the file's own data), so there the residual is only as small as that gap allows."""
    case = reference_mesh_case("kelvin_helmholtz", N=5)
    orc = O.Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    assert np.max(np.abs(orc.time_derivative())) < 1e-11
    case = reference_mesh_case("cylinder_very_coarse", N=5)
    orc = O.Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    assert np.max(np.abs(orc.time_derivative())) < 1e-3


def test_boundary_condition_kinds_accepted_and_refused(gpu_lib):
    """hg_create's mesh description takes the Euler-wall, Dirichlet, pressure and newDirichlet
    kinds (Generic/Defines.h:103-112 numbering) and refuses the Riemann kind (8), which the
    library does not provide; the channel's connectivity is the oracle's."""
    q = tuple(C.freestream_state(0.3))
    def channel(kind):
        zones = [C.Zone(C.BC_EULERWALL), C.Zone(kind, q=q), C.Zone(C.BC_EULERWALL), C.Zone(C.BC_PRESSUREINLET, q=q)]
        mesh = C.channel(5, 3, 2.0, 1.0, zones, skew=0.2)
        return C.Case("channel", mesh, np.full(mesh.K, 3, np.int32), 0.3, ic=C.uniform(q))
    for kind in (C.BC_DIRICHLET, C.BC_PRESSUREOUTLET, C.BC_FARFIELD, C.BC_NEWDIRICHLET):
        case = channel(kind)
        e1, el1, p1 = O.Oracle(case).connectivity()
        e2, el2, p2 = gpu.build_connectivity(case)
        np.testing.assert_array_equal(e1, e2)
        np.testing.assert_array_equal(el1, el2)
    for kind in (8, 10):
        with pytest.raises(RuntimeError, match="unsupported boundary condition"):
            gpu.build_connectivity(channel(kind))


def test_subdivided_edges_both_builders_agree():
    """Hanging nodes (one element's edge meets two elements, MeshFile.f90:364-490): both builders
    merge the edge (numbered after the remaining edges, in the order of the hanging nodes), link
    its three elements (big LEFT, NORTH, SOUTH) and move the hanging node onto the big edge: the
    product's O(E) builder reproduces the oracle's connectivity and geometry bit for bit, on a
    three-element mesh, skewed refined channels with rotated node lists, and the reference's
    VortexTransport mesh (34 subdivided edges, 10 curved)."""
    pts = np.array([[0, 0], [1, 0], [2, 0], [2, 1], [2, 2], [1, 2], [0, 2], [1, 1]], np.float64)
    quads = np.array([[0, 1, 5, 6], [1, 2, 3, 7], [7, 3, 4, 5]], np.int32)
    bedges = np.array([[0, 1], [1, 2], [2, 3], [3, 4], [4, 5], [5, 6], [6, 0]], np.int32)
    q = C.freestream_state(0.3)
    mesh = C.HiOMesh(pts, quads, bedges, np.ones(7, np.int32), [C.Zone(C.BC_FARFIELD, q=tuple(q))])
    case = C.Case("hanging", mesh, np.full(3, 3, np.int32), 0.3, ic=C.uniform(q))
    orc = O.Oracle(case)
    subs = [s for s in (orc.edge_subdivided(k) for k in range(orc.E)) if s is not None]
    assert len(subs) == 1 and sorted(subs[0]["elements"]) == [0, 1, 2] and subs[0]["elements"][0] == 0
    from test_oracle_subdivided import refined_case
    from test_benchmark_goldens import benchmark_case
    zones = [C.Zone(C.BC_FARFIELD, q=tuple(q)) for _ in range(4)]
    cases = [case] + [refined_case(zones, seed) for seed in range(4)] + [benchmark_case("vortex_transport")[0]]
    for c in cases:
        orc = O.Oracle(c)
        for a, b in zip(orc.connectivity(), gpu.build_connectivity(c)):
            np.testing.assert_array_equal(a, b)
        X, Ja, jac = gpu.build_geometry(c)
        Ja_o, jac_o = orc.metrics()
        np.testing.assert_array_equal(X, orc.coords())
        np.testing.assert_array_equal(Ja, Ja_o)
        np.testing.assert_array_equal(jac, jac_o)
