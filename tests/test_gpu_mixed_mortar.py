"""The mixed-precision form on meshes with several polynomial orders (p-non-conforming edges,
QuadMesh_Construct.incf:104-118) and with subdivided edges (hanging nodes), against the FP64
form, which matches the oracle within 1e-12 on the same cases (test_gpu_mortar.py,
test_gpu_subdivided.py).

The sub-handles of one mesh share the reference pressure of the whole mesh's state, so the
mortar kernels read both sides' FP32 traces (energy less Eref) and add one Eref back before the
Riemann and viscous fluxes; BR1's mortar uStar stays in the shifted energy that the gradient lifts
against.  Errors are relative L2 norms per conserved quantity (mass, momentum, energy) after 20
RK steps, as in test_gpu_mixed.py."""
import numpy as np
import pytest

from test_gpu_mixed import rel_l2, run

pytestmark = pytest.mark.gpu


def _mortar(name):
    from test_gpu_mortar import CASES as M
    return M[name]()


def _sub(name):
    from test_gpu_subdivided import CASES as S
    return S[name]()


CASES = {
    "orders_euler": lambda: _mortar("euler_seed1"),
    "orders_br1": lambda: _mortar("br1_hll"),
    "orders_sipg": lambda: _mortar("sipg_hll"),
    "orders_br1_annulus": lambda: _mortar("br1_annulus"),
    "orders_sipg_curved_annulus": lambda: _mortar("sipg_curved_annulus"),
    "orders_euler_curved_annulus": lambda: _mortar("euler_curved_annulus"),
    "subdivided_euler": lambda: _sub("euler_orders"),
    "subdivided_br1": lambda: _sub("br1_hll"),
    "subdivided_sipg": lambda: _sub("sipg_hll"),
    "subdivided_vortex_transport": lambda: _sub("vortex_transport"),
    "subdivided_vortex_transport_sipg": lambda: _sub("vortex_transport_sipg"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_mixed_form_tracks_the_fp64_form(name):
    case = CASES[name]()
    steps = 20
    a = run(case, "mixed", steps)
    b = run(case, "fp64", steps)
    err = rel_l2(a.state(), b.state())
    print(name, "rel L2 after", steps, "steps:", err)
    assert np.all(err <= 1e-6), (name, err)
    # (the residuals, max-norms of QDot, are printed, not held to a bar: a small one such as the
    # vortex's mass, 0.004 beside momentum 0.35, is a cancellation of O(1) fluxes whose FP32
    # rounding the N^2/h lift amplifies, 2.2e-4 absolute there)
    print(name, "residuals mixed", a.residuals(), "fp64", b.residuals())


def test_sub_handles_share_one_reference_pressure():
    """A state whose pressure differs between the orders' elements: the mixed form still tracks
    the FP64 one (with a reference pressure per order the two sides of a mortar edge would be
    shifted differently and the mortar fluxes would be off by O(1))."""
    from horses2d_amd import gpu
    case = _mortar("euler_seed1")
    f64, mix = gpu.DGSEM(case, "fp64"), gpu.DGSEM(case, "mixed")
    Q = case.initial_state(f64.coords()).reshape(-1, 4).copy()
    x = f64.coords().reshape(-1, 2)
    Q[:, 3] *= 1.0 + 0.2 * np.tanh(x[:, 0] - x[:, 0].mean())
    for s in (f64, mix):
        s.upload(Q.reshape(-1))
    err = rel_l2(mix.time_derivative(0.0), f64.time_derivative(0.0))
    print("QDot rel L2", err)
    assert np.all(err <= 1e-5), err


def test_curved_br1_mortars_are_refused_in_the_mixed_form():
    from horses2d_amd import gpu
    with pytest.raises(RuntimeError, match="curved mortar edges under BR1"):
        gpu.DGSEM(_mortar("br1_curved_annulus"), "mixed")
    with pytest.raises(RuntimeError, match="curved mortar edges under BR1"):
        gpu.DGSEM(_sub("vortex_transport_br1"), "mixed")
