"""FP64 parity against the CPU oracle at every kernel instantiation the benchmark
configurations run, on small meshes, and the mixed form against the oracle directly.

Each (order, equations, viscous method, Riemann solver, tile) is its own compiled kernel; the
configurations run these ones:

* Euler, Roe, N = 3..8 (the isentropic vortex and the 64-case ensemble, sides 32..128);
* Navier-Stokes BR1 + Lax-Friedrichs at N = 6 (Taylor-Green vortex, 128 x 128);
* Navier-Stokes SIPG + Roe at N = 7 (Kelvin-Helmholtz shear layer, 256 x 256);
* Navier-Stokes BR1 + Roe at N = 7 on curved (quadratic) elements with a far field and an
  adiabatic wall (cylinder O-grid, 512 x 512).

Bars: FP64 QDot and the state after 10 RK3 steps within 1e-12 of the oracle (max-norm relative to
the field's max-norm, both workgroup tiles; DGSpatialDiscretizationMethods.f90:94-235,
DGTimeIntegrator.f90:282-317) -- for QDot, or within the oracle's own rounding floor where QDot
is a small difference of large terms (the Taylor-Green vortex: a one-ulp change of the state moves
the oracle's QDot by 3e-12 of its max-norm); the mixed form's state after the same 10 steps within
1e-6 (relative L2 of mass, momentum vector and energy) of the FP64 oracle.
"""
import numpy as np
import pytest

from horses2d_amd import cases as C

pytestmark = pytest.mark.gpu

TOL = 1e-12
MIXED_TOL = 1e-6

CASES = {f"euler_roe_N{N}": (lambda N=N: C.euler_vortex_case(5, N)) for N in range(3, 9)}
CASES.update({
    "ns_br1_lf_N6_tgv": lambda: C.ns_tgv_case(4, 6, Re=1600.0),
    "ns_sipg_roe_N7_kh": lambda: C.ns_kh_case(4, 7, Re=1.0e4),
    "ns_br1_roe_N7_curved_cylinder": lambda: C.ns_cylinder_case(4, 16, N=7, Re=200.0),
})


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rel_l2_conserved(a, b):
    a = a.reshape(-1, 4); b = b.reshape(-1, 4)
    d2 = np.sum((a - b) ** 2, axis=0); n2 = np.sum(b ** 2, axis=0)
    return np.sqrt(np.array([d2[0] / n2[0], (d2[1] + d2[2]) / (n2[1] + n2[2]), d2[3] / n2[3]]))


def setup(case, precision="fp64", tile=None):
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    dg = gpu.DGSEM(case, precision, apply_ic=False)
    if tile is not None:
        dg.set_tiles(tile, tile)
    orc = Oracle(case)
    X = dg.coords()
    np.testing.assert_array_equal(X, orc.coords())
    q0 = case.initial_state(X)
    dg.upload(q0)
    orc.set_state(q0)
    return dg, orc


def step_both(dg, orc, case, steps):
    dt = orc.estimate_dt(case.dt_max, case.cfl)
    assert dg.estimate_dt(case.dt_max, case.cfl) == pytest.approx(dt, rel=1e-14)
    for it in range(steps):
        orc.step(it * dt, dt, case.scheme)
        dg.rk_step(it * dt, dt)


def rounding_floor(orc, q0, qd):
    """How far the oracle's own QDot moves when each state entry is perturbed by one ulp: where
    QDot is a small difference of large flux terms (the Taylor-Green vortex is nearly a steady
    Euler flow), this exceeds 1e-12 of its max-norm, and two correct evaluations with different
    rounding (fused multiply-adds, 1/jac) differ by as much."""
    rng = np.random.Generator(np.random.PCG64(7))
    worst = 0.0
    for _ in range(2):
        orc.set_state(q0 * (1.0 + rng.choice([-1.0, 1.0], q0.size) * np.finfo(np.float64).eps))
        worst = max(worst, rel(orc.time_derivative(0.0), qd))
    orc.set_state(q0)
    return worst


@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("name", list(CASES))
def test_config_kernels_time_derivative(name, tile):
    """QDot within 1e-12 of the oracle, or within the oracle's own one-ulp rounding floor where
    that is larger."""
    case = CASES[name]()
    dg, orc = setup(case, tile=tile)
    q0 = orc.state()
    qd = orc.time_derivative(0.0)
    e = rel(dg.time_derivative(0.0), qd)
    bar = max(TOL, rounding_floor(orc, q0, qd))
    print(name, "tile", tile, "QDot rel", e, "bar", bar)
    assert e <= bar, (name, tile, e, bar)
    assert dg.tiles()[0] == tile


@pytest.mark.parametrize("tile", [0, 1])
@pytest.mark.parametrize("name", list(CASES))
def test_config_kernels_rk3_steps(name, tile):
    case = CASES[name]()
    dg, orc = setup(case, tile=tile)
    step_both(dg, orc, case, 10)
    e = rel(dg.state(), orc.state())
    print(name, "tile", tile, "state rel after 10 steps", e)
    assert e <= TOL, (name, tile, e)
    res_o = orc.residuals()
    np.testing.assert_allclose(dg.residuals(), res_o, rtol=1e-10, atol=1e-10 * float(np.max(res_o)))


@pytest.mark.parametrize("name", list(CASES))
def test_config_mixed_form_against_the_oracle(name):
    """Mixed form (FP64 state and RK update, FP32 fluxes and contractions) after 10 RK3 steps
    against the FP64 CPU oracle on the same inputs and time steps."""
    case = CASES[name]()
    dg, orc = setup(case, "mixed")
    step_both(dg, orc, case, 10)
    a, b = dg.state(), orc.state()
    err = rel_l2_conserved(a, b)
    linf = np.max(np.abs(a.reshape(-1, 4) - b.reshape(-1, 4)), axis=0) / np.max(np.abs(b.reshape(-1, 4)), axis=0)
    print(name, "mixed vs oracle rel L2 (mass, momentum, energy)", err, "per-field rel Linf", linf)
    assert np.all(err <= MIXED_TOL), (name, err)
    assert np.all(linf <= 10 * MIXED_TOL), (name, linf)
