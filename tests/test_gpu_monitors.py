"""Monitor integrals on the device (QuadMesh_Integrals.incf:22-65 volume, :67-346 surface)
against the CPU oracle, through the C ABI.

Bar: the FP64 form within 1e-12 of the oracle, relative to the largest entry of the same
kind (the pressure-, density- and flow-type sums share one scale; the two viscous-force
entries another, at least mu times the zone's length, since a zone at rest carries only
rounding noise there); the mixed form within 1e-6 of the FP64 form on the same scales (its
traces are FP32, its sums FP64).
"""
import numpy as np
import pytest

from horses2d_amd import cases as C

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _annulus(viscous):
    c = C.ns_cylinder_case(4, 16, N=3, Re=200.0)
    c.viscous = viscous
    return c


CASES = {
    "euler_vortex": (lambda: C.euler_vortex_case(8, 4), (1, 2, 3, 4)),
    "ns_br1_annulus": (lambda: _annulus("br1"), (1, 2)),
    "ns_sipg_annulus": (lambda: _annulus("ip"), (1, 2)),
    "ns_br1_tgv": (lambda: C.ns_tgv_case(6, 4, Re=1600.0), (1, 3)),
    "ns_ip_kh": (lambda: C.ns_kh_case(6, 5, Re=1.0e4), (1, 2, 3, 4)),
    # artificial dissipation: the viscous force keeps the physical viscosity
    "ns_br1_ad_annulus": (lambda: _parity_case("ns_br1_ad_annulus"), (1, 2)),
}


def _parity_case(name):
    from test_gpu_parity import CASES as P
    return P[name]()


def _close(got, want, tol, what, mu):
    inv, vis = np.abs(want[:7]).max(), max(np.abs(want[7:]).max(), mu * want[0])
    np.testing.assert_allclose(got[:7], want[:7], rtol=0, atol=tol * inv, err_msg=what)
    np.testing.assert_allclose(got[7:], want[7:], rtol=0, atol=tol * max(vis, 1e-300), err_msg=what)


@pytest.mark.parametrize("steps", [0, 3])
@pytest.mark.parametrize("name", list(CASES))
def test_monitor_integrals_match_oracle(name, steps):
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    make, markers = CASES[name]
    case = make()
    mu = 1.0 / case.Re if case.Re else 0.0
    dg = gpu.DGSEM(case, "fp64")
    orc = Oracle(case)
    Q0 = case.initial_state(dg.coords())
    dg.upload(Q0)
    orc.set_state(Q0)
    dt = dg.estimate_dt(case.dt_max, case.cfl)
    assert dt == pytest.approx(orc.estimate_dt(case.dt_max, case.cfl), rel=1e-14)
    for it in range(steps):
        orc.step(it * dt, dt, case.scheme)
        dg.rk_step(it * dt, dt)
    vg, vo = dg.volume_integrals(), orc.volume_integrals()
    assert vg[0] == pytest.approx(vo[0], rel=TOL)
    assert vg[1] == pytest.approx(vo[1], rel=1e-10, abs=TOL * vo[0])
    for mk in markers:
        _close(dg.surface_integrals(mk), orc.surface_integrals(mk), TOL, f"{name} marker {mk}", mu)
    if case.navier_stokes:
        # the monitors leave the step untouched: continuing gives the oracle's state
        orc.step(steps * dt, dt, case.scheme)
        dg.rk_step(steps * dt, dt)
        st_o = orc.state()
        assert np.max(np.abs(dg.state() - st_o)) <= TOL * np.max(np.abs(st_o))


@pytest.mark.parametrize("name", ["ns_br1_annulus", "ns_sipg_annulus", "ns_ip_kh"])
def test_monitor_integrals_mixed_form(name):
    from horses2d_amd import gpu
    make, markers = CASES[name]
    case = make()
    mu = 1.0 / case.Re
    out = []
    for precision in ("fp64", "mixed"):
        dg = gpu.DGSEM(case, precision)
        dg.upload(case.initial_state(dg.coords()))
        dt = dg.estimate_dt(case.dt_max, case.cfl)
        for it in range(3):
            dg.rk_step(it * dt, dt)
        out.append((dg.volume_integrals(), [dg.surface_integrals(mk) for mk in markers]))
    (v64, s64), (v32, s32) = out
    np.testing.assert_allclose(v32, v64, rtol=1e-6, atol=1e-6 * v64[0])
    for mk, a, b in zip(markers, s32, s64):
        _close(a, b, 1e-6, f"{name} marker {mk}", mu)


def test_monitor_marker_out_of_range_raises():
    from horses2d_amd import gpu
    case = C.euler_vortex_case(4, 3)
    dg = gpu.DGSEM(case, "fp64")
    dg.upload(case.initial_state(dg.coords()))
    for bad in (0, 5):
        with pytest.raises(RuntimeError):
            dg.surface_integrals(bad)


VARS = ("rho", "rhou", "rhov", "rhoe", "rhot", "rhout", "rhovt", "rhoet", "u", "v", "p", "Mach")


def _probe_set(case, n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    elem = rng.integers(0, case.mesh.K, size=n)
    xi, eta = rng.uniform(0.0, 1.0, size=n), rng.uniform(0.0, 1.0, size=n)
    var = [VARS[k % len(VARS)] for k in range(n)]
    return elem, xi, eta, var


QDOT_VARS = ("rhot", "rhout", "rhovt", "rhoet")


def _probe_close(got, want, var, tol, qdot_scale=None):
    """Per variable, relative to the largest magnitude of that variable among the probes; the
    QDot variables relative to the max-norm of the whole QDot field (the scale of the QDot
    parity bar: these low-Mach QDot values are 1e-2 of the fluxes they come from)."""
    var = np.asarray(var)
    for v in set(var):
        m = var == v
        scale = qdot_scale if v in QDOT_VARS else np.abs(want[m]).max()
        np.testing.assert_allclose(got[m], want[m], rtol=0, atol=tol * scale, err_msg=v)


@pytest.mark.parametrize("name", ["euler_vortex", "ns_br1_annulus", "ns_ip_kh"])
def test_probes_match_oracle(name):
    """Probe_Update on the device: every variable, the QDot ones after hg_time_derivative; after
    steps the state variables (QDot probes then need a new time derivative)."""
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    case = CASES[name][0]()
    dg = gpu.DGSEM(case, "fp64")
    orc = Oracle(case)
    Q0 = case.initial_state(dg.coords())
    dg.upload(Q0)
    orc.set_state(Q0)
    elem, xi, eta, var = _probe_set(case, 96, 3)
    dg.probes_set(elem, xi, eta, var)
    dg.time_derivative(0.0)
    qd = orc.time_derivative(0.0)
    _probe_close(dg.probes_update(), orc.probes(elem, xi, eta, var), var, TOL, np.abs(qd).max())
    dt = dg.estimate_dt(case.dt_max, case.cfl)
    for it in range(3):
        orc.step(it * dt, dt, case.scheme)
        dg.rk_step(it * dt, dt)
    with pytest.raises(RuntimeError):
        dg.probes_update()
    state_vars = [v for v in var if v not in QDOT_VARS]
    keep = np.array([v in state_vars for v in var])
    # a smaller set replaces the first
    dg.probes_set(elem[keep], xi[keep], eta[keep], state_vars)
    _probe_close(dg.probes_update(), orc.probes(elem[keep], xi[keep], eta[keep], state_vars), state_vars, TOL)


def test_probes_mixed_form_and_errors():
    from horses2d_amd import gpu
    case = CASES["ns_ip_kh"][0]()
    elem, xi, eta, var = _probe_set(case, 40, 5)
    out = []
    for precision in ("fp64", "mixed"):
        dg = gpu.DGSEM(case, precision)
        dg.upload(case.initial_state(dg.coords()))
        dt = dg.estimate_dt(case.dt_max, case.cfl)
        for it in range(3):
            dg.rk_step(it * dt, dt)
        qd = dg.time_derivative(0.0)
        dg.probes_set(elem, xi, eta, var)
        out.append((dg.probes_update(), np.abs(qd).max()))
    (p64, qs), (p32, _) = out
    qv = np.array([v in QDOT_VARS for v in var])
    _probe_close(p32[~qv], p64[~qv], [v for v, q in zip(var, qv) if not q], 1e-6)
    # the mixed form's QDot carries the FP32 rounding of O(1) fluxes times the N^2/h lift
    # (about 1e-4 absolute); its accuracy bar is on the stepped state, above
    _probe_close(p32[qv], p64[qv], [v for v, q in zip(var, qv) if q], 1e-3, qs)
    with pytest.raises(RuntimeError):
        dg.probes_set([case.mesh.K], [0.5], [0.5], ["rho"])
    with pytest.raises(RuntimeError):
        dg.probes_set([0], [0.5], [0.5], [12])
    dg.probes_set([], [], [], [])
    assert dg.probes_update().shape == (0,)
