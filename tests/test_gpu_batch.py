"""Batched form: many cases of different mesh size and polynomial order advanced by one launch
per RK stage over all their element tiles (HIP-graph replay) must give, bit for bit, the state
and residuals of advancing each case with its own handle."""
import numpy as np
import pytest

from horses2d_amd import cases as C

pytestmark = pytest.mark.gpu


def make(cases, precision):
    from horses2d_amd import gpu
    out = []
    for c in cases:
        d = gpu.DGSEM(c, precision)
        d.upload(c.initial_state(d.coords()))
        out.append(d)
    return out


def run_pair(cases, precision, steps, **bopt):
    from horses2d_amd import gpu
    single = make(cases, precision)
    batched = make(cases, precision)
    dts = np.array([d.estimate_dt(c.dt_max, c.cfl) for d, c in zip(single, cases)])
    for d, c in zip(batched, cases):
        d.estimate_dt(c.dt_max, c.cfl)
    for it in range(steps):
        for d, dt in zip(single, dts):
            d.rk_step(it * dt, dt)
    b = gpu.Batch(batched, **bopt)
    for it in range(steps):
        b.rk_step(it * dts, dts)
    b.synchronize()
    return single, batched, b


def assert_same(single, batched):
    for s, m in zip(single, batched):
        np.testing.assert_array_equal(m.state(), s.state())
        np.testing.assert_array_equal(m.residuals(), s.residuals())


def small_ensemble(count=6):
    return C.ensemble_cases(count, seed_base=77, nmin=3, nmax=9, Nmin=1, Nmax=9)


@pytest.mark.parametrize("use_graph", [True, False])
def test_euler_ensemble_fp64_bitwise(use_graph):
    cases = small_ensemble()
    single, batched, b = run_pair(cases, "fp64", 4, use_graph=use_graph)
    assert_same(single, batched)
    tile, order = b.schedule()
    assert tile in (0, 1) and order in (0, 1, 2, 3, 4)


@pytest.mark.parametrize("tile,order", [(0, 0), (1, 1), (0, 2), (1, 0), (0, 3), (1, 3), (0, 4), (1, 4)])
def test_every_schedule_gives_the_same_bits(tile, order):
    cases = small_ensemble(4)
    single, batched, b = run_pair(cases, "mixed", 3, tile=tile, order=order)
    assert_same(single, batched)
    assert b.schedule() == (tile, order)


def test_navier_stokes_batch_and_case_parity():
    """NS (BR1) cases, one of which already took a step before the batch was formed (its trace
    buffers are at the other parity); mixed Riemann solvers across cases."""
    from horses2d_amd import gpu
    cases = [C.ns_tgv_case(4, 3), C.ns_tgv_case(5, 5, riemann="roe"), C.ns_cylinder_case(3, 12, N=4)]
    single = make(cases, "fp64")
    batched = make(cases, "fp64")
    dts = np.array([d.estimate_dt(c.dt_max, c.cfl) for d, c in zip(single, cases)])
    single[1].rk_step(0.0, dts[1])
    batched[1].rk_step(0.0, dts[1])
    for it in range(3):
        for d, dt in zip(single, dts):
            d.rk_step(it * dt, dt)
    b = gpu.Batch(batched)
    for it in range(3):
        b.rk_step(it * dts, dts)
    b.synchronize()
    assert_same(single, batched)
    # the handles stay usable on their own after the batch
    for s, m, dt in zip(single, batched, dts):
        s.rk_step(0.0, dt)
        m.rk_step(0.0, dt)
    assert_same(single, batched)


def test_ip_batch_mixed():
    cases = [C.ns_kh_case(4, 4), C.ns_kh_case(3, 6)]
    single, batched, _ = run_pair(cases, "mixed", 3)
    assert_same(single, batched)


@pytest.mark.parametrize("viscous", ["br1", "sipg"])
def test_navier_stokes_one_launch_per_order(viscous):
    """One gradient launch and one stage launch per polynomial order (BR1: the gradient
    kernels per order too; interior penalty: the stage epilogue forms the viscous traces)."""
    if viscous == "br1":
        cases = [C.ns_tgv_case(4, 3), C.ns_tgv_case(5, 5, riemann="roe"), C.ns_cylinder_case(3, 12, N=4),
                 C.ns_tgv_case(3, 7)]
    else:
        cases = [C.ns_kh_case(4, 4), C.ns_kh_case(3, 6), C.ns_kh_case(5, 4)]
    for precision in ("fp64", "mixed"):
        single, batched, b = run_pair(cases, precision, 3, order=3)
        assert_same(single, batched)
        assert b.schedule()[1] == 3


@pytest.mark.parametrize("viscous", ["br1", "sipg"])
def test_navier_stokes_persistent(viscous):
    """The persistent form: resident workgroups looping over each order's tiles, the gradient
    pass (BR1) persistent too; bit for bit the single-handle steps, both tiles."""
    if viscous == "br1":
        cases = [C.ns_tgv_case(4, 3), C.ns_tgv_case(5, 5, riemann="roe"), C.ns_cylinder_case(3, 12, N=4),
                 C.ns_tgv_case(3, 7)]
    else:
        cases = [C.ns_kh_case(4, 4), C.ns_kh_case(3, 6), C.ns_kh_case(5, 4)]
    for precision, tile in (("fp64", 1), ("mixed", 0)):
        single, batched, b = run_pair(cases, precision, 3, tile=tile, order=4)
        assert_same(single, batched)
        assert b.schedule() == (tile, 4)


def test_boundary_condition_kinds_in_a_batch():
    """Euler walls, pressure inlet/outlet, Dirichlet and newDirichlet zones (prescribed and Riemann boundary
    fluxes) in one batch, NS BR1: bit for bit the single-handle steps."""
    from test_gpu_parity import _channel, _euler_wall_annulus, _newdirichlet
    cases = [_channel(C.BC_DIRICHLET, C.BC_PRESSUREOUTLET, 0.3, 1.5, "br1", 1),
             _channel(C.BC_PRESSUREINLET, C.BC_PRESSUREOUTLET, 0.3, 1.5, "br1"),
             _euler_wall_annulus(True),
             _newdirichlet(0.3, 1.5, "br1", mode=2, outp=2)]
    single, batched, _ = run_pair(cases, "fp64", 3)
    assert_same(single, batched)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_artificial_dissipation_in_a_batch(precision):
    """Jumps-based "Physical" dissipation in a batch: each case's indicator launch before the
    batched gradient pass; BR1 cases of different orders and boundaries, bit for bit the
    single-handle steps.  A batch mixing cases with and without it is refused."""
    from test_gpu_parity import CASES as P
    from horses2d_amd import gpu
    cases = [P["ns_br1_ad_tgv"](), P["ns_br1_ad_channel"](), P["ns_br1_ad_annulus"]()]
    single, batched, _ = run_pair(cases, precision, 3)
    assert_same(single, batched)
    plain = [P["ns_br1_ad_tgv"](), C.ns_tgv_case(6, 4, Re=1600.0)]
    handles = [gpu.DGSEM(c, precision) for c in plain]
    with pytest.raises(RuntimeError, match="artificial dissipation"):
        gpu.Batch(handles)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_steps_after_direct_steps_and_uploads(precision):
    """A case stepped on its own, or given a new state, between batch steps: the next batch step
    starts from that case's current trace parity, viscous traces (interior penalty) and, in the
    FP32 form, its new reference pressure -- bit for bit the single-handle sequence."""
    from horses2d_amd import gpu
    cases = [C.ns_kh_case(4, 4), C.ns_kh_case(3, 5)]
    single = make(cases, precision)
    batched = make(cases, precision)
    dts = np.array([d.estimate_dt(c.dt_max, c.cfl) for d, c in zip(single, cases)])
    for d, c in zip(batched, cases):
        d.estimate_dt(c.dt_max, c.cfl)
    b = gpu.Batch(batched)

    def batch_step():
        b.rk_step(0.0, dts)
        b.synchronize()
        for d, dt in zip(single, dts):
            d.rk_step(0.0, dt)

    batch_step()
    batch_step()
    # case 0 steps alone (its trace parity flips)
    single[0].rk_step(0.0, dts[0])
    batched[0].rk_step(0.0, dts[0])
    batch_step()
    assert_same(single, batched)
    # case 1 gets a new state with a higher pressure (another FP32 reference pressure)
    q = single[1].state().reshape(-1, 4).copy()
    q[:, 3] *= 1.5
    single[1].upload(q.reshape(-1))
    batched[1].upload(q.reshape(-1))
    batch_step()
    batch_step()
    assert_same(single, batched)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_laplacian_dissipation_in_a_batch(precision):
    """The Laplacian dissipation (interior-penalty cases of different orders and boundaries,
    the Ramp benchmark's Re = 0 form among them) in a batch: bit for bit the single-handle
    steps."""
    from test_gpu_parity import CASES as P
    cases = [P["ns_iipg_lap_kh"](), P["ns_iipg_lap_newdirichlet_re0"]()]
    single, batched, _ = run_pair(cases, precision, 3)
    assert_same(single, batched)
