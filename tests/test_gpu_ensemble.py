"""The batched ensemble at its benchmark size: 64 isentropic-vortex Euler cases on periodic
[-5,5]^2 squares, sides drawn from 32..128, N from 3..8, vortex strength U[3,6] and centre
U[-2,2]^2 per seed (86.5 M DOF in all), each with its own dt, advanced by the batched form
(one launch per RK stage over the element tiles of all cases, HIP-graph replay).

* FP64 and mixed batches, 20 RK3 steps each, from the same initial states and time steps.
* Per case, the FP64 conservation sums (mass, momentum, energy) hold to round-off: the
  meshes are periodic, so the discrete operator conserves them exactly.
* Per case, the mixed form is within 1e-6 (relative L2 of mass, momentum vector, energy) of the
  FP64 form.
* The three smallest cases against the CPU oracle: FP64 within 1e-12 (max-norm), mixed within
  1e-6 (relative L2).
"""
import numpy as np
import pytest

from horses2d_amd import cases as C

pytestmark = pytest.mark.gpu

STEPS = 20


def rel_l2_conserved(a, b):
    a = a.reshape(-1, 4); b = b.reshape(-1, 4)
    d2 = np.sum((a - b) ** 2, axis=0); n2 = np.sum(b ** 2, axis=0)
    return np.sqrt(np.array([d2[0] / n2[0], (d2[1] + d2[2]) / (n2[1] + n2[2]), d2[3] / n2[3]]))


def test_ensemble_64_cases_fp64_and_mixed():
    from horses2d_amd import gpu
    from oracle.oracle import Oracle
    cases = C.ensemble_cases(64)
    assert len(cases) == 64
    sides = [int(round(np.sqrt(c.mesh.K))) for c in cases]
    orders = [int(c.N[0]) for c in cases]
    assert min(sides) >= 32 and max(sides) <= 128 and set(orders) <= set(range(3, 9))
    print("ensemble: %d DOF, sides %d..%d, orders %s" % (sum(c.ndof for c in cases), min(sides), max(sides),
                                                         sorted(set(orders))), flush=True)

    hs = [gpu.DGSEM(c, "fp64") for c in cases]
    q0 = [h.state() for h in hs]
    s0 = np.array([h.conservation_sums() for h in hs])
    dts = np.array([h.estimate_dt(c.dt_max, c.cfl) for h, c in zip(hs, cases)])
    b = gpu.Batch(hs)
    for it in range(STEPS):
        b.rk_step(it * dts, dts)
    b.synchronize()
    q64 = [h.state() for h in hs]
    s1 = np.array([h.conservation_sums() for h in hs])
    assert all(np.all(np.isfinite(q)) for q in q64)
    drift = np.abs(s1 - s0) / (np.abs(s0) + 1.0)
    print("max relative drift of the conservation sums:", drift.max(axis=0), flush=True)
    assert np.all(drift <= 1e-12), drift.max(axis=0)
    b.close()
    for h in hs:
        h.close()

    hm = [gpu.DGSEM(c, "mixed", apply_ic=False) for c in cases]
    for h, q in zip(hm, q0):
        h.upload(q)
    bm = gpu.Batch(hm)
    for it in range(STEPS):
        bm.rk_step(it * dts, dts)
    bm.synchronize()
    qm = [h.state() for h in hm]
    bm.close()
    for h in hm:
        h.close()
    errs = np.array([rel_l2_conserved(a, r) for a, r in zip(qm, q64)])
    print("mixed vs FP64, max over cases (mass, momentum, energy):", errs.max(axis=0), flush=True)
    assert np.all(errs <= 1e-6), errs.max(axis=0)

    smallest = np.argsort([c.ndof for c in cases])[:3]
    for c in smallest:
        case = cases[c]
        orc = Oracle(case)
        orc.set_state(q0[c])
        assert orc.estimate_dt(case.dt_max, case.cfl) == pytest.approx(dts[c], rel=1e-14)
        for it in range(STEPS):
            orc.step(it * dts[c], dts[c], case.scheme)
        qo = orc.state()
        e64 = float(np.max(np.abs(q64[c] - qo)) / np.max(np.abs(qo)))
        em = rel_l2_conserved(qm[c], qo)
        print(case.name, "FP64 vs oracle", e64, "mixed vs oracle", em, flush=True)
        assert e64 <= 1e-12, (case.name, e64)
        assert np.all(em <= 1e-6), (case.name, em)
