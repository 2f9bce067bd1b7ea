"""Runs the oracle through a whole reference benchmark case from the committed fixtures and
records where it ends (final iteration, time, residuals, lift and drag) in
tests/golden/oracle_benchmark_runs.json, for the GPU tests to compare against where the oracle's
run is too long for a test (CylinderMach2: 16079 RK5 iterations, about half an hour on 8 cores).

    python tests/golden/oracle_benchmark_runs.py cylinder_mach2
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))


def main(name):
    from oracle.oracle import Oracle
    from test_benchmark_goldens import apply_boundary_functions, benchmark_case, lift_drag
    case, _ = benchmark_case(name)
    orc = Oracle(case)
    apply_boundary_functions(orc, case)
    orc.set_state(case.initial_state(orc.coords()))
    t0 = time.time()
    t, dts, it, hist = orc.integrate(case.steps, dt_max=case.dt_max, t_end=case.t_end, scheme=case.scheme,
                                     mode=case.mode, residual_target=case.residual_target)
    out = {"final_iteration": int(it), "t": float(t), "residuals": [float(v) for v in hist[-1]],
           "first_dt": float(dts[0]), "last_dt": float(dts[-1]), "seconds": time.time() - t0}
    if len(hist) >= 20:   # an early point of the history, for a short test to start the run from
        out["residuals_at_20"] = [float(v) for v in hist[19]]
    zones = [z.kind for z in case.mesh.zones]
    if 3 in zones:   # lift and drag of the last Euler-wall zone (the body)
        marker = len(zones) - zones[::-1].index(3)
        cl, cd = lift_drag(orc.surface_integrals(marker), case.Mach)
        out.update({"marker": marker, "cl": float(cl), "cd": float(cd)})
    path = os.path.join(HERE, "oracle_benchmark_runs.json")
    data = json.load(open(path)) if os.path.exists(path) else {}
    data[name] = out
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print(name, out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "cylinder_mach2")
