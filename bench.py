"""Throughput benchmark of the DGSEM residual + Williamson RK3 step on one MI355X.

Workload (default): the Navier-Stokes Taylor-Green vortex on a periodic [0,2pi]^2 square,
128x128 quads, N=6 (P=7), Re=1600, Mach 0.1, Pr 0.72, BR1 viscous terms, Lax-Friedrichs
flux, RK3, seeded +-1e-3 velocity noise -- synthetic input of the stated shape.

A "step" is one RK3 time step (three stages) of the whole mesh.  The FP64 form and the
mixed form (FP64 state/RK, FP32 fluxes and contractions) are both timed over the same
steps from the same initial state; the mixed form's relative L2 error against the FP64
form after those steps is reported per field (nodal and quadrature-weighted) and per conserved
quantity (mass, momentum vector, energy).  ``value`` is the mixed form's rate when the
conserved-quantity errors are all <= 1e-6 and the FP64 rate otherwise.

Prints one JSON line.  With ``--gpus`` above 1 it exits: the library runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "DOF-stage updates/s on 1 MI355X: mixed/batched vs FP64 at L2 err ≤1e-6"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per DOF-stage of the stage kernel (one launch = one RK stage of the
# whole mesh): state Q read+write and RK register G read+write in FP64 (32 B), metric
# Ja (4) + jac (1) per node shared by 4 equations (10 B FP64 / 5 B FP32), and for
# Navier-Stokes the gradient dQ (2 values per DOF) read once (16 B FP64 / 8 B FP32).
STAGE_BYTES = {("euler", "fp64"): 42.0, ("euler", "mixed"): 37.0, ("ns", "fp64"): 58.0, ("ns", "mixed"): 45.0,
               # interior penalty: the stage kernel also writes the next stage's gradient share
               ("ns_ip", "fp64"): 74.0, ("ns_ip", "mixed"): 53.0}
# Whole stage (gradient + stage kernels), the per-DOF-stage figure of the path:
# adds the gradient write (16 B FP64 / 8 B FP32).
PATH_BYTES = {("euler", "fp64"): 42.0, ("euler", "mixed"): 37.0, ("ns", "fp64"): 74.0, ("ns", "mixed"): 53.0,
              ("ns_ip", "fp64"): 74.0, ("ns_ip", "mixed"): 53.0}


def make_case(name: str):
    from horses2d_amd import cases as C
    if name == "tgv":
        return C.ns_tgv_case(128, 6, steps=500), "ns"
    if name == "vortex":
        return C.euler_vortex_case(16, 4, steps=200), "euler"
    if name == "kh":
        return C.ns_kh_case(256, 7, steps=200), "ns_ip"
    if name == "cylinder":
        return C.ns_cylinder_case(512, 512, 7, steps=2000), "ns"
    raise SystemExit(f"unknown config {name}")


def describe(case, name):
    return {"workload": case.name, "config": name, "elements": int(case.mesh.K), "N": int(case.N[0]),
            "ndof": int(case.ndof), "Re": case.Re, "Mach": case.Mach, "riemann": case.riemann,
            "viscous": case.viscous if case.navier_stokes else "none", "scheme": case.scheme,
            "cfl": case.cfl}


def time_form(case, precision, steps, warmup, Q0, dt, kernel_steps=100, tiles=None, graph=True):
    """Wall time of ``steps`` RK steps (each step replayed as a captured HIP graph), then the
    per-kernel averages from HIP events around each launch over a second pass of direct launches
    from the same initial state (events cannot sit inside the replayed graph)."""
    from horses2d_amd import gpu
    dg = gpu.DGSEM(case, precision, apply_ic=False)
    if tiles is not None:
        dg.set_tiles(*tiles)
    dg.set_graph(graph)
    dg.upload(Q0)
    t = 0.0
    for _ in range(warmup):
        dg.rk_step(t, dt)
    dg.synchronize()
    dg.upload(Q0)                     # timed run starts from the initial state
    dg.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dg.rk_step(t, dt)
        t += dt
    dg.synchronize()
    wall = time.perf_counter() - t0
    Q = dg.state()
    res = dg.residuals()
    dg.upload(Q0)
    dg.profile(True)
    for _ in range(min(steps, kernel_steps)):
        dg.rk_step(0.0, dt)
    dg.synchronize()
    n_stage, ms_stage = dg.profile_read(0)
    n_grad, ms_grad = dg.profile_read(1)
    dg.close()
    return dict(wall=wall, n_stage=n_stage, ms_stage=ms_stage, n_grad=n_grad, ms_grad=ms_grad, Q=Q, res=res)


def rel_l2(a, b, w=None):
    """Per-field relative L2 error of a against b (nodal, or weighted by w per node)."""
    a = a.reshape(-1, 4); b = b.reshape(-1, 4)
    if w is None:
        w = np.ones(a.shape[0])
    num = np.sqrt(np.sum(w[:, None] * (a - b) ** 2, axis=0))
    den = np.sqrt(np.sum(w[:, None] * b ** 2, axis=0))
    return num / den


def rel_l2_conserved(a, b):
    """Relative L2 errors of mass, momentum (the vector (rho u, rho v)) and energy."""
    a = a.reshape(-1, 4); b = b.reshape(-1, 4)
    d2 = np.sum((a - b) ** 2, axis=0); n2 = np.sum(b ** 2, axis=0)
    return np.sqrt(np.array([d2[0] / n2[0], (d2[1] + d2[2]) / (n2[1] + n2[2]), d2[3] / n2[3]]))


def quadrature_weights(case):
    """w_i w_j J per node (the discrete L2 inner product of the DGSEM)."""
    from horses2d_amd import gpu
    _, _, jac = gpu.build_geometry(case)
    xg, wg = np.polynomial.legendre.leggauss(int(case.N[0]) + 1)
    w = 0.5 * wg
    return np.tile(np.outer(w, w).T.reshape(-1), case.mesh.K) * jac


def cpu_baseline(case, steps: int):
    """The oracle (a C restatement of the reference's CPU path) timed on the host cores."""
    from oracle.oracle import Oracle
    orc = Oracle(case)
    X = orc.coords()
    orc.set_state(case.initial_state(X))
    dt = orc.estimate_dt(case.dt_max, case.cfl)
    orc.step(0.0, dt, case.scheme)       # first touch
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.step(0.0, dt, case.scheme)
    wall = time.perf_counter() - t0
    stages = {"rk3": 3, "rk5": 5, "euler": 1}[case.scheme]
    return case.ndof * stages * steps / wall, wall


def ensemble_bench(steps: int, warmup: int, count: int = 64):
    """Batched ensemble: ``count`` isentropic-vortex Euler cases (sides 32..128, N 3..8, seeded
    vortices, own dt each).  Times the plain FP64 form (one handle and one stream per case),
    the batched FP64 form and the batched mixed form over the same steps from the same states."""
    from horses2d_amd import cases as C, gpu
    cases = C.ensemble_cases(count)
    ndof = sum(c.ndof for c in cases)
    stages = 3

    def build(precision):
        hs = [gpu.DGSEM(c, precision) for c in cases]
        q0 = [h.state() for h in hs]
        return hs, q0

    def reset(hs, q0):
        for h, q in zip(hs, q0):
            h.upload(q)
            h.reset_dt_state()

    hs, q0 = build("fp64")
    dts = np.array([h.estimate_dt(c.dt_max, c.cfl) for h, c in zip(hs, cases)])
    out = {"workload": f"euler_vortex_ensemble_{count}", "cases": count, "ndof": int(ndof),
           "N": [int(c.N[0]) for c in cases], "sides": [int(round(np.sqrt(c.mesh.K))) for c in cases]}

    def timed(step_fn, sync_fn):
        for _ in range(warmup):
            step_fn()
        sync_fn()
        t0 = time.perf_counter()
        for _ in range(steps):
            step_fn()
        sync_fn()
        return time.perf_counter() - t0

    # plain FP64 form: every case on its own handle and stream
    reset(hs, q0)
    wall = timed(lambda: [h.rk_step(0.0, dt) for h, dt in zip(hs, dts)], lambda: [h.synchronize() for h in hs])
    out["fp64_per_case"] = {"dof_stage_per_s": ndof * stages * steps / wall, "ms_per_step": 1e3 * wall / steps}
    # batched FP64 (warm-up steps advance the state, so reset and time again from q0 for the error)
    reset(hs, q0)
    b = gpu.Batch(hs)
    wall = timed(lambda: b.rk_step(0.0, dts), b.synchronize)
    out["batched_fp64"] = {"dof_stage_per_s": ndof * stages * steps / wall, "ms_per_step": 1e3 * wall / steps,
                           "schedule": list(b.schedule())}
    ref = [h.state() for h in hs]
    b.close()
    for h in hs:
        h.close()
    hm, q0m = build("mixed")
    for h, q in zip(hm, q0):
        h.upload(q)
    bm = gpu.Batch(hm)
    wall = timed(lambda: bm.rk_step(0.0, dts), bm.synchronize)
    errs = np.array([rel_l2_conserved(h.state(), r) for h, r in zip(hm, ref)])
    out["batched_mixed"] = {"dof_stage_per_s": ndof * stages * steps / wall, "ms_per_step": 1e3 * wall / steps,
                            "schedule": list(bm.schedule()),
                            "max_rel_l2_mass_momentum_energy_vs_fp64": [float(v) for v in errs.max(axis=0)],
                            "after_steps": steps + warmup}
    bm.close()
    for h in hm:
        h.close()
    base = out["fp64_per_case"]["dof_stage_per_s"]
    out["batched_fp64"]["speedup_vs_fp64_per_case"] = out["batched_fp64"]["dof_stage_per_s"] / base
    out["batched_mixed"]["speedup_vs_fp64_per_case"] = out["batched_mixed"]["dof_stage_per_s"] / base
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="tgv", choices=["tgv", "vortex", "kh", "cylinder"])
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble section")
    ap.add_argument("--tiles", default="auto", help="'auto' or 'S,G': fixed stage and gradient tiles (0/1)")
    ap.add_argument("--direct", action="store_true", help="launch kernel by kernel instead of graph replay")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run 5 device-to-device copies of 512 MiB (calibrates HBM byte counters)")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs on one GPU; multi-GPU runs are not provided", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")     # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    from horses2d_amd import gpu
    case, phys = make_case(args.config)
    stages = {"rk3": 3, "rk5": 5, "euler": 1}[case.scheme]
    probe = gpu.DGSEM(case, "fp64", apply_ic=False)
    X = probe.coords()
    Q0 = case.initial_state(X)
    probe.upload(Q0)
    dt = probe.estimate_dt(case.dt_max, case.cfl)
    probe.close()

    forms = [f for f in args.forms.split(",") if f]
    results = {}
    for f in forms:
        tiles = None if args.tiles == "auto" else tuple(int(v) for v in args.tiles.split(","))
        results[f] = time_form(case, f, args.steps, args.warmup, Q0, dt, tiles=tiles, graph=not args.direct)

    out_forms = {}
    for f, r in results.items():
        rate = case.ndof * stages * args.steps / r["wall"]
        avg_stage_ms = r["ms_stage"] / max(r["n_stage"], 1)
        # gradient kernel time per stage (interior penalty launches it only after an upload)
        avg_grad_ms = r["ms_grad"] / max(r["n_stage"], 1)
        sb = STAGE_BYTES[(phys, f)] * case.ndof
        pb = PATH_BYTES[(phys, f)] * case.ndof
        out_forms[f] = {
            "dof_stage_per_s": rate,
            "ms_per_step": 1e3 * r["wall"] / args.steps,
            "stage_kernel_avg_ms": avg_stage_ms,
            "gradient_kernel_avg_ms": avg_grad_ms,
            "gradient_kernel_launches_per_stage": r["n_grad"] / max(r["n_stage"], 1),
            "stage_kernel_GBps": sb / (avg_stage_ms * 1e-3) / 1e9,
            "path_GBps": pb / ((avg_stage_ms + avg_grad_ms) * 1e-3) / 1e9,
            "residuals": [float(v) for v in r["res"]],
        }
    err = None
    if "fp64" in results and "mixed" in results:
        from horses2d_amd.cases import SCHEME  # noqa: F401
        Qm, Qd = results["mixed"]["Q"], results["fp64"]["Q"]
        e_nodal = rel_l2(Qm, Qd)
        e_quad = rel_l2(Qm, Qd, quadrature_weights(case))
        e_cons = rel_l2_conserved(Qm, Qd)
        linf = np.max(np.abs(Qm.reshape(-1, 4) - Qd.reshape(-1, 4)), axis=0) / np.max(np.abs(Qd.reshape(-1, 4)), axis=0)
        err = {"after_steps": args.steps,
               "rel_l2_mass_momentum_energy": [float(v) for v in e_cons],
               "rel_l2_nodal": [float(v) for v in e_nodal], "rel_l2_quadrature": [float(v) for v in e_quad],
               "rel_linf": [float(v) for v in linf],
               "reference": "FP64 GPU form (matches the CPU oracle within 1e-12)"}
        out_forms["mixed"]["error_vs_fp64"] = err
        out_forms["mixed"]["speedup_vs_fp64"] = out_forms["mixed"]["dof_stage_per_s"] / out_forms["fp64"]["dof_stage_per_s"]

    if "mixed" in out_forms and err is not None and max(err["rel_l2_mass_momentum_energy"]) <= 1e-6:
        best = "mixed"
    else:
        best = "fp64" if "fp64" in out_forms else forms[0]
    bf = out_forms[best]
    achieved = bf["stage_kernel_GBps"]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.config}_{best}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("stage_kernel_bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": bf["dof_stage_per_s"],
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bf["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best == "fp64" else "f32+f64",
        "data": "synthetic",
        "form": best,
        "config": describe(case, args.config),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_stage", "algorithmic_bytes_per_launch": STAGE_BYTES[(phys, best)] * case.ndof,
                     "avg_launch_ms": bf["stage_kernel_avg_ms"]},
        "forms": out_forms,
    }
    if not args.no_ensemble:
        line["ensemble"] = ensemble_bench(args.ensemble_steps, 2)
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", str(min(16, os.cpu_count() or 1))))
        os.environ["OMP_NUM_THREADS"] = str(threads)
        # the oracle builds connectivity by brute force like the reference (O(E^2)); above ~70k
        # elements it times the same configuration on a coarser mesh (same N) and reports per DOF
        sample_case, how = case, "the same"
        if case.mesh.K > 70000:
            from horses2d_amd import cases as C
            sample_case, how = C.ns_cylinder_case(128, 128, int(case.N[0])), "a 128x128 version (same N) of the"
        rate, wall = cpu_baseline(sample_case, args.cpu_steps)
        line["cpu_baseline"] = {"value": rate, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
                                "sample": f"{args.cpu_steps} RK3 steps of {how} {case.name} mesh in FP64 "
                                          f"({wall:.1f} s), C oracle with OpenMP volume loops and the "
                                          f"reference's serial face loop"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
