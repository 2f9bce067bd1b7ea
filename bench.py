"""Throughput benchmark of the DGSEM residual + Williamson RK3 step on one MI355X.

Default workload: the batched ensemble, the configuration the metric ("mixed/batched vs
FP64") is quoted on and the largest one-GPU configuration -- 64 isentropic-vortex Euler cases
on periodic [-5,5]^2 squares, sides drawn from 32..128, N from 3..8, vortex strength U[3,6]
and centre U[-2,2]^2 per seed, Roe flux, RK3, each case with its own dt (86.5 M DOF in all);
synthetic input of the stated shape.

A "step" is one RK3 time step (three stages) of every case.  From the same initial states and
time steps, over the same number of steps, this times:

* the plain FP64 form, one handle (and stream) per case, each step replayed as a HIP graph;
* the batched FP64 and the batched mixed form (FP64 state and RK update, FP32 fluxes,
  contractions and metric): one launch per RK stage over the element tiles of all cases.
  Each batched form runs twice: with HIP-graph replay of the step, and kernel by kernel with
  HIP events around every launch on the batch's stream.  The event-timed region gives both
  ``value`` and the stage kernel's average launch time of ``roofline``; the graph-replay
  rate is reported beside it.

``value`` is the batched mixed rate when its error against the batched FP64 form (relative
L2 of mass, momentum vector and energy, the maximum over cases) is <= 1e-6, else the batched
FP64 rate.  ``cpu_baseline`` is the oracle (the FP64 C restatement of the reference's CPU
path) timed on the host: one RK3 step of a stated sample of the cases, each case timed on its
own and the times summed, with every host core and with one; the same leg runs the oracle on
the three smallest cases for the timed steps and reports the error of both GPU forms against it.

``--config tgv|vortex|kh|cylinder`` times one single-mesh configuration instead (the FP64
and mixed forms of one handle).  Prints one JSON line.  With ``--gpus`` above 1 it exits:
the library runs on one GPU.
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
ERR_BAR = 1e-6

# Algorithmic HBM bytes per DOF-stage (state Q and RK register G read and written in FP64,
# 32 B; metric Ja (4) + 1/jac per node shared by 4 equations, 10 B FP64 / 5 B FP32; for
# Navier-Stokes the gradient, 2 values per DOF written once and read once, 32 B FP64 / 16 B
# FP32).  STAGE_BYTES is the stage kernel's share (the gradient write belongs to the gradient
# kernel for BR1; interior penalty forms it in the stage kernel).
PATH_BYTES = {("euler", "fp64"): 42.0, ("euler", "mixed"): 37.0, ("ns", "fp64"): 74.0, ("ns", "mixed"): 53.0,
              ("ns_ip", "fp64"): 74.0, ("ns_ip", "mixed"): 53.0}
STAGE_BYTES = {("euler", "fp64"): 42.0, ("euler", "mixed"): 37.0, ("ns", "fp64"): 58.0, ("ns", "mixed"): 45.0,
               ("ns_ip", "fp64"): 74.0, ("ns_ip", "mixed"): 53.0}


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
    _, wg = np.polynomial.legendre.leggauss(int(case.N[0]) + 1)
    w = 0.5 * wg
    return np.tile(np.outer(w, w).T.reshape(-1), case.mesh.K) * jac


def host_description():
    """(CPU model, threads for the CPU baseline): every CPU this process may run on, or
    OMP_NUM_THREADS where the host sets it to the share of CPUs meant for this job."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ncpu = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return model, min(ncpu, int(omp)) if omp.isdigit() and int(omp) > 0 else ncpu


def traffic_of(name):
    """Calibrated HBM bytes per launch of a kernel, from a committed counter summary."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    return d.get("bytes_per_launch", d.get("stage_kernel_bytes_per_launch"))


# ---------------------------------------------------------------------------------------------
# Batched ensemble
# ---------------------------------------------------------------------------------------------
def ensemble_forms(cases, steps, warmup):
    """Times the three forms; returns (results, initial states, dts, final states per form)."""
    from horses2d_amd import gpu
    ndof = sum(c.ndof for c in cases)
    work = ndof * 3 * steps
    hs = [gpu.DGSEM(c, "fp64") for c in cases]
    q0 = [h.state() for h in hs]
    dts = np.array([h.estimate_dt(c.dt_max, c.cfl) for h, c in zip(hs, cases)])

    def reset(handles):
        for h, q in zip(handles, q0):
            h.upload(q)

    out, states = {}, {}

    # plain FP64 form: every case on its own handle and stream, graph replay per step
    reset(hs)
    for _ in range(warmup):
        for h, dt in zip(hs, dts):
            h.rk_step(0.0, dt)
    for h in hs:
        h.synchronize()
    reset(hs)
    t0 = time.perf_counter()
    for _ in range(steps):
        for h, dt in zip(hs, dts):
            h.rk_step(0.0, dt)
    for h in hs:
        h.synchronize()
    wall = time.perf_counter() - t0
    out["fp64_per_case"] = {"dof_stage_per_s": work / wall, "ms_per_step": 1e3 * wall / steps}
    states["fp64_per_case"] = [h.state() for h in hs]

    def batched(handles, name):
        b = gpu.Batch(handles)
        reset(handles)
        for _ in range(warmup):             # includes the autotuner and the graph capture
            b.rk_step(0.0, dts)
        b.synchronize()
        # graph replay
        reset(handles)
        t0 = time.perf_counter()
        for _ in range(steps):
            b.rk_step(0.0, dts)
        b.synchronize()
        wall_graph = time.perf_counter() - t0
        q_graph = [h.state() for h in handles]
        # kernel by kernel, HIP events around every launch on the batch stream (the timed region)
        reset(handles)
        b.profile(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            b.rk_step(0.0, dts)
        b.synchronize()
        wall = time.perf_counter() - t0
        n_st, ms_st = b.profile_read(0)
        n_rs, ms_rs = b.profile_read(2)
        b.profile(False)
        q = [h.state() for h in handles]
        same = all(np.array_equal(x, y) for x, y in zip(q, q_graph))
        tile, order = b.schedule()
        b.close()
        out[name] = {"dof_stage_per_s": work / wall, "ms_per_step": 1e3 * wall / steps,
                     "graph_replay_dof_stage_per_s": work / wall_graph,
                     "graph_replay_ms_per_step": 1e3 * wall_graph / steps,
                     "graph_and_direct_bit_identical": bool(same),
                     "stage_kernel_launches": n_st, "stage_kernel_avg_ms": ms_st / max(n_st, 1),
                     "resid_kernel_avg_ms": ms_rs / max(n_rs, 1),
                     "schedule": {"tile": ["narrow", "wide"][tile],
                                  "order": ["case-major", "costliest first", "interleaved",
                                            "one launch per order", "persistent"][order]}}
        states[name] = q

    batched(hs, "batched_fp64")
    out["batched_fp64"]["bit_identical_to_per_case"] = bool(
        all(np.array_equal(x, y) for x, y in zip(states["batched_fp64"], states["fp64_per_case"])))
    for h in hs:
        h.close()
    hm = [gpu.DGSEM(c, "mixed", apply_ic=False) for c in cases]
    batched(hm, "batched_mixed")
    for h in hm:
        h.close()
    return out, q0, dts, states


def ensemble_errors(cases, q_mixed, q_fp64):
    """Mixed against FP64: the maximum over cases of the conserved-quantity errors, and the
    whole ensemble's per-field relative L2 (nodal and quadrature-weighted) and Linf."""
    per_case = np.array([rel_l2_conserved(a, b) for a, b in zip(q_mixed, q_fp64)])
    num_n = np.zeros(4); den_n = np.zeros(4); num_q = np.zeros(4); den_q = np.zeros(4)
    dmax = np.zeros(4); fmax = np.zeros(4)
    for c, a, b in zip(cases, q_mixed, q_fp64):
        a = a.reshape(-1, 4); b = b.reshape(-1, 4)
        w = quadrature_weights(c)
        num_n += np.sum((a - b) ** 2, axis=0); den_n += np.sum(b ** 2, axis=0)
        num_q += np.sum(w[:, None] * (a - b) ** 2, axis=0); den_q += np.sum(w[:, None] * b ** 2, axis=0)
        dmax = np.maximum(dmax, np.max(np.abs(a - b), axis=0)); fmax = np.maximum(fmax, np.max(np.abs(b), axis=0))
    return {"max_over_cases_rel_l2_mass_momentum_energy": [float(v) for v in per_case.max(axis=0)],
            "rel_l2_nodal": [float(v) for v in np.sqrt(num_n / den_n)],
            "rel_l2_quadrature": [float(v) for v in np.sqrt(num_q / den_q)],
            "rel_linf": [float(v) for v in dmax / fmax]}


def ensemble_cpu_baseline(cases, q0, dts, states, steps, stride):
    """The oracle on the host cores: one RK3 step of every ``stride``-th case, each timed on its
    own, the times summed (all cores, then one thread); then the oracle's state after ``steps``
    steps of the three smallest cases, against which both GPU forms are compared."""
    from oracle import oracle
    from oracle.oracle import Oracle
    model, ncores = host_description()
    sample = list(range(0, len(cases), stride))
    legs = {}
    for threads in (ncores, 1):
        oracle.set_threads(threads)
        work = 0.0
        total = 0.0
        for c in sample:
            case = cases[c]
            orc = Oracle(case)
            orc.set_state(q0[c])
            t0 = time.perf_counter()
            orc.step(0.0, dts[c], case.scheme)
            total += time.perf_counter() - t0
            work += case.ndof * 3
            del orc
        legs[threads] = (work / total, total)
    oracle.set_threads(ncores)
    rate, wall = legs[ncores]
    rate1, wall1 = legs[1]
    checks = {}
    for c in np.argsort([cs.ndof for cs in cases])[:3]:
        case = cases[c]
        orc = Oracle(case)
        orc.set_state(q0[c])
        for _ in range(steps):
            orc.step(0.0, dts[c], case.scheme)
        qo = orc.state()
        checks[case.name] = {
            "fp64_rel_max": float(np.max(np.abs(states["batched_fp64"][c] - qo)) / np.max(np.abs(qo))),
            "mixed_rel_l2_mass_momentum_energy": [float(v) for v in rel_l2_conserved(states["batched_mixed"][c], qo)],
            "mixed_rel_l2_per_field": [float(v) for v in rel_l2(states["batched_mixed"][c], qo)]}
    return {"value": rate, "unit": "DOF-stage updates/s", "cores": ncores, "kind": "port",
            "host_cpus_visible": len(os.sched_getaffinity(0)),
            "what": "restatement: the oracle, an FP64 C/OpenMP restatement of the reference's Fortran path "
                    "(serial face loop, OpenMP element loops)",
            "cpu_model": model,
            "sample": f"one RK3 step of each of {len(sample)} of the {len(cases)} cases (every {stride}th), "
                      f"each timed on its own, times summed: {wall:.2f} s with {ncores} threads",
            "one_thread": {"value": rate1, "cores": 1, "seconds": wall1},
            "oracle_check_after_timed_steps": checks}


def ensemble_line(args):
    from horses2d_amd import cases as C
    cases = C.ensemble_cases(64)
    ndof = sum(c.ndof for c in cases)
    forms, q0, dts, states = ensemble_forms(cases, args.steps, args.warmup)
    err = ensemble_errors(cases, states["batched_mixed"], states["batched_fp64"])
    err["after_steps"] = args.steps
    err["reference"] = "batched FP64 form (bit-identical to one handle per case; the oracle check is in cpu_baseline)"
    forms["batched_mixed"]["error_vs_fp64"] = err
    base = forms["fp64_per_case"]["dof_stage_per_s"]
    for k in ("batched_fp64", "batched_mixed"):
        forms[k]["speedup_vs_fp64_per_case"] = forms[k]["dof_stage_per_s"] / base
    forms["batched_mixed"]["speedup_vs_batched_fp64"] = (forms["batched_mixed"]["dof_stage_per_s"]
                                                          / forms["batched_fp64"]["dof_stage_per_s"])
    best = "batched_mixed" if max(err["max_over_cases_rel_l2_mass_momentum_energy"]) <= ERR_BAR else "batched_fp64"
    bf = forms[best]
    prec = "mixed" if best == "batched_mixed" else "fp64"
    bpl = PATH_BYTES[("euler", prec)] * ndof
    avg_ms = bf["stage_kernel_avg_ms"]
    achieved = bpl / (avg_ms * 1e-3) / 1e9
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
        "dtype": "f32+f64" if prec == "mixed" else "f64",
        "data": "synthetic",
        "form": best,
        "config": {"workload": "euler_vortex_ensemble_64", "cases": len(cases), "ndof": int(ndof),
                   "sides": [int(round(np.sqrt(c.mesh.K))) for c in cases], "N": [int(c.N[0]) for c in cases],
                   "riemann": "roe", "scheme": "rk3", "cfl": "0.4 (N = 3) / 0.1 (N >= 4)",
                   "seeds": "PCG64(1234 + case)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_of(f"traffic_ensemble_{prec}.json"),
                     "kernel": "k_batch_stage", "algorithmic_bytes_per_dof_stage": PATH_BYTES[("euler", prec)],
                     "algorithmic_bytes_per_launch": bpl, "avg_launch_ms": avg_ms,
                     "launches": bf["stage_kernel_launches"]},
        "forms": forms,
    }
    return line, cases, q0, dts, states


# ---------------------------------------------------------------------------------------------
# One mesh (--config, and the extra Taylor-Green line)
# ---------------------------------------------------------------------------------------------
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


def time_form(case, precision, steps, warmup, Q0, dt, tiles=None):
    """Kernel by kernel with HIP events around every launch (the timed region), then the same
    steps again with graph replay."""
    from horses2d_amd import gpu
    dg = gpu.DGSEM(case, precision, apply_ic=False)
    if tiles is not None:
        dg.set_tiles(*tiles)
    dg.upload(Q0)
    for _ in range(warmup):
        dg.rk_step(0.0, dt)
    dg.synchronize()
    dg.upload(Q0)
    dg.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        dg.rk_step(0.0, dt)
    dg.synchronize()
    wall = time.perf_counter() - t0
    n_stage, ms_stage = dg.profile_read(0)
    n_grad, ms_grad = dg.profile_read(1)
    dg.profile(False)
    Q = dg.state()
    res = dg.residuals()
    dg.upload(Q0)
    t0 = time.perf_counter()
    for _ in range(steps):
        dg.rk_step(0.0, dt)
    dg.synchronize()
    wall_graph = time.perf_counter() - t0
    same = bool(np.array_equal(dg.state(), Q))
    dg.close()
    return dict(wall=wall, wall_graph=wall_graph, same=same, n_stage=n_stage, ms_stage=ms_stage, n_grad=n_grad,
                ms_grad=ms_grad, Q=Q, res=res)


def single_forms(name, steps, warmup, tiles=None):
    from horses2d_amd import gpu
    case, phys = make_case(name)
    stages = {"rk3": 3, "rk5": 5, "euler": 1}[case.scheme]
    probe = gpu.DGSEM(case, "fp64", apply_ic=False)
    Q0 = case.initial_state(probe.coords())
    probe.upload(Q0)
    dt = probe.estimate_dt(case.dt_max, case.cfl)
    probe.close()
    results = {f: time_form(case, f, steps, warmup, Q0, dt, tiles) for f in ("fp64", "mixed")}
    out = {}
    for f, r in results.items():
        avg_stage = r["ms_stage"] / max(r["n_stage"], 1)
        avg_grad = r["ms_grad"] / max(r["n_stage"], 1)
        out[f] = {"dof_stage_per_s": case.ndof * stages * steps / r["wall"],
                  "ms_per_step": 1e3 * r["wall"] / steps,
                  "graph_replay_dof_stage_per_s": case.ndof * stages * steps / r["wall_graph"],
                  "graph_and_direct_bit_identical": r["same"],
                  "stage_kernel_avg_ms": avg_stage, "gradient_kernel_avg_ms": avg_grad,
                  "gradient_kernel_launches_per_stage": r["n_grad"] / max(r["n_stage"], 1),
                  "stage_kernel_GBps": STAGE_BYTES[(phys, f)] * case.ndof / (avg_stage * 1e-3) / 1e9,
                  "path_GBps": PATH_BYTES[(phys, f)] * case.ndof / ((avg_stage + avg_grad) * 1e-3) / 1e9,
                  "residuals": [float(v) for v in r["res"]]}
    Qm, Qd = results["mixed"]["Q"], results["fp64"]["Q"]
    out["mixed"]["error_vs_fp64"] = {
        "after_steps": steps,
        "rel_l2_mass_momentum_energy": [float(v) for v in rel_l2_conserved(Qm, Qd)],
        "rel_l2_nodal": [float(v) for v in rel_l2(Qm, Qd)],
        "rel_l2_quadrature": [float(v) for v in rel_l2(Qm, Qd, quadrature_weights(case))],
        "rel_linf": [float(v) for v in (np.max(np.abs(Qm.reshape(-1, 4) - Qd.reshape(-1, 4)), axis=0)
                                        / np.max(np.abs(Qd.reshape(-1, 4)), axis=0))]}
    out["mixed"]["speedup_vs_fp64"] = out["mixed"]["dof_stage_per_s"] / out["fp64"]["dof_stage_per_s"]
    return case, phys, stages, out, Q0, dt, Qd, Qm


def single_line(args):
    tiles = None if args.tiles == "auto" else tuple(int(v) for v in args.tiles.split(","))
    case, phys, stages, forms, Q0, dt, Qd, Qm = single_forms(args.config, args.steps, args.warmup, tiles)
    err = forms["mixed"]["error_vs_fp64"]
    best = "mixed" if max(err["rel_l2_mass_momentum_energy"]) <= ERR_BAR else "fp64"
    bf = forms[best]
    achieved = bf["stage_kernel_GBps"]
    line = {
        "metric": METRIC, "value": bf["dof_stage_per_s"], "unit": "DOF-stage updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": bf["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64" if best == "mixed" else "f64",
        "data": "synthetic", "form": best,
        "config": {"workload": case.name, "elements": int(case.mesh.K), "N": int(case.N[0]), "ndof": int(case.ndof),
                   "Re": case.Re, "Mach": case.Mach, "riemann": case.riemann,
                   "viscous": case.viscous if case.navier_stokes else "none", "scheme": case.scheme, "cfl": case.cfl},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(f"traffic_{args.config}_{best}.json"),
                     "kernel": "k_stage", "algorithmic_bytes_per_launch": STAGE_BYTES[(phys, best)] * case.ndof,
                     "avg_launch_ms": bf["stage_kernel_avg_ms"]},
        "forms": forms,
    }
    if not args.no_cpu_baseline:
        from oracle import oracle
        from oracle.oracle import Oracle
        model, ncores = host_description()
        sample = case
        if case.mesh.K > 70000:      # the oracle's connectivity is O(E^2), as the reference's
            from horses2d_amd import cases as C
            sample = C.ns_cylinder_case(128, 128, int(case.N[0]))
        orc = Oracle(sample)
        orc.set_state(sample.initial_state(orc.coords()))
        dts = orc.estimate_dt(sample.dt_max, sample.cfl)
        rates = {}
        for threads in (ncores, 1):
            oracle.set_threads(threads)
            n = args.cpu_steps if threads > 1 else max(1, args.cpu_steps // 8)
            t0 = time.perf_counter()
            for _ in range(n):
                orc.step(0.0, dts, sample.scheme)
            rates[threads] = sample.ndof * stages * n / (time.perf_counter() - t0)
        oracle.set_threads(ncores)
        line["cpu_baseline"] = {"value": rates[ncores], "unit": "DOF-stage updates/s", "cores": ncores,
                                "kind": "port", "cpu_model": model,
                                "what": "restatement: the oracle, an FP64 C/OpenMP restatement of the reference's path",
                                "sample": f"{args.cpu_steps} RK3 steps of {sample.name} in FP64",
                                "one_thread": {"value": rates[1], "cores": 1}}
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--config", default="ensemble", choices=["ensemble", "tgv", "vortex", "kh", "cylinder"])
    ap.add_argument("--cpu-stride", type=int, default=4, help="ensemble CPU baseline: every n-th case")
    ap.add_argument("--cpu-steps", type=int, default=20, help="single-mesh CPU baseline steps")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the Taylor-Green line next to the ensemble")
    ap.add_argument("--tiles", default="auto", help="single mesh: 'auto' or 'S,G' fixed stage and gradient tiles")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run 5 device-to-device copies of 512 MiB (calibrates HBM byte counters)")
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

    if args.config != "ensemble":
        print(json.dumps(single_line(args)))
        return
    line, cases, q0, dts, states = ensemble_line(args)
    if not args.no_extra:
        _, _, _, forms, _, _, _, _ = single_forms("tgv", max(args.steps, 20), args.warmup)
        line["extra"] = {"ns_tgv_128x128_N6": {k: {kk: vv for kk, vv in v.items() if kk != "residuals"}
                                              for k, v in forms.items()}}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = ensemble_cpu_baseline(cases, q0, dts, states, args.steps, args.cpu_stride)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
