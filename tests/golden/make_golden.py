"""Extracts fixtures from the reference's own test cases into small data files.

* Meshes: HiOMesh (NetCDF classic) files shipped under TestCases/ of the reference tree, read
  with scipy (no reference code runs), stored as the arrays the HiOMesh reader consumes
  (Solver/src/Mesh/MeshFile.f90:99-232): points, points_of_quads, points_of_bdryedges,
  bdrymarker_of_edges, marker names and, when present, the curvilinear edges and the
  p-refinement zones.  Ids are converted to 0-based.
* Benchmark cases (TestCases/Benchmark/Simulations/*): the parameters of each case file that the
  solver reads (through horses2d_amd.hio, this project's reader of the format), as JSON, and the
  expected values its problem file's Finalize checks (residuals, final iteration, lift and drag
  coefficients and the tolerances), parsed from the numbers in that file.

Run only where the reference tree exists:

    python tests/golden/make_golden.py /root/reference
"""
import json
import os
import re
import sys

import numpy as np
from scipy.io import netcdf_file

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

MESHES = {
    "vortex_transport_16x16": "TestCases/Cases/VortexTransport/MESH/VortexTransport.HiOMesh",
    "cylinder_very_coarse": "TestCases/Cases/Cylinder/MESH/CylinderVeryCoarse.HiOMesh",
    "kelvin_helmholtz": "TestCases/Cases/KelvinHelmholtz/MESH/KelvinHelmholtz.HiOMesh",
    "ramp": "TestCases/Benchmark/Simulations/Ramp/MESH/Ramp.HiOMesh",
    "cylinder_benchmark": "TestCases/Benchmark/Simulations/CylinderMach2/MESH/Cylinder.HiOMesh",
    "cylinder_re100": "TestCases/Benchmark/Simulations/CylinderRe100/MESH/Cylinder.HiOMesh",
    "vortex_transport_benchmark": "TestCases/Benchmark/Simulations/VortexTransport/MESH/VortexTransport.HiOMesh",
    "cylinder_re1000": "TestCases/Benchmark/Simulations/CylinderRe1000/MESH/CylinderInflation.HiOMesh",
    "couette": "TestCases/Benchmark/Simulations/Couette/MESH/Couette.HiOMesh",
}

# benchmark case -> (case file, mesh fixture, problem file)
BENCHMARKS = {
    "ramp": ("TestCases/Benchmark/Simulations/Ramp/SETUP/Ramp.HiOCase", "ramp",
             "TestCases/Benchmark/Simulations/Ramp/SETUP/ProblemFile.f90"),
    "cylinder_mach2": ("TestCases/Benchmark/Simulations/CylinderMach2/SETUP/CylinderSupersonic.HiOCase",
                       "cylinder_benchmark",
                       "TestCases/Benchmark/Simulations/CylinderMach2/SETUP/ProblemFile.f90"),
    "cylinder_re100": ("TestCases/Benchmark/Simulations/CylinderRe100/SETUP/Cylinder.HiOCase", "cylinder_re100",
                       "TestCases/Benchmark/Simulations/CylinderRe100/SETUP/ProblemFile.f90"),
    "vortex_transport": ("TestCases/Benchmark/Simulations/VortexTransport/SETUP/VortexTransport.HiOCase",
                         "vortex_transport_benchmark",
                         "TestCases/Benchmark/Simulations/VortexTransport/SETUP/ProblemFile.f90"),
    "cylinder_re1000": ("TestCases/Benchmark/Simulations/CylinderRe1000/SETUP/CylinderRe1000.HiOCase",
                        "cylinder_re1000",
                        "TestCases/Benchmark/Simulations/CylinderRe1000/SETUP/ProblemFile.f90"),
    "couette": ("TestCases/Benchmark/Simulations/Couette/SETUP/Couette.HiOCase", "couette",
                "TestCases/Benchmark/Simulations/Couette/SETUP/ProblemFile.f90"),
}


def extract(path):
    nc = netcdf_file(path, "r", mmap=False)
    v = nc.variables
    out = {
        "points": np.array(v["points"][:], dtype=np.float64),
        "quads": np.array(v["points_of_quads"][:], dtype=np.int32) - 1,
        "bedges": np.array(v["points_of_bdryedges"][:], dtype=np.int32) - 1,
        "bmarker": np.array(v["bdrymarker_of_edges"][:], dtype=np.int32),
    }
    names = []
    for m in range(1, int(nc.dimensions["no_of_markers"]) + 1):
        names.append(v[f"marker{m}"][:].tobytes().decode().strip(" \x00"))
    out["marker_names"] = np.array(names)
    if "curvilinear_edges" in v:
        out["cedges"] = np.array(v["curvilinear_edges"][:], dtype=np.int32) - 1
        x = np.array(v["x_curvilinear_edges"][:], dtype=np.float64)
        y = np.array(v["y_curvilinear_edges"][:], dtype=np.float64)
        out["cxy"] = np.stack([x, y], axis=-1)
    if "pRefinement_zones" in v:
        out["pref_zones"] = np.array(v["pRefinement_zones"][:], dtype=np.int32)
    nc.close()
    return out


def case_parameters(case_file, mesh_file):
    """The solver parameters and zone states of a case file, as this project's reader forms them."""
    from horses2d_amd import hio
    c = hio.case_from_hiocase(case_file, mesh_path=mesh_file, navier_stokes=True)
    return {
        "Mach": c.Mach, "Re": c.Re, "Pr": c.Pr, "N": sorted(set(int(n) for n in c.N)),
        # per element, when the p-refinement block gives the elements different orders
        "orders": [int(n) for n in c.N] if len(set(int(n) for n in c.N)) > 1 else None,
        "riemann": c.riemann, "viscous": c.viscous, "ip_method": c.ip_method, "sigma0": c.sigma0,
        "artificial_dissipation": c.artificial_dissipation, "ad_intensity": c.ad_intensity,
        "scheme": c.scheme, "mode": c.mode, "cfl": c.cfl, "dt_max": c.dt_max, "steps": c.steps,
        "residual_target": c.residual_target, "t_end": c.t_end, "t0": c.t0,
        "zones": [{"kind": z.kind, "q": list(z.q), "outflow": z.outflow, "mode": z.mode,
                   "outflow_pressure": z.outflow_pressure, "isothermal": z.isothermal, "Tw": z.Tw,
                   "wall_velocity": list(z.wall_velocity), "riemann": z.riemann,
                   "direction": z.direction, "partner": z.partner, "function": z.function} for z in c.mesh.zones],
        "p_ref": c.p_ref, "T_ref": c.T_ref,
    }


def finalize_values(problem_file):
    """Numbers the problem file's Finalize compares against: residuals(4), finalIteration, cl, cd,
    and the tolerances of its residual and coefficient checks."""
    text = open(problem_file).read()
    num = r"([-+0-9.]+E[-+][0-9]+)_RP"
    out = {}
    m = re.search(r"[Rr]esiduals\((?:4|NCONS)\)\s*&?\s*=\s*\[([^\]]*)\]", text)
    if m:
        out["residuals"] = [float(v) for v in re.findall(num, m.group(1))]
    m = re.search(r"(?:finalIteration|expectedIter)\s*=\s*(\d+)", text)
    if m:
        out["final_iteration"] = int(m.group(1))
    m = re.search(r"expectedErrors\(NCONS\)\s*=\s*&?\s*\[([^\]]*)\]", text)
    if m:   # errors against the analytical solution (VortexTransport)
        out["errors"] = [float(v) for v in re.findall(num, m.group(1))]
    for k in ("cl", "cd"):
        m = re.search(r"::\s*" + k + r"\s*=\s*" + num, text)
        if m:
            out[k] = float(m.group(1))
    tols = [float(v) for v in re.findall(r"\.gt\.\s*([0-9.]+e-?[0-9]+)_RP", text, re.I)]
    if tols:
        out["residual_tolerance"] = tols[0]
        if len(tols) > 1:
            out["coefficient_tolerance"] = tols[1:]
    return out


def main(ref):
    for name, rel in MESHES.items():
        data = extract(os.path.join(ref, rel))
        np.savez_compressed(os.path.join(HERE, f"mesh_{name}.npz"), **data)
        print(name, {k: getattr(v, "shape", None) for k, v in data.items()})
    for name, (case_rel, mesh_name, pf_rel) in BENCHMARKS.items():
        out = {"case": case_parameters(os.path.join(ref, case_rel), os.path.join(ref, MESHES[mesh_name])),
               "mesh": f"mesh_{mesh_name}.npz", "finalize": finalize_values(os.path.join(ref, pf_rel))}
        with open(os.path.join(HERE, f"benchmark_{name}.json"), "w") as f:
            json.dump(out, f, indent=1)
        print(name, out["finalize"])


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
