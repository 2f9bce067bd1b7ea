"""Extracts mesh fixtures from the reference's own test cases into small .npz files.

Reads HiOMesh (NetCDF classic) files shipped under TestCases/ of the reference tree with
scipy (no reference code runs) and stores the arrays the HiOMesh reader consumes
(Solver/src/Mesh/MeshFile.f90:99-232): points, points_of_quads, points_of_bdryedges,
bdrymarker_of_edges, marker names and, when present, the curvilinear edges.
Ids are converted to 0-based.  Run only where the reference tree exists:

    python tests/golden/make_golden.py /root/reference
"""
import os
import sys

import numpy as np
from scipy.io import netcdf_file

MESHES = {
    "vortex_transport_16x16": "TestCases/Cases/VortexTransport/MESH/VortexTransport.HiOMesh",
    "cylinder_very_coarse": "TestCases/Cases/Cylinder/MESH/CylinderVeryCoarse.HiOMesh",
    "kelvin_helmholtz": "TestCases/Cases/KelvinHelmholtz/MESH/KelvinHelmholtz.HiOMesh",
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
        names.append(v[f"marker{m}"][:].tobytes().decode().strip())
    out["marker_names"] = np.array(names)
    if "curvilinear_edges" in v:
        out["cedges"] = np.array(v["curvilinear_edges"][:], dtype=np.int32) - 1
        x = np.array(v["x_curvilinear_edges"][:], dtype=np.float64)
        y = np.array(v["y_curvilinear_edges"][:], dtype=np.float64)
        out["cxy"] = np.stack([x, y], axis=-1)
    nc.close()
    return out


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for name, rel in MESHES.items():
        data = extract(os.path.join(ref, rel))
        np.savez_compressed(os.path.join(here, f"mesh_{name}.npz"), **data)
        print(name, {k: getattr(v, "shape", None) for k, v in data.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
