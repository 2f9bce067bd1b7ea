"""Builds the in-tree HIP library ``horses2d_amd/libhorses2d_gpu.so`` for gfx950.

Every translation unit is compiled with hipcc, in parallel, then linked into one shared
object that exports the C ABI of ``include/horses2d_gpu.h``.  Packed-FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhorses2d_gpu.so")
SOURCES = ["hg_mesh.cpp", "hg_api.cpp", "hg_device.hip", "hg_batch.hip", "hg_noalloc.hip", "hg_mortar.hip"]
ORDERS = range(2, 14)   # points per direction P = N+1: one object per order from hg_inst.hip
ARCH = os.environ.get("HG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=on: multiply-adds are fused only within one source expression, decided before
# optimisation, so every kernel that inlines the same device function (tile and solver
# variants, the batched kernels) rounds it identically.
COMMON = ["-O3", "-std=c++17", "-ffp-contract=on", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _units():
    units = [(s, os.path.splitext(s)[0] + ".o", []) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    # per order and equations (0 Euler, 1 NS BR1, 2 NS IP)
    units += [("hg_inst.hip", f"hg_inst_p{p}_{ph}.o", [f"-DHG_P={p}", f"-DHG_PHYS={ph}"])
              for p in ORDERS for ph in range(3)]
    # batched kernels: variant 2 * physics + mixed, Riemann solver + 1 (0: chosen at run time)
    units += [("hg_batch_inst.hip", f"hg_batch_v{v}_rs{r}.o", [f"-DHG_BV={v}", f"-DHG_RS1={r}"])
              for v in range(6) for r in range(5)]
    return units


def _newest_header() -> float:
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return max(os.path.getmtime(h) for h in hdrs)


def _compile(unit, incremental: bool = False) -> str:
    src, objname, defs = unit
    obj = os.path.join(BUILD, objname)
    if incremental and os.path.exists(obj):
        t = os.path.getmtime(obj)
        if t > os.path.getmtime(os.path.join(CSRC, src)) and t > _newest_header():
            return obj
    cmd = [HIPCC] + COMMON + DEVICE + defs + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(PKG, "fortran")
DRIVER = os.path.join(PKG, "bin", "horses_gpu_driver")
EVAL = os.path.join(PKG, "bin", "contraction_eval")


def build_fortran(verbose: bool = True) -> str:
    """Compiles the ISO_C_BINDING module and the Fortran host driver against the library."""
    fb = os.path.join(BUILD, "fortran")
    os.makedirs(fb, exist_ok=True)
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    objs = []
    for src in ("horses_gpu_bindings.f90", "horses_gpu_driver.f90"):
        obj = os.path.join(fb, src.replace(".f90", ".o"))
        cmd = [FLANG, "-O2", "-module-dir", fb, "-c", os.path.join(FORTRAN, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    cmd = [FLANG, "-o", DRIVER, objs[1], objs[0], "-L" + PKG, "-lhorses2d_gpu", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {DRIVER}")
    return DRIVER


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compiles every unit (incremental=True skips objects newer than their source and every
    header) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    units = _units()
    jobs = int(os.environ.get("HG_BUILD_JOBS", str(min(8, os.cpu_count() or 1))))
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda u: _compile(u, incremental), units))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    # the measurement of the volume contraction on the vector units against the matrix cores
    os.makedirs(os.path.dirname(EVAL), exist_ok=True)
    cmd = [HIPCC, "-O3", "-std=c++17"] + DEVICE + ["-o", EVAL, os.path.join(CSRC, "hg_contraction_eval.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if os.path.exists(FLANG):
        build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
