"""The Fortran host (horses2d_amd/fortran/horses_gpu_driver.f90, ISO_C_BINDING module
horses_gpu_bindings) drives the C ABI through the reference's time-loop semantics; its
state after 20 RK3 iterations of the Euler vortex configuration matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from horses2d_amd import cases as C

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "horses2d_amd", "bin", "horses_gpu_driver")


def test_fortran_driver_matches_oracle(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built: run __graft_entry__.build()")
    out = tmp_path / "q.txt"
    r = subprocess.run([DRIVER, "20", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    Qf = np.loadtxt(out)
    from oracle.oracle import Oracle
    case = C.euler_vortex_case(16, 4)
    orc = Oracle(case)
    orc.set_state(case.initial_state(orc.coords()))
    orc.integrate(20)
    Qo = orc.state()
    assert Qf.shape == Qo.shape
    assert np.max(np.abs(Qf - Qo)) / np.max(np.abs(Qo)) <= 1e-12
