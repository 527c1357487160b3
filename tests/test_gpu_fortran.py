"""Fortran callers of the library through the ISO_C_BINDING module include/mod_btdma_amd.f90,
run on the device: tests/fortran/sample_btdma_amd.f90 against the reference's published Study
signature, tests/fortran/sample_euler_amd.f90 through an euler3d ADI step (periodic residual of
every assembled system, assembled vs matrix-free sweeps), tests/fortran/sample_weno_amd.f90
through a weno2d stage's x and y sweeps (row sums, periodic residuals, conserved mass), and
tests/fortran/sample_heat3d_calls.f90, a caller using the call-compatible module
include/mod_btdma_gpu_v2.f90 with heat3d's own call sites (against host scalar solves)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "bin")


@pytest.mark.parametrize("name", ["sample_btdma_amd", "sample_euler_amd", "sample_weno_amd",
                                  "sample_heat3d_calls"])
def test_fortran_sample(cuda, name):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip("the Fortran samples are not built (amdflang absent at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout + r.stderr
