"""A Fortran caller of the library through the ISO_C_BINDING module include/mod_btdma_amd.f90
(tests/fortran/sample_btdma_amd.f90), run on the device against the reference's published
Study signature."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "bin", "sample_btdma_amd")


def test_fortran_sample(cuda):
    if not os.path.exists(EXE):
        pytest.skip("the Fortran sample is not built (amdflang absent at build time)")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASSED" in r.stdout, r.stdout + r.stderr
