"""The Fortran drop-in modules (same module and routine names as the
reference) driven by Fortran programs, on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "_build")


@pytest.mark.parametrize("program", ["test_stairstep_exact_3d", "test_fpm_3d",
                                     "test_fpm_partition_2d"])
def test_fortran_dropin_program(program):
    exe = os.path.join(BUILD, program)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -m flowfpmpp_amd.build)")
    print(f"case: Fortran program {program}", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    assert "PASS" in out.stdout
