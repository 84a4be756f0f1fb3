"""The Fortran drop-in modules (same module and routine names as the
reference) driven by Fortran programs, on the GPU."""
import os
import subprocess

import numpy as np
import pytest

import partition_cases as P
from oracle import oracle as O

pytestmark = pytest.mark.gpu

BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fortran", "_build")


def _exe(program):
    exe = os.path.join(BUILD, program)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (python -m flowfpmpp_amd.build)")
    return exe


@pytest.mark.parametrize("program", ["test_stairstep_exact_3d", "test_fpm_3d"])
def test_fortran_dropin_program(program):
    exe = _exe(program)
    print(f"case: Fortran program {program}", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    assert "PASS" in out.stdout


def test_fortran_frame_keeps_arrays_resident():
    """compat -> solve -> densities through fpm_stairstep_mpi_3d, unbound and
    inside an fpm_ffpm_frame frame: bit-identical results, no transfers
    inside the frame, one upload per bound array and one download per
    copy_out array (the program checks these and prints PASS); the unbound
    run passes use_openacc=.false., which prints the backend notice once."""
    exe = _exe("test_frame_residency")
    print("case: Fortran frame residency, 40x32x24", flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(out.stdout)
    print(out.stderr)
    assert out.returncode == 0, out.stderr
    assert "PASS" in out.stdout
    assert out.stderr.count("use_openacc=.false. requested") == 1


def _write_frame(path, c, slip):
    nx, ny = c["phi"].shape
    f64 = lambda a: np.asarray(a, dtype="<f8").ravel(order="F").tobytes()
    with open(path, "wb") as fh:
        fh.write(np.array([nx, ny, 1 if slip else 0], dtype="<i4").tobytes())
        fh.write(np.asarray(np.asarray(c["solid"]) != 0, dtype="<i4").ravel(order="F").tobytes())
        for key in ("phi", "u", "v", "xf", "yf", "g", "kin", "ring"):
            fh.write(f64(c[key]))
        fh.write(f64([c["dt"], c["nu"]]))


def _read_terms(path, nx, ny):
    a = np.fromfile(path, dtype="<f8")
    assert a.size == 4 + 2 * nx * ny
    vif = a[4:4 + nx * ny].reshape((nx, ny), order="F")
    om = a[4 + nx * ny:].reshape((nx, ny), order="F")
    return a[:4], vif, om


@pytest.mark.parametrize("frame,slip", [("random0", False), ("random1", True),
                                        ("rotation", False), ("uniform", False)])
def test_fortran_partition_module_matches_oracle(tmp_path, frame, slip):
    """fpm_partition_2d called from Fortran (logical mask, optional slip_body)
    on frames written by Python: the F_VIF density and the vorticity are
    bit-identical to the oracle's, the four terms agree to summation-order
    roundoff, and the uniform flow's terms are exact zeros."""
    exe = _exe("partition_frame_driver")
    c = {"random0": lambda: P.random_frame(0, nx=45, ny=38),
         "random1": lambda: P.random_frame(1, nx=31, ny=52),
         "rotation": P.rotation, "uniform": P.uniform_flow}[frame]()
    print(f"case: Fortran partition driver, frame {frame}, slip {slip}", flush=True)
    fin, fout = tmp_path / "frame.bin", tmp_path / "terms.bin"
    _write_frame(fin, c, slip)
    out = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(out.stdout)
    assert out.returncode == 0, out.stderr
    nx, ny = c["phi"].shape
    terms, vif, om = _read_terms(fout, nx, ny)
    ot, ov, oo = O.partition_terms_2d(**c, slip_body=slip)
    assert np.array_equal(vif, ov) and np.array_equal(om, oo)
    assert np.all(np.abs(terms - ot) <= 1e-13 * (1 + np.abs(ot)) * nx * ny)
    if frame == "uniform":
        assert np.abs(terms).max() == 0
