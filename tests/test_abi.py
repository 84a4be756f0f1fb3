"""CPU-side checks of the C ABI: the library loads and exports every symbol
include/ffpm.h declares (no GPU calls)."""
import ctypes

import pytest

from flowfpmpp_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = _lib.symbols()
    assert len(names) >= 16
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_bad_method_without_device():
    from flowfpmpp_amd import stairstep
    with pytest.raises(_lib.FFPMError, match="must be pcg or jacobi"):
        stairstep.solve_stairstep_poisson_3d(*([None] * 16), 1e-8, 10, 5, method="sor")


def test_shape_and_decomposition_errors_need_no_device():
    lib = _lib.lib()
    lib.ffpm_last_error.restype = ctypes.c_char_p
    null = ctypes.c_void_p(0)
    st = lib.ffpm_solve_stairstep_poisson_3d(*([null] * 16), 4, 4, 1, -1, -1, -1, -1,
                                             ctypes.c_double(1e-8), 10, 5, 0, null, null)
    assert st == 5 and b"nz>=2" in lib.ffpm_last_error()
    st = lib.ffpm_solve_stairstep_poisson_3d(*([null] * 16), 4, 4, 4, 2, -1, -1, -1,
                                             ctypes.c_double(1e-8), 10, 5, 0, null, null)
    assert st == 6 and b"MPI_PROC_NULL" in lib.ffpm_last_error()
    st = lib.ffpm_solve_stairstep_poisson_3d(*([null] * 16), 4, 4, 4, -1, -1, -1, -1,
                                             ctypes.c_double(1e-8), 10, 5, 7, null, null)
    assert st == 4 and b"pcg or jacobi" in lib.ffpm_last_error()
