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


def test_several_direction_densities_argument_errors_need_no_device():
    """The one-pass densities of several directions take 1 to 3 directions
    and one body gradient, six face arrays and five outputs per direction."""
    lib = _lib.lib()
    lib.ffpm_last_error.restype = ctypes.c_char_p
    null = ctypes.c_void_p(0)
    for nd in (0, 4):
        st = lib.ffpm_compute_stairstep_fpm_terms_3d_dirs_dev(nd, *([null] * 31), 4, 4, 4,
                                                              -1, -1, -1, -1)
        assert st != 0 and b"1 to 3 directions" in lib.ffpm_last_error()
    from flowfpmpp_amd import stairstep
    # phis, nine flow fields and solid, body gradients, six axes, faces, bounds
    with pytest.raises(ValueError, match="1 to 3 directions"):
        stairstep.compute_stairstep_fpm_terms_3d_dirs([], *([None] * 10), [], *([None] * 6),
                                                      [], None)
    with pytest.raises(ValueError, match="six outer face arrays"):
        stairstep.compute_stairstep_fpm_terms_3d_dirs([None], *([None] * 10), [None],
                                                      *([None] * 6), [[None] * 5], None)
