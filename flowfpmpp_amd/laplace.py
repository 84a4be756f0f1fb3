"""Python mirror of the reference's ``fpm_laplace_3d`` and ``fpm_terms_3d``
modules (no-body pure-Neumann solvers and their FPM densities) over the HIP
library's C ABI. Same routine names and argument order as
src/fpm_laplace_3d.f90:5-7 and src/fpm_terms_3d.f90:5-6; intent(out)
scalars are returned. Arrays are Fortran-ordered numpy arrays (host entry
points) or DeviceArray objects (``_dev`` entry points), as in
``flowfpmpp_amd.stairstep``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .device import DeviceArray
from .stairstep import _d, _i, _mode, _vec


def _f(x):
    return C.c_double(float(x))


def enforce_neumann_compatibility_3d(rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, dx, dy, dz):
    """src/fpm_laplace_3d.f90:9-53 -> (correction, imbalance)"""
    mode = _mode(rhs, xlo, xhi, ylo, yhi, zlo, zhi)
    corr, imb = C.c_double(), C.c_double()
    fn = getattr(lib(), "ffpm_enforce_neumann_compatibility_3d" + mode)
    check(fn(_d(rhs), _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi), nx, ny, nz, _f(dx),
             _f(dy), _f(dz), C.byref(corr), C.byref(imb)))
    return corr.value, imb.value


def solve_neumann_poisson_3d(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, dx, dy, dz,
                             tolerance, max_iterations, check_interval):
    """src/fpm_laplace_3d.f90:55-107 -> (iterations, final_residual); phi is
    (0:nx+1, 0:ny+1, 0:nz+1) and updated in place."""
    mode = _mode(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi)
    it, res = C.c_int(), C.c_double()
    fn = getattr(lib(), "ffpm_solve_neumann_poisson_3d" + mode)
    check(fn(_d(phi), _d(rhs), _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi), nx, ny, nz,
             _f(dx), _f(dy), _f(dz), _f(tolerance), int(max_iterations), int(check_interval),
             C.byref(it), C.byref(res)))
    return it.value, res.value


def enforce_neumann_compatibility_nonuniform_3d(rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz,
                                                nx, ny, nz):
    """src/fpm_laplace_3d.f90:150-217 (dx, dy, dz before nx, ny, nz)"""
    mode = _mode(rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz)
    corr, imb = C.c_double(), C.c_double()
    fn = getattr(lib(), "ffpm_enforce_neumann_compatibility_nonuniform_3d" + mode)
    check(fn(_d(rhs), _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi), _d(dx), _d(dy),
             _d(dz), nx, ny, nz, C.byref(corr), C.byref(imb)))
    return corr.value, imb.value


def solve_neumann_poisson_nonuniform_3d(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx,
                                        dy, dz, nx, ny, nz, tolerance, max_iterations,
                                        check_interval):
    """src/fpm_laplace_3d.f90:219-293 -> (iterations, final_residual)"""
    mode = _mode(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz)
    it, res = C.c_int(), C.c_double()
    fn = getattr(lib(), "ffpm_solve_neumann_poisson_nonuniform_3d" + mode)
    check(fn(_d(phi), _d(rhs), _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi), _d(xc),
             _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), nx, ny, nz, _f(tolerance),
             int(max_iterations), int(check_interval), C.byref(it), C.byref(res)))
    return it.value, res.value


def _outs(like, n3, mode):
    if mode == "_dev":
        return [DeviceArray(n3) for _ in range(5)]
    return [np.zeros(n3, order="F") for _ in range(5)]


def compute_fpm_terms_uniform_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, nx, ny, nz, dx,
                                 dy, dz):
    """src/fpm_terms_3d.f90:8-37 -> (gx, gy, gz, vortex, advection, totals)"""
    mode = _mode(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid)
    outs = _outs(u, (nx, ny, nz), mode)
    totals = np.zeros(2)
    fn = getattr(lib(), "ffpm_compute_fpm_terms_uniform_3d" + mode)
    check(fn(_d(phi), _d(u), _d(v), _d(w), _d(ox), _d(oy), _d(oz), _d(ax), _d(ay), _d(az),
             _i(fluid), *[_d(o) for o in outs], totals.ctypes.data_as(C.c_void_p), nx, ny, nz,
             _f(dx), _f(dy), _f(dz)))
    return (*outs, totals)


def compute_fpm_terms_nonuniform_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, xc, yc, zc, dx,
                                    dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds, nx, ny, nz):
    """src/fpm_terms_3d.f90:39-87 -> (gx, gy, gz, vortex, advection, totals)"""
    mode = _mode(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, xc, yc, zc, dx, dy, dz, xlo, xhi,
                 ylo, yhi, zlo, zhi)
    outs = _outs(u, (nx, ny, nz), mode)
    totals = np.zeros(2)
    bd, bdp = _vec(bounds, 6)
    fn = getattr(lib(), "ffpm_compute_fpm_terms_nonuniform_3d" + mode)
    check(fn(_d(phi), _d(u), _d(v), _d(w), _d(ox), _d(oy), _d(oz), _d(ax), _d(ay), _d(az),
             _i(fluid), _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), _d(xlo), _d(xhi),
             _d(ylo), _d(yhi), _d(zlo), _d(zhi), bdp, *[_d(o) for o in outs],
             totals.ctypes.data_as(C.c_void_p), nx, ny, nz))
    return (*outs, totals)
