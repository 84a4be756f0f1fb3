"""Python mirror of the reference's ``fpm_partition_2d`` module
(src/fpm_partition_2d.f90:35-154) over the HIP library's C ABI: the force
partition F_KINE, F_VIF, F_VIS, F_BOUND of a 2-D stair-step body with rigid
in-plane motion, with the F_VIF density and the vorticity.

Arrays are numpy (host, staged by the library) or DeviceArray objects (kept in
HBM), Fortran order: phi, u, v, solid (nx, ny); xf (nx+1), yf (ny+1); ring
(2nx+2ny, 2, 2). ``solid`` may be bool or integer (nonzero = solid)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .device import DeviceArray


def fpm_partition_terms_2d(phi, g, u, v, solid, xf, yf, kin, ring, dt, nu, slip_body=False,
                           vif_density=None, omega=None):
    """Returns (terms[4], vif_density, omega); terms = [KINE, VIF, VIS, BOUND]."""
    nx, ny = phi.shape
    gg = np.ascontiguousarray(g, dtype=np.float64).reshape(2)
    kk = np.ascontiguousarray(kin, dtype=np.float64).reshape(6)
    terms = np.zeros(4)
    dev = isinstance(phi, DeviceArray)
    if dev:
        vif = vif_density if vif_density is not None else DeviceArray((nx, ny))
        om = omega if omega is not None else DeviceArray((nx, ny))
        check(lib().ffpm_partition_terms_2d_dev(
            phi.ptr, gg.ctypes.data_as(C.c_void_p), u.ptr, v.ptr, solid.ptr, xf.ptr, yf.ptr,
            kk.ctypes.data_as(C.c_void_p), ring.ptr, C.c_double(dt), C.c_double(nu), nx, ny,
            1 if slip_body else 0, terms.ctypes.data_as(C.c_void_p), vif.ptr, om.ptr))
        return terms, vif, om
    f64 = lambda a: np.asfortranarray(a, dtype=np.float64)
    ph, uu, vv, xx, yy, rr = f64(phi), f64(u), f64(v), f64(xf), f64(yf), f64(ring)
    so = np.asfortranarray(np.asarray(solid) != 0, dtype=np.int32)
    vif = np.zeros((nx, ny), order="F")
    om = np.zeros((nx, ny), order="F")
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    check(lib().ffpm_partition_terms_2d(p(ph), p(gg), p(uu), p(vv), p(so), p(xx), p(yy), p(kk),
                                        p(rr), C.c_double(dt), C.c_double(nu), nx, ny,
                                        1 if slip_body else 0, p(terms), p(vif), p(om)))
    return terms, vif, om
