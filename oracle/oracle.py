"""ctypes binding of the CPU oracle (ffpm_oracle.c) and of the reference's own
no-body Neumann/terms build (oracle/_ref, when present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Arrays are numpy float64 / int32 in Fortran order with the reference's shapes:
phi (nx+2, ny+2, nz), solid (nx+2, ny+2, nz), rhs and fields (nx, ny, nz),
xlo/xhi (ny, nz), ylo/yhi (nx, nz), zlo/zhi (nx, ny), xc (nx+2), yc (ny+2),
zc (nz), dx (nx), dy (ny), dz (nz).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libffpm_oracle.so")
# the same restatement with the PCG path's plane loops over OpenMP threads
# (CPU baseline only: its sums are not in the reference's order)
LIB_OMP_PATH = os.path.join(HERE, "_build", "libffpm_oracle_omp.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_laplace.so")
REF_PART_PATH = os.path.join(HERE, "_ref", "libref_partition.so")

PROC_NULL = -1
METHODS = {"pcg": 0, "jacobi": 1}
ERRORS = {
    -1: "bad argument",
    -2: "stair-step fluid cell has no fluid neighbour",
    -3: "stair-step compatibility needs at least one fluid outer face",
    -4: "neighbour rank present: only one rank is modelled",
}

_lib = None
_lib_omp = None
_ref = None
_ref_part = None


class OracleError(RuntimeError):
    pass


def build() -> None:
    """Compile the oracle with gcc (and the reference checker when its
    sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir("/root/reference/src"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_operator_residual_3d.restype = C.c_double
    return _lib


def lib_omp():
    global _lib_omp
    if _lib_omp is None:
        if not os.path.exists(LIB_OMP_PATH):
            build()
        _lib_omp = C.CDLL(LIB_OMP_PATH)
        _lib_omp.orc_operator_residual_3d.restype = C.c_double
    return _lib_omp


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
    return _ref


def _p(a, dtype=np.float64):
    if not isinstance(a, np.ndarray) or a.dtype != dtype:
        raise TypeError(f"expected numpy {np.dtype(dtype).name} array")
    if not a.flags.f_contiguous:
        raise TypeError("expected a Fortran-contiguous array")
    return a.ctypes.data_as(C.c_void_p)


def _d(x):
    return C.c_double(float(x))


def _check(status):
    if status != 0:
        raise OracleError(ERRORS.get(status, f"oracle status {status}"))


def _dims(solid):
    nx, ny, nz = solid.shape[0] - 2, solid.shape[1] - 2, solid.shape[2]
    return nx, ny, nz


# --- src/fpm_stairstep_mpi_3d.f90 -------------------------------------------

def exchange_solid_halo_3d(solid, west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                           north=PROC_NULL):
    nx, ny, nz = _dims(solid)
    _check(lib().orc_exchange_solid_halo_3d(_p(solid, np.int32), nx, ny, nz, west, east,
                                            south, north))


def enforce_stairstep_compatibility_3d(rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi,
                                       dx, dy, dz, west=PROC_NULL, east=PROC_NULL,
                                       south=PROC_NULL, north=PROC_NULL):
    nx, ny, nz = _dims(solid)
    corr, imb = C.c_double(), C.c_double()
    bg = np.asfortranarray(body_grad, dtype=np.float64)
    _check(lib().orc_enforce_stairstep_compatibility_3d(
        _p(rhs), _p(solid, np.int32), _p(bg), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo),
        _p(zhi), _p(dx), _p(dy), _p(dz), nx, ny, nz, west, east, south, north, C.byref(corr),
        C.byref(imb)))
    return corr.value, imb.value


def solve_stairstep_poisson_3d(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc,
                               yc, zc, dx, dy, dz, tolerance, max_sweeps, check_interval,
                               method="pcg", west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                               north=PROC_NULL, parallel=False):
    """parallel: the OpenMP build (timing only; the checker is serial)."""
    nx, ny, nz = _dims(solid)
    sweeps, res = C.c_int(), C.c_double()
    bg = np.asfortranarray(body_grad, dtype=np.float64)
    _check((lib_omp() if parallel else lib()).orc_solve_stairstep_poisson_3d(
        _p(phi), _p(rhs), _p(solid, np.int32), _p(bg), _p(xlo), _p(xhi), _p(ylo), _p(yhi),
        _p(zlo), _p(zhi), _p(xc), _p(yc), _p(zc), _p(dx), _p(dy), _p(dz), nx, ny, nz, west,
        east, south, north, _d(tolerance), int(max_sweeps), int(check_interval),
        METHODS[method], C.byref(sweeps), C.byref(res)))
    return sweeps.value, res.value


def stairstep_residual_3d(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc,
                          zc, dx, dy, dz, west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                          north=PROC_NULL):
    nx, ny, nz = _dims(solid)
    val = C.c_double()
    bg = np.asfortranarray(body_grad, dtype=np.float64)
    _check(lib().orc_stairstep_residual_3d(
        _p(phi), _p(rhs), _p(solid, np.int32), _p(bg), _p(xlo), _p(xhi), _p(ylo), _p(yhi),
        _p(zlo), _p(zhi), _p(xc), _p(yc), _p(zc), _p(dx), _p(dy), _p(dz), nx, ny, nz, west,
        east, south, north, C.byref(val)))
    return val.value


def compute_stairstep_fpm_terms_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, solid, body_grad, xc,
                                   yc, zc, dx, dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds,
                                   west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                                   north=PROC_NULL):
    nx, ny, nz = _dims(solid)
    out = [np.zeros((nx, ny, nz), order="F") for _ in range(5)]
    totals = np.zeros(2)
    bg = np.asfortranarray(body_grad, dtype=np.float64)
    bd = np.asfortranarray(bounds, dtype=np.float64)
    _check(lib().orc_compute_stairstep_fpm_terms_3d(
        _p(phi), _p(u), _p(v), _p(w), _p(ox), _p(oy), _p(oz), _p(ax), _p(ay), _p(az),
        _p(solid, np.int32), _p(bg), _p(xc), _p(yc), _p(zc), _p(dx), _p(dy), _p(dz), _p(xlo),
        _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(bd), *[_p(o) for o in out],
        _p(totals), nx, ny, nz, west, east, south, north))
    gx, gy, gz, vortex, advection = out
    return gx, gy, gz, vortex, advection, totals


def build_operator_3d(rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy,
                      dz, for_pcg=True):
    """Operator pieces (src/fpm_stairstep_mpi_3d.f90:336-379, 499-510): returns
    src, diag, rdiag, (cw, ce, cs, cn, cb, ct), isolated."""
    nx, ny, nz = _dims(solid)
    src, diag, rdiag = (np.zeros((nx, ny, nz), order="F") for _ in range(3))
    cw, ce = np.zeros(nx), np.zeros(nx)
    cs, cn = np.zeros(ny), np.zeros(ny)
    cb, ct = np.zeros(nz), np.zeros(nz)
    iso = C.c_int()
    bg = np.asfortranarray(body_grad, dtype=np.float64)
    _check(lib().orc_build_operator_3d(
        _p(rhs), _p(solid, np.int32), _p(bg), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo),
        _p(zhi), _p(xc), _p(yc), _p(zc), _p(dx), _p(dy), _p(dz), nx, ny, nz, PROC_NULL,
        PROC_NULL, PROC_NULL, PROC_NULL, int(bool(for_pcg)), _p(src), _p(diag), _p(rdiag),
        _p(cw), _p(ce), _p(cs), _p(cn), _p(cb), _p(ct), C.byref(iso)))
    return src, diag, rdiag, (cw, ce, cs, cn, cb, ct), iso.value


def jacobi_sweep_3d(a, src, rdiag, coef):
    nx, ny, nz = a.shape[0] - 2, a.shape[1] - 2, a.shape[2]
    b = np.array(a, order="F", copy=True)
    lib().orc_jacobi_sweep_3d(_p(a), _p(b), _p(src), _p(rdiag), *[_p(c) for c in coef], nx, ny,
                              nz)
    return b


def operator_residual_3d(phi, src, diag, coef):
    nx, ny, nz = phi.shape[0] - 2, phi.shape[1] - 2, phi.shape[2]
    return lib().orc_operator_residual_3d(_p(phi), _p(src), _p(diag), *[_p(c) for c in coef],
                                          nx, ny, nz)


# --- src/fpm_laplace_3d.f90 and src/fpm_terms_3d.f90 --------------------------

def enforce_neumann_compatibility_3d(rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz):
    nx, ny, nz = rhs.shape
    corr, imb = C.c_double(), C.c_double()
    _check(lib().orc_enforce_neumann_compatibility_3d(
        _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), nx, ny, nz, _d(dx),
        _d(dy), _d(dz), C.byref(corr), C.byref(imb)))
    return corr.value, imb.value


def solve_neumann_poisson_3d(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz, tolerance,
                             max_iterations, check_interval):
    nx, ny, nz = rhs.shape
    it, res = C.c_int(), C.c_double()
    _check(lib().orc_solve_neumann_poisson_3d(
        _p(phi), _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), nx, ny, nz,
        _d(dx), _d(dy), _d(dz), _d(tolerance), int(max_iterations), int(check_interval),
        C.byref(it), C.byref(res)))
    return it.value, res.value


def enforce_neumann_compatibility_nonuniform_3d(rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz):
    nx, ny, nz = rhs.shape
    corr, imb = C.c_double(), C.c_double()
    _check(lib().orc_enforce_neumann_compatibility_nonuniform_3d(
        _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(dx), _p(dy), _p(dz),
        nx, ny, nz, C.byref(corr), C.byref(imb)))
    return corr.value, imb.value


def solve_neumann_poisson_nonuniform_3d(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc,
                                        dx, dy, dz, tolerance, max_iterations, check_interval):
    nx, ny, nz = rhs.shape
    it, res = C.c_int(), C.c_double()
    _check(lib().orc_solve_neumann_poisson_nonuniform_3d(
        _p(phi), _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(xc), _p(yc),
        _p(zc), _p(dx), _p(dy), _p(dz), nx, ny, nz, _d(tolerance), int(max_iterations),
        int(check_interval), C.byref(it), C.byref(res)))
    return it.value, res.value


def compute_fpm_terms_uniform_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, dx, dy, dz):
    nx, ny, nz = u.shape
    out = [np.zeros((nx, ny, nz), order="F") for _ in range(5)]
    totals = np.zeros(2)
    _check(lib().orc_compute_fpm_terms_uniform_3d(
        _p(phi), _p(u), _p(v), _p(w), _p(ox), _p(oy), _p(oz), _p(ax), _p(ay), _p(az),
        _p(fluid, np.int32), *[_p(o) for o in out], _p(totals), nx, ny, nz, _d(dx), _d(dy),
        _d(dz)))
    return (*out, totals)


def compute_fpm_terms_nonuniform_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, xc, yc, zc,
                                    dx, dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds):
    nx, ny, nz = u.shape
    out = [np.zeros((nx, ny, nz), order="F") for _ in range(5)]
    totals = np.zeros(2)
    bd = np.asfortranarray(bounds, dtype=np.float64)
    _check(lib().orc_compute_fpm_terms_nonuniform_3d(
        _p(phi), _p(u), _p(v), _p(w), _p(ox), _p(oy), _p(oz), _p(ax), _p(ay), _p(az),
        _p(fluid, np.int32), _p(xc), _p(yc), _p(zc), _p(dx), _p(dy), _p(dz), _p(xlo), _p(xhi),
        _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(bd), *[_p(o) for o in out], _p(totals), nx, ny,
        nz))
    return (*out, totals)


# --- the reference's own build (src/fpm_laplace_3d.f90, src/fpm_terms_3d.f90) --
# Fortran procedures take every argument by reference; explicit-shape arrays are
# passed as base addresses, so ctypes pointers suffice.

def _ri(x):
    return C.byref(C.c_int(int(x)))


def _rd(x):
    return C.byref(C.c_double(float(x)))


def ref_enforce_neumann_compatibility_3d(rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz):
    nx, ny, nz = rhs.shape
    corr, imb = C.c_double(), C.c_double()
    ref()._QMfpm_laplace_3dPenforce_neumann_compatibility_3d(
        _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _ri(nx), _ri(ny),
        _ri(nz), _rd(dx), _rd(dy), _rd(dz), C.byref(corr), C.byref(imb))
    return corr.value, imb.value


def ref_solve_neumann_poisson_3d(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz, tolerance,
                                 max_iterations, check_interval):
    nx, ny, nz = rhs.shape
    it, res = C.c_int(), C.c_double()
    ref()._QMfpm_laplace_3dPsolve_neumann_poisson_3d(
        _p(phi), _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _ri(nx),
        _ri(ny), _ri(nz), _rd(dx), _rd(dy), _rd(dz), _rd(tolerance), _ri(max_iterations),
        _ri(check_interval), C.byref(it), C.byref(res))
    return it.value, res.value


def ref_enforce_neumann_compatibility_nonuniform_3d(rhs, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy,
                                                    dz):
    nx, ny, nz = rhs.shape
    corr, imb = C.c_double(), C.c_double()
    ref()._QMfpm_laplace_3dPenforce_neumann_compatibility_nonuniform_3d(
        _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(dx), _p(dy), _p(dz),
        _ri(nx), _ri(ny), _ri(nz), C.byref(corr), C.byref(imb))
    return corr.value, imb.value


def ref_solve_neumann_poisson_nonuniform_3d(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc,
                                            dx, dy, dz, tolerance, max_iterations,
                                            check_interval):
    nx, ny, nz = rhs.shape
    it, res = C.c_int(), C.c_double()
    ref()._QMfpm_laplace_3dPsolve_neumann_poisson_nonuniform_3d(
        _p(phi), _p(rhs), _p(xlo), _p(xhi), _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(xc), _p(yc),
        _p(zc), _p(dx), _p(dy), _p(dz), _ri(nx), _ri(ny), _ri(nz), _rd(tolerance),
        _ri(max_iterations), _ri(check_interval), C.byref(it), C.byref(res))
    return it.value, res.value


def ref_compute_fpm_terms_uniform_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, dx, dy, dz):
    nx, ny, nz = u.shape
    out = [np.zeros((nx, ny, nz), order="F") for _ in range(5)]
    totals = np.zeros(2)
    ref()._QMfpm_terms_3dPcompute_fpm_terms_uniform_3d(
        _p(phi), _p(u), _p(v), _p(w), _p(ox), _p(oy), _p(oz), _p(ax), _p(ay), _p(az),
        _p(fluid, np.int32), *[_p(o) for o in out], _p(totals), _ri(nx), _ri(ny), _ri(nz),
        _rd(dx), _rd(dy), _rd(dz))
    return (*out, totals)


def ref_compute_fpm_terms_nonuniform_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, xc, yc,
                                        zc, dx, dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds):
    nx, ny, nz = u.shape
    out = [np.zeros((nx, ny, nz), order="F") for _ in range(5)]
    totals = np.zeros(2)
    bd = np.asfortranarray(bounds, dtype=np.float64)
    ref()._QMfpm_terms_3dPcompute_fpm_terms_nonuniform_3d(
        _p(phi), _p(u), _p(v), _p(w), _p(ox), _p(oy), _p(oz), _p(ax), _p(ay), _p(az),
        _p(fluid, np.int32), _p(xc), _p(yc), _p(zc), _p(dx), _p(dy), _p(dz), _p(xlo), _p(xhi),
        _p(ylo), _p(yhi), _p(zlo), _p(zhi), _p(bd), *[_p(o) for o in out], _p(totals), _ri(nx),
        _ri(ny), _ri(nz))
    return (*out, totals)


# --- src/fpm_partition_2d.f90 ------------------------------------------------

def partition_terms_2d(phi, g, u, v, solid, xf, yf, kin, ring, dt, nu, slip_body=False):
    """phi, u, v (nx, ny) float64; solid (nx, ny) int32 0/1 or bool; xf (nx+1),
    yf (ny+1); kin (6); ring (2nx+2ny, 2, 2). Returns (terms[4], vif_density,
    omega)."""
    nx, ny = phi.shape
    sol = np.asfortranarray(solid, dtype=np.int32)
    gg = np.asfortranarray(g, dtype=np.float64)
    kk = np.asfortranarray(kin, dtype=np.float64)
    terms = np.zeros(4)
    vif = np.zeros((nx, ny), order="F")
    om = np.zeros((nx, ny), order="F")
    _check(lib().orc_partition_terms_2d(
        _p(phi), _p(gg), _p(u), _p(v), _p(sol, np.int32), _p(xf), _p(yf), _p(kk), _p(ring),
        _d(dt), _d(nu), nx, ny, 1 if slip_body else 0, _p(terms), _p(vif), _p(om)))
    return terms, vif, om


def ref_partition_available() -> bool:
    return os.path.exists(REF_PART_PATH)


def ref_partition_terms_2d(phi, g, u, v, solid, xf, yf, kin, ring, dt, nu, slip_body=None):
    """The reference's own fpm_partition_terms_2d (oracle/_ref build)."""
    global _ref_part
    if _ref_part is None:
        _ref_part = C.CDLL(REF_PART_PATH)
    nx, ny = phi.shape
    sol = np.asfortranarray(np.asarray(solid) != 0, dtype=np.int32)  # logical(4): 1 / 0
    gg = np.asfortranarray(g, dtype=np.float64)
    kk = np.asfortranarray(kin, dtype=np.float64)
    terms = np.zeros(4)
    vif = np.zeros((nx, ny), order="F")
    om = np.zeros((nx, ny), order="F")
    slip = None if slip_body is None else C.byref(C.c_int32(1 if slip_body else 0))
    _ref_part._QMfpm_partition_2dPfpm_partition_terms_2d(
        _p(phi), _p(gg), _p(u), _p(v), _p(sol, np.int32), _p(xf), _p(yf), _p(kk), _p(ring),
        C.byref(C.c_double(dt)), C.byref(C.c_double(nu)), C.byref(C.c_int(nx)),
        C.byref(C.c_int(ny)), _p(terms), _p(vif), _p(om), slip)
    return terms, vif, om
