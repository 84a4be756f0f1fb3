"""Python mirror of the reference's ``fpm_stairstep_mpi_3d`` module over the
HIP library's C ABI (include/ffpm.h).

Same routine names, argument order and meaning as the Fortran module
(src/fpm_stairstep_mpi_3d.f90:21-22); the MPI ``comm`` argument is dropped
and the four neighbour ranks default to MPI_PROC_NULL (-1). Intent(out)
scalars are returned instead of passed. Arrays may be

* numpy arrays in Fortran order (float64 / int32) -- staged through host
  memory by the library's host-pointer entry points, or
* :class:`flowfpmpp_amd.device.DeviceArray` objects (Fortran-order data in
  HBM) -- passed to the ``_dev`` entry points, which keep everything
  resident on the device.

``use_openacc`` is accepted for source compatibility: this library has one
backend, the GPU. ``use_openacc=False`` still runs on the GPU and warns once
(a ``RuntimeWarning``), so a CPU-against-GPU comparison is never silently a
GPU-against-GPU one.

Every array's shape is checked against the grid (taken from ``solid``)
before the library is called: a mismatch raises ValueError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import FFPMError, check, lib
from .device import DeviceArray

PROC_NULL = -1
METHODS = {"pcg": 0, "jacobi": 1, "pcg-mixed": 2, "fcg-mg": 4}


_NOTICED = set()


def _backend_notice(use_openacc, routine):
    """The reference's use_openacc switch (src/fpm_stairstep_mpi_3d.f90:116,124,
    436,440) selects its host loops when false. This library has only the GPU
    backend: the call still runs on the GPU, and a warning says so once per
    routine."""
    if use_openacc is False and routine not in _NOTICED:
        import warnings
        _NOTICED.add(routine)
        warnings.warn(f"{routine}: use_openacc=False requested, but this library has only the "
                      "GPU backend; the call runs on the GPU", RuntimeWarning, stacklevel=3)


def _is_dev(a) -> bool:
    return isinstance(a, DeviceArray)


def _ptr(a, dtype):
    if _is_dev(a):
        if a.dtype != np.dtype(dtype):
            raise TypeError(f"device array must hold {np.dtype(dtype).name}")
        return a.ptr
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.f_contiguous:
        raise TypeError(f"expected a Fortran-ordered numpy {np.dtype(dtype).name} array")
    return a.ctypes.data_as(C.c_void_p)


def _d(a):
    return _ptr(a, np.float64)


def _i(a):
    return _ptr(a, np.int32)


def _vec(x, n):
    v = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n))
    return v, v.ctypes.data_as(C.c_void_p)


def _mode(*arrays) -> str:
    dev = [_is_dev(a) for a in arrays if a is not None]
    if all(dev):
        return "_dev"
    if not any(dev):
        return ""
    raise TypeError("mix of host and device arrays")


def _shape(solid):
    if len(solid.shape) != 3:
        raise ValueError(f"solid must be 3-D (nx+2, ny+2, nz), got shape {solid.shape}")
    nxp, nyp, nz = solid.shape
    return nxp - 2, nyp - 2, nz


# Fortran shapes of the reference's dummy arguments (src/fpm_stairstep_mpi_3d.f90:
# 106-123, 426-439), as functions of the grid taken from solid.
_SHAPES = {
    "padded": lambda nx, ny, nz: (nx + 2, ny + 2, nz),
    "cells": lambda nx, ny, nz: (nx, ny, nz),
    "xface": lambda nx, ny, nz: (ny, nz),
    "yface": lambda nx, ny, nz: (nx, nz),
    "zface": lambda nx, ny, nz: (nx, ny),
    "xc": lambda nx, ny, nz: (nx + 2,),
    "yc": lambda nx, ny, nz: (ny + 2,),
    "zc": lambda nx, ny, nz: (nz,),
    "dx": lambda nx, ny, nz: (nx,),
    "dy": lambda nx, ny, nz: (ny,),
    "dz": lambda nx, ny, nz: (nz,),
}
_FACE_KINDS = ("xface", "xface", "yface", "yface", "zface", "zface")
_FACE_NAMES = ("xlo", "xhi", "ylo", "yhi", "zlo", "zhi")
_AXIS_NAMES = ("xc", "yc", "zc", "dx", "dy", "dz")


def _check_shapes(grid, **named):
    """Raises ValueError unless every array has the shape its reference dummy
    argument has for this grid: the library reads and writes exactly that many
    elements, so a wrongly sized array would be read or written past its end.
    Each keyword is name=(array, kind); kind is a key of _SHAPES."""
    for name, (a, kind) in named.items():
        if a is None:
            continue
        want = _SHAPES[kind](*grid)
        got = tuple(int(x) for x in a.shape)
        if got != want:
            raise ValueError(f"{name} has shape {got}; the grid {grid} needs {want}")


def _check_frame(grid, faces=None, axes=None, **named):
    if faces is not None:
        named.update({n: (f, k) for n, f, k in zip(_FACE_NAMES, faces, _FACE_KINDS)})
    if axes is not None:
        named.update({n: (a, n) for n, a in zip(_AXIS_NAMES, axes) if a is not None})
    _check_shapes(grid, **named)


def exchange_solid_halo_3d(solid, west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                           north=PROC_NULL):
    """src/fpm_stairstep_mpi_3d.f90:25-36"""
    nx, ny, nz = _shape(solid)
    fn = getattr(lib(), "ffpm_exchange_solid_halo_3d" + _mode(solid))
    check(fn(_i(solid), nx, ny, nz, west, east, south, north))


def enforce_stairstep_compatibility_3d(rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, dx,
                                       dy, dz, west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                                       north=PROC_NULL):
    """src/fpm_stairstep_mpi_3d.f90:41-92; updates the face arrays in place and
    returns (correction, imbalance)."""
    nx, ny, nz = _shape(solid)
    _check_frame((nx, ny, nz), (xlo, xhi, ylo, yhi, zlo, zhi), (None, None, None, dx, dy, dz),
                 rhs=(rhs, "cells"))
    mode = _mode(rhs, solid, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz)
    bg, bgp = _vec(body_grad, 3)
    corr, imb = C.c_double(), C.c_double()
    fn = getattr(lib(), "ffpm_enforce_stairstep_compatibility_3d" + mode)
    check(fn(_d(rhs), _i(solid), bgp, _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi),
             _d(dx), _d(dy), _d(dz), nx, ny, nz, west, east, south, north, C.byref(corr),
             C.byref(imb)))
    return corr.value, imb.value


def solve_stairstep_poisson_3d(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc,
                               zc, dx, dy, dz, tolerance, max_sweeps, check_interval,
                               use_openacc=True, method="pcg", west=PROC_NULL, east=PROC_NULL,
                               south=PROC_NULL, north=PROC_NULL):
    """src/fpm_stairstep_mpi_3d.f90:106-330; phi is updated in place and
    (sweeps, residual) is returned. method "pcg-mixed" (no reference
    counterpart) runs FP64 defect correction around an FP32 PCG and stops on
    the same FP64 max-norm residual; sweeps then counts inner iterations.
    method "fcg-mg" (no reference counterpart) runs FP64 flexible CG with an
    FP32 geometric-multigrid preconditioner, stopping on the same FP64
    max-norm residual; sweeps counts FCG iterations."""
    _backend_notice(use_openacc, "solve_stairstep_poisson_3d")
    if method not in METHODS:
        raise FFPMError(4, "FFPM_ERR_METHOD: stair-step method must be pcg or jacobi")
    nx, ny, nz = _shape(solid)
    _check_frame((nx, ny, nz), (xlo, xhi, ylo, yhi, zlo, zhi), (xc, yc, zc, dx, dy, dz),
                 phi=(phi, "padded"), rhs=(rhs, "cells"))
    mode = _mode(phi, rhs, solid, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz)
    bg, bgp = _vec(body_grad, 3)
    sweeps, res = C.c_int(), C.c_double()
    fn = getattr(lib(), "ffpm_solve_stairstep_poisson_3d" + mode)
    check(fn(_d(phi), _d(rhs), _i(solid), bgp, _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo),
             _d(zhi), _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), nx, ny, nz, west, east,
             south, north, C.c_double(float(tolerance)), int(max_sweeps), int(check_interval),
             METHODS[method], C.byref(sweeps), C.byref(res)))
    return sweeps.value, res.value


def stairstep_residual_3d(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc,
                          dx, dy, dz, west=PROC_NULL, east=PROC_NULL, south=PROC_NULL,
                          north=PROC_NULL):
    """src/fpm_stairstep_mpi_3d.f90:403-419"""
    nx, ny, nz = _shape(solid)
    _check_frame((nx, ny, nz), (xlo, xhi, ylo, yhi, zlo, zhi), (xc, yc, zc, dx, dy, dz),
                 phi=(phi, "padded"), rhs=(rhs, "cells"))
    mode = _mode(phi, rhs, solid, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz)
    bg, bgp = _vec(body_grad, 3)
    val = C.c_double()
    fn = getattr(lib(), "ffpm_stairstep_residual_3d" + mode)
    check(fn(_d(phi), _d(rhs), _i(solid), bgp, _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo),
             _d(zhi), _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), nx, ny, nz, west, east,
             south, north, C.byref(val)))
    return val.value


def compute_stairstep_fpm_terms_3d(phi, u, v, w, ox, oy, oz, ax, ay, az, solid, body_grad, xc,
                                   yc, zc, dx, dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds,
                                   gx=None, gy=None, gz=None, vortex=None, advection=None,
                                   use_openacc=True, west=PROC_NULL, east=PROC_NULL,
                                   south=PROC_NULL, north=PROC_NULL):
    """src/fpm_stairstep_mpi_3d.f90:426-469; returns
    (gx, gy, gz, vortex, advection, totals). Output arrays are allocated
    when not given (numpy for host inputs, DeviceArray for device inputs)."""
    _backend_notice(use_openacc, "compute_stairstep_fpm_terms_3d")
    nx, ny, nz = _shape(solid)
    outs = [gx, gy, gz, vortex, advection]
    _check_frame((nx, ny, nz), (xlo, xhi, ylo, yhi, zlo, zhi), (xc, yc, zc, dx, dy, dz),
                 phi=(phi, "padded"),
                 **{n: (a, "cells") for n, a in zip(("u", "v", "w", "ox", "oy", "oz", "ax", "ay",
                                                     "az"), (u, v, w, ox, oy, oz, ax, ay, az))},
                 **{n: (a, "cells") for n, a in zip(("gx", "gy", "gz", "vortex", "advection"),
                                                    outs)})
    mode = _mode(phi, u, v, w, ox, oy, oz, ax, ay, az, solid, xc, yc, zc, dx, dy, dz, xlo, xhi,
                 ylo, yhi, zlo, zhi)
    for m in range(5):
        if outs[m] is None:
            outs[m] = (DeviceArray((nx, ny, nz)) if mode == "_dev"
                       else np.zeros((nx, ny, nz), order="F"))
    bg, bgp = _vec(body_grad, 3)
    bd, bdp = _vec(bounds, 6)
    totals = np.zeros(2)
    fn = getattr(lib(), "ffpm_compute_stairstep_fpm_terms_3d" + mode)
    check(fn(_d(phi), _d(u), _d(v), _d(w), _d(ox), _d(oy), _d(oz), _d(ax), _d(ay), _d(az),
             _i(solid), bgp, _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), _d(xlo), _d(xhi),
             _d(ylo), _d(yhi), _d(zlo), _d(zhi), bdp, *[_d(o) for o in outs],
             totals.ctypes.data_as(C.c_void_p), nx, ny, nz, west, east, south, north))
    return (*outs, totals)


def compute_stairstep_fpm_terms_3d_dirs(phis, u, v, w, ox, oy, oz, ax, ay, az, solid,
                                        body_grads, xc, yc, zc, dx, dy, dz, faces, bounds,
                                        outs=None, west=PROC_NULL, east=PROC_NULL,
                                        south=PROC_NULL, north=PROC_NULL, noalloc=False):
    """compute_stairstep_fpm_terms_3d (src/fpm_stairstep_mpi_3d.f90:426-469)
    for 1-3 directions of one frame in one pass over the flow fields (device
    arrays only). phis: the directions' potentials; body_grads: one 3-vector
    per direction; faces: per direction its six outer face arrays (xlo, xhi,
    ylo, yhi, zlo, zhi); outs: per direction five output arrays (gx, gy, gz,
    vortex, advection), allocated when None. Returns per direction (gx, gy,
    gz, vortex, advection, totals), each equal bit for bit to what
    compute_stairstep_fpm_terms_3d returns for that direction. noalloc: call the
    library's entry point that allocates no device memory (fixed scratch,
    nx + ny + nz <= 16384), with the same results."""
    nd = len(phis)
    if not 1 <= nd <= 3 or len(body_grads) != nd or len(faces) != nd:
        raise ValueError("1 to 3 directions, each with a body gradient and six face arrays")
    if any(len(f) != 6 for f in faces):
        raise ValueError("six outer face arrays per direction")
    nx, ny, nz = _shape(solid)
    flat = [*phis, u, v, w, ox, oy, oz, ax, ay, az, solid, xc, yc, zc, dx, dy, dz,
            *[a for f in faces for a in f]]
    if outs is None:
        outs = [[DeviceArray((nx, ny, nz)) for _ in range(5)] for _ in range(nd)]
    if len(outs) != nd or any(len(o) != 5 for o in outs):
        raise ValueError("five output arrays per direction")
    g = (nx, ny, nz)
    _check_frame(g, None, (xc, yc, zc, dx, dy, dz),
                 **{f"phi[{d}]": (phis[d], "padded") for d in range(nd)},
                 **{n: (a, "cells") for n, a in zip(("u", "v", "w", "ox", "oy", "oz", "ax", "ay",
                                                     "az"), (u, v, w, ox, oy, oz, ax, ay, az))},
                 **{f"outs[{d}][{m}]": (outs[d][m], "cells") for d in range(nd)
                    for m in range(5)})
    for d in range(nd):
        _check_shapes(g, **{f"{n}[{d}]": (f, k) for n, f, k in
                            zip(_FACE_NAMES, faces[d], _FACE_KINDS)})
    if _mode(*flat, *[a for o in outs for a in o]) != "_dev":
        raise TypeError("the multi-direction densities take device arrays")
    P = C.c_void_p * nd

    def arr(items):
        return P(*[C.c_void_p(_d(a).value) for a in items])

    bg, bgp = _vec(np.asarray(body_grads, dtype=np.float64).reshape(-1), 3 * nd)
    bd, bdp = _vec(bounds, 6)
    totals = np.zeros(2 * nd)
    ptrs = [arr(phis)] + [arr([f[m] for f in faces]) for m in range(6)] + \
        [arr([o[m] for o in outs]) for m in range(5)]
    fn = (lib().ffpm_compute_stairstep_fpm_terms_3d_dirs_noalloc_dev if noalloc
          else lib().ffpm_compute_stairstep_fpm_terms_3d_dirs_dev)
    check(fn(
        nd, ptrs[0], _d(u), _d(v), _d(w), _d(ox), _d(oy), _d(oz), _d(ax), _d(ay), _d(az),
        _i(solid), bgp, _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), *ptrs[1:7], bdp,
        *ptrs[7:12], totals.ctypes.data_as(C.c_void_p), nx, ny, nz, west, east, south, north))
    return [(*outs[d], totals[2 * d:2 * d + 2].copy()) for d in range(nd)]


def set_stream(handle) -> None:
    """Launch this thread's later calls on the given raw hipStream_t handle
    (None: the null stream). Each stream has its own workspace, so threads on
    different streams may call the library at the same time."""
    check(lib().ffpm_set_stream(C.c_void_p(0 if handle is None else int(handle))))


def stream_create() -> int:
    """A new non-blocking HIP stream of default priority (raw handle)."""
    h = C.c_void_p()
    check(lib().ffpm_stream_create(C.byref(h)))
    return int(h.value or 0)


def stream_destroy(handle) -> None:
    """Waits for the stream's work and destroys it."""
    check(lib().ffpm_stream_destroy(C.c_void_p(int(handle))))


def last_mixed_solve():
    """(outer defect corrections, inner FP32 iterations) of the last
    method="pcg-mixed" solve."""
    o, i = C.c_int(), C.c_int()
    check(lib().ffpm_last_mixed_solve(C.byref(o), C.byref(i)))
    return o.value, i.value


def last_mg_solve():
    """(levels, convergence checks) of the last "fcg-mg" solve on this thread."""
    lv, ch = C.c_int(), C.c_int()
    check(lib().ffpm_last_mg_solve(C.byref(lv), C.byref(ch)))
    return lv.value, ch.value


def timing(enable: bool, reset: bool = True) -> None:
    """Turns HIP-event timing of the hot kernels on or off; turning it on
    clears the totals unless reset is False."""
    check(lib().ffpm_timing_enable(1 if enable else 0))
    if enable and reset:
        check(lib().ffpm_timing_reset())


def timing_query(kernel: str):
    """(launches, total milliseconds) of a named hot kernel since the last reset."""
    cnt, ms = C.c_long(), C.c_double()
    check(lib().ffpm_timing_query(kernel.encode(), C.byref(cnt), C.byref(ms)))
    return cnt.value, ms.value


# ---------------------------------------------------------------- batches --
class _Problem(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("phi", C.c_void_p), ("rhs", C.c_void_p), ("solid", C.c_void_p),
                ("body_grad", C.c_double * 3),
                ("xlo", C.c_void_p), ("xhi", C.c_void_p), ("ylo", C.c_void_p),
                ("yhi", C.c_void_p), ("zlo", C.c_void_p), ("zhi", C.c_void_p),
                ("xc", C.c_void_p), ("yc", C.c_void_p), ("zc", C.c_void_p),
                ("dx", C.c_void_p), ("dy", C.c_void_p), ("dz", C.c_void_p),
                ("sweeps", C.c_int), ("residual", C.c_double), ("solve_ms", C.c_double)]


class BatchProblem:
    """One solve of a batch: the device arrays of one solve_stairstep_poisson_3d
    call. Problems that pass the same ``solid`` and axis arrays (the force
    directions of one frame) share one operator in the batch."""

    def __init__(self, phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc,
                 dx, dy, dz):
        arrays = (phi, rhs, solid, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz)
        if not all(_is_dev(a) for a in arrays):
            raise TypeError("batched solves take device arrays")
        _check_frame(_shape(solid), (xlo, xhi, ylo, yhi, zlo, zhi), (xc, yc, zc, dx, dy, dz),
                     phi=(phi, "padded"), rhs=(rhs, "cells"))
        for name, a in (("phi", phi), ("rhs", rhs), *zip(_FACE_NAMES, (xlo, xhi, ylo, yhi, zlo,
                                                                       zhi)),
                        *zip(_AXIS_NAMES, (xc, yc, zc, dx, dy, dz))):
            if a.dtype != np.float64:
                raise TypeError(f"{name} must hold float64")
        if solid.dtype != np.int32:
            raise TypeError("solid must hold int32")
        self.phi, self.rhs, self.solid = phi, rhs, solid
        self.body_grad = np.asarray(body_grad, dtype=np.float64).reshape(3)
        self.faces = (xlo, xhi, ylo, yhi, zlo, zhi)
        self.axes = (xc, yc, zc, dx, dy, dz)
        self.sweeps, self.residual, self.solve_ms = 0, float("nan"), float("nan")

    def _struct(self):
        nx, ny, nz = _shape(self.solid)
        s = _Problem()
        s.nx, s.ny, s.nz = nx, ny, nz
        s.phi, s.rhs, s.solid = _d(self.phi).value, _d(self.rhs).value, _i(self.solid).value
        s.body_grad = (C.c_double * 3)(*self.body_grad)
        (s.xlo, s.xhi, s.ylo, s.yhi, s.zlo, s.zhi) = [_d(f).value for f in self.faces]
        (s.xc, s.yc, s.zc, s.dx, s.dy, s.dz) = [_d(a).value for a in self.axes]
        return s


def _structs(problems):
    arr = (_Problem * len(problems))()
    for k, p in enumerate(problems):
        arr[k] = p._struct()
    return arr


def solve_stairstep_batch(problems, tolerance, max_sweeps, check_interval, rows=0, zchunk=0,
                          use_graph=True, method="pcg"):
    """PCG ('pcg', FP64), mixed-precision PCG ('pcg-mixed') or the
    multigrid-preconditioned CG ('fcg-mg') on many problems in one batch
    (per-problem convergence masks, HIP-graph chunks). Each problem ends as its
    own solve_stairstep_poisson_3d with the same method would; sweeps, residual
    and the time to its convergence (ms from the batch start) are stored on
    it. rows and zchunk apply to the PCG methods."""
    if method not in ("pcg", "pcg-mixed", "fcg-mg"):
        raise ValueError("batched solves support 'pcg', 'pcg-mixed' and 'fcg-mg'")
    arr = _structs(problems)
    check(lib().ffpm_solve_stairstep_batch_dev(arr, len(problems), C.c_double(tolerance),
                                                int(max_sweeps), int(check_interval),
                                                METHODS[method], int(rows), int(zchunk),
                                                1 if use_graph else 0))
    for k, p in enumerate(problems):
        p.sweeps, p.residual, p.solve_ms = arr[k].sweeps, arr[k].residual, arr[k].solve_ms
    return [(p.sweeps, p.residual) for p in problems]


def autotune_stairstep_batch(problems, iterations=20, method="pcg"):
    """Times the candidate launch shapes on the batch; returns (rows, zchunk,
    ms per batched iteration) of the fastest."""
    arr = _structs(problems)
    rows, zchunk, ms = C.c_int(), C.c_int(), C.c_double()
    check(lib().ffpm_autotune_stairstep_batch_dev(arr, len(problems), METHODS[method],
                                                   int(iterations), C.byref(rows),
                                                   C.byref(zchunk), C.byref(ms)))
    return rows.value, zchunk.value, ms.value
