"""2-D frames: the FFPM2D02 stream format of the reference's real-frame example
and its force-partition driver (examples/flapping_foil_fpm.f90), on the GPU.

Input format (little-endian Fortran stream, column-major, README.md:334-349):
``character(8) 'FFPM2D02'; int32 nx, ny; real64 xf(0:nx), yf(0:ny);
real64 u(nx,ny), v(nx,ny); int8 solid(nx,ny); real64 dt;
real64 ring(2nx+2ny, 2, 2); real64 kin(6); real64 time``.

Output format (the layout of the reference's demo_cylinder_vortices output
that examples/plot_fields.py reads): ``int32 nx, ny; real64 xc(nx), yc(ny),
omega(nx,ny), vif_x(nx,ny), vif_y(nx,ny), phi_x(nx,ny), phi_y(nx,ny),
solid(nx,ny)`` (solid as 1.0/0.0).

    python flowfpmpp_amd/frames2d.py foil <frame.bin> [--out fields.bin] [--tol 1e-10]
                                     [--method pcg|fcg-mg]
    python flowfpmpp_amd/frames2d.py cylinder [--nx 360 --ny 240] [--out fields.bin]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

MAGIC = b"FFPM2D02"


class Frame2D:
    def __init__(self, nx, ny, xf, yf, u, v, solid, dt, ring, kin, time_):
        self.nx, self.ny = nx, ny
        self.xf, self.yf, self.u, self.v = xf, yf, u, v
        self.solid, self.dt, self.ring, self.kin, self.time = solid, dt, ring, kin, time_


def read_ffpm2d02(path: str) -> Frame2D:
    """Reads an FFPM2D02 frame; raises ValueError on a bad magic string or an
    implausible size (byte order)."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != MAGIC:
        raise ValueError("unexpected input format")
    nx, ny = (int(x) for x in np.frombuffer(raw, "<i4", 2, 8))
    if not (2 <= nx <= 100000 and 2 <= ny <= 100000):
        raise ValueError("implausible grid size (byte order?)")
    off = 16

    def take(dtype, count, shape=None):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off)
        off += a.nbytes
        return np.array(a.reshape(shape, order="F") if shape else a, order="F")

    xf = take("<f8", nx + 1)
    yf = take("<f8", ny + 1)
    u = take("<f8", nx * ny, (nx, ny))
    v = take("<f8", nx * ny, (nx, ny))
    solid = take("i1", nx * ny, (nx, ny))
    dt = float(take("<f8", 1)[0])
    L = 2 * nx + 2 * ny
    ring = take("<f8", L * 4, (L, 2, 2))
    kin = take("<f8", 6)
    time_ = float(take("<f8", 1)[0])
    if off != len(raw):
        raise ValueError(f"frame size mismatch: {off} bytes read of {len(raw)}")
    return Frame2D(nx, ny, xf, yf, u, v, solid != 0, dt, ring, kin, time_)


def write_ffpm2d02(path: str, fr: Frame2D) -> None:
    """Writes a frame in the FFPM2D02 layout read by :func:`read_ffpm2d02`."""
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([fr.nx, fr.ny], "<i4").tobytes())
        for a in (fr.xf, fr.yf, fr.u, fr.v):
            f.write(np.asfortranarray(a, dtype="<f8").tobytes(order="F"))
        f.write(np.asfortranarray(np.asarray(fr.solid) != 0, dtype="i1").tobytes(order="F"))
        f.write(np.array([fr.dt], "<f8").tobytes())
        f.write(np.asfortranarray(fr.ring, dtype="<f8").tobytes(order="F"))
        f.write(np.asarray(fr.kin, dtype="<f8").tobytes())
        f.write(np.array([fr.time], "<f8").tobytes())


def write_fields(path, xc, yc, omega, vif, phi, solid):
    """Writes the output layout (vif and phi: the x and y direction fields)."""
    nx, ny = omega.shape
    with open(path, "wb") as f:
        f.write(np.array([nx, ny], "<i4").tobytes())
        for a in (xc, yc, omega, vif[0], vif[1], phi[0], phi[1],
                  np.asarray(solid, dtype=np.float64)):
            f.write(np.asfortranarray(a, dtype="<f8").tobytes(order="F"))


def _solve_2d(xf, yf, xc, yc, dx, dy, zc, dz, solid, u, v, kin, ring, dt, nu, tolerance,
              max_sweeps, check_interval, slip_body, log, method="pcg"):
    """phi_x and phi_y of a two-cell-deep stair-step frame (zero z-face data)
    by PCG (or another solve method) on the GPU, each followed by the force
    partition of its direction. ``rhs`` is shared by the two directions as in
    the reference's drivers."""
    from flowfpmpp_amd import partition as FP
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import DeviceArray, synchronize
    T = DeviceArray.from_host
    nx, ny = solid.shape
    nz = 2
    solid3 = np.zeros((nx + 2, ny + 2, nz), dtype=np.int32, order="F")
    for k in range(nz):
        solid3[1:nx + 1, 1:ny + 1, k] = solid
    d_solid = T(solid3)
    axes = [T(np.asarray(a, dtype=np.float64)) for a in (xc, yc, zc, dx, dy, dz)]
    d_rhs = T(np.zeros((nx, ny, nz), order="F"))
    sol2 = T(np.asfortranarray(solid, dtype=np.int32))
    d_u, d_v, d_xf, d_yf, d_ring = (T(np.asarray(a, dtype=np.float64))
                                    for a in (u, v, xf, yf, ring))
    out = {"iterations": [], "residuals": [], "corrections": [], "terms": np.zeros((4, 2)),
           "vif": [], "phi": []}
    synchronize()
    t0 = time.perf_counter()
    for d in range(2):
        body = np.zeros(3)
        body[d] = 1.0
        faces = [T(np.zeros(s, order="F")) for s in
                 ((ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny))]
        phi = T(np.zeros((nx + 2, ny + 2, nz), order="F"))
        corr, _ = S.enforce_stairstep_compatibility_3d(d_rhs, d_solid, body, *faces, *axes[3:])
        sw, res = S.solve_stairstep_poisson_3d(phi, d_rhs, d_solid, body, *faces, *axes,
                                               tolerance, max_sweeps, check_interval,
                                               method=method)
        phi2 = np.asfortranarray(phi.download()[1:nx + 1, 1:ny + 1, 0])
        terms, vif, om = FP.fpm_partition_terms_2d(T(phi2), body[:2], d_u, d_v, sol2, d_xf,
                                                   d_yf, kin, d_ring, dt, nu, slip_body)
        out["iterations"].append(sw)
        out["residuals"].append(res)
        out["corrections"].append(corr)
        out["terms"][:, d] = terms
        out["vif"].append(vif.download())
        out["phi"].append(phi2)
        out["omega"] = om.download()
        if log:
            log(f"direction {d + 1}: {method} iterations {sw}, residual {res:.3e}")
    synchronize()
    out["seconds"] = time.perf_counter() - t0
    out["xc"], out["yc"] = np.asarray(xc)[1:nx + 1], np.asarray(yc)[1:ny + 1]
    return out


def solve_frame(fr: Frame2D, tolerance: float = 1e-10, re: float = 1.0e4, log=None,
                method: str = "pcg"):
    """The real-frame driver (examples/flapping_foil_fpm.f90:55-86): the foil
    as a stair-step body in a two-cell-deep frame of unit span (dz = 0.5, zero
    z-face derivatives), halo centres mirrored across the outer boundary,
    phi_x and phi_y by PCG (check interval 50, cap 2,000,000), then the force
    partition of each direction with nu = 1/Re. Returns a dict with
    iterations, residuals, terms (4 x 2: KINE, VIF, VIS, BOUND by direction),
    the F_VIF densities, phi, omega and the cell centres. method "fcg-mg"
    (the multigrid-preconditioned CG) replaces the reference's PCG; both stop
    on the same FP64 max-norm residual."""
    nx, ny = fr.nx, fr.ny
    dx, dy = np.diff(fr.xf), np.diff(fr.yf)
    xc = np.empty(nx + 2)
    yc = np.empty(ny + 2)
    xc[1:nx + 1] = 0.5 * (fr.xf[:-1] + fr.xf[1:])
    yc[1:ny + 1] = 0.5 * (fr.yf[:-1] + fr.yf[1:])
    xc[0], xc[nx + 1] = xc[1] - dx[0], xc[nx] + dx[-1]
    yc[0], yc[ny + 1] = yc[1] - dy[0], yc[ny] + dy[-1]
    return _solve_2d(fr.xf, fr.yf, xc, yc, dx, dy, [0.25, 0.75], [0.5, 0.5], fr.solid, fr.u,
                     fr.v, fr.kin, fr.ring, fr.dt, 1.0 / re, tolerance, 2000000, 50, False, log,
                     method)


def cylinder_vortices(gnx: int = 360, gny: int = 240):
    """The synthetic input of examples/demo_cylinder_vortices.f90:48-79: a
    stair-step cylinder (radius 0.5 at the origin) in a uniform stream plus a
    staggered row of six Lamb-Oseen vortices (core 0.22) on stretched axes
    x in [-3, 6], y in [-3, 3] (zero amplitude), nz = 2 of width dx(1).
    Returns (xf, yf, xc, yc, dx, dy, zc, dz, solid, u, v, omega, ring)."""
    from flowfpmpp_amd.grids import local_axis, stretched_axis
    xf, gxc, gdx = stretched_axis(gnx, -3.0, 6.0, 0.0)
    yf, gyc, gdy = stretched_axis(gny, -3.0, 3.0, 0.0)
    _, zc, dz = stretched_axis(2, 0.0, 2 * gdx[0], 0.0)
    xc, dx = local_axis(gxc, gdx)
    yc, dy = local_axis(gyc, gdy)
    radius, core, nvort = 0.5, 0.22, 6
    pi = np.arccos(-1.0)
    X = xc[1:gnx + 1][:, None] * np.ones((1, gny))
    Y = np.ones((gnx, 1)) * yc[1:gny + 1][None, :]
    u = np.ones((gnx, gny), order="F")
    v = np.zeros((gnx, gny), order="F")
    om = np.zeros((gnx, gny), order="F")
    for m in range(1, nvort + 1):
        vx = 0.9 + 0.75 * (m - 1)
        vy = 0.3 if m % 2 == 1 else -0.3
        circ = (-1.6 if m % 2 == 1 else 1.6) * (1 - 0.08 * (m - 1))
        r2 = (X - vx) ** 2 + (Y - vy) ** 2
        q = circ / (2 * pi) * (1 - np.exp(-r2 / core ** 2)) / np.maximum(r2, 1e-30)
        u -= q * (Y - vy)
        v += q * (X - vx)
        om += circ / (pi * core ** 2) * np.exp(-r2 / core ** 2)
    solid = X ** 2 + Y ** 2 < radius ** 2
    ring = np.zeros((2 * gnx + 2 * gny, 2, 2), order="F")
    for d in range(2):
        ring[:, 0, d] = np.concatenate([u[0, :], u[-1, :], u[:, 0], u[:, -1]])
        ring[:, 1, d] = np.concatenate([v[0, :], v[-1, :], v[:, 0], v[:, -1]])
    return xf, yf, xc, yc, dx, dy, zc, dz, solid, u, v, om, ring


def solve_cylinder_vortices(gnx: int = 360, gny: int = 240, log=None, method: str = "pcg"):
    """The demo driver (examples/demo_cylinder_vortices.f90:80-95): PCG to
    1e-9 (check interval 500, cap 5,000,000) and the partition with the body
    faces as slip faces, zero kinematics, dt = 1 and nu = 0. The vorticity
    written out is the prescribed one, as in the reference."""
    xf, yf, xc, yc, dx, dy, zc, dz, solid, u, v, om, ring = cylinder_vortices(gnx, gny)
    r = _solve_2d(xf, yf, xc, yc, dx, dy, zc, dz, solid, u, v, np.zeros(6), ring, 1.0, 0.0,
                  1e-9, 5000000, 500, True, log, method)
    r["omega"] = om
    r["solid"] = solid
    return r


def main():  # pragma: no cover - command-line driver
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    ap = argparse.ArgumentParser(description="2-D force partition drivers on the GPU")
    sub = ap.add_subparsers(dest="case", required=True)
    f = sub.add_parser("foil", help="an FFPM2D02 frame (flapping_foil_fpm)")
    f.add_argument("frame")
    f.add_argument("--tol", type=float, default=1e-10)
    f.add_argument("--out", default=None)
    c = sub.add_parser("cylinder", help="the synthetic demo_cylinder_vortices case")
    c.add_argument("--nx", type=int, default=360)
    c.add_argument("--ny", type=int, default=240)
    c.add_argument("--out", default=None)
    for sp in (f, c):
        sp.add_argument("--method", default="pcg", choices=["pcg", "fcg-mg", "pcg-mixed"])
    args = ap.parse_args()

    def log(m):
        print(m, file=sys.stderr, flush=True)

    if args.case == "foil":
        fr = read_ffpm2d02(args.frame)
        log(f"frame: {fr.nx} x {fr.ny} cells, {int(fr.solid.sum())} solid, t = {fr.time:.3f}")
        r = solve_frame(fr, args.tol, log=log, method=args.method)
        solid, grid = fr.solid, [fr.nx, fr.ny]
    else:
        r = solve_cylinder_vortices(args.nx, args.ny, log=log, method=args.method)
        solid, grid = r["solid"], [args.nx, args.ny]
    for q, nm in enumerate(["F_KINE", "F_VIF", "F_VIS", "F_BOUND"]):
        log(f"  {nm:8s} {r['terms'][q, 0]: .5e} {r['terms'][q, 1]: .5e}")
    log(f"  sum      {r['terms'][:, 0].sum(): .5e} {r['terms'][:, 1].sum(): .5e}")
    if args.out:
        write_fields(args.out, r["xc"], r["yc"], r["omega"], r["vif"], r["phi"], solid)
    print(json.dumps({"case": args.case, "method": args.method, "grid": grid,
                      "iterations": r["iterations"],
                      "residuals": r["residuals"], "seconds": round(r["seconds"], 3),
                      "terms_x": r["terms"][:, 0].tolist(), "terms_y": r["terms"][:, 1].tolist(),
                      "sum": r["terms"].sum(axis=0).tolist()}), flush=True)


if __name__ == "__main__":  # pragma: no cover
    main()
