"""Cell-centred stretched axes.

Two families:

* ``stretched_axis`` / ``local_axis``: the reference's sinusoidal width
  weights and mirrored halo centres (tests/fpm_test_grid.f90:11-43), used by
  the reference's own tests and benchmark workload.
* ``tanh_axis``: a symmetric atanh-mapped axis whose cells cluster at the
  origin with a prescribed largest/smallest width ratio (the stretched grids
  of the 128x96x64 and 512x384x256 workloads).

All arrays are float64. Centre arrays returned by ``local_axis`` carry one
halo centre on each side, index 0 and n+1, as the solver interface expects.
"""
from __future__ import annotations

import math

import numpy as np


def stretched_axis(n: int, lo: float, hi: float, amplitude: float):
    """Faces (0:n), centres (n) and widths (n) of an axis on [lo, hi] whose
    widths follow 1 + amplitude*sin(2*pi*(m-0.5)/n).

    Follows tests/fpm_test_grid.f90:11-27 operation by operation (serial sum of
    the weights, last face pinned to ``hi`` after the centres are formed).
    """
    weight = [1.0 + amplitude * math.sin(2.0 * math.acos(-1.0) * (m - 0.5) / n)
              for m in range(1, n + 1)]
    total = 0.0
    for wgt in weight:
        total += wgt
    face = np.zeros(n + 1)
    center = np.zeros(n)
    width = np.zeros(n)
    face[0] = lo
    for m in range(1, n + 1):
        width[m - 1] = (hi - lo) * weight[m - 1] / total
        face[m] = face[m - 1] + width[m - 1]
        center[m - 1] = 0.5 * (face[m - 1] + face[m])
    face[n] = hi
    return face, center, width


def local_axis(gcenter: np.ndarray, gwidth: np.ndarray):
    """Centres (0:n+1) with halo centres mirrored across the physical
    boundaries, and widths, for a single rank holding the whole axis
    (tests/fpm_test_grid.f90:31-43 with offset 0)."""
    n = len(gcenter)
    center = np.zeros(n + 2)
    center[1:n + 1] = gcenter
    center[0] = gcenter[0] - gwidth[0]
    center[n + 1] = gcenter[n - 1] + gwidth[n - 1]
    return center, np.array(gwidth, dtype=np.float64)


def _atanh_faces(n: int, half: float, gamma: float) -> np.ndarray:
    s = -1.0 + 2.0 * np.arange(n + 1) / n
    return half * np.arctanh(gamma * s) / math.atanh(gamma)


def tanh_axis(n: int, lo: float, hi: float, ratio: float = 4.0, tol: float = 1e-13):
    """Axis on the symmetric interval [lo, hi] with faces
    f_m = c + h*atanh(gamma*s_m)/atanh(gamma), s_m = -1 + 2m/n, so the cells are
    finest at the centre c = (lo+hi)/2. gamma is found by bisection so that the
    largest/smallest cell width is exactly ``ratio`` (1 gives a uniform axis).

    Returns faces (0:n), centres (n), widths (n).
    """
    if n < 2:
        raise ValueError("tanh_axis needs n >= 2")
    c = 0.5 * (lo + hi)
    half = 0.5 * (hi - lo)

    def width_ratio(gamma: float) -> float:
        w = np.diff(_atanh_faces(n, half, gamma))
        return float(w.max() / w.min())

    if ratio <= 1.0:
        gamma = 0.0
        faces = c + half * (-1.0 + 2.0 * np.arange(n + 1) / n)
    else:
        glo, ghi = 1e-9, 1.0 - 1e-15
        if width_ratio(ghi) < ratio:
            raise ValueError(f"ratio {ratio} not reachable with n={n}")
        for _ in range(200):
            gm = 0.5 * (glo + ghi)
            if width_ratio(gm) < ratio:
                glo = gm
            else:
                ghi = gm
            if ghi - glo < tol:
                break
        gamma = 0.5 * (glo + ghi)
        faces = c + _atanh_faces(n, half, gamma)
    faces[0] = lo
    faces[n] = hi
    width = np.diff(faces)
    center = 0.5 * (faces[:-1] + faces[1:])
    return faces, center, width
