"""The 2-D stream formats (host only): the FFPM2D02 frame of
examples/flapping_foil_fpm.f90:11-20,48-54 (layout in README.md:334-349) and
the output layout read by examples/plot_fields.py:17-21."""
import numpy as np
import pytest

from flowfpmpp_amd import frames2d as F


def _frame(nx=7, ny=5, seed=0):
    rng = np.random.default_rng(seed)
    L = 2 * nx + 2 * ny
    return F.Frame2D(nx, ny, np.cumsum(rng.random(nx + 1)), np.cumsum(rng.random(ny + 1)),
                     np.asfortranarray(rng.standard_normal((nx, ny))),
                     np.asfortranarray(rng.standard_normal((nx, ny))),
                     rng.random((nx, ny)) < 0.3, 0.4,
                     np.asfortranarray(rng.standard_normal((L, 2, 2))), rng.standard_normal(6),
                     172.8)


def test_ffpm2d02_round_trip(tmp_path):
    fr = _frame()
    p = tmp_path / "f.bin"
    F.write_ffpm2d02(str(p), fr)
    # header: magic, int32 nx, ny; total size of the documented layout
    raw = p.read_bytes()
    nx, ny = fr.nx, fr.ny
    assert raw[:8] == b"FFPM2D02"
    assert np.frombuffer(raw, "<i4", 2, 8).tolist() == [nx, ny]
    assert len(raw) == 16 + 8 * (nx + 1 + ny + 1 + 2 * nx * ny + 1 + 4 * (2 * nx + 2 * ny)
                                 + 6 + 1) + nx * ny
    g = F.read_ffpm2d02(str(p))
    for a in ("xf", "yf", "u", "v", "solid", "ring", "kin"):
        assert np.array_equal(getattr(g, a), getattr(fr, a)), a
    assert (g.nx, g.ny, g.dt, g.time) == (nx, ny, fr.dt, fr.time)
    # column-major: u(2,1) is the second value after the faces
    off = 16 + 8 * (nx + 1 + ny + 1)
    assert np.frombuffer(raw, "<f8", 1, off + 8)[0] == fr.u[1, 0]


def test_ffpm2d02_rejects_bad_files(tmp_path):
    fr = _frame()
    p = tmp_path / "f.bin"
    F.write_ffpm2d02(str(p), fr)
    raw = bytearray(p.read_bytes())
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"FFPM2D01" + raw[8:])
    with pytest.raises(ValueError, match="unexpected input format"):
        F.read_ffpm2d02(str(bad))
    swapped = raw[:8] + np.array([fr.nx, fr.ny], ">i4").tobytes() + raw[16:]
    bad.write_bytes(swapped)
    with pytest.raises(ValueError, match="implausible grid size"):
        F.read_ffpm2d02(str(bad))
    bad.write_bytes(raw[:-8])
    with pytest.raises(ValueError):
        F.read_ffpm2d02(str(bad))


def test_fields_layout_matches_plot_reader(tmp_path):
    nx, ny = 6, 4
    rng = np.random.default_rng(2)
    xc, yc = rng.random(nx), rng.random(ny)
    fields = [np.asfortranarray(rng.random((nx, ny))) for _ in range(5)]
    solid = rng.random((nx, ny)) < 0.5
    p = tmp_path / "out.bin"
    F.write_fields(str(p), xc, yc, fields[0], fields[1:3], fields[3:5], solid)
    # read back the way examples/plot_fields.py:17-21 does
    with open(p, "rb") as f:
        n1, n2 = np.fromfile(f, np.int32, 2)
        rx = np.fromfile(f, float, n1)
        ry = np.fromfile(f, float, n2)
        planes = [np.fromfile(f, float, n1 * n2).reshape(n2, n1) for _ in range(6)]
        assert f.read() == b""
    assert (n1, n2) == (nx, ny) and np.array_equal(rx, xc) and np.array_equal(ry, yc)
    for got, want in zip(planes, fields + [solid.astype(float)]):
        assert np.array_equal(got, np.asarray(want).T)


def test_cylinder_vortices_input():
    """The synthetic demo input (examples/demo_cylinder_vortices.f90:48-79):
    the stream is uniform far from the vortices, the prescribed vorticity is
    the Lamb-Oseen sum, and the cylinder of radius 0.5 is stair-stepped."""
    xf, yf, xc, yc, dx, dy, zc, dz, solid, u, v, om, ring = F.cylinder_vortices(90, 60)
    assert xf[0] == -3.0 and xf[-1] == 6.0 and yf[0] == -3.0 and yf[-1] == 3.0
    assert np.allclose(dx, 0.1) and np.allclose(dy, 0.1)
    assert np.allclose(dz, 0.1) and np.allclose(zc, [0.05, 0.15])
    X, Y = np.meshgrid(xc[1:-1], yc[1:-1], indexing="ij")
    assert np.array_equal(solid, X ** 2 + Y ** 2 < 0.25)
    # circulation of each vortex: the integral of omega over the plane
    # (the six cores are well inside the box) sums to sum(circ)
    circ = sum((-1.6 if m % 2 else 1.6) * (1 - 0.08 * (m - 1)) for m in range(1, 7))
    assert abs(float((om * 0.01).sum()) - circ) < 1e-3
    assert np.array_equal(ring[:, 0, 0], ring[:, 0, 1])
    assert np.array_equal(ring[: 60, 0, 0], u[0, :])
