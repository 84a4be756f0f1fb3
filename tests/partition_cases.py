"""Inputs of the 2-D force-partition tests: random frames for the bitwise
checks, and the reference test's three flows (tests/test_fpm_partition_2d.f90:
solid-body rotation, uniform flow, a stair-step cylinder accelerating through
fluid at rest on a 160 x 160 uniform grid over [-2, 2]^2)."""
import numpy as np

N, HALF, RADIUS, W0 = 160, 2.0, 0.25, 0.7


def random_frame(seed, nx=37, ny=29):
    rng = np.random.default_rng(seed)
    xf = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, nx))])
    yf = np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, ny))])
    xc, yc = 0.5 * (xf[:-1] + xf[1:]), 0.5 * (yf[:-1] + yf[1:])
    X, Y = np.meshgrid(xc, yc, indexing="ij")
    cx, cy = xc[nx // 2], yc[ny // 3]
    solid = (((X - cx) / (0.2 * xf[-1])) ** 2 + ((Y - cy) / (0.15 * yf[-1])) ** 2 < 1)
    f = lambda: np.asfortranarray(rng.standard_normal((nx, ny)))
    phi, u, v = f(), f(), f()
    ring = np.asfortranarray(rng.standard_normal((2 * nx + 2 * ny, 2, 2)))
    kin = rng.standard_normal(6)
    g = rng.standard_normal(2)
    return dict(phi=phi, g=g, u=u, v=v, solid=np.asfortranarray(solid.astype(np.int32)),
                xf=xf, yf=yf, kin=kin, ring=ring, dt=0.37, nu=1.3e-3)


def uniform_axis(n=N, half=HALF):
    xf = np.linspace(-half, half, n + 1)
    xc = 0.5 * (xf[:-1] + xf[1:])
    dx = np.diff(xf)
    xch = np.concatenate([[xc[0] - dx[0]], xc, [xc[-1] + dx[-1]]])
    return xf, xc, dx, xch


def fill_ring(u, v):
    n1, n2 = u.shape
    ring = np.zeros((2 * n1 + 2 * n2, 2, 2), order="F")
    for t in range(2):
        ring[:, 0, t] = np.concatenate([u[0, :], u[-1, :], u[:, 0], u[:, -1]])
        ring[:, 1, t] = np.concatenate([v[0, :], v[-1, :], v[:, 0], v[:, -1]])
    return ring


def rotation():
    xf, xc, dx, _ = uniform_axis()
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    u, v = np.asfortranarray(-W0 * Y), np.asfortranarray(W0 * X)
    return dict(phi=np.ones((N, N), order="F"), g=np.array([1.0, 0.0]), u=u, v=v,
                solid=np.zeros((N, N), dtype=np.int32, order="F"), xf=xf, yf=xf,
                kin=np.zeros(6), ring=fill_ring(u, v), dt=1.0, nu=1e-3)


def uniform_flow():
    xf, xc, dx, _ = uniform_axis()
    u, v = np.ones((N, N), order="F"), np.zeros((N, N), order="F")
    phi = np.asfortranarray(xc[:, None] - xc[None, :])
    return dict(phi=phi, g=np.array([1.0, 0.0]), u=u, v=v,
                solid=np.zeros((N, N), dtype=np.int32, order="F"), xf=xf, yf=xf,
                kin=np.zeros(6), ring=fill_ring(u, v), dt=1.0, nu=1e-3)


def cylinder_problem():
    """Stair-step cylinder of radius 0.25 (10 cells) in the 160^2 box, as a
    2-cell-deep 3-D stair-step frame for the phi_x solve."""
    xf, xc, dx, xch = uniform_axis()
    X, Y = np.meshgrid(xc, xc, indexing="ij")
    solid2 = (X ** 2 + Y ** 2 < RADIUS ** 2)
    nz = 2
    solid3 = np.zeros((N + 2, N + 2, nz), dtype=np.int32, order="F")
    for k in range(nz):
        solid3[1:N + 1, 1:N + 1, k] = solid2
    return dict(xf=xf, xc=xch, dx=dx, zc=np.array([0.25, 0.75]), dz=np.array([0.5, 0.5]),
                solid2=np.asfortranarray(solid2.astype(np.int32)), solid3=solid3)


def cylinder_terms_inputs(phi2, solid2, xf):
    u, v = np.zeros((N, N), order="F"), np.zeros((N, N), order="F")
    return dict(phi=np.asfortranarray(phi2), g=np.array([1.0, 0.0]), u=u, v=v, solid=solid2,
                xf=xf, yf=xf, kin=np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0]),
                ring=fill_ring(u, v), dt=1.0, nu=1e-3)
