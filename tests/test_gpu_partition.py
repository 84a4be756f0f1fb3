"""GPU parity of the 2-D force partition (fpm_partition_terms_2d) against the
CPU oracle (which matches the reference's own build bit for bit, see
tests/test_partition_oracle.py).

Bars: the F_VIF density and the vorticity are bit-identical per cell (same
expressions, same order); the four terms are sums over the cells taken in
another order, so they agree to roundoff (|a - b| <= 1e-13 of the sum of the
magnitudes of their contributions, bounded here by 1e-13 * (1 + |b|) * cells),
and exact zeros stay exact. The reference test's criteria hold on the GPU
path: rotation, uniform flow, accelerating cylinder (added mass 1.14)."""
import numpy as np
import pytest

import partition_cases as P
from flowfpmpp_amd import partition as FP
from flowfpmpp_amd import stairstep as S
from flowfpmpp_amd.device import DeviceArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _close(a, b, cells):
    return np.all(np.abs(a - b) <= 1e-13 * (1 + np.abs(b)) * cells)


@pytest.mark.parametrize("seed,slip", [(0, False), (1, True), (5, False)])
def test_partition_matches_oracle(seed, slip):
    print(f"case: partition random frame seed {seed} slip {slip}", flush=True)
    c = P.random_frame(seed, nx=45, ny=38)
    go = FP.fpm_partition_terms_2d(**c, slip_body=slip)
    oo = O.partition_terms_2d(**c, slip_body=slip)
    assert np.array_equal(go[1], oo[1]) and np.array_equal(go[2], oo[2])
    assert _close(go[0], oo[0], c["phi"].size)


def test_partition_device_arrays_match_host():
    print("case: partition device arrays", flush=True)
    c = P.random_frame(3, nx=64, ny=40)
    h = FP.fpm_partition_terms_2d(**c)
    T = DeviceArray.from_host
    d = FP.fpm_partition_terms_2d(T(c["phi"]), c["g"], T(c["u"]), T(c["v"]), T(c["solid"]),
                                  T(c["xf"]), T(c["yf"]), c["kin"], T(c["ring"]), c["dt"], c["nu"])
    assert np.array_equal(h[0], d[0])
    assert np.array_equal(h[1], d[1].download()) and np.array_equal(h[2], d[2].download())


def test_partition_reference_criteria_on_gpu():
    print("case: partition rotation / uniform flow / cylinder on the GPU", flush=True)
    n = P.N
    t, dens, om = FP.fpm_partition_terms_2d(**P.rotation())
    err1 = max(np.abs(dens[2:n - 2, 2:n - 2] + 2 * P.W0 ** 2).max(),
               np.abs(om[2:n - 2, 2:n - 2] - 2 * P.W0).max())
    assert err1 < 1e-10
    t2, _, _ = FP.fpm_partition_terms_2d(**P.uniform_flow())
    assert np.abs(t2).max() == 0
    c = P.cylinder_problem()
    nz = 2
    rhs = np.zeros((n, n, nz), order="F")
    phi3 = np.zeros((n + 2, n + 2, nz), order="F")
    faces = [np.zeros(s, order="F") for s in ((n, nz), (n, nz), (n, nz), (n, nz), (n, n), (n, n))]
    bg = np.array([1.0, 0.0, 0.0])
    S.enforce_stairstep_compatibility_3d(rhs, c["solid3"], bg, *faces, c["dx"], c["dx"], c["dz"])
    sw, res = S.solve_stairstep_poisson_3d(phi3, rhs, c["solid3"], bg, *faces, c["xc"], c["xc"],
                                           c["zc"], c["dx"], c["dx"], c["dz"], 1e-10, 200000, 50)
    assert res <= 1e-10
    t3, _, _ = FP.fpm_partition_terms_2d(**P.cylinder_terms_inputs(phi3[1:n + 1, 1:n + 1, 0],
                                                                   c["solid2"], c["xf"]))
    cm = -t3[0] / (np.pi * P.RADIUS ** 2)
    print(f"cylinder: {sw} PCG iterations, added-mass coefficient {cm:.4f}")
    assert abs(cm - 1) < 0.2 and np.abs(t3[1:]).max() == 0


def test_partition_errors():
    from flowfpmpp_amd._lib import FFPMError
    c = P.random_frame(0, nx=8, ny=6)
    c["phi"] = np.zeros((0, 6), order="F")
    with pytest.raises(FFPMError):
        FP.fpm_partition_terms_2d(**c)
