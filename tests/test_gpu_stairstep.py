"""GPU parity of the stair-step solve and FPM densities against the CPU oracle.

Bars (tolerances written per test):
* Jacobi has no reductions inside the sweep and the max-norm residual is
  order-independent, so iteration counts are identical and the pre-gauge
  iterates are bit-identical; after the gauge the fields differ only by the
  difference of the two parallel/serial mean sums, a constant.
* The face-gradient residual and the per-cell FPM densities are bit-identical
  for identical phi; integrals agree to summation-order roundoff.
* PCG's dot products are summed in another order, so converged fields agree
  to the solver tolerance (1e-10 relative max-norm) with the same
  iteration counts at the same check cadence.
"""
import numpy as np
import pytest

import cases
from flowfpmpp_amd import stairstep as S
from flowfpmpp_amd import workloads as W
from flowfpmpp_amd._lib import FFPMError
from flowfpmpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu


def _copy(c):
    import copy
    return copy.deepcopy(c)


def _faces(c):
    return tuple(np.array(f, order="F", copy=True) for f in c.faces)


def _gpu_solve(c, faces, tol, cap, check, method):
    phi = np.array(c.phi, order="F", copy=True)
    sw, res = S.solve_stairstep_poisson_3d(phi, c.rhs, c.solid, c.body, *faces, c.xc, c.yc, c.zc,
                                           c.dx, c.dy, c.dz, tol, cap, check, method=method)
    return phi, sw, res


def _orc_solve(orc, c, faces, tol, cap, check, method):
    phi = np.array(c.phi, order="F", copy=True)
    sw, res = orc.solve_stairstep_poisson_3d(phi, c.rhs, c.solid, c.body, *faces, c.xc, c.yc,
                                             c.zc, c.dx, c.dy, c.dz, tol, cap, check, method)
    return phi, sw, res


def _fluid(c):
    return c.solid[1:-1, 1:-1, :] == 0


def _assert_constant_shift(a, b, fluid, scale):
    d = (a[1:-1, 1:-1, :] - b[1:-1, 1:-1, :])[fluid]
    assert np.abs(d).max() <= 4e-16 * scale + 1e-300
    assert d.max() - d.min() <= 4e-16 * scale + 1e-300


def test_compatibility_matches_oracle(orc):
    print("case: exact grid influence compatibility", flush=True)
    c = cases.exact_case(False)
    f1, f2 = _faces(c), _faces(c)
    cg, ig = S.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *f1, c.dx, c.dy, c.dz)
    co, io = orc.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *f2, c.dx, c.dy, c.dz)
    assert abs(cg - co) <= 1e-14 * abs(co)
    assert abs(ig) < 1e-13 and abs(io) < 1e-13
    for a, b in zip(f1, f2):
        assert np.abs(a - b).max() <= 1e-14 * abs(co)


@pytest.mark.parametrize("check", [50, 7, 1])
def test_jacobi_fixed_sweeps_bitwise(orc, check):
    print(f"case: jacobi 600 fixed sweeps, check {check}", flush=True)
    c = cases.exact_case(False)
    faces = _faces(c)
    orc.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx, c.dy, c.dz)
    pg, sg, rg = _gpu_solve(c, faces, -1.0, 600, check, "jacobi")
    po, so, ro = _orc_solve(orc, c, faces, -1.0, 600, check, "jacobi")
    assert sg == so == 600
    scale = np.abs(po).max()
    _assert_constant_shift(pg, po, _fluid(c), scale)
    assert np.all(pg[1:-1, 1:-1, :][~_fluid(c)] == 0)
    assert np.all(pg[0, :, :] == 0) and np.all(pg[-1, :, :] == 0)
    assert np.all(pg[:, 0, :] == 0) and np.all(pg[:, -1, :] == 0)
    assert abs(rg - ro) <= 1e-12 * ro


def test_jacobi_to_tolerance_same_sweeps(orc):
    print("case: exact linear solution, jacobi to 1e-12", flush=True)
    c = cases.exact_case(True)
    faces = _faces(c)
    orc.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx, c.dy, c.dz)
    pg, sg, rg = _gpu_solve(c, faces, 1e-12, 400000, 50, "jacobi")
    po, so, ro = _orc_solve(orc, c, faces, 1e-12, 400000, 50, "jacobi")
    assert sg == so
    _assert_constant_shift(pg, po, _fluid(c), np.abs(po).max())


@pytest.mark.parametrize("exact,tol", [(True, 1e-10), (False, 1e-11)])
def test_pcg_converged_matches_oracle(orc, exact, tol):
    print(f"case: exact grid pcg exact={exact} tol={tol}", flush=True)
    c = cases.exact_case(exact)
    faces = _faces(c)
    orc.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx, c.dy, c.dz)
    pg, sg, rg = _gpu_solve(c, faces, tol, 100000, 50, "pcg")
    po, so, ro = _orc_solve(orc, c, faces, tol, 100000, 50, "pcg")
    assert sg == so
    assert rg <= 2 * tol and ro <= 2 * tol
    fl = _fluid(c)
    rel = np.abs(pg - po)[1:-1, 1:-1, :][fl].max() / np.abs(po).max()
    assert rel <= 1e-10
    chk = S.stairstep_residual_3d(pg, c.rhs, c.solid, c.body, *faces, c.xc, c.yc, c.zc, c.dx,
                                  c.dy, c.dz)
    assert abs(chk - rg) <= 1e-8 * rg + 1e-12


def test_exact_case_pass_criteria_on_gpu(orc):
    """tests/test_stairstep_exact_3d.f90:46-50 through the GPU path, both methods."""
    for method, tol in (("jacobi", 1e-12), ("pcg", 1e-10)):
        print(f"case: exact linear solution criteria, {method}", flush=True)
        c = cases.exact_case(True)
        faces = _faces(c)
        corr, imb = S.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx,
                                                         c.dy, c.dz)
        phi, sw, res = _gpu_solve(c, faces, tol, 400000, 50, method)
        chk = S.stairstep_residual_3d(phi, c.rhs, c.solid, c.body, *faces, c.xc, c.yc, c.zc,
                                      c.dx, c.dy, c.dz)
        f = c.fields
        gx, gy, gz, vo, ad, tot = S.compute_stairstep_fpm_terms_3d(
            phi, f["u"], f["v"], f["w"], f["ox"], f["oy"], f["oz"], f["ax"], f["ay"], f["az"],
            c.solid, c.body, c.xc, c.yc, c.zc, c.dx, c.dy, c.dz, *faces, c.bounds)
        errs, ierr = cases.exact_expectations(c, phi, gx, gy, gz, vo, tot)
        print(f"  {method}: {sw} iterations, residual {res:.2e}; errors phi {errs[0]:.1e}, "
              f"gradient {errs[1]:.1e}, vortex {errs[2]:.1e}, integrals {ierr:.1e}", flush=True)
        assert abs(corr) < 1e-13 and abs(imb) < 1e-13
        assert res <= 2 * tol and chk < 1e-9
        assert errs[0] < 1e-9 and errs[1] < 1e-9 and errs[2] < 1e-9 and errs[3] == 0
        # roundoff, as the reference reports (README.md:114-119: <= 1e-13 phi,
        # 5.9e-13 gradient, 7e-13 vortex density)
        assert errs[0] <= 1e-12 and errs[1] <= 1e-12 and errs[2] <= 1e-12
        assert ierr < 1e-10
        assert np.all(np.isfinite(errs))


def test_terms_and_face_residual_bitwise(orc):
    print("case: terms and face residual on the oracle's phi", flush=True)
    c = cases.exact_case(False)
    faces = _faces(c)
    orc.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx, c.dy, c.dz)
    po, so, ro = _orc_solve(orc, c, faces, 1e-11, 100000, 50, "pcg")
    f = c.fields
    args = (po, f["u"], f["v"], f["w"], f["ox"], f["oy"], f["oz"], f["ax"], f["ay"], f["az"],
            c.solid, c.body, c.xc, c.yc, c.zc, c.dx, c.dy, c.dz, *faces, c.bounds)
    g = S.compute_stairstep_fpm_terms_3d(*args)
    o = orc.compute_stairstep_fpm_terms_3d(*args)
    for a, b in zip(g[:5], o[:5]):
        assert np.array_equal(a, b)
    assert np.abs(g[5] - o[5]).max() <= 1e-13 * np.abs(o[5]).max()
    rg = S.stairstep_residual_3d(po, c.rhs, c.solid, c.body, *faces, c.xc, c.yc, c.zc, c.dx,
                                 c.dy, c.dz)
    rr = orc.stairstep_residual_3d(po, c.rhs, c.solid, c.body, *faces, c.xc, c.yc, c.zc, c.dx,
                                   c.dy, c.dz)
    assert rg == rr


@pytest.mark.parametrize("n,its", [(12, 100), (24, 200), (48, 300)])
def test_sphere_convergence_on_gpu(orc, n, its):
    """README.md:127-132 through the GPU path."""
    print(f"case: sphere N={n}", flush=True)
    rms_p, far_p, cm_p = {12: (1.08e-2, 2.24e-2, 0.897), 24: (9.27e-3, 1.87e-2, 0.770),
                          48: (4.56e-3, 9.21e-3, 0.639)}[n]
    c = cases.sphere_case(n)
    faces = _faces(c)
    S.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx, c.dy, c.dz)
    phi, sw, res = _gpu_solve(c, faces, 1e-10, 2000000, 100, "pcg")
    rms, far, cm = cases.sphere_errors(c, phi)
    assert sw == its and res <= 1e-10
    assert abs(rms - rms_p) <= 0.005 * rms_p and abs(far - far_p) <= 0.005 * far_p
    assert abs(cm - cm_p) <= 0.0006


def test_frame_128x96x64_fixed_work_and_device_entry(orc):
    """128x96x64 tanh-stretched frame, all three directions: 300 fixed PCG
    iterations and 400 fixed Jacobi sweeps agree with the oracle, and the
    device-pointer entry points give the host-pointer results."""
    p = W.frame_128x96x64(with_fields=True)
    for d in range(3):
        print(f"case: 128x96x64 direction {d}", flush=True)
        rhs, bg, faces = W.influence_inputs(p, d)
        orc.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
        for method, iters in (("pcg", 300), ("jacobi", 400)):
            pg = p.phi0()
            sg, rg = S.solve_stairstep_poisson_3d(pg, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                                  p.dx, p.dy, p.dz, -1.0, iters, 100,
                                                  method=method)
            po = p.phi0()
            so, ro = orc.solve_stairstep_poisson_3d(po, rhs, p.solid, bg, *faces, p.xc, p.yc,
                                                    p.zc, p.dx, p.dy, p.dz, -1.0, iters, 100,
                                                    method)
            assert sg == so == iters
            fl = p.solid[1:-1, 1:-1, :] == 0
            if method == "jacobi":
                _assert_constant_shift(pg, po, fl, np.abs(po).max())
            else:
                assert np.abs(pg - po).max() <= 1e-10 * np.abs(po).max()
                assert abs(rg - ro) <= 1e-8 * ro
        # device entry points on arrays resident in HBM
        T = DeviceArray.from_host
        dphi = T(p.phi0())
        sd, rd = S.solve_stairstep_poisson_3d(dphi, T(rhs), T(p.solid), bg, *[T(f) for f in faces],
                                              T(p.xc), T(p.yc), T(p.zc), T(p.dx), T(p.dy),
                                              T(p.dz), -1.0, 300, 100, method="pcg")
        ph = p.phi0()
        sh, rh = S.solve_stairstep_poisson_3d(ph, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc, p.dx,
                                              p.dy, p.dz, -1.0, 300, 100, method="pcg")
        assert sd == sh and rd == rh
        assert np.array_equal(dphi.download(), ph)


def test_error_paths_on_gpu():
    print("case: error paths", flush=True)
    c = cases.exact_case(False)
    s = c.solid.copy(order="F")
    s[3:6, 3:6, 2:5] = 1
    s[4, 4, 3] = 0
    faces = _faces(c)
    with pytest.raises(FFPMError, match="no fluid neighbour"):
        S.solve_stairstep_poisson_3d(np.array(c.phi, order="F"), c.rhs, s, c.body, *faces, c.xc,
                                     c.yc, c.zc, c.dx, c.dy, c.dz, 1e-8, 10, 5)
    s2 = np.ones_like(c.solid, order="F")
    S.exchange_solid_halo_3d(s2)
    assert np.all(s2[0] == 0) and np.all(s2[:, -1] == 0) and np.all(s2[1:-1, 1:-1] == 1)
    with pytest.raises(FFPMError, match="outer face"):
        S.enforce_stairstep_compatibility_3d(c.rhs, s2, c.body, *faces, c.dx, c.dy, c.dz)
    with pytest.raises(FFPMError, match="pcg or jacobi"):
        S.solve_stairstep_poisson_3d(np.array(c.phi, order="F"), c.rhs, c.solid, c.body, *faces,
                                     c.xc, c.yc, c.zc, c.dx, c.dy, c.dz, 1e-8, 10, 5,
                                     method="sor")


def test_breakdown_restart_and_zero_problem(orc):
    """All-zero data: the true residual is 0, so rz = 0 and PCG breaks down,
    restarts once and leaves (src/fpm_stairstep_mpi_3d.f90:259-265); an exact
    initial guess converges with no iteration."""
    print("case: breakdown paths", flush=True)
    c = cases.exact_case(False)
    zero = tuple(np.zeros_like(f, order="F") for f in c.faces)
    body = np.zeros(3)
    for tol in (-1.0, 0.0):
        pg = np.zeros_like(c.phi, order="F")
        sg, rg = S.solve_stairstep_poisson_3d(pg, c.rhs, c.solid, body, *zero, c.xc, c.yc, c.zc,
                                              c.dx, c.dy, c.dz, tol, 50, 10)
        po = np.zeros_like(c.phi, order="F")
        so, ro = orc.solve_stairstep_poisson_3d(po, c.rhs, c.solid, body, *zero, c.xc, c.yc,
                                                c.zc, c.dx, c.dy, c.dz, tol, 50, 10, "pcg")
        assert sg == so == 0 and rg == ro == 0.0
        assert np.all(pg == 0)
    # exact case starting from its converged field
    e = cases.exact_case(True)
    faces = _faces(e)
    orc.enforce_stairstep_compatibility_3d(e.rhs, e.solid, e.body, *faces, e.dx, e.dy, e.dz)
    po, so, ro = _orc_solve(orc, e, faces, 1e-10, 100000, 50, "pcg")
    e.phi = po
    pg, sg, rg = _gpu_solve(e, faces, 1e-10, 100000, 50, "pcg")
    e2 = cases.exact_case(True)
    e2.phi = po
    po2, so2, ro2 = _orc_solve(orc, e2, faces, 1e-10, 100000, 50, "pcg")
    assert sg == so2 and rg <= 1e-10


def test_max_sweeps_zero_and_check_interval_zero(orc):
    print("case: max_sweeps 0 / check_interval 0", flush=True)
    c = cases.exact_case(False)
    faces = _faces(c)
    orc.enforce_stairstep_compatibility_3d(c.rhs, c.solid, c.body, *faces, c.dx, c.dy, c.dz)
    for method in ("pcg", "jacobi"):
        pg, sg, rg = _gpu_solve(c, faces, 1e-8, 0, 50, method)
        po, so, ro = _orc_solve(orc, c, faces, 1e-8, 0, 50, method)
        assert sg == so == 0 and abs(rg - ro) <= 1e-14 * ro
        pg, sg, rg = _gpu_solve(c, faces, -1.0, 13, 0, method)
        po, so, ro = _orc_solve(orc, c, faces, -1.0, 13, 0, method)
        assert sg == so == 13
        assert np.abs(pg - po).max() <= 1e-12 * np.abs(po).max()


def test_fused_and_three_kernel_pcg_agree(orc, monkeypatch):
    """The fused two-kernel iteration and the three-kernel iteration run the
    same formulas; at fixed work they agree to summation-order roundoff, and
    both match the oracle's converged answer and iteration count."""
    p = W.frame_128x96x64(with_fields=False)
    rhs, bg, faces = W.influence_inputs(p, 1)
    orc.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
    out = {}
    for variant in ("classic", "fused"):
        print(f"case: 128x96x64 direction y, {variant} PCG", flush=True)
        monkeypatch.setenv("FFPM_PCG", variant)
        ph = p.phi0()
        sw, res = S.solve_stairstep_poisson_3d(ph, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                               p.dx, p.dy, p.dz, -1.0, 250, 50, method="pcg")
        out[variant] = (ph, sw, res)
    a, b = out["classic"], out["fused"]
    assert a[1] == b[1] == 250
    assert np.abs(a[0] - b[0]).max() <= 1e-12 * np.abs(a[0]).max()
    assert abs(a[2] - b[2]) <= 1e-9 * a[2]
    po = p.phi0()
    so, ro = orc.solve_stairstep_poisson_3d(po, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                            p.dx, p.dy, p.dz, 1e-10, 100000, 100, "pcg")
    for variant in ("classic", "fused"):
        monkeypatch.setenv("FFPM_PCG", variant)
        ph = p.phi0()
        sw, res = S.solve_stairstep_poisson_3d(ph, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                               p.dx, p.dy, p.dz, 1e-10, 100000, 100,
                                               method="pcg")
        assert sw == so and res <= 1e-10
        assert np.abs(ph - po).max() <= 1e-10 * np.abs(po).max()


@pytest.mark.parametrize("nx_mult", [4, 2, 1])
def test_mixed_precision_reaches_fp64_tolerance(orc, nx_mult):
    """FP32-inner / FP64-outer PCG stops on the FP64 max-norm residual of the
    reference's formula; its field matches the FP64 oracle
    and the densities follow: within 1e-10 relative max-norm, the bound for two
    solves converged to the same residual."""
    grids = {4: (48, 36, 24), 2: (46, 36, 24), 1: (45, 35, 23)}
    nx, ny, nz = grids[nx_mult]
    print(f"case: mixed precision on {nx}x{ny}x{nz}", flush=True)
    p = W.stretched_frame(nx, ny, nz, [W.Sphere((0.0, 0.0, 0.0), 0.9),
                                       W.Box((3.0, -1.0, -0.7), (4.5, 1.0, 0.7))], seed=7)
    rhs, bg, faces = W.influence_inputs(p, 0)
    orc.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
    pm = p.phi0()
    sm, rm = S.solve_stairstep_poisson_3d(pm, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc, p.dx,
                                          p.dy, p.dz, 1e-10, 100000, 25, method="pcg-mixed")
    outer, inner = S.last_mixed_solve()
    po = p.phi0()
    so, ro = orc.solve_stairstep_poisson_3d(po, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                            p.dx, p.dy, p.dz, 1e-10, 100000, 25, "pcg")
    assert rm <= 1e-10 and ro <= 1e-10
    assert outer >= 2 and inner == sm > 0
    rel = np.abs(pm - po).max() / np.abs(po).max()
    print(f"mixed: {outer} corrections, {inner} FP32 iterations (FP64 PCG {so}); "
          f"max rel |phi - phi_FP64| = {rel:.2e}", flush=True)
    assert rel <= 1e-10
    f = p.fields
    args = lambda ph: (ph, f["u"], f["v"], f["w"], f["ox"], f["oy"], f["oz"], f["ax"], f["ay"],
                       f["az"], p.solid, bg, p.xc, p.yc, p.zc, p.dx, p.dy, p.dz, *faces, p.bounds)
    gm = S.compute_stairstep_fpm_terms_3d(*args(pm))
    go = orc.compute_stairstep_fpm_terms_3d(*args(po))
    for a, b in zip(gm[3:5], go[3:5]):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()



@pytest.mark.parametrize("grid", [(130, 37, 19), (45, 35, 23)])
def test_streaming_jacobi_matches_one_cell_kernel(monkeypatch, grid):
    """The paired-cell streaming Jacobi kernel (rdiag rebuilt from the
    neighbour code) is bit-identical to the one-cell kernel with the stored
    rdiag, for even and odd nx, at fixed work."""
    nx, ny, nz = grid
    print(f"case: streaming Jacobi {nx}x{ny}x{nz}", flush=True)
    p = W.stretched_frame(nx, ny, nz, [W.Sphere((0.0, 0.0, 0.0), 0.9)], seed=5,
                          with_fields=False)
    rhs, bg, faces = W.influence_inputs(p, 1)
    S.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
    out = {}
    for variant in ("classic", "streaming"):
        monkeypatch.setenv("FFPM_JACOBI", variant)
        ph = p.phi0()
        sw, res = S.solve_stairstep_poisson_3d(ph, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                               p.dx, p.dy, p.dz, -1.0, 301, 100,
                                               method="jacobi")
        out[variant] = (ph, sw, res)
    a, b = out["classic"], out["streaming"]
    assert a[1] == b[1] == 301 and a[2] == b[2]
    assert np.array_equal(a[0], b[0])


@pytest.mark.parametrize("nd", [1, 2, 3])
def test_terms_of_several_directions_in_one_pass(nd):
    """The directions' densities from one pass over the flow fields equal the
    per-direction compute_stairstep_fpm_terms_3d outputs and totals bit for
    bit (random potentials and outer face data per direction, a body touching
    the outer boundary, integration bounds cutting the domain)."""
    print(f"case: densities of {nd} directions in one pass, 40x30x20", flush=True)
    T = DeviceArray.from_host
    p = W.stretched_frame(40, 30, 20, [W.Sphere((0.0, 0.0, 0.0), 0.9),
                                       W.Box((3.0, -1.0, -0.7), (4.5, 1.0, 0.7))], seed=5)
    rng = np.random.default_rng(31)
    solid, axes = T(p.solid), [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
    fl = [T(p.fields[k]) for k in ("u", "v", "w", "ox", "oy", "oz", "ax", "ay", "az")]
    bounds = np.array([-2.0, 3.0, -1.5, 2.0, -1.0, 1.5])
    phis, faces, bgs = [], [], []
    for d in range(nd):
        phis.append(T(np.asfortranarray(rng.standard_normal(p.solid.shape))))
        faces.append([T(np.asfortranarray(rng.standard_normal(f.shape))) for f in p.faces()])
        bgs.append(np.eye(3)[d] * (1.0 + 0.5 * d))
    one = [S.compute_stairstep_fpm_terms_3d(phis[d], *fl, solid, bgs[d], *axes, *faces[d],
                                            bounds) for d in range(nd)]
    many = S.compute_stairstep_fpm_terms_3d_dirs(phis, *fl, solid, bgs, *axes, faces, bounds)
    assert len(many) == nd
    for d in range(nd):
        for a, b in zip(one[d][:5], many[d][:5]):
            assert np.array_equal(a.download(), b.download())
        assert np.array_equal(one[d][5], many[d][5])
        assert np.all(np.isfinite(many[d][5])) and np.any(many[d][5] != 0)
