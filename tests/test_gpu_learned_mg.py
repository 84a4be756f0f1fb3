"""The multigrid CG with the learned coarse correction (no counterpart in the
reference), on the GPU: on one level of the cycle the U-Net either takes
over the recursion below it ("replace") or corrects what the recursion
leaves ("augment").

Bars: the result is judged like every solve here, by the reference's FP64
max-norm residual (src/fpm_stairstep_mpi_3d.f90:514-530): it reaches 1e-10,
with the diagonal-PCG fallback if the learned cycle stalls, and the field
agrees with the plain multigrid CG's to 1e-10 relative max-norm (two solves
converged to the same residual). On unseen bodies at 256x192x128 the trained
network is compared with the untrained one (the ablation) and with the plain
cycle; the iteration counts and times are printed."""
import time

import numpy as np
import pytest

from flowfpmpp_amd import learned as L
from flowfpmpp_amd import stairstep as S
from flowfpmpp_amd import workloads as W
from flowfpmpp_amd._lib import FFPMError
from flowfpmpp_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu
T = DeviceArray.from_host


def _device_problem(p, d):
    rhs, bg, faces = W.influence_inputs(p, d)
    solid = T(p.solid)
    axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
    d_rhs, d_faces = T(rhs), [T(f) for f in faces]
    S.enforce_stairstep_compatibility_3d(d_rhs, solid, bg, *d_faces, *axes[3:])
    return d_rhs, solid, bg, d_faces, axes


def _solve_pair(p, d, net, level, mode="augment", tol=1e-10):
    d_rhs, solid, bg, d_faces, axes = _device_problem(p, d)
    feats = T(L.level_features(p, level, d))
    out = {}
    for name in ("fcg-mg", "learned"):
        phi = T(p.phi0())
        synchronize()
        t0 = time.perf_counter()
        if name == "learned":
            sw, res, fb = L.solve_fcg_mg_learned(phi, d_rhs, solid, bg, *d_faces, *axes, tol,
                                                 20000, 100, net, feats, level, mode)
        else:
            sw, res = S.solve_stairstep_poisson_3d(phi, d_rhs, solid, bg, *d_faces, *axes, tol,
                                                   20000, 100, method="fcg-mg")
            fb = -1
        synchronize()
        out[name] = (phi.download(), sw, res, fb, (time.perf_counter() - t0) * 1e3)
    return out


def test_learned_cycle_untrained_network_reaches_tolerance():
    p = W.stretched_frame(64, 64, 64, [W.Sphere((0.5, 0.0, 0.0), 0.8),
                                       W.Box((-2.5, -1.0, -0.5), (-1.5, 0.5, 0.6))], seed=21,
                          with_fields=False)
    level = L.learned_level(64, 64, 64)
    assert level == 1
    net = L.UNet(seed=4)
    for mode in ("augment", "replace"):
        for d in range(3):
            print(f"case: learned coarse correction ({mode}), untrained network, 64^3 "
                  f"direction {d}", flush=True)
            o = _solve_pair(p, d, net, level, mode)
            (a, sa, ra, fa, _), (b, sb, rb, _, _) = o["learned"], o["fcg-mg"]
            rel = np.abs(a - b).max() / np.abs(b).max()
            print(f"  learned cycle {sa} iterations (fallback at {fa}), residual {ra:.2e}; "
                  f"plain cycle {sb} ({rb:.2e}); rel |phi| {rel:.2e}", flush=True)
            assert ra <= 1e-10 and rb <= 1e-10
            assert rel <= 1e-10


def test_learned_cycle_argument_checks():
    p = W.stretched_frame(64, 64, 64, [W.Sphere((0.5, 0.0, 0.0), 0.8)], seed=22,
                          with_fields=False)
    d_rhs, solid, bg, d_faces, axes = _device_problem(p, 0)
    net = L.UNet(seed=4)
    feats = T(L.level_features(p, 1, 0))
    with pytest.raises(ValueError):  # features of another level
        L.solve_fcg_mg_learned(T(p.phi0()), d_rhs, solid, bg, *d_faces, *axes, 1e-10, 100, 100,
                               net, feats, 2)
    coarsest = len(L.mg_levels(64, 64, 64)) - 1
    bad = T(np.zeros((8,) + L.mg_levels(64, 64, 64)[coarsest], dtype=np.float32, order="F"))
    with pytest.raises(FFPMError):  # the coarsest level has nothing below it to replace
        L.solve_fcg_mg_learned(T(p.phi0()), d_rhs, solid, bg, *d_faces, *axes, 1e-10, 100, 100,
                               net, bad, coarsest)


def _one_step_a_norm_left(solid, axes, e, n):
    """||e - g n||_A^2 / ||e||_A^2 for the A-optimal step g along n."""
    nx, ny, nz = e.shape
    r, an = DeviceArray((nx, ny, nz)), DeviceArray((nx, ny, nz))
    L.apply_operator(solid, *axes, T(e), r)
    L.apply_operator(solid, *axes, T(n), an)
    r, an = r.download(), an.download()
    return 1.0 - (n.ravel() @ r.ravel()) ** 2 / ((n.ravel() @ an.ravel()) * (e.ravel() @ r.ravel()))


def test_trained_coarse_correction_on_unseen_bodies_256x192x128():
    """Unseen bodies (seeds 6000-6001; training used 5000-5666), all three
    directions at 256x192x128: the multigrid CG with the trained network in
    the coarse correction of level 1 (128x96x64), in both modes, and with the
    untrained network, reaches the FP64 tolerance (or falls back to diagonal
    PCG) and agrees with the plain cycle's field; iterations and times are
    printed."""
    trained = L.UNet.load()
    untrained = L.UNet(seed=1)
    level = L.learned_level(256, 192, 128)
    assert level == 1
    totals = {}
    for seed in (6000, 6001):
        rng = np.random.Generator(np.random.PCG64(seed))
        p = W.stretched_frame(256, 192, 128, L.random_bodies(rng), seed=seed, with_fields=False)
        for d in range(3):
            print(f"case: learned coarse correction, 256x192x128 seed {seed} direction {d}",
                  flush=True)
            d_rhs, solid, bg, d_faces, axes = _device_problem(p, d)
            feats = T(L.level_features(p, level, d))
            got = {}
            for name, net, mode in (("plain", None, None),
                                    ("augment trained", trained, "augment"),
                                    ("augment untrained", untrained, "augment"),
                                    ("replace trained", trained, "replace"),
                                    ("replace untrained", untrained, "replace")):
                phi = T(p.phi0())
                synchronize()
                t0 = time.perf_counter()
                if net is None:
                    sw, res = S.solve_stairstep_poisson_3d(phi, d_rhs, solid, bg, *d_faces,
                                                           *axes, 1e-10, 20000, 100,
                                                           method="fcg-mg")
                    fb = -1
                else:
                    sw, res, fb = L.solve_fcg_mg_learned(phi, d_rhs, solid, bg, *d_faces, *axes,
                                                         1e-10, 20000, 100, net, feats, level,
                                                         mode)
                synchronize()
                got[name] = (phi.download(), sw, res, fb, (time.perf_counter() - t0) * 1e3)
                totals[name] = totals.get(name, 0) + sw
            b = got["plain"][0]
            for name, (phi, sw, res, fb, ms) in got.items():
                rel = np.abs(phi - b).max() / np.abs(b).max()
                print(f"  {name}: {sw} iterations, {ms:.0f} ms, fallback {fb}, residual "
                      f"{res:.1e}, rel |phi - phi_plain| {rel:.1e}", flush=True)
                assert res <= 1e-10 and rel <= 1e-10
    print(f"iterations over the six solves: {totals}", flush=True)


def test_trained_network_gain_at_its_resolution():
    """The ablation: what training buys. On unseen 128x96x64 frames (the
    resolution of the level the network works on inside the 512x384x256 and
    256x192x128 cycles; seeds 6000-6001), for the error the multigrid CG
    leaves after one iteration and for the full solution (all three
    directions), the fraction of the squared A-norm error left after the
    A-optimal step along the trained network's output is smaller, on
    average, than along the untrained network's (the same architecture and
    initialisation seed before training)."""
    trained, untrained = L.UNet.load(), L.UNet(seed=1)
    left = {"trained": [], "untrained": []}
    for seed in (6000, 6001):
        rng = np.random.Generator(np.random.PCG64(seed))
        p = W.stretched_frame(128, 96, 64, L.random_bodies(rng), seed=seed, with_fields=False)
        nx, ny, nz = p.nx, p.ny, p.nz
        fluid = p.solid[1:nx + 1, 1:ny + 1, :] == 0
        vol = p.dx[:, None, None] * p.dy[None, :, None] * p.dz[None, None, :]
        for d in range(3):
            d_rhs, solid, bg, d_faces, axes = _device_problem(p, d)
            star = T(p.phi0())
            S.solve_stairstep_poisson_3d(star, d_rhs, solid, bg, *d_faces, *axes, 1e-10, 20000,
                                         100, method="fcg-mg")
            one = T(p.phi0())
            S.solve_stairstep_poisson_3d(one, d_rhs, solid, bg, *d_faces, *axes, -1.0, 1, 1,
                                         method="fcg-mg")
            feats = T(L.features(p, d))
            for k, x in (("solution", None), ("after 1 iteration", one)):
                e = star.download() if x is None else star.download() - x.download()
                e = e[1:nx + 1, 1:ny + 1, :] * fluid
                e = np.asfortranarray(e - e[fluid].mean() * fluid)
                r = DeviceArray((nx, ny, nz))
                L.apply_operator(solid, *axes, T(e), r)
                rv = r.download() / vol * fluid
                src = T((rv / np.abs(rv).max()).astype(np.float32))
                row = []
                for name, net in (("trained", trained), ("untrained", untrained)):
                    out = DeviceArray((nx, ny, nz), np.float32)
                    net.forward(feats, src, out)
                    n = np.asfortranarray(out.download().astype(np.float64) * fluid)
                    f = _one_step_a_norm_left(solid, axes, e, n)
                    left[name].append(f)
                    row.append(f"{name} {f:.3f}")
                print(f"seed {seed} direction {d}, {k}: A-norm error left after one step: "
                      + ", ".join(row), flush=True)
    mt, mu = np.mean(left["trained"]), np.mean(left["untrained"])
    print(f"mean A-norm error left: trained {mt:.3f}, untrained {mu:.3f}", flush=True)
    assert mt < mu
