"""GPU tests of the U-Net kernels (BF16 MFMA, FP32 accumulation) against the
CPU FP64 restatement (tests/unet_ref.py), of training, and of the
network-preconditioned flexible CG.

Tolerances (BF16): the reference rounds weights, input and stored
activations to BF16 exactly where the kernels do, so the forward agrees to
FP32-vs-FP64 accumulation roundoff (relative L2 <= 1e-2); the GPU also stores
gradient tensors in BF16, so each parameter tensor's gradient agrees to
relative L2 <= 5e-2 (cosine >= 0.998). Adam agrees with a numpy Adam step on
the GPU's own gradients to FP32 roundoff.
"""
import os

import numpy as np
import pytest

import unet_ref as R
from flowfpmpp_amd import learned as L
from flowfpmpp_amd import stairstep as S
from flowfpmpp_amd import workloads as W
from flowfpmpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu
T = DeviceArray.from_host


def _sample(nx=32, ny=24, nz=16, seed=0):
    rng = np.random.default_rng(seed)
    p = W.stretched_frame(nx, ny, nz, [W.Sphere((0.0, 0.0, 0.0), 1.0)], seed=seed,
                          with_fields=False)
    feats = L.features(p, 1)
    source = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    target = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    mask = feats[0].copy()
    return feats, source, target, mask


def test_param_count_and_roundtrip():
    print("case: unet parameters", flush=True)
    net = L.UNet(seed=3)
    assert net.nparams == R.TOTAL
    a = net.get_params()
    assert np.isfinite(a).all() and np.abs(a).max() > 0
    b = np.random.default_rng(0).standard_normal(R.TOTAL).astype(np.float32) * 0.1
    net.set_params(b)
    assert np.array_equal(net.get_params(), b)


def test_forward_matches_cpu_fp64():
    print("case: unet forward 32x24x16 vs CPU", flush=True)
    feats, source, _, _ = _sample()
    net = L.UNet(seed=5)
    flat = net.get_params()
    out = DeviceArray(source.shape, np.float32)
    net.forward(T(feats), T(source), out)
    g = out.download().astype(np.float64)
    ps, hw, hb = R.unpack(flat, True)
    ref = R.forward(ps, hw, hb, feats, source, emulate=True).detach().numpy()
    rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    print(f"forward relative L2 error {rel:.2e}")
    assert rel <= 1e-2


def test_inference_and_training_forward_identical():
    """Inference forms the decoder's skip concatenations while staging each
    convolution's input; training materialises them for the backward pass.
    Both feed the same bits to the same MFMA sequence: outputs are identical."""
    print("case: unet forward, fused vs materialised concatenation, 64x48x32", flush=True)
    feats, source, target, mask = _sample(64, 48, 32, seed=4)
    net = L.UNet(seed=9)
    out_a = DeviceArray(source.shape, np.float32)
    net.forward(T(feats), T(source), out_a)
    a = out_a.download()
    net.train_step(T(feats), T(source), T(target), T(mask), 0.0)  # lr 0: weights unchanged
    out_b = DeviceArray(source.shape, np.float32)
    net.forward(T(feats), T(source), out_b)
    assert np.array_equal(a, out_b.download())
    assert np.isfinite(a).all() and np.abs(a).max() > 0


def test_gradients_match_cpu_fp64():
    print("case: unet gradients 32x24x16 vs CPU", flush=True)
    feats, source, target, mask = _sample(seed=1)
    net = L.UNet(seed=7)
    flat = net.get_params()
    loss = net.train_step(T(feats), T(source), T(target), T(mask), 0.0)  # gradients only
    g = net.get_grads().astype(np.float64)
    rloss, _, rg = R.loss_and_grads(flat, feats, source, target, mask, emulate=True)
    assert abs(loss - rloss) <= 1e-2 * rloss
    worst = 0.0
    for l, (cin, cout) in enumerate(R.SPEC):
        for a, b in ((R.W_OFF[l], R.W_OFF[l] + 27 * cin * cout), (R.B_OFF[l], R.B_OFF[l] + cout)):
            x, y = g[a:b], rg[a:b]
            rel = np.linalg.norm(x - y) / np.linalg.norm(y)
            cos = x @ y / (np.linalg.norm(x) * np.linalg.norm(y))
            worst = max(worst, rel)
            assert rel <= 5e-2 and cos >= 0.998, (l, a, rel, cos)
    hx, hy = g[R.HEAD_W:R.TOTAL], rg[R.HEAD_W:R.TOTAL]
    assert np.linalg.norm(hx - hy) <= 1e-2 * np.linalg.norm(hy)
    print(f"worst per-tensor gradient relative L2 error {worst:.2e}")


def test_adam_step_matches_numpy():
    print("case: unet Adam", flush=True)
    feats, source, target, mask = _sample(seed=2)
    net = L.UNet(seed=9)
    p0 = net.get_params().astype(np.float64)
    lr = 1e-3
    net.train_step(T(feats), T(source), T(target), T(mask), lr)
    g = net.get_grads().astype(np.float64)
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p0 - lr * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    p1 = net.get_params().astype(np.float64)
    assert np.abs(p1 - ref).max() <= 1e-6 + 1e-5 * np.abs(ref).max()


def test_training_reduces_loss():
    """A short run on solver-produced pairs (two 32^3 frames, three
    directions, iterates after 5 and 20 iterations) lowers the loss."""
    print("case: unet training 32^3", flush=True)
    frames = L.training_frames(2, (32, 32, 32), seed0=7000)
    pairs = L.make_pairs(frames, snapshots=(5, 20))
    assert len(pairs) == 2 * 3 * 3
    net = L.UNet(seed=11)
    hist = L.train(net, pairs, 300, lr0=2e-3, lr1=2e-4)
    first, last = np.mean(hist[:30]), np.mean(hist[-30:])
    print(f"loss {first:.3e} -> {last:.3e}")
    assert np.isfinite(hist).all() and last < 0.5 * first


def _frame_problem(nx, ny, nz, seed, direction=0):
    rng = np.random.Generator(np.random.PCG64(seed))
    p = W.stretched_frame(nx, ny, nz, L.random_bodies(rng), seed=seed, with_fields=False)
    rhs, bg, faces = W.influence_inputs(p, direction)
    solid = T(p.solid)
    axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
    d_rhs, d_faces = T(rhs), [T(f) for f in faces]
    S.enforce_stairstep_compatibility_3d(d_rhs, solid, bg, *d_faces, *axes[3:])
    return p, solid, axes, d_rhs, bg, d_faces


def test_fcg_untrained_network_reaches_tolerance():
    """Whatever the network, FCG ends at the FP64 tolerance (switching to
    diagonal PCG when it stalls) and agrees with plain PCG's phi."""
    print("case: FCG with an untrained network, 48x32x32", flush=True)
    p, solid, axes, rhs, bg, faces = _frame_problem(48, 32, 32, seed=8100)
    net = L.UNet(seed=13)
    feats = T(L.features(p, 0))
    phi = T(p.phi0())
    sw, res, fb = L.solve_fcg(phi, rhs, solid, bg, *faces, *axes, 1e-10, 100000, 25, net, feats)
    ref = T(p.phi0())
    sw0, res0 = S.solve_stairstep_poisson_3d(ref, rhs, solid, bg, *faces, *axes, 1e-10, 100000,
                                             25, method="pcg")
    print(f"FCG {sw} iterations (fallback at {fb}), PCG {sw0}")
    assert res <= 1e-10 and res0 <= 1e-10
    a, b = phi.download(), ref.download()
    assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max()


def _a_norm_error_after_step(solid, axes, e, n):
    """||e - g n||_A^2 / ||e||_A^2 for the optimal step g along n (A e = r)."""
    nx, ny, nz = e.shape
    r, an = DeviceArray((nx, ny, nz)), DeviceArray((nx, ny, nz))
    L.apply_operator(solid, *axes, T(e), r)
    L.apply_operator(solid, *axes, T(n), an)
    r, an = r.download(), an.download()
    return 1.0 - (n.ravel() @ r.ravel()) ** 2 / ((n.ravel() @ an.ravel()) * (e.ravel() @ r.ravel()))


@pytest.mark.skipif(not os.path.exists(L.WEIGHTS), reason="no trained weights in the tree")
def test_trained_network_ablation():
    """What the trained network gains, on unseen frames (64^3, all three
    directions): as a correction of the exact error e = phi* (the initial
    guess from phi = 0), the A-norm error left after the optimal step along
    the trained network's output is smaller than along the untrained
    network's. The flexible-CG iteration counts to 1e-10 with either network
    and plain PCG's are printed; all three reach the FP64 tolerance."""
    print("case: trained vs untrained network, unseen 64^3 frames", flush=True)
    trained, untrained = L.UNet.load(), L.UNet(seed=1)
    gains = []
    for seed in (9100, 9101):
        for d in range(3):
            p, solid, axes, rhs, bg, faces = _frame_problem(64, 64, 64, seed=seed, direction=d)
            feats = T(L.features(p, d))
            ref = T(p.phi0())
            sw0, res0 = S.solve_stairstep_poisson_3d(ref, rhs, solid, bg, *faces, *axes, 1e-12,
                                                     100000, 25, method="pcg")
            nx, ny, nz = p.nx, p.ny, p.nz
            fluid = p.solid[1:nx + 1, 1:ny + 1, :] == 0
            vol = p.dx[:, None, None] * p.dy[None, :, None] * p.dz[None, None, :]
            e = ref.download()[1:nx + 1, 1:ny + 1, :] * fluid
            r = DeviceArray((nx, ny, nz))
            L.apply_operator(solid, *axes, T(np.asfortranarray(e)), r)
            rv = r.download() / vol * fluid
            src = T((rv / np.abs(rv).max()).astype(np.float32))
            err = {}
            for name, net in (("trained", trained), ("untrained", untrained)):
                out = DeviceArray((nx, ny, nz), np.float32)
                net.forward(feats, src, out)
                n = out.download().astype(np.float64) * fluid
                err[name] = _a_norm_error_after_step(solid, axes, np.asfortranarray(e),
                                                     np.asfortranarray(n))
            print(f"seed {seed} direction {d}: A-norm error after one step "
                  f"trained {err['trained']:.3f}, untrained {err['untrained']:.3f}")
            gains.append(err["trained"] < err["untrained"])
    assert all(gains)
    p, solid, axes, rhs, bg, faces = _frame_problem(64, 64, 64, seed=9100, direction=2)
    feats = T(L.features(p, 2))
    its = {}
    for name, net in (("trained", trained), ("untrained", untrained)):
        phi = T(p.phi0())
        its[name] = L.solve_fcg(phi, rhs, solid, bg, *faces, *axes, 1e-10, 100000, 25, net, feats)
    ref = T(p.phi0())
    sw0, res0 = S.solve_stairstep_poisson_3d(ref, rhs, solid, bg, *faces, *axes, 1e-10, 100000,
                                             25, method="pcg")
    print(f"iterations to 1e-10: FCG trained {its['trained']}, FCG untrained "
          f"{its['untrained']}, PCG {sw0}")
    assert all(v[1] <= 1e-10 for v in its.values()) and res0 <= 1e-10


@pytest.mark.skipif(not os.path.exists(L.WEIGHTS), reason="no trained weights in the tree")
def test_network_initial_guess_for_the_multigrid_cg():
    """What the trained network gains for the fastest method: the multigrid CG
    ("fcg-mg") started from the network's optimal first step (phi_1 = g n with
    n = net(A phi* / (V sigma)), the step the flexible CG takes from phi = 0)
    against the same solve from phi = 0, on unseen 64^3 frames. Both reach the
    FP64 tolerance; the iteration counts are printed (the network saves at
    most a couple of the ~20-30 iterations)."""
    print("case: network initial guess + fcg-mg, unseen 64^3 frames", flush=True)
    trained = L.UNet.load()
    for seed in (9100, 9101):
        p, solid, axes, rhs, bg, faces = _frame_problem(64, 64, 64, seed=seed, direction=0)
        nx, ny, nz = p.nx, p.ny, p.nz
        fluid = p.solid[1:nx + 1, 1:ny + 1, :] == 0
        vol = p.dx[:, None, None] * p.dy[None, :, None] * p.dz[None, None, :]
        ref = T(p.phi0())
        S.solve_stairstep_poisson_3d(ref, rhs, solid, bg, *faces, *axes, 1e-12, 100000, 25,
                                     method="pcg")
        e = np.asfortranarray(ref.download()[1:nx + 1, 1:ny + 1, :] * fluid)
        r = DeviceArray((nx, ny, nz))
        L.apply_operator(solid, *axes, T(e), r)
        rv = r.download() / vol * fluid
        out = DeviceArray((nx, ny, nz), np.float32)
        trained.forward(T(L.features(p, 0)), T((rv / np.abs(rv).max()).astype(np.float32)), out)
        n = np.asfortranarray(out.download().astype(np.float64) * fluid)
        an = DeviceArray((nx, ny, nz))
        L.apply_operator(solid, *axes, T(n), an)
        g = float(n.ravel() @ r.download().ravel()) / float(n.ravel() @ an.download().ravel())
        its = {}
        for name, start in (("zero", np.zeros_like(n)), ("network", g * n)):
            phi0 = p.phi0()
            phi0[1:nx + 1, 1:ny + 1, :] = start
            phi = T(phi0)
            sw, res = S.solve_stairstep_poisson_3d(phi, rhs, solid, bg, *faces, *axes, 1e-10,
                                                   100000, 25, method="fcg-mg")
            its[name] = (sw, res)
        print(f"seed {seed}: fcg-mg from phi = 0 {its['zero'][0]} iterations, from the network's "
              f"first step {its['network'][0]}", flush=True)
        assert its["zero"][1] <= 1e-10 and its["network"][1] <= 1e-10
