"""CPU checks of the FP64 U-Net restatement used as the gradient reference:
its autograd gradients against central finite differences, and its layout
against a direct 3x3x3 convolution sum (so a transposed tap or channel order
in the restatement fails here)."""
import numpy as np
import torch

import unet_ref as R


def _inputs(rng, nx=16, ny=8, nz=8):
    feats = rng.standard_normal((8, nx, ny, nz))
    source = rng.standard_normal((nx, ny, nz))
    target = rng.standard_normal((nx, ny, nz))
    mask = (rng.uniform(size=(nx, ny, nz)) < 0.8).astype(np.float64)
    return feats, source, target, mask


def _params(rng):
    flat = np.zeros(R.TOTAL, dtype=np.float32)
    for l, (cin, cout) in enumerate(R.SPEC):
        n = 27 * cin * cout
        flat[R.W_OFF[l]:R.W_OFF[l] + n] = rng.uniform(-1, 1, n) * np.sqrt(6 / (27 * cin))
        flat[R.B_OFF[l]:R.B_OFF[l] + cout] = rng.uniform(-0.1, 0.1, cout)
    flat[R.HEAD_W:R.HEAD_W + 16] = rng.uniform(-1, 1, 16)
    flat[R.HEAD_B] = 0.05
    return flat


def test_reference_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    feats, source, target, mask = _inputs(rng)
    flat = _params(rng).astype(np.float64)
    _, _, g = R.loss_and_grads(flat, feats, source, target, mask, emulate=False)

    # float64 finite differences need float64 parameters end to end
    def loss64(f):
        ps, hw, hb = [], None, None
        for l, (cin, cout) in enumerate(R.SPEC):
            w = f[R.W_OFF[l]:R.W_OFF[l] + 27 * cin * cout].reshape(27, cout, cin)
            w = torch.tensor(np.transpose(w, (1, 2, 0)).reshape(cout, cin, 3, 3, 3))
            ps.append([w, torch.tensor(f[R.B_OFF[l]:R.B_OFF[l] + cout])])
        hw = torch.tensor(f[R.HEAD_W:R.HEAD_W + 16])
        hb = torch.tensor(f[R.HEAD_B:R.HEAD_B + 1])
        out = R.forward(ps, hw, hb, feats, source, emulate=False)
        m = torch.tensor(mask)
        return float((m * (out - torch.tensor(target)) ** 2).sum() / m.sum())

    picks = [R.W_OFF[0] + 5, R.W_OFF[4] + 1234, R.W_OFF[6] + 777, R.W_OFF[9] + 42,
             R.B_OFF[3] + 7, R.B_OFF[8], R.HEAD_W + 3, R.HEAD_B]
    f0 = flat.astype(np.float64)  # float32 values, exactly representable
    for idx in picks:
        h = 1e-6 * max(1.0, abs(f0[idx]))
        fp, fm = f0.copy(), f0.copy()
        fp[idx] += h
        fm[idx] -= h
        fd = (loss64(fp) - loss64(fm)) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)), (idx, fd, g[idx])


def test_reference_layout_matches_direct_convolution():
    """First layer output at one cell from an explicit sum over taps and
    channels with the library's weight layout."""
    rng = np.random.default_rng(1)
    feats, source, _, _ = _inputs(rng)
    flat = _params(rng)
    ps, hw, hb = R.unpack(flat, False)
    nx, ny, nz = source.shape
    x = np.zeros((16, nx, ny, nz))
    x[:8], x[8] = feats, source
    t = torch.tensor(np.transpose(x, (0, 3, 2, 1))[None].copy())
    w, b = ps[0]
    y = torch.nn.functional.conv3d(t, w.detach(), b.detach(), padding=1)[0].numpy()
    i, j, k, co = 5, 3, 4, 7
    wl = flat[R.W_OFF[0]:R.W_OFF[0] + 27 * 256].reshape(27, 16, 16).astype(np.float64)
    s = flat[R.B_OFF[0] + co]
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)
                s += wl[tap, co] @ x[:, i + dx, j + dy, k + dz]
    assert abs(y[co, k, j, i] - s) <= 1e-10 * max(1.0, abs(s))
