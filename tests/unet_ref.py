"""CPU FP64 restatement of the U-Net for the gradient checks (test
infrastructure): torch float64 on the CPU, parameters in the library's flat
layout (conv weights [27][Cout][Cin] with tap t = (dz+1)*9 + (dy+1)*3 +
(dx+1), then the biases, then the 1x1 head's 16 weights and bias).

With ``emulate=True`` the forward rounds the conv weights, the input and every
stored activation to BF16 at the points where the HIP kernels store BF16, so
the forward is comparable to the GPU's to accumulation-order roundoff; the
gradient then flows straight through the rounding.
"""
import numpy as np
import torch
import torch.nn.functional as F

SPEC = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32),
        (48, 16), (16, 16)]


def layout():
    off, w_off, b_off = 0, [], []
    for cin, cout in SPEC:
        w_off.append(off)
        off += 27 * cin * cout
    for cin, cout in SPEC:
        b_off.append(off)
        off += cout
    head_w = off
    off += 16
    head_b = off
    off += 1
    return w_off, b_off, head_w, head_b, off


W_OFF, B_OFF, HEAD_W, HEAD_B, TOTAL = layout()


class _Round(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g


def _r(x, emulate):
    return _Round.apply(x) if emulate else x


def unpack(flat, emulate):
    """torch float64 leaf tensors (requires_grad) for every parameter."""
    flat = np.asarray(flat, dtype=np.float32)
    ps = []
    for l, (cin, cout) in enumerate(SPEC):
        w = flat[W_OFF[l]:W_OFF[l] + 27 * cin * cout].reshape(27, cout, cin)
        w = torch.tensor(np.transpose(w, (1, 2, 0)).reshape(cout, cin, 3, 3, 3), dtype=torch.float64)
        b = torch.tensor(flat[B_OFF[l]:B_OFF[l] + cout], dtype=torch.float64)
        ps.append([w.requires_grad_(), b.requires_grad_()])
    hw = torch.tensor(flat[HEAD_W:HEAD_W + 16], dtype=torch.float64).requires_grad_()
    hb = torch.tensor(flat[HEAD_B:HEAD_B + 1], dtype=torch.float64).requires_grad_()
    return ps, hw, hb


def forward(ps, hw, hb, feats, source, emulate=True):
    """feats (8, nx, ny, nz), source (nx, ny, nz) -> output (nx, ny, nz)."""
    nx, ny, nz = source.shape
    x = np.zeros((16, nx, ny, nz))
    x[:8] = feats
    x[8] = source
    t = torch.tensor(np.transpose(x, (0, 3, 2, 1))[None].copy(), dtype=torch.float64)
    t = _r(t, emulate)

    def conv(h, l):
        w, b = ps[l]
        w = _r(w, emulate)
        return _r(F.relu(F.conv3d(h, w, b, padding=1)), emulate)

    def pool(h):
        return _r(F.avg_pool3d(h, 2), emulate)

    def up(h):
        return F.interpolate(h, scale_factor=2, mode="nearest")

    a0 = conv(t, 0)
    s0 = conv(a0, 1)
    a1 = conv(pool(s0), 2)
    s1 = conv(a1, 3)
    a2 = conv(pool(s1), 4)
    b2 = conv(a2, 5)
    a1c = conv(torch.cat([up(b2), s1], 1), 6)
    d1 = conv(a1c, 7)
    a0c = conv(torch.cat([up(d1), s0], 1), 8)
    d0 = conv(a0c, 9)
    out = (d0[0] * hw[:, None, None, None]).sum(0) + hb[0]
    return out.permute(2, 1, 0)  # (nx, ny, nz)


def loss_and_grads(flat, feats, source, target, mask, emulate=True):
    ps, hw, hb = unpack(flat, emulate)
    out = forward(ps, hw, hb, feats, source, emulate)
    m = torch.tensor(mask, dtype=torch.float64)
    tg = torch.tensor(target, dtype=torch.float64)
    loss = (m * (out - tg) ** 2).sum() / m.sum()
    loss.backward()
    g = np.zeros(TOTAL)
    for l, (cin, cout) in enumerate(SPEC):
        w, b = ps[l]
        gw = w.grad.numpy().reshape(cout, cin, 27)
        g[W_OFF[l]:W_OFF[l] + 27 * cin * cout] = np.transpose(gw, (2, 0, 1)).ravel()
        g[B_OFF[l]:B_OFF[l] + cout] = b.grad.numpy()
    g[HEAD_W:HEAD_W + 16] = hw.grad.numpy()
    g[HEAD_B] = hb.grad.numpy()[0]
    return float(loss.detach()), out.detach().numpy(), g
