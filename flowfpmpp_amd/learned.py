"""Learned preconditioner: the 3-D U-Net of the HIP library (BF16 MFMA
forward, backward and Adam kernels), its input features, training on fields
the plain solver produces, and the network-preconditioned flexible CG solve.

The network maps a residual r to an approximation of A^-1 r. It sees the
frame (fluid mask, signed distance, the three cell widths, the direction
one-hot) and the source density r / V scaled by sigma = max |r / V|; its
output (scaled by sigma) is the correction, whose length flexible CG's
optimal step sets. Training pairs come from the plain PCG
solver on random stair-step bodies: the converged phi* of each direction and
the errors e_k = phi* - phi_k of its iterates after k iterations, with the
exact residual r_k = A e_k (so the network learns A^-1 on the residuals FCG
meets, from smooth early ones to the late ones).

No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import check, lib
from .device import DeviceArray, synchronize

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "unet_weights.npz")


# ------------------------------------------------------------- features ------
def features(p, direction: int, base: np.ndarray = None) -> np.ndarray:
    """FP32 input features of frame p for force direction 0/1/2, shape
    (8, nx, ny, nz) in Fortran order (channels-last in memory): fluid mask,
    signed distance to the solid/fluid interface (positive in fluid, scaled by
    0.5 domain units and clipped to [-1, 2]), dx, dy, dz relative to their
    axis means, and the direction one-hot. ``base`` (the features of another
    direction of the same frame) skips the geometry work."""
    if base is not None:
        f = np.array(base, order="F", copy=True)
        f[5:8] = 0.0
        f[5 + direction] = 1.0
        return f
    from scipy import ndimage
    nx, ny, nz = p.nx, p.ny, p.nz
    fluid = p.solid[1:nx + 1, 1:ny + 1, :] == 0
    sampling = (8.0 / nx, 6.0 / ny, 4.0 / nz)
    if fluid.all():
        sdf = np.full(fluid.shape, 2.0)
    else:
        d_out = ndimage.distance_transform_edt(fluid, sampling=sampling)
        d_in = ndimage.distance_transform_edt(~fluid, sampling=sampling)
        sdf = np.clip((d_out - d_in) / 0.5, -1.0, 2.0)
    f = np.zeros((8, nx, ny, nz), dtype=np.float32, order="F")
    f[0] = fluid
    f[1] = sdf
    f[2] = (p.dx / p.dx.mean())[:, None, None]
    f[3] = (p.dy / p.dy.mean())[None, :, None]
    f[4] = (p.dz / p.dz.mean())[None, None, :]
    f[5 + direction] = 1.0
    return f


# ---------------------------------------------------------------- network ----
class UNet:
    """Handle of a network in HBM (FP32 master weights, BF16 copies, Adam)."""

    def __init__(self, seed: int = 0):
        self._h = C.c_void_p()
        check(lib().ffpm_unet_create(C.byref(self._h), C.c_uint64(seed)))
        n = C.c_size_t()
        check(lib().ffpm_unet_param_count(self._h, C.byref(n)))
        self.nparams = n.value

    def __del__(self):
        try:
            if self._h:
                lib().ffpm_unet_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def get_params(self) -> np.ndarray:
        a = np.empty(self.nparams, dtype=np.float32)
        check(lib().ffpm_unet_get_params(self._h, a.ctypes.data_as(C.c_void_p),
                                         C.c_size_t(self.nparams)))
        return a

    def set_params(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.nparams:
            raise ValueError(f"expected {self.nparams} parameters")
        check(lib().ffpm_unet_set_params(self._h, a.ctypes.data_as(C.c_void_p),
                                         C.c_size_t(self.nparams)))

    def get_grads(self) -> np.ndarray:
        a = np.empty(self.nparams, dtype=np.float32)
        check(lib().ffpm_unet_get_grads(self._h, a.ctypes.data_as(C.c_void_p),
                                        C.c_size_t(self.nparams)))
        return a

    def forward(self, feats: DeviceArray, source: DeviceArray, out: DeviceArray) -> None:
        _, nx, ny, nz = feats.shape
        check(lib().ffpm_unet_forward_dev(self._h, nx, ny, nz, feats.ptr, source.ptr, out.ptr))

    def train_step(self, feats: DeviceArray, source: DeviceArray, target: DeviceArray,
                   mask: DeviceArray, lr: float) -> float:
        _, nx, ny, nz = feats.shape
        loss = C.c_float()
        check(lib().ffpm_unet_train_step_dev(self._h, nx, ny, nz, feats.ptr, source.ptr,
                                             target.ptr, mask.ptr, C.c_float(lr),
                                             C.byref(loss)))
        return float(loss.value)

    def save(self, path: str, **meta) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.get_params(),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = WEIGHTS) -> "UNet":
        net = cls(0)
        with np.load(path, allow_pickle=False) as z:
            net.set_params(z["params"])
        return net


# ------------------------------------------------------------- solving -------
def solve_fcg(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz,
              tolerance, max_sweeps, check_interval, net: UNet, feats: DeviceArray):
    """Network-preconditioned flexible CG on device arrays (the arguments of
    solve_stairstep_poisson_3d, one rank). Returns (sweeps, residual,
    fallback_at): iterations (FCG + fallback PCG), the FP64 max-norm residual
    and the FCG iterations before a switch to diagonal PCG (-1: none)."""
    from .stairstep import _d, _i, _shape, _vec
    nx, ny, nz = _shape(solid)
    bg, bgp = _vec(body_grad, 3)
    sw, res, fb = C.c_int(), C.c_double(), C.c_int()
    check(lib().ffpm_solve_stairstep_fcg_dev(
        _d(phi), _d(rhs), _i(solid), bgp, _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi),
        _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), nx, ny, nz, C.c_double(tolerance),
        int(max_sweeps), int(check_interval), net.handle, feats.ptr, C.byref(sw), C.byref(res),
        C.byref(fb)))
    return sw.value, res.value, fb.value


def apply_operator(solid, xc, yc, zc, dx, dy, dz, x: DeviceArray, y: DeviceArray) -> None:
    """y = A x (dense device arrays) for the frame's PCG operator."""
    from .stairstep import _d, _i, _shape
    nx, ny, nz = _shape(solid)
    check(lib().ffpm_stairstep_apply_operator_dev(_i(solid), _d(xc), _d(yc), _d(zc), _d(dx),
                                                  _d(dy), _d(dz), nx, ny, nz, x.ptr, y.ptr))


# ------------------------------------------------------------ training -------
def random_bodies(rng):
    """1-4 spheres, ellipsoids or boxes, sizes ~ U[0.2, 1.0], random centres
    and axis ratios."""
    from . import workloads as W
    bodies = []
    for _ in range(int(rng.integers(1, 5))):
        c = (rng.uniform(-3.0, 3.0), rng.uniform(-2.0, 2.0), rng.uniform(-1.2, 1.2))
        kind = rng.integers(3)
        size = rng.uniform(0.2, 1.0)
        if kind == 0:
            bodies.append(W.Sphere(c, size))
        elif kind == 1:
            bodies.append(Ellipsoid(c, tuple(size * rng.uniform(0.5, 1.5, size=3))))
        else:
            s = size * rng.uniform(0.5, 1.5, size=3)
            bodies.append(W.Box(tuple(np.array(c) - 0.5 * s), tuple(np.array(c) + 0.5 * s)))
    return bodies


class Ellipsoid:
    def __init__(self, center, semi):
        self.center, self.semi = center, semi

    def inside(self, X, Y, Z):
        c, a = self.center, self.semi
        return ((X - c[0]) / a[0]) ** 2 + ((Y - c[1]) / a[1]) ** 2 + ((Z - c[2]) / a[2]) ** 2 < 1


class Pair:
    """One training sample in HBM."""

    def __init__(self, feats, source, target, mask):
        self.feats, self.source, self.target, self.mask = feats, source, target, mask


def make_pairs(frames, snapshots=(10, 50, 200, 800), log=None):
    """Training pairs from the plain solver on the given frames (Problem
    objects): for every frame and direction the converged phi* (FP64 PCG to
    1e-10, batched) and the iterates after each snapshot count; pair =
    (source (A e)/(V sigma), target e / rms(e)) with e = phi* - phi_k and
    sigma = max |A e / V|."""
    from . import stairstep as S
    from . import workloads as W
    T = DeviceArray.from_host
    probs, meta = [], []
    for p in frames:
        solid = T(p.solid)
        axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
        for d in range(3):
            rhs, bg, faces = W.influence_inputs(p, d)
            d_rhs, d_faces = T(rhs), [T(f) for f in faces]
            S.enforce_stairstep_compatibility_3d(d_rhs, solid, bg, *d_faces, *axes[3:])
            probs.append(S.BatchProblem(T(p.phi0()), d_rhs, solid, bg, *d_faces, *axes))
            meta.append((p, d, solid, axes))
    S.solve_stairstep_batch(probs, 1e-10, 100000, 100)
    if log:
        log(f"solved {len(probs)} fields: iterations {min(q.sweeps for q in probs)}-"
            f"{max(q.sweeps for q in probs)}")
    star = [q.phi.download() for q in probs]
    iterates = {}
    for k in snapshots:
        for q in probs:
            q.phi.zero_()
        S.solve_stairstep_batch(probs, -1.0, k, k)
        iterates[k] = [q.phi.download() for q in probs]
    pairs = []
    base = {}
    for idx, (p, d, solid, axes) in enumerate(meta):
        nx, ny, nz = p.nx, p.ny, p.nz
        fluid = (p.solid[1:nx + 1, 1:ny + 1, :] == 0)
        vol = (p.dx[:, None, None] * p.dy[None, :, None] * p.dz[None, None, :])
        f = features(p, d, base.get(id(p)))
        base[id(p)] = f
        feats = T(f)
        mask = T(fluid.astype(np.float32))
        for k in (0,) + tuple(snapshots):
            x = iterates[k][idx] if k else np.zeros_like(star[idx])
            e = (star[idx] - x)[1:nx + 1, 1:ny + 1, :] * fluid
            e = e - e[fluid].mean() * fluid
            de, r = T(np.asfortranarray(e)), DeviceArray((nx, ny, nz))
            apply_operator(solid, *axes, de, r)
            rv = r.download() / vol
            sigma = np.abs(rv[fluid]).max()
            rms = np.sqrt(np.mean(e[fluid] ** 2))
            if not (sigma > 0 and rms > 0):
                continue
            # Only the direction of the correction matters to flexible CG (its
            # step length is optimal), so the target is e at unit RMS.
            pairs.append(Pair(feats, T((rv / sigma).astype(np.float32)),
                              T((e / rms).astype(np.float32)), mask))
    synchronize()
    return pairs


def train(net: UNet, pairs, steps: int, lr0: float = 1e-3, lr1: float = 1e-4, seed: int = 0,
          log=None, every: int = 100):
    """Adam over the pairs in random order; the learning rate decays
    geometrically from lr0 to lr1. Returns the loss history."""
    rng = np.random.Generator(np.random.PCG64(seed))
    hist = []
    for s in range(steps):
        q = pairs[int(rng.integers(len(pairs)))]
        lr = lr0 * (lr1 / lr0) ** (s / max(1, steps - 1))
        loss = net.train_step(q.feats, q.source, q.target, q.mask, lr)
        hist.append(loss)
        if log and (s % every == 0 or s == steps - 1):
            log(f"step {s}: loss {np.mean(hist[-every:]):.4e} (lr {lr:.2e})")
    return hist


def training_frames(count: int, grid=(128, 128, 128), seed0: int = 5000):
    """Random stair-step frames on the stretched domain (seeds seed0 + i)."""
    from . import workloads as W
    frames = []
    for i in range(count):
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        frames.append(W.stretched_frame(*grid, random_bodies(rng), seed=seed0 + i,
                                        with_fields=False))
    return frames


def main():  # pragma: no cover - the training run on the GPU
    import argparse
    import json
    import sys
    import time
    ap = argparse.ArgumentParser(description="train the U-Net preconditioner on the GPU")
    ap.add_argument("--frames", type=int, default=48)
    ap.add_argument("--grid", type=int, nargs=3, default=[128, 128, 128])
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--log", default=None)
    args = ap.parse_args()

    def log(m):
        print(f"[train] {m}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    frames = training_frames(args.frames, tuple(args.grid))
    log(f"{len(frames)} frames generated, {time.perf_counter() - t0:.1f} s")
    pairs = make_pairs(frames, log=log)
    log(f"{len(pairs)} training pairs, {time.perf_counter() - t0:.1f} s")
    net = UNet(seed=1)
    t1 = time.perf_counter()
    hist = train(net, pairs, args.steps, log=log)
    synchronize()
    t_train = time.perf_counter() - t1
    log(f"trained {args.steps} steps in {t_train:.1f} s")
    net.save(args.out, frames=args.frames, grid=np.array(args.grid), steps=args.steps,
             loss_first=np.mean(hist[:50]), loss_last=np.mean(hist[-50:]))
    summary = {"frames": args.frames, "grid": args.grid, "pairs": len(pairs),
               "steps": args.steps, "train_s": round(t_train, 2),
               "ms_per_step": round(1e3 * t_train / args.steps, 3),
               "loss_first50": float(np.mean(hist[:50])), "loss_last50": float(np.mean(hist[-50:]))}
    print(json.dumps(summary), flush=True)
    if args.log:
        np.savetxt(args.log, np.asarray(hist))


if __name__ == "__main__":  # pragma: no cover
    main()
