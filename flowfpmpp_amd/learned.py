"""Learned preconditioner: the 3-D U-Net of the HIP library (BF16 MFMA
forward, backward and Adam kernels), its input features, training on fields
the plain solver produces, and the network-preconditioned flexible CG solve.

The network maps a residual r to the direction of A^-1 r. It sees the
frame (fluid mask, signed distance, the three cell widths, the direction
one-hot) and the source density r / V scaled by sigma = max |r / V|; the
length of its output is set by the solver (the A-optimal step). It is used
two ways: as the coarse correction of one level of the multigrid CG
(solve_fcg_mg_learned: the level of about 128x96x64 cells, where it takes
over the recursion to the coarser levels), and as the preconditioner of
flexible CG on the whole grid (solve_fcg). Training pairs come from the
multigrid CG on random stair-step bodies at 128^3: the converged phi* of
each direction and the errors e_k = phi* - phi_k of its iterates after k
iterations, with the exact residual r_k = A e_k -- the errors the cycle
leaves, which the coarse correction has to remove.

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
H_REF = 8.0 / 128.0  # cell widths are given to the network in units of this length


def signed_distance(p) -> np.ndarray:
    """Signed distance to the solid/fluid interface at the cell centres of
    frame p (positive in fluid), scaled by 0.5 domain units and clipped to
    [-1, 2]; FP32 (nx, ny, nz). Distances are measured on the uniform
    index grid of the domain's extent (8 x 6 x 4 over nx x ny x nz)."""
    nx, ny, nz = p.nx, p.ny, p.nz
    return _sdf_of_mask(p.solid[1:nx + 1, 1:ny + 1, :] == 0)


def features(p, direction: int, base: np.ndarray = None, sdf: np.ndarray = None) -> np.ndarray:
    """FP32 input features of frame p for force direction 0/1/2, shape
    (8, nx, ny, nz) in Fortran order (channels-last in memory): fluid mask,
    signed distance (signed_distance), the cell widths dx, dy, dz in units of
    H_REF, and the direction one-hot. ``base`` (the features of another
    direction of the same frame) or ``sdf`` skip the geometry work."""
    if base is not None:
        f = np.array(base, order="F", copy=True)
        f[5:8] = 0.0
        f[5 + direction] = 1.0
        return f
    nx, ny, nz = p.nx, p.ny, p.nz
    fluid = p.solid[1:nx + 1, 1:ny + 1, :] == 0
    f = np.zeros((8, nx, ny, nz), dtype=np.float32, order="F")
    f[0] = fluid
    f[1] = signed_distance(p) if sdf is None else sdf
    f[2] = (p.dx / H_REF)[:, None, None]
    f[3] = (p.dy / H_REF)[None, :, None]
    f[4] = (p.dz / H_REF)[None, None, :]
    f[5 + direction] = 1.0
    return f


# ------------------------------------------------- the learned coarse level ----
def mg_levels(nx: int, ny: int, nz: int):
    """Extents of the multigrid levels the solver builds (2x2x2 aggregation,
    ceil for odd extents, down to at most 1024 cells)."""
    dims = [(nx, ny, nz)]
    while nx * ny * nz > 1024:
        nx, ny, nz = (nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2
        dims.append((nx, ny, nz))
    if len(dims) == 1:
        dims.append(((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2))
    return dims


def learned_level(nx: int, ny: int, nz: int, max_cells: int = 1 << 20) -> int:
    """The level whose coarse correction the network takes over: the first
    level below the finest with at most `max_cells` cells, extents that are
    multiples of 4 and nx >= 16, above the coarsest; -1 if there is none."""
    dims = mg_levels(nx, ny, nz)
    for lv in range(1, len(dims) - 1):
        a, b, c = dims[lv]
        if a * b * c <= max_cells:
            ok = a >= 16 and a % 4 == 0 and b % 4 == 0 and c % 4 == 0
            return lv if ok else -1
    return -1


def _block_sum(a: np.ndarray, axis: int, step: int) -> np.ndarray:
    idx = np.arange(0, a.shape[axis], step)
    return np.add.reduceat(a, idx, axis=axis)


def _aggregate(a: np.ndarray, step: int, mean: bool) -> np.ndarray:
    out = a.astype(np.float64)
    cnt = np.ones_like(out)
    for ax in range(3):
        out = _block_sum(out, ax, step)
        cnt = _block_sum(cnt, ax, step)
    return out / cnt if mean else out


def _sdf_of_mask(fluid: np.ndarray) -> np.ndarray:
    """signed_distance's scaling for a boolean fluid mask on the domain."""
    from scipy import ndimage
    nx, ny, nz = fluid.shape
    if fluid.all():
        return np.full(fluid.shape, 2.0, dtype=np.float32)
    sampling = (8.0 / nx, 6.0 / ny, 4.0 / nz)
    d_out = ndimage.distance_transform_edt(fluid, sampling=sampling)
    d_in = ndimage.distance_transform_edt(~fluid, sampling=sampling)
    return np.clip((d_out - d_in) / 0.5, -1.0, 2.0).astype(np.float32)


def level_features(p, level: int, direction: int) -> np.ndarray:
    """The network's features on multigrid level `level` of frame p, shape
    (8, nxL, nyL, nzL) FP32 in Fortran order: the fluid fraction of each level
    cell (the mean over its children), the signed distance on the level's grid
    (of the cells at least half fluid; the network works at that
    resolution, as on the 128^3 frames it was trained on), the level cells'
    widths (sums of the children's) in units of H_REF, and the direction
    one-hot."""
    step = 1 << level
    nx, ny, nz = p.nx, p.ny, p.nz
    fluid = (p.solid[1:nx + 1, 1:ny + 1, :] == 0).astype(np.float64)
    ff = _aggregate(fluid, step, True)
    sa = _sdf_of_mask(ff >= 0.5)
    wx = _block_sum(np.asarray(p.dx, dtype=np.float64), 0, step)
    wy = _block_sum(np.asarray(p.dy, dtype=np.float64), 0, step)
    wz = _block_sum(np.asarray(p.dz, dtype=np.float64), 0, step)
    f = np.zeros((8,) + ff.shape, dtype=np.float32, order="F")
    f[0] = ff
    f[1] = sa
    f[2] = (wx / H_REF)[:, None, None]
    f[3] = (wy / H_REF)[None, :, None]
    f[4] = (wz / H_REF)[None, None, :]
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


def solve_fcg_mg_learned(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc,
                         dx, dy, dz, tolerance, max_sweeps, check_interval, net: UNet,
                         feats: DeviceArray, level: int, mode: str = "augment"):
    """The multigrid-preconditioned CG with the network as the coarse
    correction of `level` (feats: level_features of that level, on the
    device), on device arrays (the arguments of solve_stairstep_poisson_3d,
    one rank). mode "replace": the network takes the place of the recursion
    below the level; "augment": the recursion runs and the network corrects
    what it leaves. Returns (sweeps, residual, fallback_at): iterations
    (learned cycle + fallback PCG), the FP64 max-norm residual and the
    iterations before a switch to diagonal PCG (-1: none)."""
    modes = {"replace": 0, "augment": 1}
    if mode not in modes:
        raise ValueError("mode must be 'replace' or 'augment'")
    from .stairstep import _check_frame, _d, _i, _shape, _vec
    nx, ny, nz = _shape(solid)
    _check_frame((nx, ny, nz), (xlo, xhi, ylo, yhi, zlo, zhi), (xc, yc, zc, dx, dy, dz),
                 phi=(phi, "padded"), rhs=(rhs, "cells"))
    want = (8,) + mg_levels(nx, ny, nz)[level]
    if tuple(feats.shape) != want or feats.dtype != np.float32:
        raise ValueError(f"features of level {level} must be float32 {want}")
    bg, bgp = _vec(body_grad, 3)
    sw, res, fb = C.c_int(), C.c_double(), C.c_int()
    check(lib().ffpm_solve_stairstep_fcg_mg_learned_dev(
        _d(phi), _d(rhs), _i(solid), bgp, _d(xlo), _d(xhi), _d(ylo), _d(yhi), _d(zlo), _d(zhi),
        _d(xc), _d(yc), _d(zc), _d(dx), _d(dy), _d(dz), nx, ny, nz, C.c_double(tolerance),
        int(max_sweeps), int(check_interval), net.handle, int(level), modes[mode], feats.ptr,
        C.byref(sw), C.byref(res), C.byref(fb)))
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
    return [_frame_and_sdf((seed0 + i, grid))[0] for i in range(count)]


def _frame_and_sdf(arg):
    """Frame `seed` (random_bodies of that seed on the stretched grid) and its
    signed distance; runs in worker processes."""
    from . import workloads as W
    seed, grid = arg
    rng = np.random.Generator(np.random.PCG64(seed))
    p = W.stretched_frame(*grid, random_bodies(rng), seed=seed, with_fields=False)
    return p, signed_distance(p)


def make_pairs_mg(frames, sdfs, snapshots=(0, 1, 3), log=None, threads=8):
    """Training pairs for the learned coarse correction from the multigrid
    CG on the given frames: every direction is solved to 1e-10 (phi*, batched)
    and again for each snapshot count k > 0 of iterations; pair = (source
    A e / (V sigma), target e / rms(e)) with e = phi* - phi_k (e = phi* for k =
    0) and sigma = max |A e / V|. These are the errors the cycle leaves behind
    after k iterations, the ones its coarse correction has to remove."""
    from concurrent.futures import ThreadPoolExecutor
    from . import stairstep as S
    from . import workloads as W
    T = DeviceArray.from_host
    probs, meta = [], []
    for p, sd in zip(frames, sdfs):
        solid = T(p.solid)
        axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
        base = features(p, 0, sdf=sd)
        nx, ny, nz = p.nx, p.ny, p.nz
        fluid = p.solid[1:nx + 1, 1:ny + 1, :] == 0
        mask = T(fluid.astype(np.float32))
        vol = (p.dx[:, None, None] * p.dy[None, :, None] * p.dz[None, None, :])
        for d in range(3):
            rhs, bg, faces = W.influence_inputs(p, d)
            d_rhs, d_faces = T(rhs), [T(f) for f in faces]
            S.enforce_stairstep_compatibility_3d(d_rhs, solid, bg, *d_faces, *axes[3:])
            probs.append(S.BatchProblem(T(p.phi0()), d_rhs, solid, bg, *d_faces, *axes))
            meta.append((p, d, solid, axes, base, mask, fluid, vol))
    S.solve_stairstep_batch(probs, 1e-10, 100000, 100, method="fcg-mg")
    if log:
        log(f"solved {len(probs)} fields by the multigrid CG: iterations "
            f"{min(q.sweeps for q in probs)}-{max(q.sweeps for q in probs)}")
    star = [q.phi.download() for q in probs]
    iterates = {}
    for k in snapshots:
        if k == 0:
            continue
        for q in probs:
            q.phi.zero_()
        S.solve_stairstep_batch(probs, -1.0, k, k, method="fcg-mg")
        iterates[k] = [q.phi.download() for q in probs]
    feats = {}
    import threading
    lock = threading.Lock()  # the library's calls share its workspaces: one at a time

    def one(job):
        idx, k = job
        p, d, solid, axes, base, mask, fluid, vol = meta[idx]
        nx, ny, nz = p.nx, p.ny, p.nz
        x = iterates[k][idx] if k else None
        e = star[idx][1:nx + 1, 1:ny + 1, :] if x is None else \
            (star[idx] - x)[1:nx + 1, 1:ny + 1, :]
        e = e * fluid
        e = e - e[fluid].mean() * fluid
        with lock:
            de, r = T(np.asfortranarray(e)), DeviceArray((nx, ny, nz))
            apply_operator(solid, *axes, de, r)
            rv = r.download()
            de.free()
            r.free()
        rv /= vol
        sigma = np.abs(rv[fluid]).max()
        rms = np.sqrt(np.mean(e[fluid] ** 2))
        if not (sigma > 0 and rms > 0):
            return None
        src, tgt = (rv / sigma).astype(np.float32), (e / rms).astype(np.float32)
        with lock:
            return (idx, T(src), T(tgt))

    jobs = [(idx, k) for idx in range(len(meta)) for k in snapshots]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        done = list(ex.map(one, jobs))
    pairs = []
    for item in done:
        if item is None:
            continue
        idx, src, tgt = item
        p, d, solid, axes, base, mask, fluid, vol = meta[idx]
        if idx not in feats:
            feats[idx] = T(features(p, d, base))
        pairs.append(Pair(feats[idx], src, tgt, mask))
    synchronize()
    return pairs


def main():  # pragma: no cover - the training run on the GPU
    import argparse
    import json
    import sys
    import time
    from multiprocessing import get_context
    ap = argparse.ArgumentParser(description="train the U-Net coarse correction on the GPU")
    ap.add_argument("--frames", type=int, default=667,
                    help="training frames; every frame gives 3 direction fields")
    ap.add_argument("--grid", type=int, nargs=3, default=[128, 128, 128])
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--chunk", type=int, default=48, help="frames per training chunk")
    ap.add_argument("--workers", type=int, default=12, help="processes building frames")
    ap.add_argument("--seed0", type=int, default=5000)
    ap.add_argument("--lr0", type=float, default=3e-4, help="first learning rate")
    ap.add_argument("--lr1", type=float, default=2e-5, help="last learning rate (geometric decay)")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--log", default=None)
    args = ap.parse_args()

    def log(m):
        print(f"[train] {m}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    grid = tuple(args.grid)
    seeds = list(range(args.seed0, args.seed0 + args.frames))
    chunks = [seeds[i:i + args.chunk] for i in range(0, len(seeds), args.chunk)]
    steps_per_chunk = [args.steps * len(c) // len(seeds) for c in chunks]
    steps_per_chunk[-1] += args.steps - sum(steps_per_chunk)
    net = UNet(seed=1)
    hist, npairs, nfields = [], 0, 0
    ctx = get_context("spawn")
    with ctx.Pool(args.workers) as pool:
        pending = pool.map_async(_frame_and_sdf, [(sd, grid) for sd in chunks[0]])
        for ci, chunk in enumerate(chunks):
            built = pending.get()
            if ci + 1 < len(chunks):
                pending = pool.map_async(_frame_and_sdf, [(sd, grid) for sd in chunks[ci + 1]])
            frames = [b[0] for b in built]
            sdfs = [b[1] for b in built]
            pairs = make_pairs_mg(frames, sdfs)
            npairs += len(pairs)
            nfields += 3 * len(frames)
            lr_lo, lr_hi = args.lr1, args.lr0
            done = sum(steps_per_chunk[:ci])
            rng = np.random.Generator(np.random.PCG64(ci))
            for s in range(steps_per_chunk[ci]):
                q = pairs[int(rng.integers(len(pairs)))]
                lr = lr_hi * (lr_lo / lr_hi) ** ((done + s) / max(1, args.steps - 1))
                hist.append(net.train_step(q.feats, q.source, q.target, q.mask, lr))
            synchronize()
            log(f"chunk {ci + 1}/{len(chunks)}: {len(frames)} frames, {len(pairs)} pairs, "
                f"{steps_per_chunk[ci]} steps, loss {np.mean(hist[-100:]):.4f}, "
                f"{time.perf_counter() - t0:.0f} s")
            del pairs, frames, sdfs, built
    net.save(args.out, frames=args.frames, fields=nfields, grid=np.array(grid), steps=args.steps,
             seed0=args.seed0, pairs=npairs, loss_first=np.mean(hist[:100]),
             loss_last=np.mean(hist[-100:]))
    summary = {"frames": args.frames, "fields": nfields, "grid": list(grid), "pairs": npairs,
               "steps": args.steps, "wall_s": round(time.perf_counter() - t0, 1),
               "loss_first100": float(np.mean(hist[:100])),
               "loss_last100": float(np.mean(hist[-100:]))}
    print(json.dumps(summary), flush=True)
    if args.log:
        np.savetxt(args.log, np.asarray(hist))


if __name__ == "__main__":  # pragma: no cover
    main()
