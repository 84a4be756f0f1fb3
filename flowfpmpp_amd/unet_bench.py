"""Times the U-Net kernels: forward passes at a given grid and training steps
at 128^3, for kernel traces (`rocprofv3 --kernel-trace --stats -- python3
$REPO/flowfpmpp_amd/unet_bench.py 256 192 128`). Prints one JSON line with
the average times."""
from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    from flowfpmpp_amd import learned as L
    from flowfpmpp_amd import workloads as W
    from flowfpmpp_amd.device import DeviceArray, synchronize
    grid = tuple(int(v) for v in (sys.argv[1:4] or (256, 192, 128)))
    reps = 5
    net = L.UNet(seed=1)
    rng = np.random.Generator(np.random.PCG64(1))
    p = W.stretched_frame(*grid, L.random_bodies(rng), seed=1, with_fields=False)
    feats = DeviceArray.from_host(L.features(p, 0))
    src = DeviceArray.from_host(rng.standard_normal(grid).astype(np.float32))
    out = DeviceArray(grid, np.float32)
    net.forward(feats, src, out)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.forward(feats, src, out)
    synchronize()
    fwd = (time.perf_counter() - t0) / reps
    tg = (128, 128, 128)
    q = W.stretched_frame(*tg, L.random_bodies(rng), seed=2, with_fields=False)
    tf = DeviceArray.from_host(L.features(q, 1))
    ts = DeviceArray.from_host(rng.standard_normal(tg).astype(np.float32))
    tt = DeviceArray.from_host(rng.standard_normal(tg).astype(np.float32))
    tm = DeviceArray.from_host(np.ones(tg, dtype=np.float32))
    net.train_step(tf, ts, tt, tm, 1e-4)
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        net.train_step(tf, ts, tt, tm, 1e-4)
    synchronize()
    step = (time.perf_counter() - t0) / reps
    cells = int(np.prod(grid))
    flop = 27 * 2 * (16 * 16 + 16 * 16 + 48 * 16 + 16 * 16) * cells \
        + 27 * 2 * (16 * 32 + 32 * 32 + 96 * 32 + 32 * 32) * cells / 8 \
        + 27 * 2 * (32 * 64 + 64 * 64) * cells / 64
    print(json.dumps({"grid": grid, "forward_ms": round(fwd * 1e3, 3),
                      "forward_tflops": round(flop / fwd / 1e12, 1),
                      "train_grid": tg, "train_step_ms": round(step * 1e3, 3)}), flush=True)


if __name__ == "__main__":
    main()
