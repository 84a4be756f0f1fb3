"""Headline benchmark: time to solve phi_x, phi_y, phi_z to a max-norm
residual of 1e-10 (plus their FPM force densities) on one MI355X.

Workload: the 512x384x256 tanh-stretched frame (4:1 width ratio) on
[-4,4]x[-3,3]x[-2,2] with four seeded blocked bodies (seed 2) and seeded
16-mode divergence-free velocity/acceleration fields. One step = for each of
the three force directions: compatibility correction, the solve to 1e-10 on
the reference's FP64 max-norm residual, gauge, and the FPM densities. Every
input is resident in HBM before the timed region; phi starts from zero each
step. The three directions' solves run at once, each on its own host thread
and HIP stream (--sequential-directions: one after another on one stream);
the densities of all three follow in one pass.

The default method is "fcg-mg" (FP64 conjugate gradients preconditioned by
the FP32 multigrid cycle); the same frame is then solved once with the
reference's FP64 diagonal PCG (check interval 100) for the comparison the
metric asks for ("fp64_pcg"), unless --no-pcg.

Prints ONE JSON line. ``value`` is milliseconds per step (lower is better).
``roofline`` covers the kernel with the largest share of device time, timed
live with HIP events on the library's stream (with concurrent directions, in
an extra step that solves them one after another, where each launch has the
device to itself); ``cpu_baseline`` times the CPU
oracle (the reference's PCG) on this machine's host on a bounded sample of
the same frame, scaled to the PCG iteration counts.

    python bench.py [--steps K] [--warmup W] [--method fcg-mg|pcg|pcg-mixed|jacobi]
                    [--grid 512 384 256] [--tol 1e-10] [--no-cpu] [--no-pcg]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = ("time-to-solve 3 phi to 1e-10 (ms), HBM GB/s; learned-FCG vs plain FP64 PCG")
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per cell and launch (FP64 fields; 1-D coefficient
# arrays are negligible).
BYTES_PER_CELL = {
    "pcg_dir_spmv": 25,  # read z, p_old, neighbour code (1 B); write p
    "pcg_update": 49,    # read p, code (1 B), phi, r; write phi, r, z
    "pcg32_dir_spmv": 13,  # FP32 inner: read z, p_old (4 B each), code; write p
    "pcg32_update": 25,    # FP32 inner: read p, code, e, r; write e, r, z
    "spmv_dot": 24,      # read p, diag; write q
    "update_dot": 56,    # read phi, p, r, q, rdiag; write phi, r
    "direction": 32,     # read rdiag, r, p; write p
    "jacobi": 32,        # read a, src, rdiag; write b (the reference's sweep; this one
                         # reads the 1-byte neighbour code instead of rdiag)
    "true_residual_1": 32,  # phi, src, diag; write r
    "true_residual_2": 32,  # r (read, write), rdiag, diag
    "residual_max": 24,  # phi, src, diag
    "terms": 124,        # phi, solid(4), 9 fields; write 5 outputs
    "terms_dirs": 220,   # the three directions in one pass: 3 phi, solid(4), 9 fields;
                         # write 15 outputs
    # CG with the multigrid cycle (level-0 kernels; FP32 arithmetic, the
    # cycle's vectors stored in FP16; the coarse levels add 1/8, 1/64, ... of
    # their own small traffic)
    "mg_update": 35,         # read p (FP32), code, r; write r, x1, q (FP16); every
                             # second update also reads/writes phi and reads the
                             # deferred direction: 25 or 45 B, 35 on average
    "mg_fine_sweep": 7.5,    # x, the first step x1 (the source), code; write x (FP16);
                             # the step after the first reads x1 once (5 B), the next
                             # ones 7 B, or 9 B with the previous iterate
    "mg_fine_restrict": 5,   # x, x1 (FP16), code per fine cell (coarse writes 1/8 of 4 B)
    "mg_fine_prolong": 5,    # x (FP16) read/write, code (coarse reads 1/8)
    "mg_fine_last": 11,      # x, previous iterate, x1, q (FP16), code; write z (FP16)
    "mg_direction": 11,      # read z (FP16), p_old (FP32), code; write p (FP32)
}


# The sources of the measured kernels' device code: the committed traffic
# figures hold only while these are the sources that were measured.
KERNEL_SOURCES = ("ffpm_common.hpp", "ffpm_kernels.hpp", "ffpm_pcg.hpp", "ffpm_pcg_body.hpp",
                  "ffpm_bodies.hpp", "ffpm_frame_body.hpp", "ffpm_mg.hpp", "ffpm_mg_body.hpp",
                  "ffpm_mg.hip", "ffpm_pcg_kernels.hip", "ffpm_stairstep_kernels.hip")


def kernel_sources_sha256():
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(HERE, "flowfpmpp_amd", "csrc", name), "rb") as f:
            h.update(name.encode() + b"\0" + f.read())
    return h.hexdigest()


def measured_traffic(kernel, grid):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE
    / WRITE_SIZE measurement of this workload (profiles/hbm_traffic.json), or
    None when the kernel or grid was not measured or the kernel sources have
    changed since the measurement (their SHA-256 is stored with it)."""
    path = os.path.join(HERE, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel_sources_sha256") != kernel_sources_sha256():
        log("profiles/hbm_traffic.json was measured on other kernel sources: traffic null")
        return None
    k = t.get("kernels", {}).get(kernel)
    return k["traffic_bytes"] if k and t.get("grid") == grid else None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def frame(grid):
    from flowfpmpp_amd import workloads as W
    nx, ny, nz = grid
    if (nx, ny, nz) == (512, 384, 256):
        return W.frame_512x384x256()
    if (nx, ny, nz) == (128, 96, 64):
        return W.frame_128x96x64()
    return W.stretched_frame(nx, ny, nz, W.seeded_bodies(2), seed=2,
                             label=f"{nx}x{ny}x{nz}")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, gpu_iters, tol, method, budget_s=20.0, parallel=False):
    """Times the CPU oracle on this host (one thread, or the OpenMP build on
    every thread OMP_NUM_THREADS allows): the marginal cost of one iteration
    from two fixed-work runs of the x-direction solve, plus the fixed
    setup/gauge cost, scaled to the GPU's iteration counts."""
    from oracle import oracle as orc
    from flowfpmpp_amd import workloads as W
    rhs, bg, faces = W.influence_inputs(p, 0)
    orc.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)

    def run(k):
        phi = p.phi0()
        t0 = time.perf_counter()
        orc.solve_stairstep_poisson_3d(phi, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc, p.dx,
                                       p.dy, p.dz, -1.0, k, k, method, parallel=parallel)
        return time.perf_counter() - t0

    t1 = run(1)
    per_guess = max(t1 / 6.0, 1e-6)
    k2 = int(max(3, min(40, budget_s / per_guess)))
    t2 = run(k2)
    per_iter = (t2 - t1) / (k2 - 1)
    setup = max(t1 - per_iter, 0.0)
    est_s = sum(setup + per_iter * it for it in gpu_iters)
    cores = threads if parallel else 1
    how = (f"OpenMP over z planes, {cores} threads" if parallel else "1 thread")
    return {
        "value": est_s * 1e3,
        "unit": "ms",
        "cores": cores,
        "kind": "port",
        "sample": (f"CPU oracle (gcc -O2, {how}; {cpu_model()}, nproc {os.cpu_count()}), "
                   f"{method} on this frame, direction x: 1 and {k2} fixed iterations -> "
                   f"{per_iter * 1e3:.1f} ms/iteration + {setup * 1e3:.0f} ms setup/gauge, "
                   f"scaled to the GPU's iteration counts {list(gpu_iters)} over the three "
                   f"directions"),
    }


def cpu_time_to_solution_c1(tol, check):
    """A measured (not scaled) CPU baseline: the 128x96x64 frame's three
    directions solved to the tolerance, compatibility included, by the CPU
    oracle's FP64 diagonal PCG on every thread OMP_NUM_THREADS allows, timed
    end to end; beside it the same three solves on the GPU with FP64 PCG and
    with fcg-mg (inputs resident in HBM, phi from zero). Iterations and
    residuals of each are reported."""
    from oracle import oracle as orc
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd import workloads as W
    from flowfpmpp_amd.device import DeviceArray, synchronize
    T = DeviceArray.from_host
    p = W.frame_128x96x64(with_fields=False)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    inputs = [W.influence_inputs(p, d) for d in range(3)]
    cpu_it, cpu_res = [], []
    t0 = time.perf_counter()
    for rhs, bg, faces in inputs:
        rhs, faces = rhs.copy(order="F"), [f.copy(order="F") for f in faces]
        orc.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
        phi = p.phi0()
        it, res = orc.solve_stairstep_poisson_3d(phi, rhs, p.solid, bg, *faces, p.xc, p.yc,
                                                 p.zc, p.dx, p.dy, p.dz, tol, 100000, check,
                                                 "pcg", parallel=True)
        cpu_it.append(int(it))
        cpu_res.append(float(res))
    cpu_ms = (time.perf_counter() - t0) * 1e3
    solid = T(p.solid)
    axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
    dev = [(T(rhs), bg, [T(f) for f in faces]) for rhs, bg, faces in inputs]
    out = {"cpu": {"ms": round(cpu_ms, 1), "iterations": cpu_it, "residuals": cpu_res,
                   "cores": threads, "kind": "port",
                   "sample": f"CPU oracle FP64 diagonal PCG (gcc -O2, OpenMP over z planes, "
                             f"{threads} threads; {cpu_model()}), compatibility + solve of the "
                             f"three directions, measured end to end"}}
    for method in ("pcg", "fcg-mg"):
        times, its, rs = [], [], []
        for rep in range(2):  # the first run includes the library's setup
            work = [(T(r.download()), bg, [T(f.download()) for f in fs]) for r, bg, fs in dev]
            phis = [T(p.phi0()) for _ in range(3)]
            synchronize()
            t0 = time.perf_counter()
            its, rs = [], []
            for (r, bg, fs), ph in zip(work, phis):
                S.enforce_stairstep_compatibility_3d(r, solid, bg, *fs, *axes[3:])
                it, res = S.solve_stairstep_poisson_3d(ph, r, solid, bg, *fs, *axes, tol, 100000,
                                                       check, method=method)
                its.append(int(it))
                rs.append(float(res))
            synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        out["gpu_" + method.replace("-", "_")] = {"ms": round(times[-1], 3), "iterations": its,
                                                  "residuals": rs,
                                                  "speedup_vs_cpu": round(cpu_ms / times[-1], 1)}
    out["workload"] = ("128x96x64 frame (seed 1, sphere + box), 3 directions: compatibility + "
                       f"solve to {tol:g} at check interval {check}, phi from zero")
    return out


def plain_stencil(p, args, d_rhs, d_solid, axes, faces, phi, bg, sweeps=200):
    """The plain FP64 stencil kernels on this frame against the HBM peak: the
    fused PCG iteration's two kernels and the weighted-Jacobi sweep, timed
    here over `sweeps` fixed iterations or sweeps of the x direction (HIP
    events on the library's stream around every launch).
    GB/s are given by the kernels' own algorithmic bytes and by the
    reference's (its three-kernel PCG iteration: 112 B/cell; its Jacobi
    sweep: 32 B/cell), with the rocprofv3-measured traffic where the
    committed measurement covers the current sources."""
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import synchronize
    cells = p.cells
    faces_now = faces
    phi.zero_()
    for f in faces_now:
        f.zero_()
    S.enforce_stairstep_compatibility_3d(d_rhs, d_solid, bg, *faces_now, *axes[3:])
    synchronize()
    S.timing(True)
    S.solve_stairstep_poisson_3d(phi, d_rhs, d_solid, bg, *faces_now, *axes, -1.0, sweeps,
                                 sweeps, method="pcg")
    phi.zero_()
    S.solve_stairstep_poisson_3d(phi, d_rhs, d_solid, bg, *faces_now, *axes, -1.0, sweeps,
                                 sweeps, method="jacobi")
    synchronize()
    S.timing(False)
    kern = {n: S.timing_query(n) for n in ("pcg_dir_spmv", "pcg_update", "jacobi")}
    out = {}
    for name, (cnt, ms) in kern.items():
        if not cnt:
            continue
        avg_us = ms * 1e3 / cnt
        # this Jacobi kernel reads the 1-byte neighbour code instead of rdiag
        own = 25 if name == "jacobi" else BYTES_PER_CELL[name]
        e = {"launches": cnt, "avg_us": round(avg_us, 2), "bytes_per_cell": own,
             "gbps": round(own * cells / avg_us / 1e3, 1),
             "frac": round(own * cells / avg_us / 1e3 / HBM_PEAK_GBS, 4)}
        traffic = measured_traffic(name, list(args.grid))
        if traffic:
            e["traffic_gbps"] = round(traffic / avg_us / 1e3, 1)
        out[name] = e
    if "jacobi" in out:
        out["jacobi"]["reference_bytes_per_cell"] = 32
        out["jacobi"]["reference_frac"] = round(
            32 * cells / out["jacobi"]["avg_us"] / 1e3 / HBM_PEAK_GBS, 4)
    if "pcg_dir_spmv" in out and "pcg_update" in out:
        it_us = out["pcg_dir_spmv"]["avg_us"] + out["pcg_update"]["avg_us"]
        out["pcg_iteration"] = {"us": round(it_us, 2), "reference_bytes_per_cell": 112,
                                "reference_frac": round(112 * cells / it_us / 1e3 / HBM_PEAK_GBS,
                                                        4)}
    log(f"plain stencil: {json.dumps(out)}")
    return out


def step_learned_args(p, d_phi, d_rhs, d_solid, bgs, axes, d_faces, dev, outs):
    return dict(p=p, d_phi=d_phi, d_rhs=d_rhs, d_solid=d_solid, bgs=bgs, axes=axes,
                d_faces=d_faces, dev=dev, outs=outs)


def learned_summary(p, args, a, fcg_mg_ms, pcg_ms):
    """The headline step (3 directions: compatibility, solve to the tolerance,
    gauge, densities) with the multigrid CG whose coarse correction on level 2
    (128x96x64 here) includes the trained U-Net (flowfpmpp_amd/data/
    unet_weights.npz): "augment" adds the network's A-optimal correction to
    the recursive one, "replace" lets the network stand in for everything
    below the level. Features on the level are built on the host first
    (reported apart); the network's activations are allocated by a warm-up
    solve."""
    from flowfpmpp_amd import learned as L
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import DeviceArray, synchronize
    net = L.UNet.load()
    level = L.learned_level(p.nx, p.ny, p.nz)
    t0 = time.perf_counter()
    feats = [DeviceArray.from_host(L.level_features(p, level, d)) for d in range(3)]
    feat_ms = (time.perf_counter() - t0) * 1e3
    fields = [a["dev"][k] for k in ("u", "v", "w", "ox", "oy", "oz", "ax", "ay", "az")]

    def step(mode):
        its, res, fbs = [], [], []
        for d in range(3):
            a["d_phi"][d].zero_()
            for f in a["d_faces"][d]:
                f.zero_()
            S.enforce_stairstep_compatibility_3d(a["d_rhs"], a["d_solid"], a["bgs"][d],
                                                 *a["d_faces"][d], *a["axes"][3:])
            sw, r, fb = L.solve_fcg_mg_learned(a["d_phi"][d], a["d_rhs"], a["d_solid"],
                                               a["bgs"][d], *a["d_faces"][d], *a["axes"],
                                               args.tol, args.max_iters, args.check, net,
                                               feats[d], level, mode)
            its.append(sw)
            res.append(r)
            fbs.append(fb)
        S.compute_stairstep_fpm_terms_3d_dirs(a["d_phi"], *fields, a["d_solid"], a["bgs"],
                                              *a["axes"], a["d_faces"], p.bounds, a["outs"])
        return its, res, fbs

    out = {"network": "U-Net 16/32/64, BF16 MFMA, trained on 2001 fields at 128^3 "
                      "(flowfpmpp_amd/data/unet_weights.npz)",
           "level": level, "level_grid": list(L.mg_levels(p.nx, p.ny, p.nz)[level]),
           "features_host_ms": round(feat_ms, 1), "fcg_mg_ms_per_step": fcg_mg_ms}
    for mode in ("augment", "replace"):
        if mode == "augment":
            step(mode)  # warm-up: the network's buffers at this level (shared by both modes)
        synchronize()
        t1 = time.perf_counter()
        its, res, fbs = step(mode)
        synchronize()
        ms = (time.perf_counter() - t1) * 1e3
        log(f"learned coarse correction ({mode}): {ms:.1f} ms, iterations {its}, fallback {fbs}")
        out[mode] = {"ms_per_step": round(ms, 3), "iterations": its, "residuals": res,
                     "fallback_at": fbs, "converged": all(r <= args.tol for r in res),
                     "speedup_vs_fp64_pcg": round(pcg_ms / ms, 2) if pcg_ms else None,
                     "time_vs_fcg_mg": round(ms / fcg_mg_ms, 3)}
    return out


def batch64_frames():
    """The 64 uneven frames x 3 directions of the batched workload, resident on
    the device with the compatibility correction applied; returns
    (problems, cells)."""
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd import workloads as W
    from flowfpmpp_amd.device import DeviceArray, synchronize
    T = DeviceArray.from_host
    probs, cells = [], 0
    for i in range(64):
        p = W.batch_instance(i)
        solid = T(p.solid)
        axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
        for d in range(3):
            rhs, bg, faces = W.influence_inputs(p, d)
            d_rhs, d_faces = T(rhs), [T(f) for f in faces]
            S.enforce_stairstep_compatibility_3d(d_rhs, solid, bg, *d_faces, *axes[3:])
            probs.append(S.BatchProblem(T(p.phi0()), d_rhs, solid, bg, *d_faces, *axes))
        cells += 3 * p.cells
    synchronize()
    return probs, cells


def run_batch(probs, method, tol, max_iters, check, steps, warmup, rows=0, zchunk=0):
    """Times `steps` batched solves of all problems from phi = 0 (after
    `warmup` untimed ones); returns solves/s, p50/p99 of the per-solve time
    to convergence, iteration range and residuals."""
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import synchronize

    def step():
        for q in probs:
            q.phi.zero_()
        S.solve_stairstep_batch(probs, tol, max_iters, check, rows=rows, zchunk=zchunk,
                                method=method)
    for _ in range(warmup):
        step()
    synchronize()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        step()
        synchronize()
        times.append(time.perf_counter() - t0)
    ms = 1e3 * sum(times) / len(times)
    solve_ms = np.array([q.solve_ms for q in probs])
    sweeps = np.array([q.sweeps for q in probs])
    res = np.array([q.residual for q in probs])
    log(f"batched {method}: {ms:.1f} ms for {len(probs)} solves; iterations "
        f"{sweeps.min()}-{sweeps.max()}")
    return {"ms": round(ms, 3), "solves_per_s": round(len(probs) * 1e3 / ms, 3),
            "solve_ms_p50": round(float(np.percentile(solve_ms, 50)), 3),
            "solve_ms_p99": round(float(np.percentile(solve_ms, 99)), 3),
            "iterations_min_max": [int(sweeps.min()), int(sweeps.max())],
            "max_residual": float(res.max()), "converged": bool((res <= tol).all())}


def batch64_summary(tol, max_iters, check, steps=2, warmup=1):
    """The batched workload as one object of the headline line: 192 solves by
    the batched multigrid CG with the autotuned launch shape."""
    from flowfpmpp_amd import stairstep as S
    t0 = time.perf_counter()
    probs, cells = batch64_frames()
    log(f"batch64: 64 frames x 3 directions resident ({cells} cells, "
        f"{time.perf_counter() - t0:.1f} s)")
    rows, zchunk, ms_it = S.autotune_stairstep_batch(probs, iterations=10, method="fcg-mg")
    out = run_batch(probs, "fcg-mg", tol, max_iters, check, steps, warmup, rows, zchunk)
    out.update({"workload": "64 frames (seeds 4000-4063), grids 64^3/96x80x64/128x96x64/"
                            "160x128x96, 1-3 bodies, 3 directions each: 192 solves in one "
                            "batched fcg-mg solve to 1e-10",
                "method": "fcg-mg", "cells": cells, "rows": rows, "zchunk": zchunk,
                "steps": steps, "warmup": warmup})
    return out


def bench_batch(args):
    """64 uneven frames (grids from 64^3 to 160x128x96, 1-3 bodies each) x 3
    force directions = 192 solves to the tolerance in one batched solve
    (shared operator per frame, per-solve masks, HIP-graph chunks). A step
    solves all 192 from phi = 0. With --method fcg-mg (the default) the line's
    value is the batched multigrid-preconditioned CG, and the same 192 solves
    are also timed as one batched mixed-precision PCG (autotuned launch shape,
    "pcg_mixed_batched") and as 192 fcg-mg solves one after another
    ("fcg_mg_sequential"); --method pcg / pcg-mixed time that batched PCG."""
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import synchronize
    t_gen = time.perf_counter()
    probs, cells = batch64_frames()
    log(f"64 frames x 3 directions resident: {cells} cells, {time.perf_counter() - t_gen:.1f} s")
    bmethod = args.method if args.method in ("pcg", "pcg-mixed", "fcg-mg") else "pcg-mixed"
    if args.shape:
        (rows, zchunk), ms_it = args.shape, None
    else:
        rows, zchunk, ms_it = S.autotune_stairstep_batch(
            probs, iterations=10 if bmethod == "fcg-mg" else 20, method=bmethod)
        log(f"autotuned launch shape: {rows} rows, z chunk {zchunk} ({ms_it:.3f} ms/iteration)")
    main = run_batch(probs, bmethod, args.tol, args.max_iters, args.check, args.steps,
                     args.warmup, rows, zchunk)
    out = {
        "metric": f"batched solves/s: 64 uneven frames x 3 directions, {bmethod} to 1e-10",
        "value": main["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": main["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"pcg": "f64", "pcg-mixed": "f32+f64", "fcg-mg": "f64+f32"}[bmethod],
        "data": "synthetic",
        "config": {"workload": "64 frames (seeds 4000-4063), grids 64^3/96x80x64/128x96x64/"
                               "160x128x96, 1-3 bodies, 3 directions each; one batched "
                               f"{bmethod} solve ({'given' if args.shape else 'autotuned'} "
                               f"launch shape), check {args.check}",
                   "cells": cells, "tolerance": args.tol, "rows": rows, "zchunk": zchunk},
    }
    out.update({k: main[k] for k in ("solve_ms_p50", "solve_ms_p99", "iterations_min_max",
                                     "max_residual", "converged")})
    if ms_it is not None:
        out["autotune_ms_per_iteration"] = round(ms_it, 4)
    if bmethod == "fcg-mg" and not args.no_compare:
        rows, zchunk, ms_it = S.autotune_stairstep_batch(probs, iterations=20,
                                                         method="pcg-mixed")
        mixed = run_batch(probs, "pcg-mixed", args.tol, args.max_iters, args.check, 1, 1, rows,
                          zchunk)
        mixed.update(rows=rows, zchunk=zchunk)
        out["pcg_mixed_batched"] = mixed
        for q in probs[:3]:
            q.phi.zero_()
            S.solve_stairstep_poisson_3d(q.phi, q.rhs, q.solid, q.body_grad, *q.faces,
                                         *q.axes, args.tol, args.max_iters, args.check,
                                         method="fcg-mg")
        synchronize()
        per, its, rs = [], [], []
        t0 = time.perf_counter()
        for q in probs:
            t1 = time.perf_counter()
            q.phi.zero_()
            sw, r = S.solve_stairstep_poisson_3d(q.phi, q.rhs, q.solid, q.body_grad, *q.faces,
                                                 *q.axes, args.tol, args.max_iters, args.check,
                                                 method="fcg-mg")
            per.append((time.perf_counter() - t1) * 1e3)
            its.append(sw)
            rs.append(r)
        synchronize()
        seq = time.perf_counter() - t0
        out["fcg_mg_sequential"] = {
            "ms": round(seq * 1e3, 3), "solves_per_s": round(192 / seq, 3),
            "solve_ms_p50": round(float(np.percentile(per, 50)), 3),
            "solve_ms_p99": round(float(np.percentile(per, 99)), 3),
            "iterations_min_max": [int(min(its)), int(max(its))],
            "max_residual": float(max(rs)), "converged": bool(max(rs) <= args.tol)}
        log(f"fcg-mg one after another: {seq * 1e3:.1f} ms for 192 solves, iterations "
            f"{min(its)}-{max(its)}")
        out["batch_speedup_vs_sequential"] = round(seq * 1e3 / main["ms"], 3)
    elif args.sequential:
        t0 = time.perf_counter()
        for q in probs:
            q.phi.zero_()
            S.solve_stairstep_poisson_3d(q.phi, q.rhs, q.solid, q.body_grad, *q.faces,
                                         *q.axes, args.tol, args.max_iters, args.check,
                                         method=bmethod)
        synchronize()
        seq = time.perf_counter() - t0
        out["sequential_ms"] = round(seq * 1e3, 3)
        out["batch_speedup_vs_sequential"] = round(seq * 1e3 / main["ms"], 3)
    print(json.dumps(out), flush=True)


def unet_flops(cells):
    """Multiply-add FLOPs of one forward pass of the U-Net on `cells` cells: the
    3x3x3 convolutions of its three levels (channels in x out per level:
    16x16, 16x16, 48x16, 16x16 at full resolution; 16x32, 32x32, 96x32, 32x32
    at 1/8; 32x64, 64x64 at 1/64)."""
    return 27 * 2 * cells * ((16 * 16 + 16 * 16 + 48 * 16 + 16 * 16)
                             + (16 * 32 + 32 * 32 + 96 * 32 + 32 * 32) / 8
                             + (32 * 64 + 64 * 64) / 64)


def bench_unet(args):
    """The learned preconditioner's network: forward passes on a frame of
    --grid (default 256x192x128) and training steps (forward, backward, Adam)
    at 128^3; MFMA TFLOP/s of the forward pass against the dense BF16 peak."""
    from flowfpmpp_amd import learned as L
    from flowfpmpp_amd import workloads as W
    from flowfpmpp_amd.device import DeviceArray, synchronize
    grid = tuple(args.grid) if tuple(args.grid) != (512, 384, 256) else (256, 192, 128)
    reps = max(1, args.steps)
    net = L.UNet(seed=1)
    rng = np.random.Generator(np.random.PCG64(1))
    p = W.stretched_frame(*grid, L.random_bodies(rng), seed=1, with_fields=False)
    feats = DeviceArray.from_host(L.features(p, 0))
    src = DeviceArray.from_host(rng.standard_normal(grid).astype(np.float32))
    out = DeviceArray(grid, np.float32)
    for _ in range(max(1, args.warmup)):
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
    flop = unet_flops(int(np.prod(grid)))
    tflops = flop / fwd / 1e12
    out = {"metric": "U-Net forward time (ms) and MFMA TFLOP/s", "value": round(fwd * 1e3, 3),
           "unit": "ms", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
           "ms_per_step": round(fwd * 1e3, 3), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
           "config": {"workload": f"U-Net forward on {grid[0]}x{grid[1]}x{grid[2]}, training "
                                  "steps on 128^3", "grid": list(grid)},
           "forward_tflop": round(flop / 1e12, 4), "forward_tflops": round(tflops, 1),
           "roofline": {"bound": "mfma", "achieved": round(tflops, 1), "peak": 2500.0,
                        "unit": "TFLOP/s", "frac": round(tflops / 2500.0, 4), "traffic": None},
           "train_grid": list(tg), "train_step_ms": round(step * 1e3, 3)}
    print(json.dumps(out), flush=True)


def bench_laplace(args):
    """The no-body Neumann solvers of fpm_laplace_3d on --grid (512x384x256 by
    default): the uniform solver (undamped Jacobi with wall ghosts) and the
    non-uniform one (0.2/0.8-damped Jacobi on the frame's stretched axes),
    each for a fixed --steps x 100 sweeps from zero on a random source and
    random face data (compatibility enforced), the max-norm residual checked
    every --check sweeps. Per-sweep times come from HIP events on the
    library's stream, for the streaming sweeps and for the one-cell-per-thread
    sweeps (FFPM_LAP=classic), whose gauged fields must be bit-identical."""
    from flowfpmpp_amd import laplace as LP
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import DeviceArray, synchronize
    T = DeviceArray.from_host
    from flowfpmpp_amd.grids import tanh_axis
    nx, ny, nz = args.grid
    cells = nx * ny * nz
    # the stretched axes of the influence-potential frames (width ratio 4)
    ax = [tanh_axis(n, -e, e, 4.0) for n, e in ((nx, 4.0), (ny, 3.0), (nz, 2.0))]
    rng = np.random.Generator(np.random.PCG64(7))
    rhs = np.asfortranarray(rng.standard_normal((nx, ny, nz)))
    faces = [np.asfortranarray(rng.standard_normal(s)) for s in
             ((ny, nz), (ny, nz), (nx, nz), (nx, nz), (nx, ny), (nx, ny))]
    centres = [np.ascontiguousarray(a[1], dtype=np.float64) for a in ax]
    widths = [np.ascontiguousarray(a[2], dtype=np.float64) for a in ax]
    h = (float(widths[0].mean()), float(widths[1].mean()), float(widths[2].mean()))
    sweeps = 100 * max(1, args.steps)
    check = max(1, min(args.check, sweeps))
    d = {}
    for kind in ("uniform", "nonuniform"):
        r, fs = T(rhs), [T(f) for f in faces]
        if kind == "uniform":
            LP.enforce_neumann_compatibility_3d(r, *fs, nx, ny, nz, *h)
        else:
            LP.enforce_neumann_compatibility_nonuniform_3d(r, *fs, *[T(w) for w in widths],
                                                           nx, ny, nz)
        d[kind] = (r, fs)
    axes = [T(a) for a in centres] + [T(w) for w in widths]
    res = {}
    prev = os.environ.get("FFPM_LAP")
    try:
        for mode in ("streaming", "classic"):
            if mode == "classic":
                os.environ["FFPM_LAP"] = "classic"
            else:
                os.environ.pop("FFPM_LAP", None)
            for kind in ("uniform", "nonuniform"):
                r, fs = d[kind]
                shape = (nx + 2, ny + 2, nz + 2) if kind == "uniform" else (nx, ny, nz)
                phi = DeviceArray(shape)
                phi.zero_()
                synchronize()
                S.timing(True)
                t0 = time.perf_counter()
                if kind == "uniform":
                    it, rr = LP.solve_neumann_poisson_3d(phi, r, *fs, nx, ny, nz, *h, -1.0,
                                                         sweeps, check)
                else:
                    it, rr = LP.solve_neumann_poisson_nonuniform_3d(phi, r, *fs, *axes, nx, ny,
                                                                    nz, -1.0, sweeps, check)
                synchronize()
                wall = time.perf_counter() - t0
                cnt, ms = S.timing_query("lap_sweep")
                S.timing(False)
                res[(mode, kind)] = (phi.download(), it, rr, ms / cnt, wall)
                log(f"laplace {kind} ({mode} sweeps): {it} sweeps, {ms / cnt * 1e3:.1f} us per "
                    f"sweep, {wall * 1e3:.1f} ms for the solve, residual {rr:.3e}")
    finally:
        if prev is None:
            os.environ.pop("FFPM_LAP", None)
        else:
            os.environ["FFPM_LAP"] = prev
    per = {}
    for kind in ("uniform", "nonuniform"):
        a, b = res[("streaming", kind)], res[("classic", kind)]
        same = bool(np.array_equal(a[0], b[0]) and a[1] == b[1] and a[2] == b[2])
        if not same:
            raise SystemExit(f"laplace {kind}: streaming and classic sweeps differ")
        us = a[3] * 1e3
        per[kind] = {"us_per_sweep": round(us, 1), "classic_us_per_sweep": round(b[3] * 1e3, 1),
                     "solve_ms": round(a[4] * 1e3, 1), "classic_solve_ms": round(b[4] * 1e3, 1),
                     "gbs": round(24 * cells / (us * 1e-6) / 1e9, 0),
                     "residual": a[2], "bit_identical_to_classic": same}
    worst = max(per[k]["us_per_sweep"] for k in per)
    gbs = 24 * cells / (worst * 1e-6) / 1e9
    out = {"metric": "no-body Jacobi sweep time (us, the slower of the two solvers)",
           "value": worst, "unit": "us", "n_gpus": 1, "steps": sweeps, "warmup": 0,
           "ms_per_step": round(worst / 1e3, 4), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"fpm_laplace_3d no-body solvers on {nx}x{ny}x{nz} "
                                  f"(uniform spacing {h[0]:.4g}/{h[1]:.4g}/{h[2]:.4g}; tanh-stretched "
                                  f"axes, width ratio 4), {sweeps} sweeps from zero, "
                                  f"residual every {check}", "grid": list(args.grid)},
           "solvers": per,
           "roofline": {"bound": "hbm", "achieved": round(gbs, 0), "peak": 8000.0,
                        "unit": "GB/s", "frac": round(gbs / 8000.0, 4), "traffic": None}}
    print(json.dumps(out), flush=True)


# The reference's published A100 times (seconds) for the same frames:
# benchmarks/pcg_time_to_solution.csv:5-9 and benchmarks/stairstep_benchmark.csv:9-12.
A100_PCG_1E8 = {(128, 128, 32): (1000, 0.1759), (256, 256, 64): (1900, 1.249),
                (512, 512, 64): (3600, 7.112)}
A100_JACOBI_FRAME = {(64, 64, 16): (0.00702, 19.9), (128, 128, 32): (0.0279, 26.1),
                     (256, 256, 64): (0.1862, 115.2), (512, 512, 64): (0.7039, 426.0)}
A100_JACOBI_1E8 = {(128, 128, 32): (226000, 5.329)}


def bench_jacobi(args):
    """Weighted Jacobi to 1e-10 on the 512x384x256 frame (x direction), capped:
    the solve runs in chunks of --check sweeps continuing from the previous
    phi (the sweep is memoryless and commutes with the gauge's constant
    shift), the FP64 max-norm residual is recorded after every chunk, and the
    sweeps still needed are extrapolated from the asymptotic rate of the last
    chunks; the time is that count times the measured per-sweep cost."""
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd import workloads as W
    from flowfpmpp_amd.device import DeviceArray, synchronize
    T = DeviceArray.from_host
    p = frame(tuple(args.grid))
    rhs, bg, faces = W.influence_inputs(p, 0)
    d_solid, d_rhs = T(p.solid), T(rhs)
    d_faces = [T(f) for f in faces]
    axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
    S.enforce_stairstep_compatibility_3d(d_rhs, d_solid, bg, *d_faces, *axes[3:])
    phi = T(p.phi0())
    chunk, total = args.check, 0
    hist = []
    synchronize()
    t0 = time.perf_counter()
    while total < args.max_iters:
        sw, res = S.solve_stairstep_poisson_3d(phi, d_rhs, d_solid, bg, *d_faces, *axes, -1.0,
                                               chunk, chunk, method="jacobi")
        total += sw
        hist.append((total, res))
        log(f"jacobi: {total} sweeps, residual {res:.3e}")
        if res <= args.tol:
            break
    synchronize()
    elapsed = time.perf_counter() - t0
    per_sweep_ms = elapsed * 1e3 / total
    tail = hist[-4:]
    rate = (np.log(tail[-1][1]) - np.log(tail[0][1])) / (tail[-1][0] - tail[0][0])
    res_now = hist[-1][1]
    more = 0.0 if res_now <= args.tol else (np.log(args.tol) - np.log(res_now)) / rate
    sweeps = total + more
    out = {
        "metric": "weighted Jacobi time to solve phi_x to 1e-10 (ms), extrapolated",
        "value": round(sweeps * per_sweep_ms, 1), "unit": "ms", "n_gpus": 1,
        "steps": 1, "warmup": 0, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{p.label} frame, x direction, weighted Jacobi (0.2/0.8), "
                               f"{total} sweeps run in chunks of {chunk}, then the sweeps left "
                               "to the tolerance from the rate of the last three chunks",
                   "grid": list(args.grid), "tolerance": args.tol},
        "sweeps_run": total, "residual_history": [[int(a), float(b)] for a, b in hist],
        "rate_per_sweep": float(rate), "sweeps_extrapolated": int(round(sweeps)),
        "ms_per_sweep": round(per_sweep_ms, 4),
    }
    print(json.dumps(out), flush=True)


def bench_reference(args):
    """The reference's benchmark (tests/benchmark_stairstep_3d.f90:51-77: a
    stair-step sphere on a sinusoidally stretched grid, x influence field,
    then the FPM terms) timed as the reference times it
    (tests/benchmark_stairstep_3d.f90:142-177): compatibility + solve +
    gauge + FPM terms including every host-device transfer, median of 3
    after a warm-up; host numpy arrays through the host-pointer entry
    points. PCG to 1e-8 (check 100), the 250-sweep Jacobi frame, the
    marginal Jacobi sweep (solve at 2500 minus 250 sweeps, / 2250) and Jacobi
    to 1e-8 at 128x128x32."""
    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd import workloads as W

    def frame_time(p, method, tol, cap, check):
        rhs, bg, faces = W.influence_inputs(p, 0)
        f = p.fields

        def once():
            fc = [np.array(a, order="F", copy=True) for a in faces]
            phi = p.phi0()
            t0 = time.perf_counter()
            S.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *fc, p.dx, p.dy, p.dz)
            sw, res = S.solve_stairstep_poisson_3d(phi, rhs, p.solid, bg, *fc, p.xc, p.yc, p.zc,
                                                   p.dx, p.dy, p.dz, tol, cap, check,
                                                   method=method)
            S.compute_stairstep_fpm_terms_3d(phi, f["u"], f["v"], f["w"], f["ox"], f["oy"],
                                             f["oz"], f["ax"], f["ay"], f["az"], p.solid, bg,
                                             p.xc, p.yc, p.zc, p.dx, p.dy, p.dz, *fc, p.bounds)
            return time.perf_counter() - t0, sw, res

        once()
        runs = [once() for _ in range(3)]
        t = sorted(r[0] for r in runs)[1]
        return t, runs[0][1], runs[0][2]

    rows = []
    for g, (it_ref, a100) in A100_PCG_1E8.items():
        p = W.reference_benchmark(*g)
        t, sw, res = frame_time(p, "pcg", 1e-8, 1000000, 100)
        rows.append({"grid": list(g), "method": "pcg", "tol": 1e-8, "iterations": sw,
                     "reference_iterations": it_ref, "residual": res, "frame_s": round(t, 4),
                     "a100_frame_s": a100, "speedup_vs_a100": round(a100 / t, 2)})
        log(f"PCG {g}: {sw} iterations, {t:.3f} s (A100 {a100} s)")
        # the same frame to the same tolerance with the multigrid CG (no
        # counterpart in the reference): the time a caller of this library
        # waits for the frame
        t, sw, res = frame_time(p, "fcg-mg", 1e-8, 1000000, 100)
        rows.append({"grid": list(g), "method": "fcg-mg", "tol": 1e-8, "iterations": sw,
                     "residual": res, "frame_s": round(t, 4), "a100_pcg_frame_s": a100,
                     "speedup_vs_a100_pcg": round(a100 / t, 2)})
        log(f"fcg-mg {g}: {sw} iterations, {t:.3f} s (A100 PCG {a100} s)")
    for g, (a100, a100_sweep) in A100_JACOBI_FRAME.items():
        p = W.reference_benchmark(*g)
        t, sw, res = frame_time(p, "jacobi", -1.0, 250, 250)
        t250 = frame_time(p, "jacobi", -1.0, 250, 2500)[0]
        t2500 = frame_time(p, "jacobi", -1.0, 2500, 2500)[0]
        sweep_us = (t2500 - t250) / 2250 * 1e6
        rows.append({"grid": list(g), "method": "jacobi", "sweeps": 250, "frame_s": round(t, 4),
                     "a100_frame_s": a100, "speedup_vs_a100": round(a100 / t, 2),
                     "sweep_us": round(sweep_us, 2), "a100_sweep_us": a100_sweep,
                     "sweep_speedup_vs_a100": round(a100_sweep / sweep_us, 2)})
        log(f"Jacobi {g}: 250-sweep frame {t:.4f} s (A100 {a100}), sweep {sweep_us:.1f} us "
            f"(A100 {a100_sweep})")
    for g, (sw_ref, a100) in A100_JACOBI_1E8.items():
        p = W.reference_benchmark(*g)
        t, sw, res = frame_time(p, "jacobi", 1e-8, 10000000, 100)
        rows.append({"grid": list(g), "method": "jacobi", "tol": 1e-8, "sweeps": sw,
                     "reference_sweeps": sw_ref, "residual": res, "frame_s": round(t, 4),
                     "a100_frame_s": a100, "speedup_vs_a100": round(a100 / t, 2)})
        log(f"Jacobi {g} to 1e-8: {sw} sweeps, {t:.3f} s (A100 {a100} s)")
    pcg = [r for r in rows if r["method"] == "pcg"]
    out = {"metric": "reference benchmark frame time (s): compatibility + solve + gauge + FPM "
                     "terms with host-device transfers", "value": pcg[-1]["frame_s"], "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "tests/benchmark_stairstep_3d.f90 frames (sphere, sinusoidal "
                                  "stretching), x influence field; value = 512x512x64 PCG to 1e-8"},
           "vs_baseline": None, "rows": rows}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--method", default="fcg-mg", choices=["pcg", "jacobi", "pcg-mixed", "fcg-mg"])
    ap.add_argument("--no-pcg", action="store_true",
                    help="skip the FP64 diagonal PCG comparison solve")
    ap.add_argument("--grid", type=int, nargs=3, default=[512, 384, 256])
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--check", type=int, default=100)
    ap.add_argument("--max-iters", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="frame workload: skip the batch64 object (192 batched solves)")
    ap.add_argument("--no-learned", action="store_true",
                    help="frame workload: skip the learned_fcg object (the multigrid CG "
                         "with the trained network in its coarse correction)")
    ap.add_argument("--workload", default="frame",
                    choices=["frame", "batch64", "reference", "jacobi", "unet", "laplace"],
                    help="frame: the 3-direction solve of one frame (the headline); batch64: "
                         "64 uneven frames x 3 directions in one batched solve; reference: "
                         "the reference's own benchmark frames, timed as it times them; "
                         "jacobi: weighted Jacobi on the frame's x direction, capped at "
                         "--max-iters sweeps in chunks of --check, extrapolated to --tol; "
                         "unet: the learned preconditioner's network, forward passes on "
                         "--grid (256x192x128 by default) and training steps at 128^3; "
                         "laplace: the no-body Neumann solvers (uniform and non-uniform), "
                         "--steps x 100 sweeps on --grid, streaming against classic sweeps")
    ap.add_argument("--sequential", action="store_true",
                    help="batch64: also time the 192 solves one after another")
    ap.add_argument("--shape", type=int, nargs=2, metavar=("ROWS", "ZCHUNK"),
                    help="batch64: this launch shape instead of the autotuned one")
    ap.add_argument("--sequential-directions", action="store_true",
                    help="frame workload: solve the three directions one after another on one "
                         "stream (default: at once, one host thread and stream each)")
    ap.add_argument("--no-compare", action="store_true",
                    help="batch64 with fcg-mg: skip the batched mixed-PCG and sequential runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the whole frame lives on one MI355X and this build "
              "has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from flowfpmpp_amd import stairstep as S
    from flowfpmpp_amd.device import DeviceArray, synchronize

    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "reference":
        return bench_reference(args)
    if args.workload == "jacobi":
        return bench_jacobi(args)
    if args.workload == "unet":
        return bench_unet(args)
    if args.workload == "laplace":
        return bench_laplace(args)

    t_gen = time.perf_counter()
    p = frame(tuple(args.grid))
    T = DeviceArray.from_host
    dev = {k: T(v) for k, v in p.fields.items()}
    d_solid = T(p.solid)
    axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
    d_rhs = T(p.zeros())
    d_phi = [T(p.phi0()) for _ in range(3)]
    faces0 = p.faces()
    d_faces = [[T(f) for f in faces0] for _ in range(3)]
    outs = [[DeviceArray((p.nx, p.ny, p.nz)) for _ in range(5)] for _ in range(3)]
    bgs = [np.eye(3)[d] for d in range(3)]
    synchronize()
    gen_s = time.perf_counter() - t_gen
    log(f"inputs resident on the device: {p.label}, {p.cells} cells, {gen_s:.1f} s")

    mixed_outer = []
    # The three directions are independent solves: by default each runs on
    # its own host thread and HIP stream (the library keeps one workspace per
    # stream), so one direction's launch-bound coarse levels and launch gaps
    # overlap the others' level-0 kernels; the results are bit-identical to
    # solving them one after another.
    concurrent = not args.sequential_directions
    streams, pool = [None, None, None], None
    if concurrent:
        from concurrent.futures import ThreadPoolExecutor
        streams = [S.stream_create() for _ in range(3)]
        pool = ThreadPoolExecutor(max_workers=3)

    def solve_dir(d, method, stream):
        if stream is not None:
            S.set_stream(stream)
        d_phi[d].zero_()
        for f in d_faces[d]:
            f.zero_()
        S.enforce_stairstep_compatibility_3d(d_rhs, d_solid, bgs[d], *d_faces[d], *axes[3:])
        sw, r = S.solve_stairstep_poisson_3d(d_phi[d], d_rhs, d_solid, bgs[d], *d_faces[d],
                                             *axes, args.tol, args.max_iters, args.check,
                                             method=method)
        mo = S.last_mixed_solve()[0] if method == "pcg-mixed" else None
        if stream is not None:
            synchronize()
        log(f"{method} direction {d}: {sw} iterations, residual {r:.3e}")
        return sw, r, mo

    def step(method=args.method, conc=concurrent):
        if conc:
            futs = [pool.submit(solve_dir, d, method, streams[d]) for d in range(3)]
            rs = [f.result() for f in futs]
        else:
            rs = [solve_dir(d, method, None) for d in range(3)]
        iters, res = [x[0] for x in rs], [x[1] for x in rs]
        mixed_outer[:] = [x[2] for x in rs] if method == "pcg-mixed" else []
        # the three directions' densities in one pass over the flow fields
        # (each equal bit for bit to its own compute_stairstep_fpm_terms_3d)
        o = S.compute_stairstep_fpm_terms_3d_dirs(
            d_phi, *[dev[k] for k in ("u", "v", "w", "ox", "oy", "oz", "ax", "ay", "az")],
            d_solid, bgs, *axes, d_faces, p.bounds, outs)
        totals = [[float(x) for x in o[d][5]] for d in range(3)]
        return iters, res, totals

    for _ in range(args.warmup):
        step()
    synchronize()
    if not concurrent:
        S.timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        iters, res, totals = step()
    synchronize()
    elapsed = time.perf_counter() - t0
    seq_ms = None
    if concurrent:
        # One more step with the directions one after another: its time (the
        # like-for-like reference for the learned variant, which runs on one
        # stream), then once more with every hot launch timed by HIP events
        # for the roofline (in the concurrent step the launches of three
        # streams overlap, so per-launch event times there would measure the
        # sharing, not the kernel).
        t1 = time.perf_counter()
        step(conc=False)
        synchronize()
        seq_ms = (time.perf_counter() - t1) * 1e3
        S.timing(True)
        step(conc=False)
        synchronize()
    S.timing(False)

    ms_per_step = elapsed * 1e3 / max(1, args.steps)
    cells = p.cells
    kern = {}
    for name, bpc in BYTES_PER_CELL.items():
        cnt, ms = S.timing_query(name)
        if cnt:
            avg_ms = ms / cnt
            kern[name] = {"launches": cnt, "avg_us": avg_ms * 1e3, "total_ms": ms,
                          "gbps": bpc * cells / (avg_ms * 1e-3) / 1e9}
    dom = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None
    roofline = None
    if dom:
        ach = kern[dom]["gbps"]
        traffic = measured_traffic(dom, list(args.grid))
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": traffic,
                    "bytes_per_cell": BYTES_PER_CELL[dom], "cells_per_launch": cells,
                    "avg_launch_us": round(kern[dom]["avg_us"], 2)}
        if traffic:
            roofline["traffic_gbps"] = round(traffic / (kern[dom]["avg_us"] * 1e-6) / 1e9, 1)
    out = {
        "metric": METRIC,
        "value": round(ms_per_step, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.method in ("pcg", "jacobi") else "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{p.label} tanh-stretched 4:1 frame, seeded bodies and Fourier fields; "
                         f"3 directions x (compatibility + {args.method} to {args.tol:g} "
                         f"max-norm, check {args.check} + gauge + FPM densities)"),
            "grid": list(args.grid), "cells": cells, "method": args.method,
            "tolerance": args.tol, "check_interval": args.check,
            "directions": ("at once (three host threads, one HIP stream each)" if concurrent
                           else "one after another (one stream)"),
        },
        "iterations": iters,
        "mixed_outer_corrections": mixed_outer if args.method == "pcg-mixed" else None,
        "multigrid_levels": S.last_mg_solve()[0] if args.method == "fcg-mg" else None,
        "residuals": res,
        "converged": all(r <= args.tol for r in res) if args.tol > 0 else None,
        "solves_per_s": round(3 * 1e3 / ms_per_step, 4),
        "kernels": {k: {kk: round(vv, 3) for kk, vv in v.items()} for k, v in kern.items()},
        "roofline": roofline,
        "input_setup_s": round(gen_s, 2),
    }
    if concurrent:
        out["sequential_directions_ms"] = round(seq_ms, 3)
        out["kernels_timed_in"] = "a step with the directions one after another"
    import ctypes
    from flowfpmpp_amd._lib import lib
    gb, us = ctypes.c_double(), ctypes.c_double()
    lib().ffpm_copy_bandwidth(ctypes.c_size_t(cells * 8), 20, ctypes.byref(gb), ctypes.byref(us))
    out["copy_kernel_gbps"] = round(gb.value, 1)
    if roofline:
        roofline["frac_of_copy"] = round(roofline["achieved"] / gb.value, 4)
    pcg_iters = iters if args.method in ("pcg", "jacobi") else None
    if args.method != "pcg" and not args.no_pcg:
        log("FP64 diagonal PCG on the same frame (one step, directions as above)")
        synchronize()
        t1 = time.perf_counter()
        pcg_iters, pcg_res, pcg_totals = step("pcg")
        synchronize()
        pcg_ms = (time.perf_counter() - t1) * 1e3
        dtot = max(abs(a - b) / max(abs(b), 1e-300) for ta, tb in zip(totals, pcg_totals)
                   for a, b in zip(ta, tb))
        out["fp64_pcg"] = {"ms_per_step": round(pcg_ms, 3), "iterations": pcg_iters,
                           "residuals": pcg_res, "check_interval": args.check,
                           "speedup": round(pcg_ms / ms_per_step, 2),
                           "max_rel_diff_of_force_totals": dtot}
        out["plain_stencil"] = plain_stencil(p, args, d_rhs, d_solid, axes, d_faces[0],
                                             d_phi[0], bgs[0])
    if not args.no_learned and tuple(args.grid) == (512, 384, 256):
        log("multigrid CG with the learned coarse correction on the same frame")
        out["learned_fcg"] = learned_summary(p, args, step_learned_args(
            p, d_phi, d_rhs, d_solid, bgs, axes, d_faces, dev, outs),
            seq_ms if seq_ms else ms_per_step, out.get("fp64_pcg", {}).get("ms_per_step"))
    if not args.no_batch and tuple(args.grid) == (512, 384, 256):
        log("batched workload: 64 uneven frames x 3 directions")
        out["batch64"] = batch64_summary(args.tol, args.max_iters, args.check)
    if not args.no_cpu and pcg_iters:
        # the reference's CPU path runs many MPI ranks on a node: the baseline
        # is the all-thread OpenMP build of the oracle, the serial one beside it
        cm = "jacobi" if args.method == "jacobi" else "pcg"
        if cm == "pcg":
            log("timing the CPU oracle baseline (all threads)")
            out["cpu_baseline"] = cpu_baseline(p, pcg_iters, args.tol, cm, parallel=True)
            log("timing the CPU oracle baseline (one thread)")
            out["cpu_baseline_serial"] = cpu_baseline(p, pcg_iters, args.tol, cm)
        else:
            log("timing the CPU oracle baseline")
            out["cpu_baseline"] = cpu_baseline(p, pcg_iters, args.tol, cm)
        if cm == "pcg" and tuple(args.grid) == (512, 384, 256):
            log("measured CPU and GPU time to solution on the 128x96x64 frame")
            out["cpu_time_to_solution_128x96x64"] = cpu_time_to_solution_c1(args.tol,
                                                                            args.check)
    if concurrent:
        pool.shutdown()
        for h in streams:
            S.stream_destroy(h)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
