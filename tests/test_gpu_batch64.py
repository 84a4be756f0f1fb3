"""The batched-uneven-frames workload at its own size: 64 frames (seeds
4000-4063, grids 64^3 / 96x80x64 / 128x96x64 / 160x128x96, 1-3 bodies each)
x 3 force directions = 192 solves in one batched multigrid-preconditioned CG
to the FP64 max-norm residual 1e-10 (src/fpm_stairstep_mpi_3d.f90:514-530).

Bars: every residual <= 1e-10; twelve sampled solves bit-identical to their
standalone solves (same iterations, residual and phi); one instance of each
grid size within 1e-10 relative max-norm of the oracle's FP64 PCG (OpenMP
build, converged to the same residual)."""
import numpy as np
import pytest

from flowfpmpp_amd import stairstep as S
from flowfpmpp_amd import workloads as W
from flowfpmpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

T = DeviceArray.from_host
SAMPLED = (0, 7, 19, 33, 48, 63, 95, 120, 141, 160, 177, 191)


def test_batch_of_64_uneven_frames_fcg_mg():
    print("case: batch64, 64 frames x 3 directions, batched fcg-mg to 1e-10", flush=True)
    frames, probs = [], []
    for i in range(64):
        p = W.batch_instance(i)
        solid = T(p.solid)
        axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
        for d in range(3):
            rhs, bg, faces = W.influence_inputs(p, d)
            d_rhs, d_faces = T(rhs), [T(f) for f in faces]
            S.enforce_stairstep_compatibility_3d(d_rhs, solid, bg, *d_faces, *axes[3:])
            probs.append(S.BatchProblem(T(p.phi0()), d_rhs, solid, bg, *d_faces, *axes))
        frames.append(p)
    S.solve_stairstep_batch(probs, 1e-10, 100000, 100, method="fcg-mg")
    res = np.array([q.residual for q in probs])
    its = np.array([q.sweeps for q in probs])
    print(f"  192 solves: iterations {its.min()}-{its.max()}, max residual {res.max():.2e}",
          flush=True)
    assert np.all(res <= 1e-10)
    for k in SAMPLED:
        q = probs[k]
        print(f"case: batch64 solve {k} (frame {k // 3}, direction {k % 3}) standalone",
              flush=True)
        phi = T(frames[k // 3].phi0())
        sw, r = S.solve_stairstep_poisson_3d(phi, q.rhs, q.solid, q.body_grad, *q.faces,
                                             *q.axes, 1e-10, 100000, 100, method="fcg-mg")
        assert sw == q.sweeps and r == q.residual
        assert np.array_equal(phi.download(), q.phi.download())


def test_batch64_one_frame_per_grid_against_oracle(orc):
    seen = {}
    for i in range(64):
        g = tuple(int(v) for v in W.batch_instance_grid(i))
        seen.setdefault(g, i)
    assert len(seen) == len(W.BATCH_GRIDS)
    for g, i in sorted(seen.items()):
        p = W.batch_instance(i)
        d = i % 3
        print(f"case: batch64 frame {i} ({g[0]}x{g[1]}x{g[2]}), direction {d}: batched fcg-mg "
              "against the oracle's PCG", flush=True)
        rhs, bg, faces = W.influence_inputs(p, d)
        orc.enforce_stairstep_compatibility_3d(rhs, p.solid, bg, *faces, p.dx, p.dy, p.dz)
        q = S.BatchProblem(T(p.phi0()), T(rhs), T(p.solid), bg, *[T(f) for f in faces],
                           *[T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)])
        # a one-frame batch beside a second frame keeps the batched path honest
        other = W.batch_instance((i + 1) % 64)
        r2, b2, f2 = W.influence_inputs(other, 0)
        S.enforce_stairstep_compatibility_3d(r2, other.solid, b2, *f2, other.dx, other.dy,
                                             other.dz)
        q2 = S.BatchProblem(T(other.phi0()), T(r2), T(other.solid), b2, *[T(f) for f in f2],
                            *[T(a) for a in (other.xc, other.yc, other.zc, other.dx, other.dy,
                                             other.dz)])
        S.solve_stairstep_batch([q, q2], 1e-10, 100000, 100, method="fcg-mg")
        po = p.phi0()
        so, ro = orc.solve_stairstep_poisson_3d(po, rhs, p.solid, bg, *faces, p.xc, p.yc, p.zc,
                                                p.dx, p.dy, p.dz, 1e-10, 200000, 100, "pcg",
                                                parallel=True)
        ph = q.phi.download()
        e = float(np.abs(ph - po).max() / np.abs(po).max())
        print(f"  batched fcg-mg {q.sweeps} iterations ({q.residual:.2e}), oracle PCG {so} "
              f"({ro:.2e}); rel |phi| {e:.2e}", flush=True)
        assert q.residual <= 1e-10 and ro <= 1e-10 and q2.residual <= 1e-10
        assert e <= 1e-10
