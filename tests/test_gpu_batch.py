"""GPU tests of the batched (multi-right-hand-side, uneven-frame) FP64 and
mixed-precision PCG.

Bar: every problem of a batch computes bit-for-bit what its standalone
solve_stairstep_poisson_3d with the same method computes -- the batched kernels run
the standalone kernels' bodies with each problem's own launch shape, so the
iterates, the reductions and the convergence decisions are identical: same
sweeps, same residual, identical phi.
"""
import numpy as np
import pytest

from flowfpmpp_amd import stairstep as S
from flowfpmpp_amd import workloads as W
from flowfpmpp_amd._lib import FFPMError
from flowfpmpp_amd.device import DeviceArray

pytestmark = pytest.mark.gpu

T = DeviceArray.from_host


class Frame:
    """Device-resident frame with the three direction problems (compatibility
    applied on the device)."""

    def __init__(self, nx, ny, nz, bodies, seed, dirs=(0, 1, 2)):
        p = W.stretched_frame(nx, ny, nz, bodies, seed=seed, with_fields=False)
        self.p = p
        self.solid = T(p.solid)
        self.axes = [T(a) for a in (p.xc, p.yc, p.zc, p.dx, p.dy, p.dz)]
        self.dirs = []
        for d in dirs:
            rhs, bg, faces = W.influence_inputs(p, d)
            d_rhs = T(rhs)
            d_faces = [T(f) for f in faces]
            S.enforce_stairstep_compatibility_3d(d_rhs, self.solid, bg, *d_faces,
                                                 *self.axes[3:])
            self.dirs.append((d_rhs, bg, d_faces))

    def problems(self):
        out = []
        for d_rhs, bg, d_faces in self.dirs:
            out.append(S.BatchProblem(T(self.p.phi0()), d_rhs, self.solid, bg, *d_faces,
                                      *self.axes))
        return out

    def standalone(self, tol, cap, check, method="pcg"):
        res = []
        for d_rhs, bg, d_faces in self.dirs:
            phi = T(self.p.phi0())
            sw, r = S.solve_stairstep_poisson_3d(phi, d_rhs, self.solid, bg, *d_faces,
                                                 *self.axes, tol, cap, check, method=method)
            res.append((phi.download(), sw, r))
        return res


def _assert_same(problems, ref):
    for q, (phi, sw, r) in zip(problems, ref):
        assert q.sweeps == sw
        assert q.residual == r
        assert np.array_equal(q.phi.download(), phi)


def _uneven_frames():
    return [
        Frame(48, 36, 24, [W.Sphere((0.0, 0.0, 0.0), 0.9)], seed=11),
        Frame(40, 30, 20, [W.Box((-1.0, -1.0, -0.5), (0.5, 0.8, 0.6))], seed=12, dirs=(0,)),
        Frame(33, 25, 17, [W.Sphere((0.5, 0.2, 0.0), 1.0)], seed=13, dirs=(1, 2)),
        Frame(64, 48, 32, [W.Sphere((-1.0, 0.0, 0.0), 0.7),
                           W.Box((3.0, -1.0, -0.7), (4.5, 1.0, 0.7))], seed=14),
    ]


@pytest.mark.parametrize("tol,cap,check", [(-1.0, 120, 40), (1e-10, 100000, 25)])
def test_multi_rhs_equals_standalone(tol, cap, check):
    """The three directions of one frame share one operator in the batch and
    each matches its own standalone solve exactly (fixed work and converged)."""
    print(f"case: multi-RHS 48x36x24, tol {tol}, cap {cap}, check {check}", flush=True)
    f = Frame(48, 36, 24, [W.Sphere((0.0, 0.0, 0.0), 0.9),
                           W.Box((3.0, -1.0, -0.7), (4.5, 1.0, 0.7))], seed=7)
    ref = f.standalone(tol, cap, check)
    probs = f.problems()
    S.solve_stairstep_batch(probs, tol, cap, check)
    _assert_same(probs, ref)
    if tol > 0:
        assert all(q.residual <= tol for q in probs)


@pytest.mark.parametrize("use_graph", [True, False])
def test_uneven_batch_equals_standalone(use_graph):
    """Frames of different sizes (nx even and odd, so both cells-per-lane
    widths) with one to three directions each, converged to 1e-10."""
    print(f"case: uneven batch, graph {use_graph}", flush=True)
    frames = _uneven_frames()
    ref = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25)]
    probs = [q for fr in frames for q in fr.problems()]
    S.solve_stairstep_batch(probs, 1e-10, 100000, 25, use_graph=use_graph)
    _assert_same(probs, ref)
    ms = [q.solve_ms for q in probs]
    assert all(m > 0 for m in ms)
    # the largest problem needs the most iterations and finishes last
    big = max(range(len(probs)), key=lambda k: probs[k].sweeps)
    assert ms[big] == max(ms)


def test_max_sweeps_quantised_like_standalone():
    """A cap that is not a multiple of the check interval ends every problem
    after the same number of iterations as its standalone solve."""
    print("case: uneven batch, cap 70, check 25", flush=True)
    frames = _uneven_frames()[:2]
    ref = [r for fr in frames for r in fr.standalone(1e-12, 70, 25)]
    probs = [q for fr in frames for q in fr.problems()]
    S.solve_stairstep_batch(probs, 1e-12, 70, 25)
    _assert_same(probs, ref)
    assert all(q.sweeps == 70 for q in probs)


def test_autotuned_shape_converges():
    """The autotuner picks one of its candidate shapes; the batch solved with
    it converges to the same answer (other reduction order: 1e-10 relative)."""
    print("case: autotune uneven batch", flush=True)
    frames = _uneven_frames()
    probs = [q for fr in frames for q in fr.problems()]
    rows, zchunk, ms = S.autotune_stairstep_batch(probs, iterations=10)
    assert rows in (4, 8) and zchunk in (8, 16, 32) and ms > 0
    assert all(not np.any(q.phi.download()) for q in probs)  # phi untouched
    S.solve_stairstep_batch(probs, 1e-10, 100000, 25, rows=rows, zchunk=zchunk)
    ref = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25)]
    for q, (phi, sw, r) in zip(probs, ref):
        assert q.residual <= 1e-10
        assert np.abs(q.phi.download() - phi).max() <= 1e-10 * np.abs(phi).max()


@pytest.mark.parametrize("use_graph", [True, False])
def test_mixed_batch_equals_standalone_mixed(use_graph):
    """Mixed precision in a batch: each problem runs the standalone mixed
    solve's defect-correction loop on the device and ends bit-identical to it
    (same FP32 inner iteration count, same residual, identical phi). The
    uneven frames cover nx a multiple of 4, even only, and odd."""
    print(f"case: uneven mixed batch, graph {use_graph}", flush=True)
    frames = _uneven_frames()
    ref = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25, method="pcg-mixed")]
    probs = [q for fr in frames for q in fr.problems()]
    S.solve_stairstep_batch(probs, 1e-10, 100000, 25, use_graph=use_graph, method="pcg-mixed")
    _assert_same(probs, ref)
    assert all(q.residual <= 1e-10 for q in probs)


def test_batch_errors():
    print("case: batch argument errors", flush=True)
    with pytest.raises(FFPMError):
        S.solve_stairstep_batch([], 1e-10, 100, 10)
    f = Frame(32, 24, 16, [W.Sphere((0.0, 0.0, 0.0), 0.9)], seed=3, dirs=(0,))
    with pytest.raises(FFPMError):
        S.solve_stairstep_batch(f.problems(), 1e-10, 100, 10, rows=3)
    with pytest.raises(ValueError):
        S.solve_stairstep_batch(f.problems(), 1e-10, 100, 10, method="jacobi")


def _multigrid_frames():
    # hierarchy depths 1 (a frame of at most 1024 cells), 2 and 3; nx a
    # multiple of 4, even only, and odd (all level-0 launch widths and both
    # level-1 column variants)
    return _uneven_frames() + [Frame(12, 10, 8, [W.Sphere((0.0, 0.0, 0.0), 1.2)], seed=15,
                                     dirs=(0, 2))]


@pytest.mark.parametrize("use_graph", [True, False])
def test_multigrid_batch_equals_standalone(use_graph):
    """The multigrid-preconditioned CG in a batch: frames of different sizes
    and hierarchy depths, one launch per cycle step over every solve that has
    the level; each problem ends bit-identical to its standalone fcg-mg solve
    (same iterations, same residual, identical phi)."""
    print(f"case: uneven fcg-mg batch, graph {use_graph}", flush=True)
    frames = _multigrid_frames()
    ref = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25, method="fcg-mg")]
    probs = [q for fr in frames for q in fr.problems()]
    S.solve_stairstep_batch(probs, 1e-10, 100000, 25, use_graph=use_graph, method="fcg-mg")
    _assert_same(probs, ref)
    assert all(q.residual <= 1e-10 for q in probs)
    assert all(0 < q.sweeps < 200 for q in probs)


@pytest.mark.parametrize("method", ["fcg-mg", "pcg"])
def test_lane_fit_rule_batch_equals_standalone(method, monkeypatch):
    """FFPM_LANE_FIT=1 gives the frames of these tests one cell per lane (in
    place of the widest vector dividing nx); the batched and standalone solves
    use the same rule, so each problem stays bit-identical to its standalone
    solve, and the fields agree with the default rule's to 1e-10 relative."""
    print(f"case: lane-fit rule, {method} batch", flush=True)
    frames = _multigrid_frames()
    wide = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25, method=method)]
    monkeypatch.setenv("FFPM_LANE_FIT", "1")
    ref = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25, method=method)]
    probs = [q for fr in frames for q in fr.problems()]
    S.solve_stairstep_batch(probs, 1e-10, 100000, 25, method=method)
    monkeypatch.delenv("FFPM_LANE_FIT")
    _assert_same(probs, ref)
    for q, (phi, sw, r) in zip(probs, wide):
        assert q.residual <= 1e-10
        assert np.abs(q.phi.download() - phi).max() <= 1e-10 * np.abs(phi).max()


def test_multigrid_batch_fixed_work():
    """Fixed work (negative tolerance) with a cap that is not a multiple of the
    two-iteration chunk: every problem runs exactly the cap, as standalone."""
    print("case: fcg-mg batch, fixed work 7 iterations", flush=True)
    frames = _multigrid_frames()[:3]
    ref = [r for fr in frames for r in fr.standalone(-1.0, 7, 25, method="fcg-mg")]
    probs = [q for fr in frames for q in fr.problems()]
    S.solve_stairstep_batch(probs, -1.0, 7, 25, method="fcg-mg")
    _assert_same(probs, ref)
    assert all(q.sweeps == 7 for q in probs)


def test_multigrid_batch_autotuned_shape():
    """The autotuner times the multigrid CG's launch shapes too; a batch solved
    with the chosen shape (other reduction order than the standalone solves)
    converges to the same fields within 1e-10 relative."""
    print("case: autotune uneven fcg-mg batch", flush=True)
    frames = _multigrid_frames()
    probs = [q for fr in frames for q in fr.problems()]
    rows, zchunk, ms = S.autotune_stairstep_batch(probs, iterations=6, method="fcg-mg")
    assert rows in (4, 8) and zchunk in (8, 16, 32) and ms > 0
    assert all(not np.any(q.phi.download()) for q in probs)  # phi untouched
    S.solve_stairstep_batch(probs, 1e-10, 100000, 25, rows=rows, zchunk=zchunk,
                            method="fcg-mg")
    ref = [r for fr in frames for r in fr.standalone(1e-10, 100000, 25, method="fcg-mg")]
    for q, (phi, sw, r) in zip(probs, ref):
        assert q.residual <= 1e-10
        assert abs(q.sweeps - sw) <= 2
        assert np.abs(q.phi.download() - phi).max() <= 1e-10 * np.abs(phi).max()
