// Batched stair-step PCG: many independent solves advance together, one grid
// per kernel for the whole batch. A batch holds uneven frames, and the three
// force-direction right-hand sides of one frame share that frame's operator
// (neighbour codes, diagonals and 1-D coefficients are built once).
//
// Every workgroup of a batched kernel looks up its solve and its index inside
// that solve's own launch shape and runs the same device body as the
// standalone kernel (ffpm_bodies.hpp, ffpm_pcg_body.hpp), so each solve
// computes bit-for-bit what its standalone solve computes. The convergence
// logic of pcg_iterate (src/fpm_stairstep_mpi_3d.f90:235-293, as in the
// standalone fused iteration) runs per solve on the device, with a per-solve
// active mask: a converged, stalled or failed solve turns every later launch
// of its workgroups into a no-op. One chunk of check_interval iterations plus
// the convergence check is captured once as a HIP graph and replayed until no
// solve is active; the host reads one counter per replay. In mixed precision
// each solve also runs the FP64 defect-correction loop of the standalone mixed
// solve (FP32 inner PCG to a relative rz tolerance, FP64 accumulation and true
// residual) as a per-solve phase on the device.
//
// The multigrid-preconditioned CG (FFPM_METHOD_FCG_MG) batches the same way:
// every solve gets its own hierarchy state (the level operators are shared by
// the solves of one frame), each level's cycle step runs as one launch over
// every solve that has that level, the coarsest levels (at different depths)
// run in one launch of one workgroup per solve, and the host loop of the
// standalone fcg_mg_iterate (restart, gated true residual, stall exit) runs
// per solve on the device.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ffpm.h"
#include "ffpm_bodies.hpp"
#include "ffpm_frame_body.hpp"
#include "ffpm_mg.hpp"
#include "ffpm_mg_body.hpp"
#include "ffpm_pcg_body.hpp"
#include "ffpm_runtime.hpp"

namespace ffpm {

namespace {

constexpr int kThreads = 256;

// Device descriptor of one solve.
struct BSolve {
  GeomT<double> g;  // 1-D coefficients and widths (shared by the frame's solves)
  Geom geom;        // the same, for the true-residual body
  Tiling tl;        // launch shape of the PCG kernels (CL cells per lane)
  Tiling tt;        // launch shape of the true-residual kernel
  size_t n;         // cells
  int eb;           // workgroups of the elementwise kernels
  int pb;           // workgroups of the gauge shift (halo layout)
  int own;          // first solve of its operator (builds the shared arrays)
  const uint8_t *code;
  const double *src, *diag, *rdiag;
  double *x, *r, *z, *p[2];
  double *partials;
  Scalars *sc;
  // FP32 inner solve of mixed precision; g32 and tl32 also serve the
  // multigrid cycle's level 0
  GeomT<float> g32;
  Tiling tl32;
  float *e32, *r32, *z32, *p32[2];
  // multigrid CG: the hierarchy and the K chunk of the level-1 column launches
  Multigrid mg;
  int kc1;
};

// Per-solve phase of the loop: the PCG iterations (FP64, or the FP32 inner
// solve), and for mixed precision the start of a correction, its FP64
// accumulation and the outer check.
enum Phase : int { kPhIter = 0, kPhBegin = 1, kPhAcc = 2, kPhCheck = 3 };

// Fields of the value type T's iteration (FP64 solve or FP32 inner solve).
template <typename T>
struct Fields;
template <>
struct Fields<double> {
  __device__ static const GeomT<double> &g(const BSolve &d) { return d.g; }
  __device__ static const Tiling &tl(const BSolve &d) { return d.tl; }
  __device__ static double *z(const BSolve &d) { return d.z; }
  __device__ static double *p(const BSolve &d, int i) { return d.p[i]; }
  __device__ static double *x(const BSolve &d) { return d.x; }
  __device__ static double *r(const BSolve &d) { return d.r; }
};
template <>
struct Fields<float> {
  __device__ static const GeomT<float> &g(const BSolve &d) { return d.g32; }
  __device__ static const Tiling &tl(const BSolve &d) { return d.tl32; }
  __device__ static float *z(const BSolve &d) { return d.z32; }
  __device__ static float *p(const BSolve &d, int i) { return d.p32[i]; }
  __device__ static float *x(const BSolve &d) { return d.e32; }
  __device__ static float *r(const BSolve &d) { return d.r32; }
};

// Per-solve state of the convergence logic (the host loop of pcg_iterate).
struct BState {
  int active, it, sweeps, restarts, stalled, mode, cur, n_chunk;
  int need_tr;   // 0 none, 1 after a breakdown restart, 2 convergence check, 3 initial
  int need_dir;  // deferred direction update before the check
  int finished_at;
  int phase;         // Phase
  int outer, inner;  // mixed precision: corrections and FP32 iterations
  int first_ready;   // multigrid CG: the cycle's first step is in place for r
  int last_check;    // multigrid CG: iteration of the last true residual
  double res, best;
  double target;     // mixed precision: inner stop rz <= target
};

__device__ inline bool iterating(const BState &st, int slot) {
  return st.active && st.phase == kPhIter && slot < st.n_chunk;
}

// Workgroup -> (solve, index within the solve's launch) tables of one grid.
struct Grid {
  int *solve = nullptr, *local = nullptr;
  int blocks = 0;
};

// ------------------------------------------------------------ kernels ------
template <typename T, int CL>
__global__ __launch_bounds__(512) void k_b_dir(const BSolve *solves, Grid grid,
                                               const BState *state, int slot) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!iterating(st, slot)) return;
  const BSolve &d = solves[s];
  using F = Fields<T>;
  const int lb = grid.local[blockIdx.x];
  const T *pin = F::p(d, st.cur);
  T *pout = F::p(d, st.cur ^ 1);
  const Tiling &t = F::tl(d);
  if (st.mode == kDirUpdate)
    pcg_dir_body<T, CL, kDirUpdate>(t, lb, t.blocks, F::g(d), F::z(d), d.code, pin, pout,
                                    d.partials, d.sc, lds);
  else if (st.mode == kDirReset)
    pcg_dir_body<T, CL, kDirReset>(t, lb, t.blocks, F::g(d), F::z(d), d.code, pin, pout,
                                   d.partials, d.sc, lds);
  else
    pcg_dir_body<T, CL, kDirCurrent>(t, lb, t.blocks, F::g(d), F::z(d), d.code, pin, pout,
                                     d.partials, d.sc, lds);
}

template <typename T, int CL>
__global__ __launch_bounds__(512) void k_b_upd(const BSolve *solves, Grid grid,
                                               const BState *state, int slot) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!iterating(st, slot)) return;
  const BSolve &d = solves[s];
  using F = Fields<T>;
  const Tiling &t = F::tl(d);
  pcg_upd_body<T, CL>(t, grid.local[blockIdx.x], t.blocks, F::g(d), F::p(d, st.cur), d.code,
                      F::x(d), F::r(d), F::z(d), d.partials, d.sc, lds);
}

// One workgroup per solve: the finish of the preceding reduction, plus the
// bookkeeping the standalone host loop does between launches (which direction
// buffer is current; the mode of the next direction kernel). `prec` 1: the
// iteration ran on the FP32 fields. Which solves take part follows the op:
// iteration ops the iterating solves, FIN_TR1 and FIN_RZ the solves that need
// a true residual, FIN_RZ with `begin` the solves starting a correction.
__global__ __launch_bounds__(256) void k_b_finish(const BSolve *solves, BState *state, int op,
                                                  int slot, int prec, int begin) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = blockIdx.x;
  const BState st = state[s];
  const BSolve &d = solves[s];
  int nblocks, nslots = 1;
  if (op == FIN_PQ || op == FIN_RZNEW) {
    if (!iterating(st, slot)) return;
    nblocks = prec ? d.tl32.blocks : d.tl.blocks;
  } else if (begin) {
    if (!st.active || st.phase != kPhBegin) return;
    nblocks = d.eb;
  } else {
    if (!st.need_tr) return;
    nblocks = op == FIN_TR1 ? d.tt.blocks : d.eb;
    if (op == FIN_TR1) nslots = 4;
  }
  finish_body(op, nslots, d.partials, nblocks, d.sc, lds);
  if (threadIdx.x == 0) {
    if (op == FIN_PQ && st.mode != kDirCurrent) state[s].cur = st.cur ^ 1;
    if (op == FIN_RZNEW) state[s].mode = kDirUpdate;
  }
}

// After a chunk: the host logic of the standalone loop that follows a chunk
// (iteration count, breakdown restart, negative rz exit, check points).
__global__ void k_b_check_begin(const BSolve *solves, BState *state, int count, int ci,
                                int max_sweeps) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  st.need_tr = 0;
  st.need_dir = 0;
  if (st.active) {
    Scalars *sc = solves[s].sc;
    const int iters = sc->iters;
    st.sweeps += iters;
    st.it += iters;
    if (iters > 0) st.restarts = 0;
    if (sc->status == kNegative) {
      st.active = 0;
    } else if (sc->status == kBreakdown) {
      st.restarts += 1;
      if (st.restarts > 1) {
        st.active = 0;
      } else {
        sc->status = 0;
        sc->iters = 0;
        sc->sig = 0.0;  // the cycle scale starts afresh, as k_clear_status
        st.need_tr = 1;
        st.mode = kDirReset;
      }
    } else if (st.it % ci == 0 || st.sweeps == max_sweeps) {
      st.need_dir = 1;
      st.need_tr = 2;
      st.mode = kDirCurrent;
    }
  }
  state[s] = st;
}

__global__ __launch_bounds__(256) void k_b_dir_final(const BSolve *solves, Grid grid,
                                                     const BState *state) {
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!st.need_dir) return;
  const BSolve &d = solves[s];
  dir_update_body<double>(d.n, grid.local[blockIdx.x], d.eb, d.p[st.cur], d.z, d.sc);
}

__global__ __launch_bounds__(512) void k_b_tr1(const BSolve *solves, Grid grid,
                                               const BState *state) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  if (!state[s].need_tr) return;
  const BSolve &d = solves[s];
  true_residual_1_body(d.tt, grid.local[blockIdx.x], d.tt.blocks, d.geom, d.x, d.src, d.diag,
                       d.r, d.partials, lds);
}

__global__ __launch_bounds__(256) void k_b_tr2(const BSolve *solves, Grid grid,
                                               const BState *state) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  if (!state[s].need_tr) return;
  const BSolve &d = solves[s];
  true_residual_2_body(d.n, grid.local[blockIdx.x], d.eb, d.r, d.rdiag, d.diag, nullptr, 0, d.z,
                       d.partials, d.sc, lds);
}

// After the true residual: stall test, loop condition, next chunk length.
__global__ void k_b_check_end(const BSolve *solves, BState *state, int count, int ci,
                              int max_sweeps, double tolerance, const int *replay) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  const int was_active = st.active;
  Scalars *sc = solves[s].sc;
  if (st.need_tr) {
    st.res = sc->res;
    if (st.need_tr == 3) {
      st.best = st.res;
    } else if (st.need_tr == 2) {
      if (st.res < 0.99 * st.best) {
        st.best = st.res;
        st.stalled = 0;
      } else {
        st.stalled += 1;
      }
      if (st.stalled >= 50 || !(st.res == st.res)) st.active = 0;
    }
  }
  if (st.active && !(st.sweeps < max_sweeps && st.res > tolerance)) st.active = 0;
  if (st.active) {
    int n = ci - st.it % ci;
    if (max_sweeps - st.sweeps < n) n = max_sweeps - st.sweeps;
    st.n_chunk = n;
    sc->iters = 0;
  }
  if (was_active && !st.active) st.finished_at = *replay;
  st.need_tr = 0;
  state[s] = st;
}

// ---- mixed precision: the standalone mixed solve's outer loop per solve ----
// Start of a correction: r32 = r, z32 = M32^-1 r32, e32 = 0, rz partials.
__global__ __launch_bounds__(256) void k_bm_begin(const BSolve *solves, Grid grid,
                                                  const BState *state) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!st.active || st.phase != kPhBegin) return;
  const BSolve &d = solves[s];
  mixed_begin_body(d.g32, d.r, d.code, d.r32, d.z32, d.e32, d.partials, grid.local[blockIdx.x],
                   d.eb, lds);
}

// After the begin reduction: rz0, the inner target, the first inner chunk.
__global__ void k_bm_after_begin(const BSolve *solves, BState *state, int count, int ci,
                                 int max_inner, double rtol) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  if (!st.active || st.phase != kPhBegin) return;
  Scalars *sc = solves[s].sc;
  sc->status = 0;
  sc->iters = 0;
  const double rz0 = sc->rz;
  if (!(rz0 > 0)) {
    st.active = 0;
  } else {
    st.target = rz0 * rtol * rtol;
    st.mode = kDirReset;
    st.cur = 0;
    st.phase = kPhIter;
    st.n_chunk = ci < max_inner - st.inner ? ci : max_inner - st.inner;
  }
  state[s] = st;
}

// After an inner chunk: count, stop test of the inner solve, next chunk.
__global__ void k_bm_after_inner(const BSolve *solves, BState *state, int count, int ci,
                                 int max_inner) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  if (!st.active || st.phase != kPhIter) return;
  Scalars *sc = solves[s].sc;
  st.inner += sc->iters;
  if (sc->status != 0 || sc->rz <= st.target || sc->iters == 0 || st.inner >= max_inner) {
    st.phase = kPhAcc;
  } else {
    st.n_chunk = ci < max_inner - st.inner ? ci : max_inner - st.inner;
    sc->iters = 0;
  }
  state[s] = st;
}

__global__ __launch_bounds__(256) void k_bm_acc(const BSolve *solves, Grid grid,
                                                const BState *state) {
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!st.active || st.phase != kPhAcc) return;
  const BSolve &d = solves[s];
  mixed_accumulate_body(d.n, d.x, d.e32, grid.local[blockIdx.x], d.eb);
}

__global__ void k_bm_after_acc(const BSolve *solves, BState *state, int count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  if (!st.active || st.phase != kPhAcc) return;
  Scalars *sc = solves[s].sc;
  sc->status = 0;
  sc->iters = 0;
  st.outer += 1;
  st.need_tr = 2;
  st.phase = kPhCheck;
  state[s] = st;
}

// After the FP64 true residual: the outer stall test and loop condition.
__global__ void k_bm_check_end(const BSolve *solves, BState *state, int count, int max_inner,
                               double tolerance, const int *replay) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  const int was_active = st.active;
  if (st.need_tr) {
    st.res = solves[s].sc->res;
    if (st.need_tr == 3) {
      st.best = st.res;
    } else if (st.res < 0.99 * st.best) {
      st.best = st.res;
      st.stalled = 0;
    } else if (++st.stalled >= 3) {
      st.active = 0;
    }
    if (st.active && !(st.res > tolerance && st.inner < max_inner && st.outer < 100))
      st.active = 0;
    if (st.active) st.phase = kPhBegin;
  }
  if (was_active && !st.active) st.finished_at = *replay;
  st.need_tr = 0;
  state[s] = st;
}

// ---- multigrid CG: the standalone fcg_mg_iterate per solve ----------------
// Cell and column launches of level-l arrays: each workgroup decodes its
// (x, y, z) coordinates in its own solve's launch from its local index.
__device__ __forceinline__ mgk::Blk cell_blk(const MgLevel &v, int local) {
  return mgk::blk_of(local, mgk::plane_blocks(v.nx, v.ny), 1);
}

// The cycle's source and first smoothing step after a new r (the first slot
// of a chunk after a true residual or a restart), with the true residual's
// projection of r: a solve's first step is pending exactly when a first pass
// has run since its last one.
__global__ __launch_bounds__(256) void k_bg_first(const BSolve *solves, Grid grid,
                                                  const BState *state) {
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!iterating(st, 0) || st.first_ready) return;
  const BSolve &d = solves[s];
  const Multigrid &mg = d.mg;
  mgk::mgf_jac0_body(cell_blk(mg.lv[0], grid.local[blockIdx.x]), mg.g32, mg.code, d.r,
                     mg.buf0(0), mg.first, d.sc);
}

// One level-0 smoothing step (MgStep: buffers ia, ip, io, coefficients);
// LAST writes the cycle's output z with the z.r, z.q (z.q only after the
// first direction, as the standalone cycle) and max |z| partials.
template <int CL, bool LAST, bool NT, bool CHEB>
__global__ __launch_bounds__(1024) void k_bg_sweep(const BSolve *solves, Grid grid,
                                                   const BState *state, int slot, MgStep step) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!iterating(st, slot)) return;
  const BSolve &d = solves[s];
  const Multigrid &mg = d.mg;
  const MgStore *q = LAST && st.mode == kDirUpdate ? mg.q16 : nullptr;
  mgk::mg_sweep_body<CL, LAST, NT, CHEB>(grid.local[blockIdx.x], mg.tl.blocks, lds, mg.tl,
                                         mg.g32, mg.buf0(step.a), mg.buf0(0), mg.code,
                                         mg.buf0(step.o), mg.partials, step.c2, q,
                                         step.p >= 0 ? mg.buf0(step.p) : nullptr, step.c1,
                                         mg.sscale);
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_bg_restrict0(const BSolve *solves, Grid grid,
                                                      const BState *state, int slot, int ia) {
  const int s = grid.solve[blockIdx.x];
  if (!iterating(state[s], slot)) return;
  const BSolve &d = solves[s];
  const Multigrid &mg = d.mg;
  const MgLevel &c = mg.lv[1];
  const mgk::CoarseOut out{c.b, c.x, c.rd};
  mgk::mgf_restrict_z_body<VEC>(cell_blk(c, grid.local[blockIdx.x]), mg.g32, mg.code,
                                mg.buf0(0), mg.buf0(ia), c.nx, c.ny, c.nz, d.kc1, out,
                                mg.sscale);
}

// The streaming restriction from level 0's iterate (buffer ia) onto level 1,
// over each solve's restriction tiles (the standalone mg_restrict_body).
template <int CL, bool NT>
__global__ __launch_bounds__(512) void k_bg_restrict_t(const BSolve *solves, Grid grid,
                                                       const BState *state, int slot, int ia) {
  extern __shared__ float mg_rs_lds[];  // mg_restrict_lds(ty, CL) bytes
  const int s = grid.solve[blockIdx.x];
  if (!iterating(state[s], slot)) return;
  const Multigrid &mg = solves[s].mg;
  const MgLevel &c = mg.lv[1];
  const mgk::CoarseOut out{c.b, c.x, c.rd};
  mgk::mg_restrict_body<CL, NT>(grid.local[blockIdx.x], mg_rs_lds, mg.tlr, mg.g32, mg.buf0(ia),
                                mg.buf0(0), mg.code, c.nx, c.ny, out, mg.sscale);
}

// Prolongation from level 1's result (buffer ic; the coarsest level writes
// the same buffer as the others) into level 0's iterate (buffer ia).
template <bool VEC>
__global__ __launch_bounds__(256) void k_bg_prolong0(const BSolve *solves, Grid grid,
                                                     const BState *state, int slot, int ia,
                                                     int ic) {
  const int s = grid.solve[blockIdx.x];
  if (!iterating(state[s], slot)) return;
  const BSolve &d = solves[s];
  const Multigrid &mg = d.mg;
  const MgLevel &f0 = mg.lv[0], &c = mg.lv[1];
  mgk::mgf_prolong_z_body<VEC>(cell_blk(c, grid.local[blockIdx.x]), f0.nx, f0.ny, f0.nz,
                               mg.code, mg.buf0(ia), c.nx, c.ny, c.nz, Multigrid::bufc(c, ic),
                               c.m, d.kc1);
}

// Coarse level l (not a solve's coarsest): one smoothing step.
__global__ __launch_bounds__(256) void k_bg_csweep(const BSolve *solves, Grid grid,
                                                   const BState *state, int slot, int l,
                                                   MgStep step) {
  const int s = grid.solve[blockIdx.x];
  if (!iterating(state[s], slot)) return;
  const Multigrid &mg = solves[s].mg;
  if (l >= mg.levels - 1) return;
  const MgLevel &v = mg.lv[l];
  mgk::mgc_sweep_body(cell_blk(v, grid.local[blockIdx.x]), v.nx, v.ny, v.nz, v.gE, v.gN, v.gT,
                      v.b, Multigrid::bufc(v, step.a),
                      step.p >= 0 ? Multigrid::bufc(v, step.p) : nullptr,
                      Multigrid::bufc(v, step.o), step.c1, step.c2);
}

// Restriction of level l's residual (buffer ia) into level l + 1 (grid of
// level l + 1).
__global__ __launch_bounds__(256) void k_bg_crestrict(const BSolve *solves, Grid grid,
                                                      const BState *state, int slot, int l,
                                                      int ia) {
  const int s = grid.solve[blockIdx.x];
  if (!iterating(state[s], slot)) return;
  const Multigrid &mg = solves[s].mg;
  if (l >= mg.levels - 1) return;
  const MgLevel &v = mg.lv[l], &c = mg.lv[l + 1];
  const mgk::CoarseOut out{c.b, c.x, c.rd};
  mgk::mgc_restrict_body(cell_blk(c, grid.local[blockIdx.x]), v.nx, v.ny, v.nz, v.gE, v.gN,
                         v.gT, v.b, Multigrid::bufc(v, ia), c.nx, c.ny, out);
}

// Prolongation from level l + 1's result (buffer ic) into level l's iterate
// (buffer ia; grid of level l).
__global__ __launch_bounds__(256) void k_bg_cprolong(const BSolve *solves, Grid grid,
                                                     const BState *state, int slot, int l,
                                                     int ia, int ic) {
  const int s = grid.solve[blockIdx.x];
  if (!iterating(state[s], slot)) return;
  const Multigrid &mg = solves[s].mg;
  if (l >= mg.levels - 1) return;
  const MgLevel &v = mg.lv[l], &c = mg.lv[l + 1];
  mgk::mgc_prolong_body(cell_blk(v, grid.local[blockIdx.x]), v.nx, v.ny, v.m,
                        Multigrid::bufc(v, ia), c.nx, c.ny, c.nz, Multigrid::bufc(c, ic), c.m);
}

// Every solve's coarsest level, one workgroup per solve, into buffer io.
__global__ __launch_bounds__(1024) void k_bg_coarsest(const BSolve *solves, const BState *state,
                                                      int slot, int io) {
  __shared__ float xs[kMgCoarsestCells];
  const int s = blockIdx.x;
  if (!iterating(state[s], slot)) return;
  const Multigrid &mg = solves[s].mg;
  const MgLevel &c = mg.lv[mg.levels - 1];
  mgk::mgc_coarsest_body(xs, c.nx, c.ny, c.nz, c.gE, c.gN, c.gT, c.b, c.rd,
                         Multigrid::bufc(c, io), mg.omega_coarse, mg.coarse_sweeps);
}

// The direction step on the cycle's output (p = z or z + beta p_old,
// q = A p, p.q), as the standalone mg_direction.
template <int CL>
__global__ __launch_bounds__(512) void k_bg_dir(const BSolve *solves, Grid grid,
                                                const BState *state, int slot) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!iterating(st, slot)) return;
  const BSolve &d = solves[s];
  const int lb = grid.local[blockIdx.x];
  // FP32 directions in the FP64 buffers' memory, as the standalone solve
  const float *pin = reinterpret_cast<const float *>(d.p[st.cur]);
  float *pout = reinterpret_cast<float *>(d.p[st.cur ^ 1]);
  // on the FP32 tiling, as the standalone solve
  const Tiling &t = d.mg.tld;
  if (st.mode == kDirUpdate)
    pcg_dir_body<float, CL, kDirUpdate, MgStore, float>(t, lb, t.blocks, d.mg.g32,
                                                        d.mg.buf0(3), d.code, pin, pout,
                                                        d.partials, d.sc, lds);
  else
    pcg_dir_body<float, CL, kDirReset, MgStore, float>(t, lb, t.blocks, d.mg.g32,
                                                       d.mg.buf0(3), d.code, pin, pout,
                                                       d.partials, d.sc, lds);
}

// x += alpha p, r -= alpha A p and the next cycle's first step; the slot that
// ends a solve's chunk also forms the max |r / V| partials.
template <int CL>
__global__ __launch_bounds__(1024) void k_bg_upd(const BSolve *solves, Grid grid,
                                                 const BState *state, int slot) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BState &st = state[s];
  if (!iterating(st, slot)) return;
  const BSolve &d = solves[s];
  const Multigrid &mg = d.mg;
  double *rmax = slot == st.n_chunk - 1 ? d.partials : nullptr;
  mgk::mg_upd_body<CL>(grid.local[blockIdx.x], d.tl.blocks, lds, d.tl, d.g,
                       reinterpret_cast<const float *>(d.p[st.cur]), d.code, d.x, d.r,
                       mg.buf0(0), d.sc, (double)mg.first, rmax, mg.q16);
}

// One workgroup per solve: the finish of the preceding reduction for the
// solves that ran it (FIN_FCG1, FIN_PQ: the slot's iterating solves; FIN_MAX:
// those whose chunk ends at the slot; FIN_TR1, FIN_RZ: those that need a
// true residual). FIN_PQ also switches the current direction buffer.
__global__ __launch_bounds__(256) void k_bg_finish(const BSolve *solves, BState *state, int op,
                                                   int slot) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = blockIdx.x;
  const BState st = state[s];
  const BSolve &d = solves[s];
  int nblocks = d.tl.blocks, nslots = 1;
  if (op == FIN_FCG1 || op == FIN_PQ || op == FIN_MAX) {
    if (!iterating(st, slot)) return;
    if (op == FIN_MAX && slot != st.n_chunk - 1) return;
    if (op == FIN_FCG1) {
      nblocks = d.mg.tl.blocks;
      nslots = 3;
    } else if (op == FIN_PQ) {
      nblocks = d.mg.tld.blocks;  // the direction runs on the FP32 tiling
    }
  } else {
    if (!st.need_tr) return;
    nblocks = op == FIN_TR1 ? d.tt.blocks : d.eb;
    if (op == FIN_TR1) nslots = 4;
  }
  finish_body(op, nslots, d.partials, nblocks, d.sc, lds);
  if (op == FIN_PQ && threadIdx.x == 0) state[s].cur = st.cur ^ 1;
}

// After a slot: the cycle's first step is in place; later directions update.
__global__ void k_bg_slot_end(BState *state, int count, int slot) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count || !iterating(state[s], slot)) return;
  state[s].first_ready = 1;
  state[s].mode = kDirUpdate;
}

// After a chunk: iteration count, breakdown restart, and whether the FP64
// true residual is formed (near the tolerance by the chunk's max |r / V|
// estimate, 20 iterations after the last one, or at the iteration cap).
__global__ void k_bg_check_begin(const BSolve *solves, BState *state, int count, int ci,
                                 int max_sweeps, double tolerance) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  st.need_tr = 0;
  if (st.active) {
    Scalars *sc = solves[s].sc;
    const int iters = sc->iters;
    st.sweeps += iters;
    st.it += iters;
    if (iters > 0) st.restarts = 0;
    if (sc->status != 0) {
      st.restarts += 1;
      if (st.restarts > 1) {
        st.active = 0;
      } else {
        sc->status = 0;
        sc->iters = 0;
        sc->sig = 0.0;  // the cycle scale starts afresh, as k_clear_status
        st.need_tr = 1;
        st.first_ready = 0;
        st.mode = kDirReset;
      }
    } else {
      const bool near = !(sc->res > 4.0 * tolerance);
      if ((st.it % ci == 0 && (near || st.it - st.last_check >= 20)) ||
          st.sweeps == max_sweeps) {
        st.last_check = st.it;
        st.need_tr = 2;
        st.first_ready = 0;
      }
    }
  }
  state[s] = st;
}

// After the true residual: stall test (ten checks without a 1% gain), loop
// condition, next chunk length.
__global__ void k_bg_check_end(const BSolve *solves, BState *state, int count, int ci,
                               int max_sweeps, double tolerance, const int *replay) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st = state[s];
  const int was_active = st.active;
  Scalars *sc = solves[s].sc;
  if (st.need_tr) {
    st.res = sc->res;
    if (st.need_tr == 3) {
      st.best = st.res;
    } else if (st.need_tr == 2) {
      if (st.res < 0.99 * st.best) {
        st.best = st.res;
        st.stalled = 0;
      } else {
        st.stalled += 1;
      }
      if (st.stalled >= 10 || !(st.res == st.res)) st.active = 0;
    }
  }
  if (st.active && !(st.sweeps < max_sweeps && st.res > tolerance)) st.active = 0;
  if (st.active) {
    int n = ci - st.it % ci;
    if (max_sweeps - st.sweeps < n) n = max_sweeps - st.sweeps;
    st.n_chunk = n;
    sc->iters = 0;
  }
  if (was_active && !st.active) st.finished_at = *replay;
  st.need_tr = 0;
  state[s] = st;
}

// Active solves after a replay; advances the replay counter.
__global__ __launch_bounds__(256) void k_b_count(const BState *state, int count, int *replay,
                                                 int *n_active) {
  __shared__ int part[256];
  int a = 0;
  for (int s = threadIdx.x; s < count; s += blockDim.x) a += state[s].active;
  part[threadIdx.x] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < (int)blockDim.x; ++w) t += part[w];
    *n_active = t;
    *replay = *replay + 1;
  }
}

// Initial state of every solve (before the first true residual).
__global__ void k_b_init(BState *state, int count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= count) return;
  BState st{};
  st.active = 1;
  st.mode = kDirReset;
  st.need_tr = 3;
  st.finished_at = 0;
  st.phase = kPhIter;
  state[s] = st;
}

__global__ void k_b_zero_int(int *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// ---- set-up and finish of every solve (the standalone solve's steps, one
// launch each over the batch) ------------------------------------------------
// 1-D coefficients of every operator (solve blockIdx.y, entry blockIdx.x *
// 256 + threadIdx.x), then their FP32 copies and those of dx, dy, dz.
__global__ __launch_bounds__(256) void k_bs_axes(const BSolve *solves) {
  const BSolve &d = solves[blockIdx.y];
  if (!d.own) return;
  const Geom &g = d.geom;
  axis_coefficients_body(blockIdx.x * blockDim.x + threadIdx.x, g.nx, g.ny, g.nz, g.xc, g.yc,
                         g.zc, g.dx, g.dy, g.dz, (double *)g.cw, (double *)g.ce,
                         (double *)g.cs, (double *)g.cn, (double *)g.cb, (double *)g.ct);
}

__global__ __launch_bounds__(256) void k_bs_axes32(const BSolve *solves) {
  const BSolve &d = solves[blockIdx.y];
  if (!d.own) return;
  const Geom &g = d.geom;
  const int tot = g.nx + g.ny + g.nz, t = blockIdx.x * blockDim.x + threadIdx.x;
  float *c32 = (float *)d.g32.cw;  // coefficients, then dx, dy, dz
  if (t < 2 * tot) c32[t] = (float)g.cw[t];
  else if (t < 2 * tot + g.nx) c32[t] = (float)g.dx[t - 2 * tot];
  else if (t < 2 * tot + g.nx + g.ny) c32[t] = (float)g.dy[t - 2 * tot - g.nx];
  else if (t < 3 * tot) c32[t] = (float)g.dz[t - 2 * tot - g.nx - g.ny];
}

// build_operator for every solve; the first solve of a frame also writes the
// shared diag, rdiag and neighbour code.
__global__ __launch_bounds__(256) void k_bs_build(const BSolve *solves,
                                                  const ffpm_stairstep_problem *probs, Grid grid,
                                                  int *isolated) {
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  const ffpm_stairstep_problem &p = probs[s];
  build_operator_body(grid.local[blockIdx.x], d.eb, d.geom, p.solid, p.rhs, p.xlo, p.xhi, p.ylo,
                      p.yhi, p.zlo, p.zhi, p.body_grad[0], p.body_grad[1], p.body_grad[2], 1,
                      (double *)d.src, (double *)d.diag, (double *)d.rdiag, (uint8_t *)d.code,
                      isolated, d.own != 0);
}

__global__ __launch_bounds__(256) void k_bs_pack(const BSolve *solves,
                                                 const ffpm_stairstep_problem *probs, Grid grid) {
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  const ffpm_stairstep_problem &p = probs[s];
  pack_phi_body(grid.local[blockIdx.x], d.eb, p.nx, p.ny, p.nz, p.solid, p.phi, d.x);
}

__global__ void k_bs_zero_scalars(const BSolve *solves, int count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < count) *solves[s].sc = Scalars{};
}

// Each multigrid solve's operator-norm bound (one workgroup per solve, as the
// standalone k_mg_anorm).
__global__ __launch_bounds__(256) void k_bs_anorm(const BSolve *solves) {
  __shared__ double lds[16];
  mgk::mg_anorm_body(solves[blockIdx.x].mg.g64, solves[blockIdx.x].sc, lds);
}

// Multigrid level operators of every frame, level l (grid of level l).
__global__ __launch_bounds__(256) void k_bs_coarsen(const BSolve *solves, Grid grid, int l) {
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  if (!d.own) return;
  const Multigrid &mg = d.mg;
  const MgLevel &c = mg.lv[l];
  const mgk::Blk bk = cell_blk(c, grid.local[blockIdx.x]);
  if (l == 1) {
    mgk::mg_coarsen0_body(bk, mg.g64, mg.code, c.nx, c.ny, mg.scale, c.gE, c.gN, c.gT);
  } else {
    const MgLevel &f = mg.lv[l - 1];
    mgk::mg_coarsen_body(bk, f.nx, f.ny, f.nz, f.gE, f.gN, f.gT, c.nx, c.ny, mg.scale, c.gE,
                         c.gN, c.gT);
  }
}

__global__ __launch_bounds__(256) void k_bs_mask(const BSolve *solves, Grid grid, int l) {
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  if (!d.own) return;
  const Multigrid &mg = d.mg;
  const MgLevel &c = mg.lv[l];
  mgk::mg_mask_body(cell_blk(c, grid.local[blockIdx.x]), c.nx, c.ny, c.gE, c.gN, c.gT,
                    mg.first_coarse, c.m, c.rd);
}

// Gauge (src/fpm_stairstep_mpi_3d.f90:170-185) and the final residual
// (src/fpm_stairstep_mpi_3d.f90:186-190) of every solve.
__global__ __launch_bounds__(256) void k_bf_gauge_sums(const BSolve *solves,
                                                       const ffpm_stairstep_problem *probs,
                                                       Grid grid) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  gauge_sums_body(grid.local[blockIdx.x], d.eb, d.geom, probs[s].solid, d.x, d.partials, lds);
}

__global__ __launch_bounds__(256) void k_bf_gauge_apply(const BSolve *solves,
                                                        const ffpm_stairstep_problem *probs,
                                                        Grid grid) {
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  const ffpm_stairstep_problem &p = probs[s];
  gauge_apply_body(grid.local[blockIdx.x], d.pb, p.nx, p.ny, p.nz, p.solid, d.x, p.phi, d.sc);
}

__global__ __launch_bounds__(512) void k_bf_residual_max(const BSolve *solves, Grid grid) {
  __shared__ double lds[16 * kReduceSlots];
  const int s = grid.solve[blockIdx.x];
  const BSolve &d = solves[s];
  residual_max_body(d.tt, grid.local[blockIdx.x], d.tt.blocks, d.geom, d.x, d.src, d.diag,
                    d.partials, lds);
}

// One workgroup per solve: FIN_SUMS of the gauge (eb partials, two slots) or
// FIN_MAX of the final residual (tt.blocks partials); with res, the residual
// of every solve is gathered there.
__global__ __launch_bounds__(256) void k_bf_finish(const BSolve *solves, int op, double *res) {
  __shared__ double lds[16 * kReduceSlots];
  const BSolve &d = solves[blockIdx.x];
  if (op == FIN_SUMS) finish_body(op, 2, d.partials, d.eb, d.sc, lds);
  else finish_body(op, 1, d.partials, d.tt.blocks, d.sc, lds);
  if (res) {
    __syncthreads();
    if (threadIdx.x == 0) res[blockIdx.x] = d.sc->res;
  }
}

// ------------------------------------------------------------- host --------
// Cells per lane as the standalone solves choose them (so that each solve's
// tiles and reduction order are its standalone solve's).
int cells_per_lane_f64(int nx) { return cells_per_lane(nx, 2); }
int cells_per_lane_f32(int nx) { return cells_per_lane(nx, 4); }

double inner_rtol() {
  const char *e = std::getenv("FFPM_MIXED_RTOL");
  return e && *e ? std::atof(e) : 1e-4;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bump allocation inside one device arena.
struct Arena {
  char *base = nullptr;
  size_t used = 0;
  template <typename P>
  P *take(size_t bytes) {
    P *p = reinterpret_cast<P *>(base + used);
    used += align256(bytes);
    return p;
  }
};

bool same_operator(const ffpm_stairstep_problem &a, const ffpm_stairstep_problem &b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.solid == b.solid && a.xc == b.xc &&
         a.yc == b.yc && a.zc == b.zc && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz;
}

// Launch shape of the PCG kernels for one solve: the standalone solve's
// defaults unless the caller (or the autotuner) fixes rows and z chunk.
Tiling pcg_tiling(int nx, int ny, int nz, int cl, int rows, int zchunk) {
  const int r = rows > 0 ? rows : (std::getenv("FFPM_TILE_Y") ? 0 : 8);
  return make_tiling(nx, ny, nz, zchunk > 0 ? zchunk : 0, r, kTileX * cl);
}

struct Batch {
  int count = 0;
  bool mixed = false;
  std::vector<BSolve> h;    // host copies of the descriptors
  std::vector<int> cl;      // cells per lane of each solve (FP64 kernels)
  std::vector<int> cl32;    // the same for the FP32 kernels
  std::vector<int> owner;   // operator owner (first solve of the frame)
  BSolve *d_solves = nullptr;
  BState *d_state = nullptr;
  int *d_counters = nullptr;  // [0] replay, [1] active count
  Grid g_pcg[3], g_pcg32[5], g_tr, g_ew;  // g_pcg[CL], g_pcg32[CL]
  int *d_isolated = nullptr;
  // multigrid CG: level-0 cell grid, level-1 column grids (odd / even nx),
  // cell grids of coarse level l over the solves that have it, and the
  // largest coarsest-level index
  bool mgm = false;
  Grid g_cell0, g_col[2], g_lvl[kMgMaxLevels];
  Grid g_res[5];  // streaming restriction tiles, by cells per lane (1, 2, 4)
  int mg_top = 0;
  // the problems on the device, the final residuals, the gauge-shift grid
  ffpm_stairstep_problem *d_probs = nullptr;
  double *d_res = nullptr;
  Grid g_pad;
  int max_axis = 0;  // largest nx + ny + nz
};

int build_grid(Arena &a, Grid &g, const std::vector<int> &blocks, const std::vector<int> &use,
               std::vector<int> &h_solve, std::vector<int> &h_local) {
  h_solve.clear();
  h_local.clear();
  for (size_t s = 0; s < blocks.size(); ++s) {
    if (!use[s]) continue;
    for (int b = 0; b < blocks[s]; ++b) {
      h_solve.push_back((int)s);
      h_local.push_back(b);
    }
  }
  g.blocks = (int)h_solve.size();
  if (g.blocks == 0) return OK;
  g.solve = a.take<int>(sizeof(int) * g.blocks);
  g.local = a.take<int>(sizeof(int) * g.blocks);
  FFPM_HIP(hipMemcpyAsync(g.solve, h_solve.data(), sizeof(int) * g.blocks,
                          hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(g.local, h_local.data(), sizeof(int) * g.blocks,
                          hipMemcpyHostToDevice, stream()));
  return OK;
}

// Builds every solve's operator and state; arrays live in one arena kept in
// the workspace (grow-only).
int batch_setup(Batch &B, ffpm_stairstep_problem *P, int count, int rows, int zchunk,
                bool mixed, bool mgm = false) {
  B.count = count;
  B.mixed = mixed;
  B.mgm = mgm;
  B.h.assign(count, BSolve{});
  B.cl.assign(count, 1);
  B.cl32.assign(count, 1);
  B.owner.assign(count, -1);
  for (int s = 0; s < count; ++s) {
    const ffpm_stairstep_problem &p = P[s];
    if (p.nx < 1 || p.ny < 1 || p.nz < 2) {
      set_error(ERR_SHAPE, "stair-step solve needs nz>=2 and nonempty ranks");
      return ERR_SHAPE;
    }
    if (!p.phi || !p.rhs || !p.solid || !p.xlo || !p.xhi || !p.ylo || !p.yhi || !p.zlo ||
        !p.zhi || !p.xc || !p.yc || !p.zc || !p.dx || !p.dy || !p.dz) {
      set_error(ERR_ARGUMENT, "null array argument");
      return ERR_ARGUMENT;
    }
    B.owner[s] = s;
    for (int o = 0; o < s; ++o)
      if (B.owner[o] == o && same_operator(P[o], p)) {
        B.owner[s] = o;
        break;
      }
  }
  // sizes
  size_t total = align256(sizeof(BSolve) * count) + align256(sizeof(BState) * count) +
                 align256(sizeof(int) * 4) + align256(sizeof(int)) +
                 align256(sizeof(ffpm_stairstep_problem) * count) +
                 align256(sizeof(double) * count);
  std::vector<Tiling> tl(count), tl32(count), tt(count);
  std::vector<int> eb(count);
  std::vector<Multigrid> plan(mgm ? count : 0);
  size_t grid_blocks = 0, all_cells = 0;
  auto cells3 = [](const dim3 &g) { return (size_t)g.x * g.y * g.z; };
  for (int s = 0; s < count; ++s) {
    const ffpm_stairstep_problem &p = P[s];
    const size_t n = (size_t)p.nx * p.ny * p.nz;
    all_cells += n;
    B.cl[s] = cells_per_lane_f64(p.nx);
    B.cl32[s] = cells_per_lane_f32(p.nx);
    tl[s] = pcg_tiling(p.nx, p.ny, p.nz, B.cl[s], rows, zchunk);
    // FP32 kernels: the mixed inner PCG's, or the multigrid cycle's (their
    // own rows unless a shape is given, as the standalone solves choose)
    tl32[s] = pcg_tiling(p.nx, p.ny, p.nz, B.cl32[s], mgm && rows <= 0 ? mg_rows() : rows,
                         zchunk);
    tt[s] = make_tiling(p.nx, p.ny, p.nz, 0);
    eb[s] = elementwise_blocks(n);
    const size_t nb = (size_t)std::max(std::max(std::max(tl[s].blocks, tl32[s].blocks),
                                                tt[s].blocks), eb[s]);
    grid_blocks += (size_t)tl[s].blocks + tl32[s].blocks + tt[s].blocks + eb[s] +
                   elementwise_blocks((size_t)(p.nx + 2) * (p.ny + 2) * p.nz);
    B.max_axis = std::max(B.max_axis, p.nx + p.ny + p.nz);
    total += 6 * align256(sizeof(double) * n) + align256(sizeof(double) * kReduceSlots * nb) +
             align256(sizeof(Scalars));
    if (mixed) total += 5 * align256(sizeof(float) * n);
    if (mgm) {
      // extents only: the pointers are bound once the arrays exist
      Multigrid &mg = plan[s];
      const GeomT<double> g64{p.nx, p.ny, p.nz, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr, nullptr};
      const GeomT<float> g32{p.nx, p.ny, p.nz, nullptr, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr};
      int st = mg_plan(mg, g64, g32, nullptr, tl32[s], B.cl32[s], nullptr);
      if (st) return st;
      if (mg.cl != B.cl32[s] || mg.tl.zchunk != tl32[s].zchunk) {
        // the batched sweeps run on the FP32 launch grids
        set_error(ERR_ARGUMENT,
                  "batched multigrid solves need the default FFPM_MG_CL and FFPM_MG_ZCHUNK");
        return ERR_ARGUMENT;
      }
      total += align256(mg_state_bytes(mg));
      if (B.owner[s] == s) total += align256(mg_operator_bytes(mg));
      const MgLevel &c1 = mg.lv[1];
      grid_blocks += (size_t)mg.tlr.blocks;
      grid_blocks += cells3(mgk::cell_grid(p.nx, p.ny, p.nz)) +
                     cells3(mgk::column_grid(c1.nx, c1.ny, c1.nz,
                                             mgk::column_chunk(c1.nx, c1.ny, c1.nz)));
      for (int l = 1; l < mg.levels; ++l)
        grid_blocks += cells3(mgk::cell_grid(mg.lv[l].nx, mg.lv[l].ny, mg.lv[l].nz));
      B.mg_top = std::max(B.mg_top, mg.levels - 1);
    }
    if (B.owner[s] == s)
      total += 2 * align256(sizeof(double) * n) + align256(n) +
               align256(sizeof(double) * 2 * (p.nx + p.ny + p.nz)) +
               align256(sizeof(float) * 3 * (p.nx + p.ny + p.nz));
  }
  // workgroup tables: two int arrays per grid (at most nine grids, plus
  // 6 + kMgMaxLevels for the multigrid), each 256-aligned
  total += 2 * align256(sizeof(int) * grid_blocks) + 2 * 256 * (9 + 6 + kMgMaxLevels);
  int st = 0;
  Arena a;
  a.base = (char *)workspace().get(S_BATCH, total, &st);
  if (!a.base) return st;
  B.d_solves = a.take<BSolve>(sizeof(BSolve) * count);
  B.d_state = a.take<BState>(sizeof(BState) * count);
  B.d_counters = a.take<int>(sizeof(int) * 4);
  B.d_isolated = a.take<int>(sizeof(int));
  B.d_probs = a.take<ffpm_stairstep_problem>(sizeof(ffpm_stairstep_problem) * count);
  B.d_res = a.take<double>(sizeof(double) * count);
  k_b_zero_int<<<1, 64, 0, stream()>>>(B.d_isolated, 1);
  FFPM_LAUNCHED("k_b_zero_int");
  for (int s = 0; s < count; ++s) {
    const ffpm_stairstep_problem &p = P[s];
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const size_t n = (size_t)nx * ny * nz;
    BSolve &d = B.h[s];
    d.n = n;
    d.tl = tl[s];
    d.tl32 = tl32[s];
    d.tt = tt[s];
    d.eb = eb[s];
    d.pb = elementwise_blocks((size_t)(nx + 2) * (ny + 2) * nz);
    const int o = B.owner[s];
    d.own = o == s;
    if (o == s) {
      double *coef = a.take<double>(sizeof(double) * 2 * (nx + ny + nz));
      Geom g{};
      g.nx = nx; g.ny = ny; g.nz = nz;
      g.cw = coef; g.ce = coef + nx; g.cs = coef + 2 * nx; g.cn = coef + 2 * nx + ny;
      g.cb = coef + 2 * nx + 2 * ny; g.ct = coef + 2 * nx + 2 * ny + nz;
      g.dx = p.dx; g.dy = p.dy; g.dz = p.dz;
      g.xc = p.xc; g.yc = p.yc; g.zc = p.zc;
      const int tot = nx + ny + nz;
      d.geom = g;
      d.g = GeomT<double>{nx, ny, nz, g.cw, g.ce, g.cs, g.cn, g.cb, g.ct, g.dx, g.dy, g.dz};
      // FP32 copies for the mixed inner solve and the multigrid (the
      // standalone solves' conversions): coefficients, then dx, dy, dz
      float *c32 = a.take<float>(sizeof(float) * 3 * tot);
      const int ncoef = 2 * tot;
      d.g32 = GeomT<float>{nx, ny, nz, c32, c32 + nx, c32 + 2 * nx, c32 + 2 * nx + ny,
                           c32 + 2 * nx + 2 * ny, c32 + 2 * nx + 2 * ny + nz, c32 + ncoef,
                           c32 + ncoef + nx, c32 + ncoef + nx + ny};
      d.diag = a.take<double>(sizeof(double) * n);
      d.rdiag = a.take<double>(sizeof(double) * n);
      d.code = a.take<uint8_t>(n);
    } else {
      const BSolve &od = B.h[o];
      d.geom = od.geom;
      d.g = od.g;
      d.g32 = od.g32;
      d.diag = od.diag;
      d.rdiag = od.rdiag;
      d.code = od.code;
    }
    d.x = a.take<double>(sizeof(double) * n);
    d.r = a.take<double>(sizeof(double) * n);
    d.z = a.take<double>(sizeof(double) * n);
    d.p[0] = a.take<double>(sizeof(double) * n);
    d.p[1] = a.take<double>(sizeof(double) * n);
    double *src = a.take<double>(sizeof(double) * n);
    d.src = src;
    if (mixed) {
      d.e32 = a.take<float>(sizeof(float) * n);
      d.r32 = a.take<float>(sizeof(float) * n);
      d.z32 = a.take<float>(sizeof(float) * n);
      d.p32[0] = a.take<float>(sizeof(float) * n);
      d.p32[1] = a.take<float>(sizeof(float) * n);
    }
    const size_t nb = (size_t)std::max(std::max(std::max(tl[s].blocks, tl32[s].blocks),
                                                tt[s].blocks), eb[s]);
    d.partials = a.take<double>(sizeof(double) * kReduceSlots * nb);
    d.sc = a.take<Scalars>(sizeof(Scalars));
    if (mgm) {
      // the hierarchy: level operators of the frame's first solve, own state
      Multigrid mg = plan[s];
      mg.g64 = d.g;
      mg.g32 = d.g32;
      mg.code = d.code;
      mg.partials = d.partials;
      if (o == s) {
        char *ob = a.take<char>(mg_operator_bytes(mg));
        mg_bind_operator(mg, ob);
      } else {
        const Multigrid &om = B.h[o].mg;
        for (int l = 1; l < mg.levels; ++l) {
          mg.lv[l].gE = om.lv[l].gE;
          mg.lv[l].gN = om.lv[l].gN;
          mg.lv[l].gT = om.lv[l].gT;
          mg.lv[l].m = om.lv[l].m;
          mg.lv[l].rd = om.lv[l].rd;
        }
      }
      char *sb = a.take<char>(mg_state_bytes(mg));
      mg_bind_state(mg, sb);
      d.mg = mg;
      // the column launches' wavefront target, shared among the solves by
      // cells (the K chunk does not change any value)
      const long target = std::max(1L, (long)((double)mgk::column_waves() * (double)n /
                                              (double)all_cells));
      d.kc1 = mgk::column_chunk(mg.lv[1].nx, mg.lv[1].ny, mg.lv[1].nz, target);
    }
  }
  FFPM_HIP(hipMemcpyAsync(B.d_solves, B.h.data(), sizeof(BSolve) * count,
                          hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(B.d_probs, P, sizeof(ffpm_stairstep_problem) * count,
                          hipMemcpyHostToDevice, stream()));
  std::vector<int> blocks(count), use(count), hs, hl;
  for (int c = 1; c <= 2; ++c) {
    for (int s = 0; s < count; ++s) {
      blocks[s] = tl[s].blocks;
      use[s] = B.cl[s] == c;
    }
    if ((st = build_grid(a, B.g_pcg[c], blocks, use, hs, hl))) return st;
  }
  if (mixed || mgm) {
    for (int c : {1, 2, 4}) {
      for (int s = 0; s < count; ++s) {
        blocks[s] = tl32[s].blocks;
        use[s] = B.cl32[s] == c;
      }
      if ((st = build_grid(a, B.g_pcg32[c], blocks, use, hs, hl))) return st;
    }
  }
  for (int s = 0; s < count; ++s) {
    blocks[s] = tt[s].blocks;
    use[s] = 1;
  }
  if ((st = build_grid(a, B.g_tr, blocks, use, hs, hl))) return st;
  for (int s = 0; s < count; ++s) blocks[s] = eb[s];
  if ((st = build_grid(a, B.g_ew, blocks, use, hs, hl))) return st;
  if (mgm) {
    for (int s = 0; s < count; ++s) {
      const MgLevel &f0 = B.h[s].mg.lv[0];
      blocks[s] = (int)cells3(mgk::cell_grid(f0.nx, f0.ny, f0.nz));
      use[s] = 1;
    }
    if ((st = build_grid(a, B.g_cell0, blocks, use, hs, hl))) return st;
    for (int v = 0; v < 2; ++v) {
      for (int s = 0; s < count; ++s) {
        const MgLevel &c1 = B.h[s].mg.lv[1];
        blocks[s] = (int)cells3(mgk::column_grid(c1.nx, c1.ny, c1.nz, B.h[s].kc1));
        use[s] = (B.h[s].mg.lv[0].nx % 2 == 0) == (v == 1);
      }
      if ((st = build_grid(a, B.g_col[v], blocks, use, hs, hl))) return st;
    }
    for (int c : {1, 2, 4}) {
      for (int s = 0; s < count; ++s) {
        blocks[s] = B.h[s].mg.tlr.blocks;
        use[s] = B.cl32[s] == c;
      }
      if ((st = build_grid(a, B.g_res[c], blocks, use, hs, hl))) return st;
    }
    for (int l = 1; l <= B.mg_top; ++l) {
      for (int s = 0; s < count; ++s) {
        const Multigrid &mg = B.h[s].mg;
        use[s] = mg.levels - 1 >= l;
        blocks[s] = use[s] ? (int)cells3(mgk::cell_grid(mg.lv[l].nx, mg.lv[l].ny, mg.lv[l].nz))
                           : 0;
      }
      if ((st = build_grid(a, B.g_lvl[l], blocks, use, hs, hl))) return st;
    }
  }
  for (int s = 0; s < count; ++s) {
    blocks[s] = B.h[s].pb;
    use[s] = 1;
  }
  if ((st = build_grid(a, B.g_pad, blocks, use, hs, hl))) return st;
  // the standalone solve's set-up, one launch per step over the batch:
  // coefficients of every operator, build_operator
  // (src/fpm_stairstep_mpi_3d.f90:336-379), the dense phi, zero scalars, and
  // the multigrid level operators
  const unsigned ax = (unsigned)((3 * B.max_axis + 255) / 256);
  k_bs_axes<<<dim3(ax, count), 256, 0, stream()>>>(B.d_solves);
  FFPM_LAUNCHED("k_bs_axes");
  k_bs_axes32<<<dim3(ax, count), 256, 0, stream()>>>(B.d_solves);
  FFPM_LAUNCHED("k_bs_axes32");
  k_bs_build<<<B.g_ew.blocks, kThreads, 0, stream()>>>(B.d_solves, B.d_probs, B.g_ew,
                                                       B.d_isolated);
  FFPM_LAUNCHED("k_bs_build");
  k_bs_pack<<<B.g_ew.blocks, kThreads, 0, stream()>>>(B.d_solves, B.d_probs, B.g_ew);
  FFPM_LAUNCHED("k_bs_pack");
  k_bs_zero_scalars<<<(count + 63) / 64, 64, 0, stream()>>>(B.d_solves, count);
  FFPM_LAUNCHED("k_bs_zero_scalars");
  if (mgm) {
    for (int l = 1; l <= B.mg_top; ++l)
      k_bs_coarsen<<<B.g_lvl[l].blocks, mgk::kMt, 0, stream()>>>(B.d_solves, B.g_lvl[l], l);
    FFPM_LAUNCHED("k_bs_coarsen");
    for (int l = 1; l <= B.mg_top; ++l)
      k_bs_mask<<<B.g_lvl[l].blocks, mgk::kMt, 0, stream()>>>(B.d_solves, B.g_lvl[l], l);
    FFPM_LAUNCHED("k_bs_mask");
    k_bs_anorm<<<count, 256, 0, stream()>>>(B.d_solves);
    FFPM_LAUNCHED("k_bs_anorm");
  }
  int isolated = 0;
  FFPM_HIP(hipMemcpyAsync(&isolated, B.d_isolated, sizeof(int), hipMemcpyDeviceToHost,
                          stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  if (isolated > 0) {
    set_error(ERR_ISOLATED, "stair-step fluid cell has no fluid neighbour");
    return ERR_ISOLATED;
  }
  return OK;
}

int launch_check(const Batch &B, hipStream_t s, int ci, int max_sweeps, double tol) {
  const int cb = (B.count + 63) / 64;
  if (B.g_ew.blocks) {
    k_b_dir_final<<<B.g_ew.blocks, kThreads, 0, s>>>(B.d_solves, B.g_ew, B.d_state);
    FFPM_LAUNCHED("k_b_dir_final");
  }
  k_b_tr1<<<B.g_tr.blocks, dim3(kTileX, B.h[0].tt.ty), 0, s>>>(B.d_solves, B.g_tr, B.d_state);
  FFPM_LAUNCHED("k_b_tr1");
  k_b_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_TR1, 0, 0, 0);
  FFPM_LAUNCHED("k_b_finish");
  k_b_tr2<<<B.g_ew.blocks, kThreads, 0, s>>>(B.d_solves, B.g_ew, B.d_state);
  FFPM_LAUNCHED("k_b_tr2");
  k_b_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_RZ, 0, 0, 0);
  FFPM_LAUNCHED("k_b_finish");
  if (B.mixed)
    k_bm_check_end<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, max_sweeps, tol,
                                     B.d_counters);
  else
    k_b_check_end<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, ci, max_sweeps, tol,
                                    B.d_counters);
  FFPM_LAUNCHED("k_b_check_end");
  return OK;
}

// The iteration slots of one chunk, in value type T: per slot the direction
// kernel, its finish, the update kernel and its finish.
template <typename T>
int launch_slots(const Batch &B, hipStream_t s, int ci, int rows) {
  const dim3 blk(kTileX, rows);
  const Grid *grids = sizeof(T) == 8 ? B.g_pcg : B.g_pcg32;
  const int prec = sizeof(T) == 8 ? 0 : 1;
  const int widths[3] = {1, 2, 4};
  for (int m = 0; m < ci; ++m) {
    for (int c : widths) {
      if ((sizeof(T) == 8 && c == 4) || !grids[c].blocks) continue;
      if (c == 4) {
        if constexpr (sizeof(T) == 4)
          k_b_dir<T, 4><<<grids[c].blocks, blk, 0, s>>>(B.d_solves, grids[c], B.d_state, m);
      } else if (c == 2) {
        k_b_dir<T, 2><<<grids[c].blocks, blk, 0, s>>>(B.d_solves, grids[c], B.d_state, m);
      } else {
        k_b_dir<T, 1><<<grids[c].blocks, blk, 0, s>>>(B.d_solves, grids[c], B.d_state, m);
      }
      FFPM_LAUNCHED("k_b_dir");
    }
    k_b_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_PQ, m, prec, 0);
    FFPM_LAUNCHED("k_b_finish");
    for (int c : widths) {
      if ((sizeof(T) == 8 && c == 4) || !grids[c].blocks) continue;
      if (c == 4) {
        if constexpr (sizeof(T) == 4)
          k_b_upd<T, 4><<<grids[c].blocks, blk, 0, s>>>(B.d_solves, grids[c], B.d_state, m);
      } else if (c == 2) {
        k_b_upd<T, 2><<<grids[c].blocks, blk, 0, s>>>(B.d_solves, grids[c], B.d_state, m);
      } else {
        k_b_upd<T, 1><<<grids[c].blocks, blk, 0, s>>>(B.d_solves, grids[c], B.d_state, m);
      }
      FFPM_LAUNCHED("k_b_upd");
    }
    k_b_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_RZNEW, m, prec, 0);
    FFPM_LAUNCHED("k_b_finish");
  }
  return OK;
}

// One chunk: FP64 PCG slots and the convergence check; in mixed precision the
// start of corrections, FP32 inner slots, accumulation and the outer check.
int launch_chunk(const Batch &B, hipStream_t s, int ci, int max_sweeps, double tol, int rows) {
  const int cb = (B.count + 63) / 64;
  int st;
  if (!B.mixed) {
    if ((st = launch_slots<double>(B, s, ci, rows))) return st;
    k_b_check_begin<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, ci, max_sweeps);
    FFPM_LAUNCHED("k_b_check_begin");
  } else {
    k_bm_begin<<<B.g_ew.blocks, kThreads, 0, s>>>(B.d_solves, B.g_ew, B.d_state);
    FFPM_LAUNCHED("k_bm_begin");
    k_b_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_RZ, 0, 0, 1);
    FFPM_LAUNCHED("k_b_finish");
    k_bm_after_begin<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, ci, max_sweeps,
                                       inner_rtol());
    FFPM_LAUNCHED("k_bm_after_begin");
    if ((st = launch_slots<float>(B, s, ci, rows))) return st;
    k_bm_after_inner<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, ci, max_sweeps);
    FFPM_LAUNCHED("k_bm_after_inner");
    k_bm_acc<<<B.g_ew.blocks, kThreads, 0, s>>>(B.d_solves, B.g_ew, B.d_state);
    FFPM_LAUNCHED("k_bm_acc");
    k_bm_after_acc<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count);
    FFPM_LAUNCHED("k_bm_after_acc");
  }
  if ((st = launch_check(B, s, ci, max_sweeps, tol))) return st;
  k_b_count<<<1, 256, 0, s>>>(B.d_state, B.count, B.d_counters, B.d_counters + 1);
  FFPM_LAUNCHED("k_b_count");
  return OK;
}

// ---- multigrid CG launches -------------------------------------------------
// The true residual of the solves that need one, then the loop logic after it.
int launch_check_mg(const Batch &B, hipStream_t s, int ci, int max_sweeps, double tol) {
  const int cb = (B.count + 63) / 64;
  k_b_tr1<<<B.g_tr.blocks, dim3(kTileX, B.h[0].tt.ty), 0, s>>>(B.d_solves, B.g_tr, B.d_state);
  FFPM_LAUNCHED("k_b_tr1");
  k_bg_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_TR1, 0);
  FFPM_LAUNCHED("k_bg_finish");
  // the projection of r (the true residual's second pass) runs in the
  // solve's next k_bg_first
  k_bg_check_end<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, ci, max_sweeps, tol,
                                   B.d_counters);
  FFPM_LAUNCHED("k_bg_check_end");
  return OK;
}

// One V-cycle of every solve iterating in slot m: each level's step is one
// launch over the solves that have the level. The steps (buffers and
// coefficients) are the same for every solve.
template <bool NT, bool CHEB>
int launch_cycle_mg(const Batch &B, hipStream_t s, int m) {
  const Multigrid &h = B.h[0].mg;
  const MgSchedule &s0 = h.s0, &sc = h.sc;
  const dim3 cblk(mgk::kMt), sblk(kTileX, B.h[0].tl32.ty);
  auto step0 = [&](const MgStep &stp, bool last) -> int {
    for (int c : {1, 2, 4}) {
      const Grid &g = B.g_pcg32[c];
      if (!g.blocks) continue;
      auto go = [&](auto kern) {
        kern<<<g.blocks, sblk, 0, s>>>(B.d_solves, g, B.d_state, m, stp);
      };
      if (c == 4) last ? go(k_bg_sweep<4, true, NT, CHEB>) : go(k_bg_sweep<4, false, NT, CHEB>);
      else if (c == 2) last ? go(k_bg_sweep<2, true, NT, CHEB>) : go(k_bg_sweep<2, false, NT, CHEB>);
      else last ? go(k_bg_sweep<1, true, NT, CHEB>) : go(k_bg_sweep<1, false, NT, CHEB>);
    }
    FFPM_LAUNCHED("k_bg_sweep");
    return OK;
  };
  int st;
  for (int k = 0; k < s0.npre; ++k)
    if ((st = step0(s0.pre[k], false))) return st;
  if (h.stream_restrict) {
    const dim3 rblk(kTileX, h.tlr.ty);
    for (int c : {1, 2, 4}) {
      const Grid &g = B.g_res[c];
      if (!g.blocks) continue;
      const size_t lb = mgk::mg_restrict_lds(h.tlr.ty, c);
      auto go = [&](auto kern) {
        kern<<<g.blocks, rblk, lb, s>>>(B.d_solves, g, B.d_state, m, s0.mid);
      };
      if (c == 4) go(k_bg_restrict_t<4, NT>);
      else if (c == 2) go(k_bg_restrict_t<2, NT>);
      else go(k_bg_restrict_t<1, NT>);
    }
  } else {
    if (B.g_col[0].blocks)
      k_bg_restrict0<false><<<B.g_col[0].blocks, cblk, 0, s>>>(B.d_solves, B.g_col[0],
                                                               B.d_state, m, s0.mid);
    if (B.g_col[1].blocks)
      k_bg_restrict0<true><<<B.g_col[1].blocks, cblk, 0, s>>>(B.d_solves, B.g_col[1],
                                                              B.d_state, m, s0.mid);
  }
  FFPM_LAUNCHED("k_bg_restrict0");
  for (int l = 1; l < B.mg_top; ++l) {
    for (int k = 0; k < sc.npre; ++k)
      k_bg_csweep<<<B.g_lvl[l].blocks, cblk, 0, s>>>(B.d_solves, B.g_lvl[l], B.d_state, m, l,
                                                     sc.pre[k]);
    k_bg_crestrict<<<B.g_lvl[l + 1].blocks, cblk, 0, s>>>(B.d_solves, B.g_lvl[l + 1],
                                                          B.d_state, m, l, sc.mid);
  }
  FFPM_LAUNCHED("k_bg_crestrict");
  k_bg_coarsest<<<B.count, kMgCoarsestCells, 0, s>>>(B.d_solves, B.d_state, m, sc.out);
  FFPM_LAUNCHED("k_bg_coarsest");
  for (int l = B.mg_top - 1; l >= 1; --l) {
    k_bg_cprolong<<<B.g_lvl[l].blocks, cblk, 0, s>>>(B.d_solves, B.g_lvl[l], B.d_state, m, l,
                                                     sc.mid, sc.out);
    for (int k = 0; k < sc.npost; ++k)
      k_bg_csweep<<<B.g_lvl[l].blocks, cblk, 0, s>>>(B.d_solves, B.g_lvl[l], B.d_state, m, l,
                                                     sc.post[k]);
  }
  FFPM_LAUNCHED("k_bg_csweep");
  if (B.g_col[0].blocks)
    k_bg_prolong0<false><<<B.g_col[0].blocks, cblk, 0, s>>>(B.d_solves, B.g_col[0], B.d_state,
                                                            m, s0.mid, sc.out);
  if (B.g_col[1].blocks)
    k_bg_prolong0<true><<<B.g_col[1].blocks, cblk, 0, s>>>(B.d_solves, B.g_col[1], B.d_state,
                                                           m, s0.mid, sc.out);
  FFPM_LAUNCHED("k_bg_prolong0");
  for (int k = 0; k < s0.npost; ++k)
    if ((st = step0(s0.post[k], k == s0.npost - 1))) return st;
  return OK;
}

// One chunk: ci slots (first step, cycle, FCG(1) beta, direction and SpMV,
// alpha, update), then the check logic and the true residual where needed.
int launch_chunk_mg(const Batch &B, hipStream_t s, int ci, int max_sweeps, double tol) {
  const int cb = (B.count + 63) / 64;
  const dim3 cblk(mgk::kMt), pblk(kTileX, B.h[0].tl.ty), sblk(kTileX, B.h[0].tl32.ty);
  int st;
  for (int m = 0; m < ci; ++m) {
    if (m == 0) {
      k_bg_first<<<B.g_cell0.blocks, cblk, 0, s>>>(B.d_solves, B.g_cell0, B.d_state);
      FFPM_LAUNCHED("k_bg_first");
    }
    const Multigrid &h = B.h[0].mg;
    if (h.cheb)
      st = h.nontemporal ? launch_cycle_mg<true, true>(B, s, m)
                         : launch_cycle_mg<false, true>(B, s, m);
    else
      st = h.nontemporal ? launch_cycle_mg<true, false>(B, s, m)
                         : launch_cycle_mg<false, false>(B, s, m);
    if (st) return st;
    k_bg_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_FCG1, m);
    FFPM_LAUNCHED("k_bg_finish");
    for (int c : {1, 2, 4}) {
      const Grid &g = B.g_pcg32[c];
      if (!g.blocks) continue;
      if (c == 4) k_bg_dir<4><<<g.blocks, sblk, 0, s>>>(B.d_solves, g, B.d_state, m);
      else if (c == 2) k_bg_dir<2><<<g.blocks, sblk, 0, s>>>(B.d_solves, g, B.d_state, m);
      else k_bg_dir<1><<<g.blocks, sblk, 0, s>>>(B.d_solves, g, B.d_state, m);
      FFPM_LAUNCHED("k_bg_dir");
    }
    k_bg_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_PQ, m);
    FFPM_LAUNCHED("k_bg_finish");
    for (int c : {1, 2}) {
      const Grid &g = B.g_pcg[c];
      if (!g.blocks) continue;
      if (c == 2) k_bg_upd<2><<<g.blocks, pblk, 0, s>>>(B.d_solves, g, B.d_state, m);
      else k_bg_upd<1><<<g.blocks, pblk, 0, s>>>(B.d_solves, g, B.d_state, m);
      FFPM_LAUNCHED("k_bg_upd");
    }
    k_bg_finish<<<B.count, kThreads, 0, s>>>(B.d_solves, B.d_state, FIN_MAX, m);
    FFPM_LAUNCHED("k_bg_finish");
    k_bg_slot_end<<<cb, 64, 0, s>>>(B.d_state, B.count, m);
    FFPM_LAUNCHED("k_bg_slot_end");
  }
  k_bg_check_begin<<<cb, 64, 0, s>>>(B.d_solves, B.d_state, B.count, ci, max_sweeps, tol);
  FFPM_LAUNCHED("k_bg_check_begin");
  if ((st = launch_check_mg(B, s, ci, max_sweeps, tol))) return st;
  k_b_count<<<1, 256, 0, s>>>(B.d_state, B.count, B.d_counters, B.d_counters + 1);
  FFPM_LAUNCHED("k_b_count");
  return OK;
}

struct StreamGuard {
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~StreamGuard() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (s) (void)hipStreamDestroy(s);
  }
};

// Runs the batch. Returns per-solve states in `out` (host), the host time of
// each replay's end in `replay_ms` (entry 0: the initial residuals) and the
// time the first replay was issued (after the graph is built) in loop_ms.
int batch_run(Batch &B, double tolerance, int max_sweeps, int check_interval, int use_graph,
              std::vector<BState> &out, std::vector<double> &replay_ms, double *loop_ms) {
  const int ci = B.mgm ? mg_check_interval(check_interval)
                       : (check_interval > 1 ? check_interval : 1);
  const int count = B.count;
  int st;
  // All solves must share the PCG workgroup height (one launch per grid);
  // the multigrid cycle's sweeps have their own (sblk), the mixed inner
  // PCG's FP32 kernels launch with the FP64 one.
  const int ty = B.h[0].tl.ty, ty32 = B.mgm ? B.h[0].tl32.ty : ty;
  for (int s = 0; s < count; ++s)
    if (B.h[s].tl.ty != ty || B.h[s].tl32.ty != ty32 || B.h[s].tt.ty != B.h[0].tt.ty) {
      set_error(ERR_ARGUMENT, "batched solves need one workgroup height");
      return ERR_ARGUMENT;
    }
  StreamGuard g;
  FFPM_HIP(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
  FFPM_HIP(hipStreamSynchronize(stream()));
  const auto t0 = std::chrono::steady_clock::now();
  k_b_init<<<(count + 63) / 64, 64, 0, g.s>>>(B.d_state, count);
  FFPM_LAUNCHED("k_b_init");
  k_b_zero_int<<<1, 64, 0, g.s>>>(B.d_counters, 4);
  FFPM_LAUNCHED("k_b_zero_int");
  // the initial true residual of every solve
  st = B.mgm ? launch_check_mg(B, g.s, ci, max_sweeps, tolerance)
             : launch_check(B, g.s, ci, max_sweeps, tolerance);
  if (st) return st;
  k_b_count<<<1, 256, 0, g.s>>>(B.d_state, count, B.d_counters, B.d_counters + 1);
  FFPM_LAUNCHED("k_b_count");
  int active = 0;
  FFPM_HIP(hipMemcpyAsync(&active, B.d_counters + 1, sizeof(int), hipMemcpyDeviceToHost, g.s));
  FFPM_HIP(hipStreamSynchronize(g.s));
  replay_ms.assign(1, std::chrono::duration<double, std::milli>(
                          std::chrono::steady_clock::now() - t0).count());
  if (use_graph && active > 0) {
    FFPM_HIP(hipStreamBeginCapture(g.s, hipStreamCaptureModeThreadLocal));
    st = B.mgm ? launch_chunk_mg(B, g.s, ci, max_sweeps, tolerance)
               : launch_chunk(B, g.s, ci, max_sweeps, tolerance, ty);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(g.s, &graph);
    g.graph = graph;
    if (st) return st;
    FFPM_HIP(e);
    FFPM_HIP(hipGraphInstantiate(&g.exec, g.graph, nullptr, nullptr, 0));
  }
  *loop_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                 .count();
  while (active > 0) {
    if (use_graph) {
      FFPM_HIP(hipGraphLaunch(g.exec, g.s));
    } else if ((st = B.mgm ? launch_chunk_mg(B, g.s, ci, max_sweeps, tolerance)
                           : launch_chunk(B, g.s, ci, max_sweeps, tolerance, ty))) {
      return st;
    }
    FFPM_HIP(hipMemcpyAsync(&active, B.d_counters + 1, sizeof(int), hipMemcpyDeviceToHost,
                            g.s));
    FFPM_HIP(hipStreamSynchronize(g.s));
    replay_ms.push_back(std::chrono::duration<double, std::milli>(
                            std::chrono::steady_clock::now() - t0).count());
  }
  out.resize(count);
  FFPM_HIP(hipMemcpyAsync(out.data(), B.d_state, sizeof(BState) * count, hipMemcpyDeviceToHost,
                          g.s));
  FFPM_HIP(hipStreamSynchronize(g.s));
  return OK;
}

// Gauge, unpack and the final residual of every solve (as the standalone
// solve ends, src/fpm_stairstep_mpi_3d.f90:170-190), one launch per step;
// the residuals land in res (host).
int batch_finalize(const Batch &B, std::vector<double> &res) {
  k_bf_gauge_sums<<<B.g_ew.blocks, kThreads, 0, stream()>>>(B.d_solves, B.d_probs, B.g_ew);
  FFPM_LAUNCHED("k_bf_gauge_sums");
  k_bf_finish<<<B.count, 256, 0, stream()>>>(B.d_solves, FIN_SUMS, nullptr);
  FFPM_LAUNCHED("k_bf_finish");
  k_bf_gauge_apply<<<B.g_pad.blocks, kThreads, 0, stream()>>>(B.d_solves, B.d_probs, B.g_pad);
  FFPM_LAUNCHED("k_bf_gauge_apply");
  k_bf_residual_max<<<B.g_tr.blocks, dim3(kTileX, B.h[0].tt.ty), 0, stream()>>>(B.d_solves,
                                                                                 B.g_tr);
  FFPM_LAUNCHED("k_bf_residual_max");
  k_bf_finish<<<B.count, 256, 0, stream()>>>(B.d_solves, FIN_MAX, B.d_res);
  FFPM_LAUNCHED("k_bf_finish");
  res.resize(B.count);
  FFPM_HIP(hipMemcpyAsync(res.data(), B.d_res, sizeof(double) * B.count, hipMemcpyDeviceToHost,
                          stream()));
  return OK;
}

}  // namespace

}  // namespace ffpm

using namespace ffpm;

extern "C" {

int ffpm_solve_stairstep_batch_dev(ffpm_stairstep_problem *problems, int count,
                                   double tolerance, int max_sweeps, int check_interval,
                                   int method, int rows, int zchunk, int use_graph) {
  if (!problems || count < 1) {
    set_error(ERR_ARGUMENT, "empty batch");
    return ERR_ARGUMENT;
  }
  if (method != FFPM_METHOD_PCG && method != FFPM_METHOD_PCG_MIXED &&
      method != FFPM_METHOD_FCG_MG) {
    set_error(ERR_METHOD, "batched solves support pcg, pcg-mixed and fcg-mg");
    return ERR_METHOD;
  }
  if (rows != 0 && rows != 1 && rows != 2 && rows != 4 && rows != 8) {
    set_error(ERR_ARGUMENT, "rows must be 0, 1, 2, 4 or 8");
    return ERR_ARGUMENT;
  }
  Batch B;
  int st = batch_setup(B, problems, count, rows, zchunk, method == FFPM_METHOD_PCG_MIXED,
                       method == FFPM_METHOD_FCG_MG);
  if (st) return st;
  std::vector<BState> out;
  std::vector<double> replay_ms;
  double loop_ms = 0;
  if ((st = batch_run(B, tolerance, max_sweeps, check_interval, use_graph, out, replay_ms,
                      &loop_ms)))
    return st;
  std::vector<double> res;
  if ((st = batch_finalize(B, res))) return st;
  FFPM_HIP(hipStreamSynchronize(stream()));
  for (int s = 0; s < count; ++s) {
    problems[s].sweeps = B.mixed ? out[s].inner : out[s].sweeps;
    problems[s].residual = res[s];
    const int r = std::min(std::max(out[s].finished_at, 0), (int)replay_ms.size() - 1);
    problems[s].solve_ms = replay_ms[r];
  }
  return FFPM_OK;
}

int ffpm_autotune_stairstep_batch_dev(ffpm_stairstep_problem *problems, int count,
                                      int method, int iterations, int *rows, int *zchunk,
                                      double *ms_per_iteration) {
  if (!problems || count < 1 || !rows || !zchunk || !ms_per_iteration || iterations < 1) {
    set_error(ERR_ARGUMENT, "autotune needs a batch, an iteration count and outputs");
    return ERR_ARGUMENT;
  }
  if (method != FFPM_METHOD_PCG && method != FFPM_METHOD_PCG_MIXED &&
      method != FFPM_METHOD_FCG_MG) {
    set_error(ERR_METHOD, "batched solves support pcg, pcg-mixed and fcg-mg");
    return ERR_METHOD;
  }
  static const int kRows[] = {4, 8};
  static const int kChunks[] = {8, 16, 32};
  double best = 1e300;
  int br = 8, bz = 16;
  for (int r : kRows) {
    for (int zc : kChunks) {
      Batch B;
      int st = batch_setup(B, problems, count, r, zc, method == FFPM_METHOD_PCG_MIXED,
                           method == FFPM_METHOD_FCG_MG);
      if (st) return st;
      std::vector<BState> out;
      std::vector<double> replay_ms;
      // Fixed work: every solve runs `iterations` iterations in one chunk.
      double loop_ms = 0;
      if ((st = batch_run(B, -1.0, iterations, iterations, 1, out, replay_ms, &loop_ms)))
        return st;
      const double ms = replay_ms.back() - loop_ms;
      if (ms < best) {
        best = ms;
        br = r;
        bz = zc;
      }
    }
  }
  *rows = br;
  *zchunk = bz;
  *ms_per_iteration = best / iterations;
  return FFPM_OK;
}

}  // extern "C"
