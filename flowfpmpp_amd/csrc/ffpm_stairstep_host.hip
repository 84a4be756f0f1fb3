// Host orchestration of the stair-step solve (src/fpm_stairstep_mpi_3d.f90)
// and its C ABI. The iteration logic (convergence checks, breakdown restart,
// stall exit) follows the reference; the per-iteration scalars stay on the
// device, and the host reads them once per convergence check.
#include <cfloat>
#include <cstdlib>
#include <initializer_list>
#include <mutex>
#include <cstring>
#include <string>

#include "../../include/ffpm.h"
#include "ffpm_mg.hpp"
#include "ffpm_pcg.hpp"
#include "ffpm_pcg_body.hpp"
#include "ffpm_runtime.hpp"
#include "ffpm_unet.hpp"

namespace ffpm {

namespace {

constexpr int kThreads = 256;

bool any_neighbour(int west, int east, int south, int north) {
  return west >= 0 || east >= 0 || south >= 0 || north >= 0;
}

// Every array argument of a device entry point must be present.
bool all_present(std::initializer_list<const void *> ptrs) {
  for (const void *p : ptrs)
    if (!p) return false;
  return true;
}

int null_error() {
  set_error(ERR_ARGUMENT, "null array argument");
  return ERR_ARGUMENT;
}

int decomposition_error() {
  set_error(ERR_DECOMPOSITION,
            "stair-step solver holds the whole domain on one GPU: every neighbour rank must be "
            "MPI_PROC_NULL");
  return ERR_DECOMPOSITION;
}

// Device buffers of one solve; fields are dense (nx,ny,nz).
struct Frame {
  int nx = 0, ny = 0, nz = 0;
  size_t n = 0;
  Geom g{};
  Tiling tiling{};
  Tiling tl[5]{};  // launch shapes of the CL-cells-per-lane kernels, CL = 1, 2, 4
  double *phi = nullptr, *next = nullptr, *src = nullptr, *diag = nullptr, *rdiag = nullptr;
  double *r = nullptr, *q = nullptr, *z = nullptr, *partials = nullptr;
  uint8_t *code = nullptr;
  Scalars *sc = nullptr;
  int *isolated = nullptr;
};

int max_blocks_for(const Tiling &t, size_t n) {
  int b = elementwise_blocks(n);
  return t.blocks > b ? t.blocks : b;
}

// Allocates the frame's work fields and 1-D coefficients and fills the
// coefficient arrays from the axes.
int frame_setup(Frame &f, int nx, int ny, int nz, const double *xc, const double *yc,
                const double *zc, const double *dx, const double *dy, const double *dz,
                bool fields, bool pcg_fields) {
  int st = 0;
  Workspace &ws = workspace();
  f.nx = nx;
  f.ny = ny;
  f.nz = nz;
  f.n = (size_t)nx * ny * nz;
  f.tiling = make_tiling(nx, ny, nz, 0);
  size_t nb = (size_t)max_blocks_for(f.tiling, f.n);
  // Rows per workgroup of the pipelined PCG and Jacobi kernels (built for up
  // to 1024 threads): FFPM_PCG_ROWS (1, 2, 4, 8, 16), else FFPM_TILE_Y, else
  // 8, measured best on a 512x384x256 frame.
  const char *pr = std::getenv("FFPM_PCG_ROWS");
  const int pcg_rows = pr ? std::atoi(pr) : (std::getenv("FFPM_TILE_Y") ? 0 : 8);
  for (int cl : {1, 2, 4}) {
    f.tl[cl] = make_tiling(nx, ny, nz, 0, pcg_rows, kTileX * cl);
    if ((size_t)f.tl[cl].blocks > nb) nb = (size_t)f.tl[cl].blocks;
    // the multigrid direction's FP32 level-0 tiling (its partials)
    const Tiling tm = make_tiling(nx, ny, nz, 0, mg_rows(), kTileX * cl);
    if ((size_t)tm.blocks > nb) nb = (size_t)tm.blocks;
  }
  {
    // the cycle's level-0 tiling (its last sweep's partials)
    const Tiling tc = mg_cycle_tiling(nx, ny, nz, mg_rows(), mg_cells_per_lane(nx));
    if ((size_t)tc.blocks > nb) nb = (size_t)tc.blocks;
  }
  double *coef = (double *)ws.get(S_COEF, sizeof(double) * 2 * (nx + ny + nz), &st);
  if (!coef) return st;
  f.partials = (double *)ws.get(S_PARTIALS, sizeof(double) * kReduceSlots * nb, &st);
  if (!f.partials) return st;
  f.sc = (Scalars *)ws.get(S_SCALARS, sizeof(Scalars), &st);
  if (!f.sc) return st;
  f.isolated = (int *)ws.get(S_ISOLATED, sizeof(int) * 4, &st);
  if (!f.isolated) return st;
  f.g.nx = nx;
  f.g.ny = ny;
  f.g.nz = nz;
  double *cw = coef, *ce = cw + nx, *cs = ce + nx, *cn = cs + ny, *cb = cn + ny, *ct = cb + nz;
  f.g.cw = cw; f.g.ce = ce; f.g.cs = cs; f.g.cn = cn; f.g.cb = cb; f.g.ct = ct;
  f.g.dx = dx; f.g.dy = dy; f.g.dz = dz;
  f.g.xc = xc; f.g.yc = yc; f.g.zc = zc;
  FFPM_HIP(hipMemsetAsync(f.sc, 0, sizeof(Scalars), stream()));
  if (!dx || !dy || !dz) {
    set_error(ERR_ARGUMENT, "null cell-width array");
    return ERR_ARGUMENT;
  }
  // The stencil coefficients need the centres; callers without centres
  // (the compatibility correction) use only the widths.
  if (xc && yc && zc) {
    const int tot = nx + ny + nz;
    k_axis_coefficients<<<(tot + 255) / 256, 256, 0, stream()>>>(nx, ny, nz, xc, yc, zc, dx, dy,
                                                                 dz, cw, ce, cs, cn, cb, ct);
    FFPM_LAUNCHED("k_axis_coefficients");
  } else if (fields) {
    set_error(ERR_ARGUMENT, "null cell-centre array");
    return ERR_ARGUMENT;
  }
  if (fields) {
    const size_t bytes = sizeof(double) * f.n;
    if (!(f.phi = (double *)ws.get(S_PHI, bytes, &st))) return st;
    if (!(f.next = (double *)ws.get(S_NEXT, bytes, &st))) return st;
    if (!(f.src = (double *)ws.get(S_SRC, bytes, &st))) return st;
    if (!(f.diag = (double *)ws.get(S_DIAG, bytes, &st))) return st;
    if (!(f.rdiag = (double *)ws.get(S_RDIAG, bytes, &st))) return st;
    if (!(f.code = (uint8_t *)ws.get(S_CODE, f.n, &st))) return st;
    if (pcg_fields) {
      if (!(f.r = (double *)ws.get(S_R, bytes, &st))) return st;
      if (!(f.q = (double *)ws.get(S_Q, bytes, &st))) return st;
      if (!(f.z = (double *)ws.get(S_Z, bytes, &st))) return st;
    }
  }
  return OK;
}

// Totals the partials of the reduction kernel just launched (k_finish).
int finish(const Frame &f, int op, int nslots, int nblocks) {
  k_finish<<<1, 256, 0, stream()>>>(op, nslots, f.partials, nblocks, f.sc);
  FFPM_LAUNCHED("k_finish");
  return OK;
}

int read_scalars(const Frame &f, Scalars *h) {
  FFPM_HIP(hipMemcpyAsync(h, f.sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  return OK;
}

// true_residual (src/fpm_stairstep_mpi_3d.f90:297-329); res read back. With
// reset the direction p = M^-1 r is written to f.next (the three-kernel
// iteration); the fused iteration applies the reset in its next SpMV instead.
// true_residual's first pass: r = V (src - A phi) unprojected, the max-norm
// residual in sc->res and the projection sums in sc->red1, sc->red2.
int true_residual_pass1(Frame &f) {
  {
    TimedLaunch tl("true_residual_1");
    k_true_residual_1<<<f.tiling.blocks, dim3(kTileX, f.tiling.ty), 0, stream()>>>(
        f.tiling, f.g, f.phi, f.src, f.diag, f.r, f.partials, f.sc);
    FFPM_LAUNCHED("k_true_residual_1");
  }
  return finish(f, FIN_TR1, 4, f.tiling.blocks);
}

int true_residual(Frame &f, bool reset, double *res) {
  const int eb = elementwise_blocks(f.n);
  if (int st = true_residual_pass1(f)) return st;
  {
    TimedLaunch tl("true_residual_2");
    k_true_residual_2<<<eb, kThreads, 0, stream()>>>(f.n, f.r, f.rdiag, f.diag, f.next,
                                                     reset ? 1 : 0, f.z, f.partials, f.sc);
    FFPM_LAUNCHED("k_true_residual_2");
  }
  finish(f, FIN_RZ, 1, eb);
  Scalars h;
  int st = read_scalars(f, &h);
  if (st) return st;
  *res = h.res;
  return OK;
}

int residual_max(Frame &f, const double *phi, double *res) {
  {
    TimedLaunch tl("residual_max");
    k_residual_max<<<f.tiling.blocks, dim3(kTileX, f.tiling.ty), 0, stream()>>>(
        f.tiling, f.g, phi, f.src, f.diag, f.partials, f.sc);
    FFPM_LAUNCHED("k_residual_max");
  }
  finish(f, FIN_MAX, 1, f.tiling.blocks);
  Scalars h;
  int st = read_scalars(f, &h);
  if (st) return st;
  *res = h.res;
  return OK;
}

// pcg_iterate (src/fpm_stairstep_mpi_3d.f90:235-293). Iterations between two
// convergence checks run as one chunk of launches without a host sync; a
// breakdown or a negative rz_new inside a chunk turns the rest of the chunk
// into no-ops, and the host resumes exactly where the reference would.
int pcg_iterate(Frame &f, double tolerance, int max_sweeps, int check_interval, int *sweeps,
                double *residual) {
  const int ci = check_interval > 1 ? check_interval : 1;
  const int eb = elementwise_blocks(f.n);
  double res;
  int st = true_residual(f, true, &res);
  if (st) return st;
  int it = 0, restarts = 0, stalled = 0;
  double best = res;
  while (*sweeps < max_sweeps && res > tolerance) {
    int n = ci - it % ci;
    if (max_sweeps - *sweeps < n) n = max_sweeps - *sweeps;
    k_chunk_begin<<<1, 64, 0, stream()>>>(f.sc);
    FFPM_LAUNCHED("k_chunk_begin");
    for (int m = 0; m < n; ++m) {
      {
        TimedLaunch tl("spmv_dot");
        k_spmv_dot<<<f.tiling.blocks, dim3(kTileX, f.tiling.ty), 0, stream()>>>(
            f.tiling, f.g, f.next, f.q, f.diag, f.partials, f.sc);
      }
      finish(f, FIN_PQ, 1, f.tiling.blocks);
      {
        TimedLaunch tl("update_dot");
        k_update_dot<<<eb, kThreads, 0, stream()>>>(f.n, f.phi, f.next, f.r, f.q, f.rdiag,
                                                    f.partials, f.sc);
      }
      finish(f, FIN_RZNEW, 1, eb);
      {
        TimedLaunch tl("direction");
        k_direction<<<eb, kThreads, 0, stream()>>>(f.n, f.next, f.r, f.rdiag, f.sc);
      }
    }
    FFPM_LAUNCHED("pcg chunk");
    Scalars h;
    if ((st = read_scalars(f, &h))) return st;
    *sweeps += h.iters;
    it += h.iters;
    if (h.iters > 0) restarts = 0;
    if (h.status == kNegative) break;
    if (h.status == kBreakdown) {
      // Breakdown (converged to roundoff or lost conjugacy): restart once.
      restarts = restarts + 1;
      if (restarts > 1) break;
      k_clear_status<<<1, 64, 0, stream()>>>(f.sc);
      FFPM_LAUNCHED("k_clear_status");
      if ((st = true_residual(f, true, &res))) return st;
      continue;
    }
    if (it % ci == 0 || *sweeps == max_sweeps) {
      if ((st = true_residual(f, false, &res))) return st;
      // Stop at the roundoff floor: no 1% improvement over 50 checks.
      if (res < 0.99 * best) {
        best = res;
        stalled = 0;
      } else {
        stalled = stalled + 1;
      }
      if (stalled >= 50 || !(res == res)) break;
    }
  }
  *residual = res;
  return OK;
}

// Cells per lane of the fused kernels: up to a 16-byte vector.
template <typename T>
int cells_per_lane(int nx) {
  return ffpm::cells_per_lane(nx, (int)(16 / sizeof(T)));
}

template <typename T>
int launch_dir(const Frame &f, int cl, const GeomT<T> &g, const T *z, const T *pin, T *pout,
               int mode) {
  const Tiling &t = f.tl[cl];
  const dim3 blk(kTileX, t.ty);
  {
    TimedLaunch tl(sizeof(T) == 8 ? "pcg_dir_spmv" : "pcg32_dir_spmv");
    auto go = [&](auto kern) {
      kern<<<t.blocks, blk, 0, stream()>>>(t, g, z, f.code, pin, pout, f.partials, f.sc);
    };
    auto by_mode = [&](auto k_update, auto k_reset, auto k_current) {
      if (mode == kDirUpdate) go(k_update);
      else if (mode == kDirReset) go(k_reset);
      else go(k_current);
    };
    constexpr int CLV = sizeof(T) == 4 ? 4 : 2;
    if (cl == 4)
      by_mode(k_pcg_dir_t<T, CLV, kDirUpdate>, k_pcg_dir_t<T, CLV, kDirReset>,
              k_pcg_dir_t<T, CLV, kDirCurrent>);
    else if (cl == 2)
      by_mode(k_pcg_dir_t<T, 2, kDirUpdate>, k_pcg_dir_t<T, 2, kDirReset>,
              k_pcg_dir_t<T, 2, kDirCurrent>);
    else
      by_mode(k_pcg_dir_t<T, 1, kDirUpdate>, k_pcg_dir_t<T, 1, kDirReset>,
              k_pcg_dir_t<T, 1, kDirCurrent>);
  }
  FFPM_LAUNCHED("k_pcg_dir_t");
  return finish(f, FIN_PQ, 1, t.blocks);
}

template <typename T>
int launch_upd(const Frame &f, int cl, const GeomT<T> &g, const T *p, T *x, T *r, T *z) {
  const Tiling &t = f.tl[cl];
  const dim3 blk(kTileX, t.ty);
  {
    TimedLaunch tl(sizeof(T) == 8 ? "pcg_update" : "pcg32_update");
    if (cl == 4)
      k_pcg_upd_t<T, (sizeof(T) == 4 ? 4 : 2)><<<t.blocks, blk, 0, stream()>>>(
          t, g, p, f.code, x, r, z, f.partials, f.sc);
    else if (cl == 2)
      k_pcg_upd_t<T, 2><<<t.blocks, blk, 0, stream()>>>(t, g, p, f.code, x, r, z, f.partials,
                                                        f.sc);
    else
      k_pcg_upd_t<T, 1><<<t.blocks, blk, 0, stream()>>>(t, g, p, f.code, x, r, z, f.partials,
                                                        f.sc);
  }
  FFPM_LAUNCHED("k_pcg_upd_t");
  return finish(f, FIN_RZNEW, 1, t.blocks);
}

// pcg_iterate with the fused two-kernel iteration (k_pcg_dir_t +
// k_pcg_upd_t): the same formulas and the same convergence logic; the
// direction update p = M^-1 r + beta p of iteration k runs inside the SpMV of
// iteration k+1, except before a convergence check, where it runs on its own
// with the recursively updated r exactly as the reference orders it
// (src/fpm_stairstep_mpi_3d.f90:279-286), and after a restart, where the next
// SpMV starts from p = M^-1 r.
int pcg_iterate_fused(Frame &f, double tolerance, int max_sweeps, int check_interval,
                      int *sweeps, double *residual) {
  const int ci = check_interval > 1 ? check_interval : 1;
  const int eb = elementwise_blocks(f.n);
  const int cl = cells_per_lane<double>(f.nx);
  const GeomT<double> g = geom_t(f.g);
  double *pbuf[2] = {f.next, f.q};
  int cur = 0;
  double res;
  int st = true_residual(f, false, &res);
  if (st) return st;
  int mode = kDirReset;
  int it = 0, restarts = 0, stalled = 0;
  double best = res;
  while (*sweeps < max_sweeps && res > tolerance) {
    int n = ci - it % ci;
    if (max_sweeps - *sweeps < n) n = max_sweeps - *sweeps;
    k_chunk_begin<<<1, 64, 0, stream()>>>(f.sc);
    FFPM_LAUNCHED("k_chunk_begin");
    for (int m = 0; m < n; ++m) {
      const int dst = cur ^ 1;
      if ((st = launch_dir<double>(f, cl, g, f.z, pbuf[cur], pbuf[dst], mode))) return st;
      if (mode != kDirCurrent) cur = dst;
      if ((st = launch_upd<double>(f, cl, g, pbuf[cur], f.phi, f.r, f.z))) return st;
      mode = kDirUpdate;
    }
    Scalars h;
    if ((st = read_scalars(f, &h))) return st;
    *sweeps += h.iters;
    it += h.iters;
    if (h.iters > 0) restarts = 0;
    if (h.status == kNegative) break;
    if (h.status == kBreakdown) {
      restarts = restarts + 1;
      if (restarts > 1) break;
      k_clear_status<<<1, 64, 0, stream()>>>(f.sc);
      FFPM_LAUNCHED("k_clear_status");
      if ((st = true_residual(f, false, &res))) return st;
      mode = kDirReset;
      continue;
    }
    if (it % ci == 0 || *sweeps == max_sweeps) {
      {
        TimedLaunch tl("direction");
        k_dir_t<double><<<eb, kThreads, 0, stream()>>>(f.n, pbuf[cur], f.z, f.sc);
      }
      FFPM_LAUNCHED("k_dir_t");
      mode = kDirCurrent;
      if ((st = true_residual(f, false, &res))) return st;
      if (res < 0.99 * best) {
        best = res;
        stalled = 0;
      } else {
        stalled = stalled + 1;
      }
      if (stalled >= 50 || !(res == res)) break;
    }
  }
  *residual = res;
  return OK;
}

// Mixed precision: FP64 defect correction around an FP32 PCG. Each outer step
// forms the projected FP64 true residual r (src/fpm_stairstep_mpi_3d.f90:
// 297-329), solves A e = r approximately with the fused PCG in FP32 (FP32
// stencil, coefficients and preconditioner; dot products summed in FP64)
// until the preconditioned residual norm has dropped by inner_rtol, and adds
// e to phi in FP64. The outer loop stops on the reference's FP64 max-norm
// residual test, so the result is judged exactly like the FP64 solve.
// FP32 copies of the frame's 1-D coefficients and cell widths (slot
// S_COEF32), for the FP32 inner solves and the multigrid cycle.
int make_geom32(const Frame &f, GeomT<float> *g) {
  int st = 0;
  const int nx = f.nx, ny = f.ny, nz = f.nz;
  float *c32 = (float *)workspace().get(S_COEF32, sizeof(float) * 3 * (nx + ny + nz), &st);
  if (!c32) return st;
  const int ncoef = 2 * (nx + ny + nz);
  k_to_float<<<(ncoef + 255) / 256, 256, 0, stream()>>>(ncoef, f.g.cw, c32);
  k_to_float<<<(nx + 255) / 256, 256, 0, stream()>>>(nx, f.g.dx, c32 + ncoef);
  k_to_float<<<(ny + 255) / 256, 256, 0, stream()>>>(ny, f.g.dy, c32 + ncoef + nx);
  k_to_float<<<(nz + 255) / 256, 256, 0, stream()>>>(nz, f.g.dz, c32 + ncoef + nx + ny);
  FFPM_LAUNCHED("k_to_float");
  g->nx = nx; g->ny = ny; g->nz = nz;
  g->cw = c32; g->ce = c32 + nx; g->cs = c32 + 2 * nx; g->cn = c32 + 2 * nx + ny;
  g->cb = c32 + 2 * nx + 2 * ny; g->ct = c32 + 2 * nx + 2 * ny + nz;
  g->dx = c32 + ncoef; g->dy = c32 + ncoef + nx; g->dz = c32 + ncoef + nx + ny;
  return OK;
}

struct MixedStats {
  int outer = 0, inner = 0;
};
thread_local MixedStats g_mixed;

int pcg_mixed(Frame &f, double tolerance, int max_inner, int check_interval, double inner_rtol,
              int *sweeps, double *residual) {
  int st = 0;
  Workspace &ws = workspace();
  const size_t n = f.n;
  const int eb = elementwise_blocks(n);
  GeomT<float> g;
  if ((st = make_geom32(f, &g))) return st;
  float *e = (float *)ws.get(S_E32, sizeof(float) * n, &st);
  float *r32 = e ? (float *)ws.get(S_R32, sizeof(float) * n, &st) : nullptr;
  float *z32 = r32 ? (float *)ws.get(S_Z32, sizeof(float) * n, &st) : nullptr;
  float *p0 = z32 ? (float *)ws.get(S_P32A, sizeof(float) * n, &st) : nullptr;
  float *p1 = p0 ? (float *)ws.get(S_P32B, sizeof(float) * n, &st) : nullptr;
  if (!p1) return st;
  float *pbuf[2] = {p0, p1};
  const int cl = cells_per_lane<float>(f.nx);
  const int ci = check_interval > 1 ? check_interval : 1;
  double res;
  if ((st = true_residual(f, false, &res))) return st;
  double best = res;
  int stalled = 0, outer = 0, inner = 0;
  while (res > tolerance && inner < max_inner && outer < 100) {
    k_mixed_begin<<<eb, kThreads, 0, stream()>>>(g, f.r, f.code, r32, z32, e, f.partials);
    FFPM_LAUNCHED("k_mixed_begin");
    finish(f, FIN_RZ, 1, eb);
    k_clear_status<<<1, 64, 0, stream()>>>(f.sc);
    Scalars h;
    if ((st = read_scalars(f, &h))) return st;
    const double rz0 = h.rz;
    if (!(rz0 > 0)) break;
    const double target = rz0 * inner_rtol * inner_rtol;
    int mode = kDirReset, cur = 0;
    while (inner < max_inner) {
      int m_n = ci;
      if (max_inner - inner < m_n) m_n = max_inner - inner;
      k_chunk_begin<<<1, 64, 0, stream()>>>(f.sc);
      for (int m = 0; m < m_n; ++m) {
        const int dst = cur ^ 1;
        if ((st = launch_dir<float>(f, cl, g, z32, pbuf[cur], pbuf[dst], mode))) return st;
        cur = dst;
        if ((st = launch_upd<float>(f, cl, g, pbuf[cur], e, r32, z32))) return st;
        mode = kDirUpdate;
      }
      if ((st = read_scalars(f, &h))) return st;
      inner += h.iters;
      if (h.status != 0 || h.rz <= target || h.iters == 0) break;
    }
    k_mixed_accumulate<<<eb, kThreads, 0, stream()>>>(n, f.phi, e);
    FFPM_LAUNCHED("k_mixed_accumulate");
    k_clear_status<<<1, 64, 0, stream()>>>(f.sc);
    ++outer;
    if ((st = true_residual(f, false, &res))) return st;
    if (res < 0.99 * best) {
      best = res;
      stalled = 0;
    } else if (++stalled >= 3) {
      break;
    }
  }
  g_mixed.outer = outer;
  g_mixed.inner = inner;
  *sweeps = inner;
  *residual = res;
  return OK;
}

double mixed_inner_rtol() {
  const char *e = std::getenv("FFPM_MIXED_RTOL");
  return e && *e ? std::atof(e) : 1e-4;
}

bool classic_pcg() {
  const char *e = std::getenv("FFPM_PCG");
  return e && std::strcmp(e, "classic") == 0;
}

// The one-cell Jacobi kernel with the stored rdiag (FFPM_JACOBI=classic), for
// comparison.
bool classic_jacobi() {
  const char *e = std::getenv("FFPM_JACOBI");
  return e && std::strcmp(e, "classic") == 0;
}

// jacobi_iterate (src/fpm_stairstep_mpi_3d.f90:200-225) with ping-pong
// buffers: the current iterate is tracked instead of copied back.

int jacobi_iterate(Frame &f, double tolerance, int max_sweeps, int check_interval, int *sweeps,
                   double *residual, double **current) {
  double *cur = f.phi, *oth = f.next;
  while (*sweeps < max_sweeps && *residual > tolerance) {
    int nsteps = check_interval > 1 ? check_interval : 1;
    if (max_sweeps - *sweeps < nsteps) nsteps = max_sweeps - *sweeps;
    const int cl = cells_per_lane<double>(f.nx);
    const Tiling &t = f.tl[cl];
    const GeomT<double> g = geom_t(f.g);
    for (int s = 0; s < nsteps; ++s) {
      {
        TimedLaunch tl("jacobi");
        if (classic_jacobi())
          k_jacobi<<<f.tiling.blocks, dim3(kTileX, f.tiling.ty), 0, stream()>>>(
              f.tiling, f.g, cur, oth, f.src, f.rdiag);
        else if (cl == 2)
          k_jacobi_t<2><<<t.blocks, dim3(kTileX, t.ty), 0, stream()>>>(t, g, cur, oth, f.src,
                                                                        f.code);
        else
          k_jacobi_t<1><<<t.blocks, dim3(kTileX, t.ty), 0, stream()>>>(t, g, cur, oth, f.src,
                                                                        f.code);
      }
      double *t = cur;
      cur = oth;
      oth = t;
      *sweeps += 1;
    }
    FFPM_LAUNCHED("k_jacobi");
    int st = residual_max(f, cur, residual);
    if (st) return st;
  }
  *current = cur;
  return OK;
}

}  // namespace

// ------------------------------------------------ learned flexible CG ------
// Flexible CG with truncated orthogonalisation (Notay's FCG(m), m = 8), FP64
// except inside the network, with a two-term preconditioner: z = a u + g n,
// u = M_D^-1 r (the diagonal preconditioner of the plain PCG) and n = sigma *
// net(features, (r / V) / sigma) (sigma = max |r / V| over fluid cells, 0 in
// solids), where (a, g) minimise the A-norm error of z as a correction (a 2x2
// Galerkin system from u.Au, u.An, n.An, u.r, n.r). Then p = z minus its
// A-projections on the last m directions, q = A p formed from Au, An and the
// stored q's, alpha = p.r / p.q.
// With the initial phi the first step is the network-assisted initial guess.
// Every check_interval iterations the FP64 true residual (the reference's
// projected residual and max-norm) replaces r; three checks without a 1%
// improvement switch to diagonal PCG from the current phi. The answer is
// always judged by the FP64 max-norm residual.
__global__ __launch_bounds__(256) void k_fcg_absmax(Geom g, const double *__restrict__ r,
                                                    const double *__restrict__ diag,
                                                    double *partials) {
  __shared__ double lds[16 * kReduceSlots];
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  double m = 0.0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    if (!(diag[c] > 0)) continue;
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((size_t)g.nx * g.ny));
    m = fmax(m, fabs(r[c] / (g.dx[i] * g.dy[j] * g.dz[k])));
  }
  block_max(m, lds);
  store_partials(partials, 1, &m, gridDim.x);
}

__global__ void k_fcg_source(Geom g, const double *__restrict__ r, const double *__restrict__ diag,
                             const Scalars *sc, float *__restrict__ src) {
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  const double inv = sc->res > 0 ? 1.0 / sc->res : 0.0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % g.nx), j = (int)((c / g.nx) % g.ny), k = (int)(c / ((size_t)g.nx * g.ny));
    src[c] = diag[c] > 0 ? (float)(r[c] / (g.dx[i] * g.dy[j] * g.dz[k]) * inv) : 0.f;
  }
}

__global__ void k_fcg_correction(size_t n, const float *__restrict__ out,
                                 const double *__restrict__ diag, const Scalars *sc,
                                 double *__restrict__ z) {
  const double sigma = sc->res;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    z[c] = diag[c] > 0 ? sigma * (double)out[c] : 0.0;
}

__global__ __launch_bounds__(256) void k_dot2(size_t n, const double *__restrict__ a,
                                              const double *__restrict__ b,
                                              const double *__restrict__ c,
                                              const double *__restrict__ d, double *partials) {
  __shared__ double lds[16 * kReduceSlots];
  double s[2] = {0.0, 0.0};
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    s[0] += a[e] * b[e];
    s[1] += c[e] * d[e];
  }
  block_sum<2>(s, lds);
  store_partials(partials, 2, s, gridDim.x);
}


__global__ void k_fcg_update(size_t n, double alpha, const double *__restrict__ p,
                             const double *__restrict__ q, double *__restrict__ x,
                             double *__restrict__ r) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    x[c] = x[c] + alpha * p[c];
    r[c] = r[c] - alpha * q[c];
  }
}

struct FcgArgs {
  UNet *net;
  const float *features;
  int *fallback_at;  // FCG iterations before the switch to PCG, or -1
};

// Truncated orthogonalisation of flexible CG: p = z - sum_j c_j P_j and
// q = a Au + g An - sum_j c_j Q_j over the stored directions.
constexpr int kFcgMax = 8;
struct Combo {
  const double *v[kFcgMax];
  double c[kFcgMax];
  int m;
};

__global__ void k_fcg_orth(size_t n, const double *__restrict__ z, Combo P,
                           double *__restrict__ p) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    double v = z[e];
    for (int j = 0; j < P.m; ++j) v -= P.c[j] * P.v[j][e];
    p[e] = v;
  }
}

__global__ void k_fcg_orth_q(size_t n, double a, const double *__restrict__ au, double g,
                             const double *__restrict__ an, Combo Q, double *__restrict__ q) {
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    double v = a * au[e] + g * an[e];
    for (int j = 0; j < Q.m; ++j) v -= Q.c[j] * Q.v[j][e];
    q[e] = v;
  }
}

__global__ void k_fcg_combine(size_t n, double a, const double *__restrict__ u, double g,
                              const double *__restrict__ nn, double *__restrict__ z) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    z[c] = a * u[c] + g * nn[c];
}

__global__ void k_fcg_jacobi(size_t n, const double *__restrict__ rdiag,
                             const double *__restrict__ r, double *__restrict__ u) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    u[c] = rdiag[c] * r[c];
}

// Directions kept by the truncated orthogonalisation (FFPM_FCG_M, 1..8).
int fcg_depth() {
  const char *e = std::getenv("FFPM_FCG_M");
  int m = e ? std::atoi(e) : kFcgMax;
  return m < 1 ? 1 : (m > kFcgMax ? kFcgMax : m);
}

int fcg_iterate(Frame &f, const FcgArgs &a, double tolerance, int max_sweeps,
                int check_interval, int *sweeps, double *residual) {
  const int ci = check_interval > 1 ? check_interval : 1;
  const int eb = elementwise_blocks(f.n);
  const int m = fcg_depth(), ring = m + 1;
  const size_t bytes = sizeof(double) * f.n;
  int st = 0;
  Workspace &ws = workspace();
  float *src = (float *)ws.get(S_FCG_SRC, sizeof(float) * f.n, &st);
  if (!src) return st;
  double *u = (double *)ws.get(S_FCG_U, bytes, &st);
  if (!u) return st;
  double *au = (double *)ws.get(S_FCG_AU, bytes, &st);
  if (!au) return st;
  double *an = (double *)ws.get(S_FCG_AN, bytes, &st);
  if (!an) return st;
  double *nn = (double *)ws.get(S_FCG_N, bytes, &st);
  if (!nn) return st;
  double *pr_ring = (double *)ws.get(S_FCG_RING, 2 * ring * bytes, &st);
  if (!pr_ring) return st;
  std::vector<double *> P(ring), Q(ring);
  for (int j = 0; j < ring; ++j) {
    P[j] = pr_ring + (size_t)j * f.n;
    Q[j] = pr_ring + (size_t)(ring + j) * f.n;
  }
  std::vector<double> pq(ring, 0.0);
  double *z = f.z;  // r in f.r, x in f.phi
  double res;
  if ((st = true_residual(f, false, &res))) return st;
  double best = res;
  int stalled = 0, k = 0;
  *a.fallback_at = -1;
  auto sums = [&](int nslots, int nblocks, double *out) -> int {
    finish(f, FIN_SUMS, nslots, nblocks);
    Scalars h;
    int e = read_scalars(f, &h);
    for (int s2 = 0; s2 < nslots; ++s2) out[s2] = h.sums[s2];
    return e;
  };
  auto spmv = [&](const double *x, double *y, double *xy) -> int {
    k_spmv_dot<<<f.tiling.blocks, dim3(kTileX, f.tiling.ty), 0, stream()>>>(
        f.tiling, f.g, x, y, f.diag, f.partials, f.sc);
    FFPM_LAUNCHED("k_spmv_dot");
    return sums(1, f.tiling.blocks, xy);
  };
  auto dot2 = [&](const double *x1, const double *y1, const double *x2, const double *y2,
                  double *out) -> int {
    k_dot2<<<eb, kThreads, 0, stream()>>>(f.n, x1, y1, x2, y2, f.partials);
    FFPM_LAUNCHED("k_dot2");
    return sums(2, eb, out);
  };
  while (*sweeps < max_sweeps && res > tolerance) {
    // u = M_D^-1 r and n = sigma net(r / (V sigma))
    k_fcg_jacobi<<<eb, kThreads, 0, stream()>>>(f.n, f.rdiag, f.r, u);
    k_fcg_absmax<<<eb, kThreads, 0, stream()>>>(f.g, f.r, f.diag, f.partials);
    FFPM_LAUNCHED("k_fcg_absmax");
    finish(f, FIN_MAX, 1, eb);  // sigma -> sc->res
    k_fcg_source<<<eb, kThreads, 0, stream()>>>(f.g, f.r, f.diag, f.sc, src);
    FFPM_LAUNCHED("k_fcg_source");
    const float *out = nullptr;
    if ((st = unet_apply(*a.net, f.nx, f.ny, f.nz, a.features, src, &out))) return st;
    k_fcg_correction<<<eb, kThreads, 0, stream()>>>(f.n, out, f.diag, f.sc, nn);
    FFPM_LAUNCHED("k_fcg_correction");
    // Galerkin weights of the two terms
    double uau, nan_, v[2], w[2];
    if ((st = spmv(u, au, &uau))) return st;
    if ((st = spmv(nn, an, &nan_))) return st;
    if ((st = dot2(u, an, u, f.r, v))) return st;   // u.An, u.r
    if ((st = dot2(nn, f.r, nn, nn, w))) return st;  // n.r
    const double uan = v[0], ur = v[1], nr = w[0];
    double ca = uau > 0 ? ur / uau : 0.0, cg = 0.0;
    const double det = uau * nan_ - uan * uan;
    if (nan_ > 0 && det > 1e-12 * uau * nan_) {
      ca = (nan_ * ur - uan * nr) / det;
      cg = (uau * nr - uan * ur) / det;
    }
    k_fcg_combine<<<eb, kThreads, 0, stream()>>>(f.n, ca, u, cg, nn, z);
    FFPM_LAUNCHED("k_fcg_combine");
    // orthogonalise against the last min(k, m) directions (A-inner product)
    Combo cp{}, cq{};
    const int used = k < m ? k : m;
    for (int t = 0; t < used; t += 2) {
      const int j0 = (k - 1 - t) % ring, j1 = (k - 2 - t + ring) % ring;
      double zq[2] = {0.0, 0.0};
      if ((st = dot2(z, Q[j0], z, t + 1 < used ? Q[j1] : Q[j0], zq))) return st;
      for (int h = 0; h < 2 && t + h < used; ++h) {
        const int j = h ? j1 : j0;
        const double c = pq[j] > 0 ? zq[h] / pq[j] : 0.0;
        cp.v[cp.m] = P[j];
        cp.c[cp.m++] = c;
        cq.v[cq.m] = Q[j];
        cq.c[cq.m++] = c;
      }
    }
    const int slot = k % ring;
    k_fcg_orth<<<eb, kThreads, 0, stream()>>>(f.n, z, cp, P[slot]);
    k_fcg_orth_q<<<eb, kThreads, 0, stream()>>>(f.n, ca, au, cg, an, cq, Q[slot]);
    FFPM_LAUNCHED("k_fcg_orth");
    double prq[2];
    if ((st = dot2(P[slot], f.r, P[slot], Q[slot], prq))) return st;
    const double pqv = prq[1];
    if (!(pqv > 0)) break;  // breakdown: the final check decides
    const double alpha = prq[0] / pqv;
    pq[slot] = pqv;
    k_fcg_update<<<eb, kThreads, 0, stream()>>>(f.n, alpha, P[slot], Q[slot], f.phi, f.r);
    FFPM_LAUNCHED("k_fcg_update");
    *sweeps += 1;
    k += 1;
    if (k % ci == 0 || *sweeps == max_sweeps) {
      if ((st = true_residual(f, false, &res))) return st;
      if (res < 0.99 * best) {
        best = res;
        stalled = 0;
      } else if (++stalled >= 3) {
        break;
      }
      if (!(res == res)) break;
    }
  }
  if ((st = true_residual(f, false, &res))) return st;
  if (*sweeps < max_sweeps && res > tolerance) {
    // fall back to diagonal PCG from the current phi
    *a.fallback_at = k;
    int more = 0;
    if ((st = pcg_iterate_fused(f, tolerance, max_sweeps - *sweeps, check_interval, &more,
                                &res)))
      return st;
    *sweeps += more;
  }
  *residual = res;
  return OK;
}

struct MgStats {
  int levels = 0;
  int checks = 0;
};
thread_local MgStats g_mg;

// Flexible conjugate gradients with one stored direction (Notay's FCG(1)) in
// FP64, preconditioned by the FP32 multigrid cycle: z = B r;
// beta = -z.A p_old / p_old.A p_old; p = z + beta p_old; alpha = z.r / p.A p;
// x += alpha p; r -= alpha A p. (The cycle is not exactly symmetric: the
// Fletcher-Reeves beta of plain PCG stagnates with it on strongly stretched
// grids.) The direction and SpMV run in the fused PCG direction kernel
// (p = z + beta p_old, q = A p, p.q), the update in the fused PCG update body,
// which also forms the cycle's first smoothing step and stores q = A p in FP16
// (times a power of two set from a bound on |A| max |p|) for the next beta. Every
// `check` iterations the FP64 projected true residual replaces r
// (src/fpm_stairstep_mpi_3d.f90:297-329) and the max-norm residual is tested
// against the tolerance, as in the reference's PCG (formed only when the
// recursive residual is near the tolerance, see below); a breakdown restarts
// once from p = z, and ten checks without a 1% gain stop at the roundoff
// floor.
// The learned coarse correction of the multigrid CG (no reference
// counterpart): the network, its features on the level and the level.
struct MgLearnedArgs {
  UNet *net;
  const float *features;
  int level;
  int mode;          // 0 replace the recursion below the level, 1 augment it
  int *fallback_at;  // iterations before the switch to diagonal PCG, or -1
};

int fcg_mg_iterate(Frame &f, double tolerance, int max_sweeps, int check_interval, int *sweeps,
                   double *residual, const MgLearnedArgs *lrn = nullptr) {
  int st = 0;
  GeomT<float> g32;
  if ((st = make_geom32(f, &g32))) return st;
  const GeomT<double> g = geom_t(f.g);
  const int cl = cells_per_lane<double>(f.nx), cl32 = cells_per_lane<float>(f.nx);
  Multigrid mg;
  const Tiling tm = make_tiling(f.nx, f.ny, f.nz, 0, mg_rows(), kTileX * cl32);
  if ((st = mg_setup(mg, g, g32, f.code, tm, cl32, f.partials))) return st;
  if ((st = mg_anorm(mg, f.sc))) return st;
  if (lrn && (st = mg_learned_setup(mg, lrn->net, lrn->features, lrn->level, lrn->mode)))
    return st;
  if (lrn) *lrn->fallback_at = -1;
  // the learned cycle is not a fixed linear map; three convergence checks
  // without a 1% gain hand the solve to diagonal PCG
  const int stall_limit = lrn ? 3 : 10;
  g_mg = MgStats();
  g_mg.levels = mg.levels;
  const int ci = mg_check_interval(check_interval);
  // FFPM_MG_DEFER_X=0: x += alpha p at every iteration
  const char *dx_env = std::getenv("FFPM_MG_DEFER_X");
  const bool defer_x = !(dx_env && std::strcmp(dx_env, "0") == 0);
  // the directions, stored in FP32 (in the FP64 buffers' memory)
  float *pbuf[2] = {reinterpret_cast<float *>(f.next), reinterpret_cast<float *>(f.q)};
  int cur = 0, mode = kDirReset;
  bool first_ready = false;  // the cycle's first step holds sig w D^-1 r for the current r
  // The true residual's second pass (projection of r) runs inside the next
  // first smoothing step: its z and r.z are not used here, and the projected
  // r is needed only by that step and what follows it.
  auto check = [&](double *res_out) -> int {
    Scalars h;
    int s2 = true_residual_pass1(f);
    if (s2 || (s2 = read_scalars(f, &h))) return s2;
    *res_out = h.res;
    first_ready = false;
    return OK;
  };
  double res;
  if ((st = check(&res))) return st;
  double best = res;
  int it = 0, restarts = 0, stalled = 0, last_check = 0;
  while (*sweeps < max_sweeps && res > tolerance) {
    int n = ci - it % ci;
    if (max_sweeps - *sweeps < n) n = max_sweeps - *sweeps;
    k_chunk_begin<<<1, 64, 0, stream()>>>(f.sc);
    FFPM_LAUNCHED("k_chunk_begin");
    for (int m = 0; m < n; ++m) {
      if (!first_ready && (st = mg_first(mg, f.r, f.sc))) return st;
      if ((st = mg_vcycle(mg, mode == kDirUpdate))) return st;
      k_finish<<<1, 256, 0, stream()>>>(FIN_FCG1, 3, f.partials, mg.tl.blocks, f.sc);
      FFPM_LAUNCHED("k_finish");
      const int dst = cur ^ 1;
      // the direction's streams are FP32: it runs on the FP32 tiling
      if ((st = mg_direction(mg, mg.tld, mg.cld, pbuf[cur], pbuf[dst], mode, f.sc))) return st;
      if ((st = finish(f, FIN_PQ, 1, mg.tld.blocks))) return st;
      cur = dst;
      // the chunk's last update also estimates the max-norm residual from
      // the recursively updated r
      const bool last = m == n - 1;
      // x moves every second iteration: the even iterations of a chunk defer
      // their step to the odd ones (whose direction's buffer leaves the
      // deferred one intact), so x is current at every chunk end
      const int xm = !defer_x ? XM_NOW : (m & 1) ? XM_TWO : last ? XM_NOW : XM_DEFER;
      if ((st = mg_update(mg, f.tl[cl], cl, pbuf[cur], f.phi, f.r, f.sc,
                          last ? f.partials : nullptr, xm, pbuf[cur ^ 1])))
        return st;
      if (last) {
        k_finish<<<1, 256, 0, stream()>>>(FIN_MAX, 1, f.partials, f.tl[cl].blocks, f.sc);
        FFPM_LAUNCHED("k_finish");
      }
      first_ready = true;
      mode = kDirUpdate;
    }
    Scalars h;
    if ((st = read_scalars(f, &h))) return st;
    *sweeps += h.iters;
    it += h.iters;
    if (h.iters > 0) restarts = 0;
    if (h.status != 0) {
      // breakdown (roundoff floor or lost conjugacy): restart once from p = z
      restarts = restarts + 1;
      if (restarts > 1) break;
      k_clear_status<<<1, 64, 0, stream()>>>(f.sc);
      FFPM_LAUNCHED("k_clear_status");
      if ((st = check(&res))) return st;
      mode = kDirReset;
      continue;
    }
    // The FP64 true residual (the stopping test) is formed when the
    // recursive estimate is within 4x of the tolerance, at the iteration cap,
    // and at least every 20 iterations (which also refreshes r).
    const bool near = !(h.res > 4.0 * tolerance);
    if ((it % ci == 0 && (near || it - last_check >= 20)) || *sweeps == max_sweeps) {
      last_check = it;
      if ((st = check(&res))) return st;
      g_mg.checks += 1;
      if (res < 0.99 * best) {
        best = res;
        stalled = 0;
      } else {
        stalled = stalled + 1;
      }
      if (stalled >= stall_limit || !(res == res)) break;
    }
  }
  if (lrn && *sweeps < max_sweeps && res > tolerance) {
    // fall back to diagonal PCG from the current phi (its first step forms
    // the true residual afresh)
    *lrn->fallback_at = *sweeps;
    int more = 0;
    if ((st = pcg_iterate_fused(f, tolerance, max_sweeps - *sweeps, check_interval, &more,
                                &res)))
      return st;
    *sweeps += more;
  }
  *residual = res;
  return OK;
}

// The whole solve on device pointers (src/fpm_stairstep_mpi_3d.f90:106-330).
int solve_stairstep_dev(double *phi_ref, const double *rhs, const int32_t *solid,
                        const double *bg, const double *xlo, const double *xhi,
                        const double *ylo, const double *yhi, const double *zlo,
                        const double *zhi, const double *xc, const double *yc, const double *zc,
                        const double *dx, const double *dy, const double *dz, int nx, int ny,
                        int nz, int west, int east, int south, int north, double tolerance,
                        int max_sweeps, int check_interval, int method, int *sweeps,
                        double *residual, const FcgArgs *fcg = nullptr,
                        const MgLearnedArgs *lrn = nullptr) {
  if (method != FFPM_METHOD_PCG && method != FFPM_METHOD_JACOBI &&
      method != FFPM_METHOD_PCG_MIXED && method != FFPM_METHOD_FCG_MG &&
      !(method == FFPM_METHOD_FCG && fcg)) {
    set_error(ERR_METHOD, "stair-step method must be pcg or jacobi");
    return ERR_METHOD;
  }
  if (nx < 1 || ny < 1 || nz < 2) {
    set_error(ERR_SHAPE, "stair-step solve needs nz>=2 and nonempty ranks");
    return ERR_SHAPE;
  }
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  if (!sweeps || !residual || !bg) {
    set_error(ERR_ARGUMENT, "null output or body_grad pointer");
    return ERR_ARGUMENT;
  }
  if (!all_present({phi_ref, rhs, solid, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz}))
    return null_error();
  const bool pcg = method == FFPM_METHOD_PCG || method == FFPM_METHOD_PCG_MIXED ||
                   method == FFPM_METHOD_FCG || method == FFPM_METHOD_FCG_MG;
  TraceRange trace("stair-step setup");
  Frame f;
  int st = frame_setup(f, nx, ny, nz, xc, yc, zc, dx, dy, dz, true, pcg);
  if (st) return st;
  *sweeps = 0;
  *residual = DBL_MAX;
  const int eb = elementwise_blocks(f.n);
  FFPM_HIP(hipMemsetAsync(f.isolated, 0, sizeof(int), stream()));
  // the multigrid CG reads no rdiag (unless its learned variant may hand the
  // solve to diagonal PCG)
  double *rdiag = method == FFPM_METHOD_FCG_MG && !lrn ? nullptr : f.rdiag;
  k_build_operator<<<eb, kThreads, 0, stream()>>>(f.g, solid, rhs, xlo, xhi, ylo, yhi, zlo, zhi,
                                                  bg[0], bg[1], bg[2], pcg ? 1 : 0, f.src, f.diag,
                                                  rdiag, f.code, f.isolated);
  FFPM_LAUNCHED("k_build_operator");
  int isolated = 0;
  FFPM_HIP(hipMemcpyAsync(&isolated, f.isolated, sizeof(int), hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  if (isolated > 0) {
    set_error(ERR_ISOLATED, "stair-step fluid cell has no fluid neighbour");
    return ERR_ISOLATED;
  }
  k_pack_phi<<<eb, kThreads, 0, stream()>>>(nx, ny, nz, solid, phi_ref, f.phi);
  FFPM_LAUNCHED("k_pack_phi");
  double *cur = f.phi;
  g_mixed = MixedStats();
  trace.next("stair-step iterations");
  if (method == FFPM_METHOD_FCG) {
    st = fcg_iterate(f, *fcg, tolerance, max_sweeps, check_interval, sweeps, residual);
  } else if (method == FFPM_METHOD_FCG_MG) {
    st = fcg_mg_iterate(f, tolerance, max_sweeps, check_interval, sweeps, residual, lrn);
  } else if (method == FFPM_METHOD_PCG_MIXED) {
    st = pcg_mixed(f, tolerance, max_sweeps, check_interval, mixed_inner_rtol(), sweeps,
                   residual);
  } else if (pcg) {
    st = classic_pcg()
             ? pcg_iterate(f, tolerance, max_sweeps, check_interval, sweeps, residual)
             : pcg_iterate_fused(f, tolerance, max_sweeps, check_interval, sweeps, residual);
  } else {
    st = jacobi_iterate(f, tolerance, max_sweeps, check_interval, sweeps, residual, &cur);
  }
  if (st) return st;
  // gauge (src/fpm_stairstep_mpi_3d.f90:170-187)
  trace.next("stair-step gauge");
  k_gauge_sums<<<eb, kThreads, 0, stream()>>>(f.g, solid, cur, f.partials, f.sc);
  FFPM_LAUNCHED("k_gauge_sums");
  finish(f, FIN_SUMS, 2, eb);
  const int pb = elementwise_blocks((size_t)(nx + 2) * (ny + 2) * nz);
  k_gauge_apply<<<pb, kThreads, 0, stream()>>>(nx, ny, nz, solid, cur, phi_ref, f.sc);
  FFPM_LAUNCHED("k_gauge_apply");
  return residual_max(f, cur, residual);
}

int compat_dev(const double *rhs, const int32_t *solid, const double *bg, double *xlo,
               double *xhi, double *ylo, double *yhi, double *zlo, double *zhi,
               const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
               int west, int east, int south, int north, double *correction,
               double *imbalance) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  if (!all_present({rhs, solid, bg, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz, correction,
                    imbalance}))
    return null_error();
  Frame f;
  int st = frame_setup(f, nx, ny, nz, nullptr, nullptr, nullptr, dx, dy, dz, false, false);
  if (st) return st;
  const int eb = elementwise_blocks(f.n);
  k_compat_sums<<<eb, kThreads, 0, stream()>>>(f.g, solid, rhs, xlo, xhi, ylo, yhi, zlo, zhi,
                                               bg[0], bg[1], bg[2], f.partials, f.sc);
  FFPM_LAUNCHED("k_compat_sums");
  finish(f, FIN_SUMS, 3, eb);
  Scalars h;
  if ((st = read_scalars(f, &h))) return st;
  const double source = h.sums[0], flux = h.sums[1], area = h.sums[2];
  if (area <= 0) {
    set_error(ERR_NO_OUTER_FACE, "stair-step compatibility needs at least one fluid outer face");
    return ERR_NO_OUTER_FACE;
  }
  const double c = -(flux - source) / area;
  const size_t nf = (size_t)ny * nz + (size_t)nx * nz + (size_t)nx * ny;
  k_shift_faces<<<elementwise_blocks(nf), kThreads, 0, stream()>>>(nx, ny, nz, xlo, xhi, ylo,
                                                                   yhi, zlo, zhi, c);
  FFPM_LAUNCHED("k_shift_faces");
  // the flux after the correction (the source integral is kept)
  k_compat_sums<<<eb, kThreads, 0, stream()>>>(f.g, solid, nullptr, xlo, xhi, ylo, yhi, zlo,
                                               zhi, bg[0], bg[1], bg[2], f.partials, f.sc);
  FFPM_LAUNCHED("k_compat_sums");
  finish(f, FIN_SUMS, 3, eb);
  if ((st = read_scalars(f, &h))) return st;
  *correction = c;
  // the reference keeps the first source integral and recomputes only the flux
  *imbalance = h.sums[1] - source;
  return OK;
}

int residual_face_dev(const double *phi, const double *rhs, const int32_t *solid,
                      const double *bg, const double *xlo, const double *xhi, const double *ylo,
                      const double *yhi, const double *zlo, const double *zhi,
                      const double *xc, const double *yc, const double *zc, const double *dx,
                      const double *dy, const double *dz, int nx, int ny, int nz, int west,
                      int east, int south, int north, double *value) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  if (!all_present({phi, rhs, solid, bg, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy, dz,
                    value}))
    return null_error();
  Frame f;
  int st = frame_setup(f, nx, ny, nz, xc, yc, zc, dx, dy, dz, false, false);
  if (st) return st;
  k_face_residual<<<elementwise_blocks(f.n), kThreads, 0, stream()>>>(
      f.g, phi, rhs, solid, xlo, xhi, ylo, yhi, zlo, zhi, bg[0], bg[1], bg[2], f.partials, f.sc);
  FFPM_LAUNCHED("k_face_residual");
  finish(f, FIN_MAX, 1, elementwise_blocks(f.n));
  Scalars h;
  if ((st = read_scalars(f, &h))) return st;
  *value = h.res;
  return OK;
}

int terms_dev(const double *phi, const double *u, const double *v, const double *w,
              const double *ox, const double *oy, const double *oz, const double *ax,
              const double *ay, const double *az, const int32_t *solid, const double *bg,
              const double *xc, const double *yc, const double *zc, const double *dx,
              const double *dy, const double *dz, const double *xlo, const double *xhi,
              const double *ylo, const double *yhi, const double *zlo, const double *zhi,
              const double *bounds, double *gx, double *gy, double *gz, double *vortex,
              double *advection, double *totals, int nx, int ny, int nz, int west, int east,
              int south, int north) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  if (!all_present({phi, u, v, w, ox, oy, oz, ax, ay, az, solid, bg, xc, yc, zc, dx, dy, dz, xlo,
                    xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex, advection, totals}))
    return null_error();
  TraceRange trace("stair-step FPM terms");
  Frame f;
  int st = frame_setup(f, nx, ny, nz, xc, yc, zc, dx, dy, dz, false, false);
  if (st) return st;
  Bounds bd;
  for (int i = 0; i < 6; ++i) bd.b[i] = bounds[i];
  {
    TimedLaunch tl("terms");
    k_terms<<<elementwise_blocks(f.n), kThreads, 0, stream()>>>(
        f.g, phi, solid, u, v, w, ox, oy, oz, ax, ay, az, xlo, xhi, ylo, yhi, zlo, zhi, bg[0],
        bg[1], bg[2], bd, gx, gy, gz, vortex, advection, f.partials, f.sc);
  }
  FFPM_LAUNCHED("k_terms");
  finish(f, FIN_SUMS, 2, elementwise_blocks(f.n));
  Scalars h;
  if ((st = read_scalars(f, &h))) return st;
  totals[0] = h.sums[0];
  totals[1] = h.sums[1];
  return OK;
}

// compute_stairstep_fpm_terms_3d for nd directions (1-3) of one frame in one
// pass over the flow fields; per-direction arguments are host arrays of
// device pointers, body_grad holds 3 nd values and totals 2 nd.
// Checks the arguments of the nd-direction densities and gathers the
// per-direction pointers and body gradients.
int pack_terms_dirs(int nd, const double *const *phi, const double *u, const double *v,
                    const double *w, const double *ox, const double *oy, const double *oz,
                    const double *ax, const double *ay, const double *az, const int32_t *solid,
                    const double *bg, const double *xc, const double *yc, const double *zc,
                    const double *dx, const double *dy, const double *dz,
                    const double *const *xlo, const double *const *xhi,
                    const double *const *ylo, const double *const *yhi,
                    const double *const *zlo, const double *const *zhi, const double *bounds,
                    double *const *gx, double *const *gy, double *const *gz,
                    double *const *vortex, double *const *advection, double *totals, int nx,
                    int ny, int nz, int west, int east, int south, int north, TermsDirs &a) {
  if (nd < 1 || nd > 3) {
    set_error(ERR_ARGUMENT, "densities: 1 to 3 directions");
    return ERR_ARGUMENT;
  }
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  if (!all_present({phi, u, v, w, ox, oy, oz, ax, ay, az, solid, bg, xc, yc, zc, dx, dy, dz, xlo,
                    xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex, advection, totals}))
    return null_error();
  a = TermsDirs{};
  for (int d = 0; d < nd; ++d) {
    if (!all_present({phi[d], xlo[d], xhi[d], ylo[d], yhi[d], zlo[d], zhi[d], gx[d], gy[d],
                      gz[d], vortex[d], advection[d]}))
      return null_error();
    a.phi[d] = phi[d];
    a.xlo[d] = xlo[d];
    a.xhi[d] = xhi[d];
    a.ylo[d] = ylo[d];
    a.yhi[d] = yhi[d];
    a.zlo[d] = zlo[d];
    a.zhi[d] = zhi[d];
    for (int e = 0; e < 3; ++e) a.bg[d][e] = bg[3 * d + e];
    a.gx[d] = gx[d];
    a.gy[d] = gy[d];
    a.gz[d] = gz[d];
    a.vortex[d] = vortex[d];
    a.advection[d] = advection[d];
  }
  return OK;
}

// The densities pass and its integral finish: per-workgroup partials (6 per
// workgroup of elementwise_blocks(n)) in `part`, the 2 nd totals in `out`
// (device), copied to the host array `totals`.
int launch_terms_dirs(int nd, const TermsDirs &a, const Geom &g, size_t n, const int32_t *solid,
                      const double *u, const double *v, const double *w, const double *ox,
                      const double *oy, const double *oz, const double *ax, const double *ay,
                      const double *az, const double *bounds, double *part, double *out,
                      double *totals, bool timed) {
  TraceRange trace("stair-step FPM terms");
  Bounds bd;
  for (int i = 0; i < 6; ++i) bd.b[i] = bounds[i];
  const int eb = elementwise_blocks(n);
  auto go = [&](auto kern) {
    kern<<<eb, kThreads, 0, stream()>>>(g, a, solid, u, v, w, ox, oy, oz, ax, ay, az, bd, part);
  };
  auto launch = [&]() {
    if (nd == 3) go(k_terms_dirs<3>);
    else if (nd == 2) go(k_terms_dirs<2>);
    else go(k_terms_dirs<1>);
  };
  if (timed) {
    TimedLaunch tl("terms_dirs");
    launch();
  } else {
    launch();
  }
  FFPM_LAUNCHED("k_terms_dirs");
  k_terms_dirs_finish<<<1, 256, 0, stream()>>>(nd, part, eb, out);
  FFPM_LAUNCHED("k_terms_dirs_finish");
  FFPM_HIP(hipMemcpyAsync(totals, out, sizeof(double) * 2 * nd, hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  return OK;
}

int terms_dirs_dev(int nd, const double *const *phi, const double *u, const double *v,
                   const double *w, const double *ox, const double *oy, const double *oz,
                   const double *ax, const double *ay, const double *az, const int32_t *solid,
                   const double *bg, const double *xc, const double *yc, const double *zc,
                   const double *dx, const double *dy, const double *dz,
                   const double *const *xlo, const double *const *xhi,
                   const double *const *ylo, const double *const *yhi,
                   const double *const *zlo, const double *const *zhi, const double *bounds,
                   double *const *gx, double *const *gy, double *const *gz,
                   double *const *vortex, double *const *advection, double *totals, int nx,
                   int ny, int nz, int west, int east, int south, int north) {
  TermsDirs a;
  int st = pack_terms_dirs(nd, phi, u, v, w, ox, oy, oz, ax, ay, az, solid, bg, xc, yc, zc, dx,
                           dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex,
                           advection, totals, nx, ny, nz, west, east, south, north, a);
  if (st) return st;
  Frame f;
  if ((st = frame_setup(f, nx, ny, nz, xc, yc, zc, dx, dy, dz, false, false))) return st;
  const int eb = elementwise_blocks(f.n);
  double *part = (double *)workspace().get(S_TD_PART, sizeof(double) * 6 * eb, &st);
  if (!part) return st;
  double *out = (double *)workspace().get(S_TD_OUT, sizeof(double) * 6, &st);
  if (!out) return st;
  return launch_terms_dirs(nd, a, f.g, f.n, solid, u, v, w, ox, oy, oz, ax, ay, az, bounds, part,
                           out, totals, true);
}

}  // namespace ffpm

// ---- densities without allocation --------------------------------------------
// The scratch memory of ffpm_compute_stairstep_fpm_terms_3d_dirs_noalloc_dev:
// the 1-D stencil coefficients (2 (nx+ny+nz) doubles), the per-workgroup
// partials of the integrals (6 per workgroup, at most 8192 workgroups) and the
// 6 totals. Each GPU holds its own copy at its own address; host code uses only
// the address hipGetSymbolAddress returns for the current device.
namespace {
constexpr int kNoAllocMaxAxisCells = 16384;  // nx + ny + nz
constexpr int kNoAllocMaxBlocks = 8192;      // the cap of ffpm::elementwise_blocks
constexpr size_t kNoAllocCoef = 2 * (size_t)kNoAllocMaxAxisCells;
constexpr size_t kNoAllocPart = 6 * (size_t)kNoAllocMaxBlocks;
constexpr size_t kNoAllocDoubles = kNoAllocCoef + kNoAllocPart + 32;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace ffpm {
namespace {

std::mutex g_scratch_mutex;
double *g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device.
int scratch_on_current_device(double **ptr) {
  int dev = 0;
  FFPM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_error(ERR_DEVICE, "device index beyond the library's scratch table");
    return ERR_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                hipGetErrorString(e));
      return ERR_DEVICE;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return OK;
}

int terms_dirs_noalloc_dev(int nd, const double *const *phi, const double *u, const double *v,
                           const double *w, const double *ox, const double *oy,
                           const double *oz, const double *ax, const double *ay,
                           const double *az, const int32_t *solid, const double *bg,
                           const double *xc, const double *yc, const double *zc,
                           const double *dx, const double *dy, const double *dz,
                           const double *const *xlo, const double *const *xhi,
                           const double *const *ylo, const double *const *yhi,
                           const double *const *zlo, const double *const *zhi,
                           const double *bounds, double *const *gx, double *const *gy,
                           double *const *gz, double *const *vortex,
                           double *const *advection, double *totals, int nx, int ny, int nz,
                           int west, int east, int south, int north) {
  TermsDirs a;
  int st = pack_terms_dirs(nd, phi, u, v, w, ox, oy, oz, ax, ay, az, solid, bg, xc, yc, zc, dx,
                           dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex,
                           advection, totals, nx, ny, nz, west, east, south, north, a);
  if (st) return st;
  if ((long long)nx + ny + nz > kNoAllocMaxAxisCells) {
    set_error(ERR_ARGUMENT, "densities without allocation: nx+ny+nz exceeds 16384, the size of "
                            "the library's fixed scratch memory");
    return ERR_ARGUMENT;
  }
  double *base = nullptr;
  if ((st = scratch_on_current_device(&base))) return st;
  double *cw = base, *ce = cw + nx, *cs = ce + nx, *cn = cs + ny, *cb = cn + ny, *ct = cb + nz;
  double *part = base + kNoAllocCoef;
  double *out = part + kNoAllocPart;
  const int tot = nx + ny + nz;
  k_axis_coefficients<<<(tot + 255) / 256, 256, 0, stream()>>>(nx, ny, nz, xc, yc, zc, dx, dy, dz,
                                                               cw, ce, cs, cn, cb, ct);
  FFPM_LAUNCHED("k_axis_coefficients");
  Geom g{};
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.cw = cw; g.ce = ce; g.cs = cs; g.cn = cn; g.cb = cb; g.ct = ct;
  g.dx = dx; g.dy = dy; g.dz = dz;
  g.xc = xc; g.yc = yc; g.zc = zc;
  return launch_terms_dirs(nd, a, g, (size_t)nx * ny * nz, solid, u, v, w, ox, oy, oz, ax, ay,
                           az, bounds, part, out, totals, false);
}

}  // namespace

// ---- host-pointer staging ---------------------------------------------------
namespace {

// A host array bound to a frame (ffpm_frame_bind) is already resident: its
// device copy is used in place and nothing moves, as inside the caller's
// OpenACC data region (src/fpm_stairstep_mpi_3d.f90:145-148).
template <class T>
int stage_in(int slot, const T *host, size_t count, T **dev) {
  int st = 0;
  if ((*dev = (T *)present_device(host, sizeof(T) * count, &st))) return OK;
  if (st) return st;
  *dev = (T *)workspace().get(slot, sizeof(T) * count, &st);
  if (!*dev) return st;
  FFPM_HIP(hipMemcpyAsync(*dev, host, sizeof(T) * count, hipMemcpyHostToDevice, stream()));
  count_h2d(sizeof(T) * count);
  return OK;
}

template <class T>
int stage_out(T *host, const T *dev, size_t count) {
  int st = 0;
  if (T *p = (T *)present_device(host, sizeof(T) * count, &st)) {
    if (p != dev)
      FFPM_HIP(hipMemcpyAsync(p, dev, sizeof(T) * count, hipMemcpyDeviceToDevice, stream()));
    return OK;
  }
  if (st) return st;
  FFPM_HIP(hipMemcpyAsync(host, dev, sizeof(T) * count, hipMemcpyDeviceToHost, stream()));
  count_d2h(sizeof(T) * count);
  return OK;
}

// an output array: its bound device copy, else a workspace slot
int alloc_out(int slot, const double *host, size_t count, double **dev) {
  int st = 0;
  if ((*dev = (double *)present_device(host, sizeof(double) * count, &st))) return OK;
  if (st) return st;
  *dev = (double *)workspace().get(slot, sizeof(double) * count, &st);
  return *dev ? OK : st;
}

struct Staged {
  double *xlo, *xhi, *ylo, *yhi, *zlo, *zhi;
  double *xc, *yc, *zc, *dx, *dy, *dz;
};

int stage_faces(int nx, int ny, int nz, const double *xlo, const double *xhi, const double *ylo,
                const double *yhi, const double *zlo, const double *zhi, Staged &s) {
  int st;
  const size_t nyz = (size_t)ny * nz, nxz = (size_t)nx * nz, nxy = (size_t)nx * ny;
  if ((st = stage_in(S_U_XLO, xlo, nyz, &s.xlo))) return st;
  if ((st = stage_in(S_U_XHI, xhi, nyz, &s.xhi))) return st;
  if ((st = stage_in(S_U_YLO, ylo, nxz, &s.ylo))) return st;
  if ((st = stage_in(S_U_YHI, yhi, nxz, &s.yhi))) return st;
  if ((st = stage_in(S_U_ZLO, zlo, nxy, &s.zlo))) return st;
  if ((st = stage_in(S_U_ZHI, zhi, nxy, &s.zhi))) return st;
  return OK;
}

int stage_axes(int nx, int ny, int nz, const double *xc, const double *yc, const double *zc,
               const double *dx, const double *dy, const double *dz, Staged &s) {
  int st;
  if (xc && (st = stage_in(S_U_XC, xc, (size_t)nx + 2, &s.xc))) return st;
  if (yc && (st = stage_in(S_U_YC, yc, (size_t)ny + 2, &s.yc))) return st;
  if (zc && (st = stage_in(S_U_ZC, zc, (size_t)nz, &s.zc))) return st;
  if ((st = stage_in(S_U_DX, dx, (size_t)nx, &s.dx))) return st;
  if ((st = stage_in(S_U_DY, dy, (size_t)ny, &s.dy))) return st;
  if ((st = stage_in(S_U_DZ, dz, (size_t)nz, &s.dz))) return st;
  return OK;
}

int sync() {
  FFPM_HIP(hipStreamSynchronize(stream()));
  return OK;
}

bool bad_dims(int nx, int ny, int nz) { return nx < 1 || ny < 1 || nz < 1; }

int dims_error() {
  set_error(ERR_SHAPE, "empty grid");
  return ERR_SHAPE;
}

// Physical-boundary halo of the solid flags (src/fpm_stairstep_mpi_3d.f90:29-32).
__global__ void k_zero_solid_halo(int nx, int ny, int nz, int32_t *solid) {
  const size_t sx = (size_t)nx + 2, sy = (size_t)ny + 2;
  const size_t n = sx * sy * nz;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const size_t I = p % sx, J = (p / sx) % sy;
    if (I == 0 || I == sx - 1 || J == 0 || J == sy - 1) solid[p] = 0;
  }
}

}  // namespace
}  // namespace ffpm

using namespace ffpm;

extern "C" {

int ffpm_exchange_solid_halo_3d_dev(int32_t *solid, int nx, int ny, int nz, int west, int east,
                                    int south, int north);

int ffpm_last_mixed_solve(int *outer, int *inner) {
  *outer = g_mixed.outer;
  *inner = g_mixed.inner;
  return FFPM_OK;
}

int ffpm_last_mg_solve(int *levels, int *checks) {
  *levels = g_mg.levels;
  *checks = g_mg.checks;
  return FFPM_OK;
}

int ffpm_exchange_solid_halo_3d(int32_t *solid, int nx, int ny, int nz, int west, int east,
                                int south, int north) {
  // src/fpm_stairstep_mpi_3d.f90:29-32: zero the halo at physical boundaries.
  if (bad_dims(nx, ny, nz)) return dims_error();
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  const size_t sx = (size_t)nx + 2, sy = (size_t)ny + 2;
  for (int k = 0; k < nz; ++k) {
    int32_t *plane = solid + sx * sy * k;
    for (size_t j = 0; j < sy; ++j) {
      plane[sx * j] = 0;
      plane[sx * j + nx + 1] = 0;
    }
    for (size_t i = 0; i < sx; ++i) {
      plane[i] = 0;
      plane[sx * (ny + 1) + i] = 0;
    }
  }
  // a solid array bound to a frame gets the same halo in its device copy
  int st = 0;
  const size_t np = sx * sy * nz;
  if (int32_t *d = (int32_t *)present_device(solid, sizeof(int32_t) * np, &st))
    return ffpm_exchange_solid_halo_3d_dev(d, nx, ny, nz, west, east, south, north);
  return st;
}

int ffpm_exchange_solid_halo_3d_dev(int32_t *solid, int nx, int ny, int nz, int west, int east,
                                    int south, int north) {
  if (bad_dims(nx, ny, nz)) return dims_error();
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  const size_t n = ((size_t)nx + 2) * ((size_t)ny + 2) * nz;
  k_zero_solid_halo<<<elementwise_blocks(n), 256, 0, stream()>>>(nx, ny, nz, solid);
  FFPM_LAUNCHED("k_zero_solid_halo");
  return sync();
}

int ffpm_enforce_stairstep_compatibility_3d_dev(
    const double *rhs, const int32_t *solid, const double *body_grad, double *xlo, double *xhi,
    double *ylo, double *yhi, double *zlo, double *zhi, const double *dx, const double *dy,
    const double *dz, int nx, int ny, int nz, int west, int east, int south, int north,
    double *correction, double *imbalance) {
  return compat_dev(rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz, nx, ny, nz,
                    west, east, south, north, correction, imbalance);
}

int ffpm_enforce_stairstep_compatibility_3d(
    const double *rhs, const int32_t *solid, const double *body_grad, double *xlo, double *xhi,
    double *ylo, double *yhi, double *zlo, double *zhi, const double *dx, const double *dy,
    const double *dz, int nx, int ny, int nz, int west, int east, int south, int north,
    double *correction, double *imbalance) {
  if (bad_dims(nx, ny, nz)) return dims_error();
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  int st;
  const size_t n = (size_t)nx * ny * nz, np = ((size_t)nx + 2) * ((size_t)ny + 2) * nz;
  double *d_rhs;
  int32_t *d_solid;
  Staged s{};
  if ((st = stage_in(S_U_RHS, rhs, n, &d_rhs))) return st;
  if ((st = stage_in(S_U_SOLID, solid, np, &d_solid))) return st;
  if ((st = stage_faces(nx, ny, nz, xlo, xhi, ylo, yhi, zlo, zhi, s))) return st;
  if ((st = stage_axes(nx, ny, nz, nullptr, nullptr, nullptr, dx, dy, dz, s))) return st;
  st = compat_dev(d_rhs, d_solid, body_grad, s.xlo, s.xhi, s.ylo, s.yhi, s.zlo, s.zhi, s.dx,
                  s.dy, s.dz, nx, ny, nz, west, east, south, north, correction, imbalance);
  if (st) return st;
  const size_t nyz = (size_t)ny * nz, nxz = (size_t)nx * nz, nxy = (size_t)nx * ny;
  if ((st = stage_out(xlo, s.xlo, nyz)) || (st = stage_out(xhi, s.xhi, nyz)) ||
      (st = stage_out(ylo, s.ylo, nxz)) || (st = stage_out(yhi, s.yhi, nxz)) ||
      (st = stage_out(zlo, s.zlo, nxy)) || (st = stage_out(zhi, s.zhi, nxy)))
    return st;
  return sync();
}

int ffpm_solve_stairstep_poisson_3d_dev(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double tolerance, int max_sweeps, int check_interval,
    int method, int *sweeps, double *residual) {
  int st = solve_stairstep_dev(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc,
                               zc, dx, dy, dz, nx, ny, nz, west, east, south, north, tolerance,
                               max_sweeps, check_interval, method, sweeps, residual);
  if (st) return st;
  return sync();
}

int ffpm_stairstep_apply_operator_dev(const int32_t *solid, const double *xc, const double *yc,
                                      const double *zc, const double *dx, const double *dy,
                                      const double *dz, int nx, int ny, int nz, const double *x,
                                      double *y) {
  if (!all_present({solid, xc, yc, zc, dx, dy, dz, x, y})) return null_error();
  if (nx < 1 || ny < 1 || nz < 2) {
    set_error(ERR_SHAPE, "stair-step solve needs nz>=2 and nonempty ranks");
    return ERR_SHAPE;
  }
  Frame f;
  int st = frame_setup(f, nx, ny, nz, xc, yc, zc, dx, dy, dz, true, false);
  if (st) return st;
  // the operator only: zero right-hand side and face data
  Workspace &ws = workspace();
  const size_t nf = (size_t)nx * ny * nz;
  double *zero = (double *)ws.get(S_U_F8, sizeof(double) * nf, &st);
  if (!zero) return st;
  FFPM_HIP(hipMemsetAsync(zero, 0, sizeof(double) * nf, stream()));
  const int eb = elementwise_blocks(f.n);
  FFPM_HIP(hipMemsetAsync(f.isolated, 0, sizeof(int), stream()));
  k_build_operator<<<eb, kThreads, 0, stream()>>>(f.g, solid, zero, zero, zero, zero, zero, zero,
                                                  zero, 0.0, 0.0, 0.0, 1, f.src, f.diag, f.rdiag,
                                                  f.code, f.isolated);
  FFPM_LAUNCHED("k_build_operator");
  k_spmv_dot<<<f.tiling.blocks, dim3(kTileX, f.tiling.ty), 0, stream()>>>(
      f.tiling, f.g, x, y, f.diag, f.partials, f.sc);
  FFPM_LAUNCHED("k_spmv_dot");
  return sync();
}

int ffpm_solve_stairstep_fcg_dev(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double tolerance, int max_sweeps, int check_interval, ffpm_unet *net, const float *features,
    int *sweeps, double *residual, int *fallback_at) {
  if (!net || !features || !fallback_at) {
    set_error(ERR_ARGUMENT, "flexible CG needs a network, its features and an output");
    return ERR_ARGUMENT;
  }
  FcgArgs a{reinterpret_cast<UNet *>(net), features, fallback_at};
  int st = solve_stairstep_dev(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc,
                               zc, dx, dy, dz, nx, ny, nz, -1, -1, -1, -1, tolerance, max_sweeps,
                               check_interval, FFPM_METHOD_FCG, sweeps, residual, &a);
  if (st) return st;
  return sync();
}

int ffpm_solve_stairstep_fcg_mg_learned_dev(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double tolerance, int max_sweeps, int check_interval, ffpm_unet *net, int level, int mode,
    const float *features, int *sweeps, double *residual, int *fallback_at) {
  if (!net || !features || !fallback_at) {
    set_error(ERR_ARGUMENT, "learned multigrid CG needs a network, its features and an output");
    return ERR_ARGUMENT;
  }
  MgLearnedArgs a{reinterpret_cast<UNet *>(net), features, level, mode, fallback_at};
  int st = solve_stairstep_dev(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc,
                               zc, dx, dy, dz, nx, ny, nz, -1, -1, -1, -1, tolerance, max_sweeps,
                               check_interval, FFPM_METHOD_FCG_MG, sweeps, residual, nullptr, &a);
  if (st) return st;
  return sync();
}

int ffpm_solve_stairstep_poisson_3d(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double tolerance, int max_sweeps, int check_interval,
    int method, int *sweeps, double *residual) {
  if (method != FFPM_METHOD_PCG && method != FFPM_METHOD_JACOBI &&
      method != FFPM_METHOD_PCG_MIXED && method != FFPM_METHOD_FCG_MG) {
    set_error(ERR_METHOD, "stair-step method must be pcg or jacobi");
    return ERR_METHOD;
  }
  if (nx < 1 || ny < 1 || nz < 2) {
    set_error(ERR_SHAPE, "stair-step solve needs nz>=2 and nonempty ranks");
    return ERR_SHAPE;
  }
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  int st;
  const size_t n = (size_t)nx * ny * nz, np = ((size_t)nx + 2) * ((size_t)ny + 2) * nz;
  double *d_phi, *d_rhs;
  int32_t *d_solid;
  Staged s{};
  if ((st = stage_in(S_U_PHI, phi, np, &d_phi))) return st;
  if ((st = stage_in(S_U_RHS, rhs, n, &d_rhs))) return st;
  if ((st = stage_in(S_U_SOLID, solid, np, &d_solid))) return st;
  if ((st = stage_faces(nx, ny, nz, xlo, xhi, ylo, yhi, zlo, zhi, s))) return st;
  if ((st = stage_axes(nx, ny, nz, xc, yc, zc, dx, dy, dz, s))) return st;
  st = solve_stairstep_dev(d_phi, d_rhs, d_solid, body_grad, s.xlo, s.xhi, s.ylo, s.yhi, s.zlo,
                           s.zhi, s.xc, s.yc, s.zc, s.dx, s.dy, s.dz, nx, ny, nz, west, east,
                           south, north, tolerance, max_sweeps, check_interval, method, sweeps,
                           residual);
  if (st) return st;
  if ((st = stage_out(phi, d_phi, np))) return st;
  return sync();
}

int ffpm_stairstep_residual_3d_dev(
    const double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double *value) {
  return residual_face_dev(phi, rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc,
                           dx, dy, dz, nx, ny, nz, west, east, south, north, value);
}

int ffpm_stairstep_residual_3d(
    const double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double *value) {
  if (bad_dims(nx, ny, nz)) return dims_error();
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  int st;
  const size_t n = (size_t)nx * ny * nz, np = ((size_t)nx + 2) * ((size_t)ny + 2) * nz;
  double *d_phi, *d_rhs;
  int32_t *d_solid;
  Staged s{};
  if ((st = stage_in(S_U_PHI, phi, np, &d_phi))) return st;
  if ((st = stage_in(S_U_RHS, rhs, n, &d_rhs))) return st;
  if ((st = stage_in(S_U_SOLID, solid, np, &d_solid))) return st;
  if ((st = stage_faces(nx, ny, nz, xlo, xhi, ylo, yhi, zlo, zhi, s))) return st;
  if ((st = stage_axes(nx, ny, nz, xc, yc, zc, dx, dy, dz, s))) return st;
  return residual_face_dev(d_phi, d_rhs, d_solid, body_grad, s.xlo, s.xhi, s.ylo, s.yhi, s.zlo,
                           s.zhi, s.xc, s.yc, s.zc, s.dx, s.dy, s.dz, nx, ny, nz, west, east,
                           south, north, value);
}

int ffpm_compute_stairstep_fpm_terms_3d_dev(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *solid, const double *body_grad, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, const double *bounds, double *gx, double *gy, double *gz, double *vortex,
    double *advection, double *totals, int nx, int ny, int nz, int west, int east, int south,
    int north) {
  return terms_dev(phi, u, v, w, ox, oy, oz, ax, ay, az, solid, body_grad, xc, yc, zc, dx, dy, dz,
                   xlo, xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex, advection, totals,
                   nx, ny, nz, west, east, south, north);
}

int ffpm_compute_stairstep_fpm_terms_3d_dirs_dev(
    int ndir, const double *const *phi, const double *u, const double *v, const double *w,
    const double *ox, const double *oy, const double *oz, const double *ax, const double *ay,
    const double *az, const int32_t *solid, const double *body_grad, const double *xc,
    const double *yc, const double *zc, const double *dx, const double *dy, const double *dz,
    const double *const *xlo, const double *const *xhi, const double *const *ylo,
    const double *const *yhi, const double *const *zlo, const double *const *zhi,
    const double *bounds, double *const *gx, double *const *gy, double *const *gz,
    double *const *vortex, double *const *advection, double *totals, int nx, int ny, int nz,
    int west, int east, int south, int north) {
  return terms_dirs_dev(ndir, phi, u, v, w, ox, oy, oz, ax, ay, az, solid, body_grad, xc, yc, zc,
                        dx, dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex,
                        advection, totals, nx, ny, nz, west, east, south, north);
}

int ffpm_compute_stairstep_fpm_terms_3d_dirs_noalloc_dev(
    int ndir, const double *const *phi, const double *u, const double *v, const double *w,
    const double *ox, const double *oy, const double *oz, const double *ax, const double *ay,
    const double *az, const int32_t *solid, const double *body_grad, const double *xc,
    const double *yc, const double *zc, const double *dx, const double *dy, const double *dz,
    const double *const *xlo, const double *const *xhi, const double *const *ylo,
    const double *const *yhi, const double *const *zlo, const double *const *zhi,
    const double *bounds, double *const *gx, double *const *gy, double *const *gz,
    double *const *vortex, double *const *advection, double *totals, int nx, int ny, int nz,
    int west, int east, int south, int north) {
  return terms_dirs_noalloc_dev(ndir, phi, u, v, w, ox, oy, oz, ax, ay, az, solid, body_grad, xc,
                                yc, zc, dx, dy, dz, xlo, xhi, ylo, yhi, zlo, zhi, bounds, gx, gy,
                                gz, vortex, advection, totals, nx, ny, nz, west, east, south,
                                north);
}

int ffpm_compute_stairstep_fpm_terms_3d(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *solid, const double *body_grad, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, const double *bounds, double *gx, double *gy, double *gz, double *vortex,
    double *advection, double *totals, int nx, int ny, int nz, int west, int east, int south,
    int north) {
  if (bad_dims(nx, ny, nz)) return dims_error();
  if (any_neighbour(west, east, south, north)) return decomposition_error();
  int st;
  const size_t n = (size_t)nx * ny * nz, np = ((size_t)nx + 2) * ((size_t)ny + 2) * nz;
  double *d_phi, *f[9], *o[5];
  int32_t *d_solid;
  Staged s{};
  const double *fin[9] = {u, v, w, ox, oy, oz, ax, ay, az};
  if ((st = stage_in(S_U_PHI, phi, np, &d_phi))) return st;
  if ((st = stage_in(S_U_SOLID, solid, np, &d_solid))) return st;
  for (int m = 0; m < 9; ++m)
    if ((st = stage_in(S_U_F0 + m, fin[m], n, &f[m]))) return st;
  double *hout[5] = {gx, gy, gz, vortex, advection};
  for (int m = 0; m < 5; ++m)
    if ((st = alloc_out(S_U_O0 + m, hout[m], n, &o[m]))) return st;
  if ((st = stage_faces(nx, ny, nz, xlo, xhi, ylo, yhi, zlo, zhi, s))) return st;
  if ((st = stage_axes(nx, ny, nz, xc, yc, zc, dx, dy, dz, s))) return st;
  st = terms_dev(d_phi, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], d_solid, body_grad,
                 s.xc, s.yc, s.zc, s.dx, s.dy, s.dz, s.xlo, s.xhi, s.ylo, s.yhi, s.zlo, s.zhi,
                 bounds, o[0], o[1], o[2], o[3], o[4], totals, nx, ny, nz, west, east, south,
                 north);
  if (st) return st;
  for (int m = 0; m < 5; ++m)
    if ((st = stage_out(hout[m], o[m], n))) return st;
  return sync();
}

}  // extern "C"
