// Device bodies shared by the single-solve kernels and the batched kernels
// (which run many solves in one grid). Each body takes its workgroup's index
// and the workgroup count of its own solve's launch explicitly, so a batched
// workgroup computes exactly what the same workgroup of a standalone launch
// computes, and the reductions sum in the same order.
#pragma once

#include "ffpm_kernels.hpp"

namespace ffpm {

// Scale of the multigrid cycle's FP16 level-0 vectors (Scalars::sig): the
// largest power of two s with s max <= target, times the current scale cur
// (max measured in units of cur), within 2^-120..2^120; cur when max is 0
// or not finite (a non-finite output ends in a breakdown, whose restart
// starts the scale afresh). A cycle's output is the next cycle's times its
// contraction (0.1-0.4 measured), so kSigTarget = 2^10 leaves the next
// output 2^6 below the FP16 limit 65504 with its values at full precision
// (11 bits) down to 2^-14; a solve's first cycle is scaled from its first
// step x1 (FIN_TR1), whose output exceeds x1 by at most the condition of
// D^-1 A (20-45 measured on stretched frames, at most a few times 10^4 at
// 512 cells per axis): kSigFirst = 2^-4 keeps it below 2^12 times that.
constexpr double kSigTarget = 1024.0, kSigFirst = 0.0625;
// The stored q = A p: tau puts the bound anorm max |p| (above every |q|) at
// 2^13, so q never leaves the FP16 range.
constexpr double kTauTarget = 8192.0;

__host__ __device__ inline double cycle_scale(double cur, double target, double mx) {
  if (!(mx > 0.0) || !(mx <= 1.7976931348623157e308)) return cur;
  int e = 0;
  frexp(target / mx, &e);  // target / mx in [2^(e-1), 2^e)
  const int e0 = ilogb(cur);
  int en = e0 + e - 1;
  en = en < -120 ? -120 : en > 120 ? 120 : en;
  return ldexp(1.0, en);
}

// Totals the partials of the preceding reduction kernel and applies the PCG
// scalar logic (src/fpm_stairstep_mpi_3d.f90:257-278). One workgroup.
__device__ inline void finish_body(int op, int nslots, const double *partials, int nblocks,
                                   Scalars *sc, double *lds) {
  if ((op == FIN_PQ || op == FIN_RZNEW || op == FIN_FCG1) && sc->status != 0) return;
  const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
  if (op == FIN_MAX) {
    const double m = grid_finish_max(partials, 0, nblocks, lds);
    if (lead) sc->res = m;
  } else if (op == FIN_TR1) {
    const double m = grid_finish_max(partials, 2, nblocks, lds);
    const double mx = grid_finish_max(partials, 3, nblocks, lds + 16);
    double v[2];
    grid_finish_sum<2>(partials, nblocks, v, lds + 32);
    if (lead) {
      sc->res = m;
      sc->red1 = v[0];
      sc->red2 = v[1];
      // a solve's first cycle (or the first after a restart): its first step
      // is about mx (times first < 1)
      if (!(sc->sig > 0)) sc->sig = cycle_scale(1.0, kSigFirst, mx);
    }
  } else if (op == FIN_FCG1) {
    // CG preconditioned by the multigrid cycle (flexible CG with one stored
    // direction): slot 0 z.r, slot 1 z.(A p_old); p = z + beta p_old is made
    // A-orthogonal to p_old, beta = -z.A p_old / p_old.A p_old, and
    // alpha = z.r / p.A p follows in FIN_PQ. The slots hold the cycle's
    // scaled output (sig z; slot 0 also the scaled source): exact powers of
    // two divide them out.
    const double zmax = grid_finish_max(partials, 2, nblocks, lds);
    double v[2];
    grid_finish_sum<2>(partials, nblocks, v, lds + 16);
    if (lead) {
      const double sg = sc->sig, inv = 1.0 / sg;
      const double tq = sc->tau > 0.0 ? sc->tau : 1.0;  // the scale the stored q has
      sc->beta = sc->pq > 0 ? -(v[1] * inv / tq) / sc->pq : 0.0;
      sc->rz = v[0] * inv * inv;
      sc->sig_next = cycle_scale(sg, kSigTarget, zmax);
      // the direction about to be formed, p = z + beta p_old: max |p| <=
      // max |z| + |beta| max |p_old| (1% for the FP32 rounding of p)
      const double pm = 1.01 * (zmax * inv + fabs(sc->beta) * sc->pmax);
      sc->pmax = pm;
      sc->tau = cycle_scale(1.0, kTauTarget, sc->anorm * pm);
    }
  } else if (op == FIN_SUMS) {
    double v[kReduceSlots];
    grid_finish_sum<kReduceSlots>(partials, nblocks, v, lds);
    if (lead)
      for (int s = 0; s < nslots && s < kReduceSlots; ++s) sc->sums[s] = v[s];
  } else {
    double v[1];
    grid_finish_sum<1>(partials, nblocks, v, lds);
    if (!lead) return;
    if (op == FIN_RZ) {
      sc->rz = v[0];
    } else if (op == FIN_PQ) {
      const double pq = v[0], rz = sc->rz;
      sc->pq = pq;
      // The negated tests are also true for NaN.
      if (!(pq > 0) || !(rz > 0)) sc->status = kBreakdown;
      else sc->alpha = rz / pq;
    } else if (op == FIN_RZNEW) {
      const double rz_new = v[0];
      sc->rz_new = rz_new;
      if (!(rz_new >= 0)) {
        sc->status = kNegative;
      } else {
        sc->beta = rz_new / sc->rz;
        sc->rz = rz_new;
        sc->iters += 1;
      }
    }
  }
}

// true_residual, first pass (src/fpm_stairstep_mpi_3d.f90:303-319): the
// max-norm residual and r = V (src + sum c phi_nb - diag phi) in fluid cells
// share one evaluation of the bracket; sums of r and of the fluid count, and
// max |r / V| / diag (the scale of a multigrid cycle's first step).
__device__ inline void true_residual_1_body(const Tiling &t, int block, int nblocks,
                                            const Geom &g, const double *__restrict__ phi,
                                            const double *__restrict__ src,
                                            const double *__restrict__ diag,
                                            double *__restrict__ r, double *partials,
                                            double *lds) {
  const TileCoord tc = tile_coord(t, g.nx, g.ny, g.nz, block);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  double m = 0.0, mx = 0.0;
  double s[2] = {0.0, 0.0};
  if (tc.valid && tc.i < nx && tc.j < ny) {
    const int i = tc.i, j = tc.j;
    const size_t nxy = (size_t)nx * ny;
    size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * tc.k0);
    const double cw = g.cw[i], ce = g.ce[i], cs = g.cs[j], cn = g.cn[j];
    const double dxdy = g.dx[i] * g.dy[j];
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    double p_cur = phi[c];
    double p_prev = tc.k0 > 0 ? phi[c - nxy] : p_cur;
    for (int k = tc.k0; k < tc.k1; ++k) {
      const double p_next = (k + 1 < nz) ? phi[c + nxy] : p_cur;
      const double d = diag[c];
      double rv = 0.0;
      if (d > 0) {
        const double pw = hw ? phi[c - 1] : 0.0;
        const double pe = he ? phi[c + 1] : 0.0;
        const double ps = hs ? phi[c - nx] : 0.0;
        const double pn = hn ? phi[c + nx] : 0.0;
        const double v = src[c] + cw * pw + ce * pe + cs * ps + cn * pn + g.cb[k] * p_prev +
                         g.ct[k] * p_next - d * p_cur;
        m = fmax(m, fabs(v));
        mx = fmax(mx, fabs(v) / d);
        rv = dxdy * g.dz[k] * v;
        s[0] += rv;
        s[1] += 1.0;
      }
      r[c] = rv;
      p_prev = p_cur;
      p_cur = p_next;
      c += nxy;
    }
  }
  block_max(m, lds);
  block_max(mx, lds + 16);
  block_sum<2>(s, lds + 32);
  const double vals[4] = {s[0], s[1], m, mx};
  store_partials(partials, 4, vals, nblocks, block);
}

// true_residual, second pass (src/fpm_stairstep_mpi_3d.f90:319-328): project
// r off the constants, optionally reset p = M^-1 r, and rz = r.M^-1 r.
// Grid-stride over the n cells with nblocks workgroups of 256 threads.
__device__ inline void true_residual_2_body(size_t n, int block, int nblocks,
                                            double *__restrict__ r,
                                            const double *__restrict__ rdiag,
                                            const double *__restrict__ diag,
                                            double *__restrict__ p, int reset,
                                            double *__restrict__ z, double *partials,
                                            const Scalars *sc, double *lds) {
  const double shift = sc->red1 / sc->red2;
  double s[1] = {0.0};
  const size_t nt = (size_t)blockDim.x * blockDim.y;
  const size_t tid = threadIdx.x + (size_t)blockDim.x * threadIdx.y;
  for (size_t c = block * nt + tid; c < n; c += (size_t)nblocks * nt) {
    double rv = r[c];
    if (diag[c] > 0) {
      rv = rv - shift;
      r[c] = rv;
    }
    const double zv = rdiag[c] * rv;
    if (reset) p[c] = zv;
    if (z) z[c] = zv;
    s[0] += zv * rv;
  }
  block_sum<1>(s, lds);
  store_partials(partials, 1, s, nblocks, block);
}

// Deferred direction update p = z + beta p (before a convergence check);
// grid-stride over n cells.
template <typename T>
__device__ inline void dir_update_body(size_t n, int block, int nblocks, T *__restrict__ p,
                                       const T *__restrict__ z, const Scalars *sc) {
  if (sc->status != 0) return;
  const T beta = (T)sc->beta;
  const size_t nt = (size_t)blockDim.x * blockDim.y;
  const size_t tid = threadIdx.x + (size_t)blockDim.x * threadIdx.y;
  for (size_t c = block * nt + tid; c < n; c += (size_t)nblocks * nt) p[c] = z[c] + beta * p[c];
}

}  // namespace ffpm
