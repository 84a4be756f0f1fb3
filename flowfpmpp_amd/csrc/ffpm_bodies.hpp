// Device bodies shared by the single-solve kernels and the batched kernels
// (which run many solves in one grid). Each body takes its workgroup's index
// and the workgroup count of its own solve's launch explicitly, so a batched
// workgroup computes exactly what the same workgroup of a standalone launch
// computes, and the reductions sum in the same order.
#pragma once

#include "ffpm_kernels.hpp"

namespace ffpm {

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
    double v[2];
    grid_finish_sum<2>(partials, nblocks, v, lds + 16);
    if (lead) {
      sc->res = m;
      sc->red1 = v[0];
      sc->red2 = v[1];
    }
  } else if (op == FIN_FCG1) {
    // CG preconditioned by the multigrid cycle (flexible CG with one stored
    // direction): slot 0 z.r, slot 1 z.(A p_old); p = z + beta p_old is made
    // A-orthogonal to p_old, beta = -z.A p_old / p_old.A p_old, and
    // alpha = z.r / p.A p follows in FIN_PQ.
    double v[2];
    grid_finish_sum<2>(partials, nblocks, v, lds);
    if (lead) {
      sc->beta = sc->pq > 0 ? -v[1] / sc->pq : 0.0;
      sc->rz = v[0];
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
// share one evaluation of the bracket; sums of r and of the fluid count.
__device__ inline void true_residual_1_body(const Tiling &t, int block, int nblocks,
                                            const Geom &g, const double *__restrict__ phi,
                                            const double *__restrict__ src,
                                            const double *__restrict__ diag,
                                            double *__restrict__ r, double *partials,
                                            double *lds) {
  const TileCoord tc = tile_coord(t, g.nx, g.ny, g.nz, block);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  double m = 0.0;
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
  block_sum<2>(s, lds + 16);
  const double vals[3] = {s[0], s[1], m};
  store_partials(partials, 3, vals, nblocks, block);
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
