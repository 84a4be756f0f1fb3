/*
 * CPU oracle: FP64 restatement of the reference's serial arithmetic for the
 * stair-step solve, the FPM densities and the no-body Neumann solvers.
 *
 * TEST INFRASTRUCTURE ONLY (see ffpm_oracle.h). Loops keep the reference's
 * order (k outer, then j, then i) and its serial summation order; every
 * floating-point expression is written in the reference's left-to-right
 * association, and the file is built with -ffp-contract=off so that no
 * multiply-add is fused. Array layouts are the reference's Fortran layouts.
 */
#include "ffpm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* (0:nx+1, 0:ny+1, 1:nz) */
#define PH(i, j, k) ((size_t)(i) + (size_t)(nx + 2) * ((size_t)(j) + (size_t)(ny + 2) * (size_t)((k) - 1)))
/* (1:nx, 1:ny, 1:nz) */
/* The parallel CPU-baseline build (-fopenmp -DFFPM_ORACLE_OMP) splits the
 * z-plane loops of the PCG path and of the Jacobi sweep over threads; the
 * checker build is serial, so its sums keep the reference's order (the Jacobi
 * sweep has no sums: both builds write the same bits). */
#ifdef FFPM_ORACLE_OMP
#define ORC_OMP(x) _Pragma(#x)
#else
#define ORC_OMP(x)
#endif
#define CL(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))
/* (0:nx+1, 0:ny+1, 0:nz+1) */
#define PL(i, j, k) ((size_t)(i) + (size_t)(nx + 2) * ((size_t)(j) + (size_t)(ny + 2) * (size_t)(k)))
#define FXI(j, k) ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1))
#define FYI(i, k) ((size_t)((i) - 1) + (size_t)nx * (size_t)((k) - 1))
#define FZI(i, j) ((size_t)((i) - 1) + (size_t)nx * (size_t)((j) - 1))

static int is_null(int rank) { return rank < 0; }

static int single_rank(int west, int east, int south, int north) {
  return is_null(west) && is_null(east) && is_null(south) && is_null(north);
}

/* src/fpm_stairstep_mpi_3d.f90:25-36. With no neighbours the exchange itself
 * returns at once (src/fpm_halo_mpi_3d.f90:55). */
int orc_exchange_solid_halo_3d(int32_t *solid, int nx, int ny, int nz, int west, int east,
                               int south, int north) {
  if (!single_rank(west, east, south, north)) return -4;
  for (int k = 1; k <= nz; ++k)
    for (int j = 0; j <= ny + 1; ++j) {
      solid[PH(0, j, k)] = 0;
      solid[PH(nx + 1, j, k)] = 0;
    }
  for (int k = 1; k <= nz; ++k)
    for (int i = 0; i <= nx + 1; ++i) {
      solid[PH(i, 0, k)] = 0;
      solid[PH(i, ny + 1, k)] = 0;
    }
  return 0;
}

/* total_flux, src/fpm_stairstep_mpi_3d.f90:69-91 */
static void stairstep_total_flux(const int32_t *solid, const double *bg, const double *xlo,
                                 const double *xhi, const double *ylo, const double *yhi,
                                 const double *zlo, const double *zhi, const double *dx,
                                 const double *dy, const double *dz, int nx, int ny, int nz,
                                 int west, int east, int south, int north, double *flux,
                                 double *area) {
  double f = 0, ar = 0, a;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        if (solid[PH(i, j, k)] != 0) continue;
        a = dy[j - 1] * dz[k - 1];
        if (i == 1 && is_null(west)) { f = f - xlo[FXI(j, k)] * a; ar = ar + a; }
        else if (solid[PH(i - 1, j, k)] != 0) { f = f - bg[0] * a; }
        if (i == nx && is_null(east)) { f = f + xhi[FXI(j, k)] * a; ar = ar + a; }
        else if (solid[PH(i + 1, j, k)] != 0) { f = f + bg[0] * a; }
        a = dx[i - 1] * dz[k - 1];
        if (j == 1 && is_null(south)) { f = f - ylo[FYI(i, k)] * a; ar = ar + a; }
        else if (solid[PH(i, j - 1, k)] != 0) { f = f - bg[1] * a; }
        if (j == ny && is_null(north)) { f = f + yhi[FYI(i, k)] * a; ar = ar + a; }
        else if (solid[PH(i, j + 1, k)] != 0) { f = f + bg[1] * a; }
        a = dx[i - 1] * dy[j - 1];
        if (k == 1) { f = f - zlo[FZI(i, j)] * a; ar = ar + a; }
        else if (solid[PH(i, j, k - 1)] != 0) { f = f - bg[2] * a; }
        if (k == nz) { f = f + zhi[FZI(i, j)] * a; ar = ar + a; }
        else if (solid[PH(i, j, k + 1)] != 0) { f = f + bg[2] * a; }
      }
  *flux = f;
  *area = ar;
}

/* src/fpm_stairstep_mpi_3d.f90:41-92 */
int orc_enforce_stairstep_compatibility_3d(
    const double *rhs, const int32_t *solid, const double body_grad[3], double *xlo,
    double *xhi, double *ylo, double *yhi, double *zlo, double *zhi, const double *dx,
    const double *dy, const double *dz, int nx, int ny, int nz, int west, int east,
    int south, int north, double *correction, double *imbalance) {
  if (!single_rank(west, east, south, north)) return -4;
  double source = 0, flux, area;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        if (solid[PH(i, j, k)] == 0)
          source = source + rhs[CL(i, j, k)] * dx[i - 1] * dy[j - 1] * dz[k - 1];
  stairstep_total_flux(solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz, nx, ny, nz,
                       west, east, south, north, &flux, &area);
  if (area <= 0) return -3;
  const double c = -(flux - source) / area;
  *correction = c;
  for (size_t n = 0; n < (size_t)ny * nz; ++n) { xlo[n] = xlo[n] - c; xhi[n] = xhi[n] + c; }
  for (size_t n = 0; n < (size_t)nx * nz; ++n) { ylo[n] = ylo[n] - c; yhi[n] = yhi[n] + c; }
  for (size_t n = 0; n < (size_t)nx * ny; ++n) { zlo[n] = zlo[n] - c; zhi[n] = zhi[n] + c; }
  stairstep_total_flux(solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, dx, dy, dz, nx, ny, nz,
                       west, east, south, north, &flux, &area);
  *imbalance = flux - source;
  return 0;
}

/* axis_coefficients, src/fpm_stairstep_mpi_3d.f90:499-510, and the z
 * coefficients of src/fpm_stairstep_mpi_3d.f90:138-142. */
static void axis_coefficients(const double *center, const double *width, int n, int lower,
                              int upper, double *cl, double *cu) {
  for (int m = 1; m <= n; ++m) {
    cl[m - 1] = 1 / (width[m - 1] * (center[m] - center[m - 1]));
    cu[m - 1] = 1 / (width[m - 1] * (center[m + 1] - center[m]));
  }
  if (is_null(lower)) cl[0] = 0;
  if (is_null(upper)) cu[n - 1] = 0;
}

static void z_coefficients(const double *zc, const double *dz, int nz, double *cb, double *ct) {
  for (int k = 1; k <= nz; ++k) {
    cb[k - 1] = 0;
    ct[k - 1] = 0;
    if (k > 1) cb[k - 1] = 1 / (dz[k - 1] * (zc[k - 1] - zc[k - 2]));
    if (k < nz) ct[k - 1] = 1 / (dz[k - 1] * (zc[k] - zc[k - 1]));
  }
}

/* build_operator, src/fpm_stairstep_mpi_3d.f90:336-379 */
static int build_operator(const double *rhs, const int32_t *solid, const double *bg,
                          const double *xlo, const double *xhi, const double *ylo,
                          const double *yhi, const double *zlo, const double *zhi,
                          const double *cw, const double *ce, const double *cs,
                          const double *cn, const double *cb, const double *ct,
                          const double *dx, const double *dy, const double *dz, int nx, int ny,
                          int nz, int west, int east, int south, int north, int for_pcg,
                          double *src, double *diag, double *rdiag) {
  int isolated = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double s = 0, d = 0;
        if (solid[PH(i, j, k)] == 0) {
          s = -rhs[CL(i, j, k)];
          if (i == 1 && is_null(west)) s = s - xlo[FXI(j, k)] / dx[i - 1];
          else if (solid[PH(i - 1, j, k)] != 0) s = s - bg[0] / dx[i - 1];
          else d = d + cw[i - 1];
          if (i == nx && is_null(east)) s = s + xhi[FXI(j, k)] / dx[i - 1];
          else if (solid[PH(i + 1, j, k)] != 0) s = s + bg[0] / dx[i - 1];
          else d = d + ce[i - 1];
          if (j == 1 && is_null(south)) s = s - ylo[FYI(i, k)] / dy[j - 1];
          else if (solid[PH(i, j - 1, k)] != 0) s = s - bg[1] / dy[j - 1];
          else d = d + cs[j - 1];
          if (j == ny && is_null(north)) s = s + yhi[FYI(i, k)] / dy[j - 1];
          else if (solid[PH(i, j + 1, k)] != 0) s = s + bg[1] / dy[j - 1];
          else d = d + cn[j - 1];
          if (k == 1) s = s - zlo[FZI(i, j)] / dz[k - 1];
          else if (solid[PH(i, j, k - 1)] != 0) s = s - bg[2] / dz[k - 1];
          else d = d + cb[k - 1];
          if (k == nz) s = s + zhi[FZI(i, j)] / dz[k - 1];
          else if (solid[PH(i, j, k + 1)] != 0) s = s + bg[2] / dz[k - 1];
          else d = d + ct[k - 1];
          if (d <= 0) isolated = isolated + 1;
        }
        const size_t c = CL(i, j, k);
        src[c] = s;
        diag[c] = d;
        if (d <= 0) rdiag[c] = 0;
        else if (for_pcg) rdiag[c] = 1 / (dx[i - 1] * dy[j - 1] * dz[k - 1] * d);
        else rdiag[c] = 0.8 / d;
      }
  return isolated;
}

/* jacobi_sweep, src/fpm_stairstep_mpi_3d.f90:384-399 */
int orc_jacobi_sweep_3d(const double *a, double *b, const double *src, const double *rdiag,
                        const double *cw, const double *ce, const double *cs, const double *cn,
                        const double *cb, const double *ct, int nx, int ny, int nz) {
  /* every b cell is formed from a alone, so the threaded build writes the
   * same bits as the serial one */
  ORC_OMP(omp parallel for schedule(static))
  for (int k = 1; k <= nz; ++k) {
    const int km = k > 1 ? k - 1 : 1, kp = k < nz ? k + 1 : nz;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        b[PH(i, j, k)] =
            0.2 * a[PH(i, j, k)] +
            rdiag[c] * (src[c] + cw[i - 1] * a[PH(i - 1, j, k)] + ce[i - 1] * a[PH(i + 1, j, k)] +
                        cs[j - 1] * a[PH(i, j - 1, k)] + cn[j - 1] * a[PH(i, j + 1, k)] +
                        cb[k - 1] * a[PH(i, j, km)] + ct[k - 1] * a[PH(i, j, kp)]);
      }
  }
  return 0;
}

/* operator_residual, src/fpm_stairstep_mpi_3d.f90:514-530 */
double orc_operator_residual_3d(const double *phi, const double *src, const double *diag,
                                const double *cw, const double *ce, const double *cs,
                                const double *cn, const double *cb, const double *ct, int nx,
                                int ny, int nz) {
  double local = 0;
  ORC_OMP(omp parallel for reduction(max : local) schedule(static))
  for (int k = 1; k <= nz; ++k) {
    const int km = k > 1 ? k - 1 : 1, kp = k < nz ? k + 1 : nz;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        if (diag[c] > 0) {
          const double v = fabs(src[c] + cw[i - 1] * phi[PH(i - 1, j, k)] +
                                ce[i - 1] * phi[PH(i + 1, j, k)] + cs[j - 1] * phi[PH(i, j - 1, k)] +
                                cn[j - 1] * phi[PH(i, j + 1, k)] + cb[k - 1] * phi[PH(i, j, km)] +
                                ct[k - 1] * phi[PH(i, j, kp)] - diag[c] * phi[PH(i, j, k)]);
          /* Fortran max(local, NaN) keeps local on this build; a NaN
           * residual cannot be produced from finite inputs. */
          if (v > local) local = v;
        }
      }
  }
  return local;
}

int orc_build_operator_3d(const double *rhs, const int32_t *solid, const double body_grad[3],
                          const double *xlo, const double *xhi, const double *ylo,
                          const double *yhi, const double *zlo, const double *zhi,
                          const double *xc, const double *yc, const double *zc, const double *dx,
                          const double *dy, const double *dz, int nx, int ny, int nz, int west,
                          int east, int south, int north, int for_pcg, double *src, double *diag,
                          double *rdiag, double *cw, double *ce, double *cs, double *cn,
                          double *cb, double *ct, int *isolated) {
  if (!single_rank(west, east, south, north)) return -4;
  axis_coefficients(xc, dx, nx, west, east, cw, ce);
  axis_coefficients(yc, dy, ny, south, north, cs, cn);
  z_coefficients(zc, dz, nz, cb, ct);
  *isolated = build_operator(rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, cw, ce, cs, cn,
                             cb, ct, dx, dy, dz, nx, ny, nz, west, east, south, north, for_pcg,
                             src, diag, rdiag);
  return 0;
}

typedef struct {
  int nx, ny, nz;
  double *phi, *next, *src, *diag, *rdiag, *r, *q;
  const double *cw, *ce, *cs, *cn, *cb, *ct, *dx, *dy, *dz;
} solve_state;

/* true_residual, src/fpm_stairstep_mpi_3d.f90:297-329 */
static void true_residual(solve_state *st, int reset, double *rz_out, double *res_out) {
  const int nx = st->nx, ny = st->ny, nz = st->nz;
  const double *phi = st->phi, *src = st->src, *diag = st->diag, *rdiag = st->rdiag;
  double *r = st->r, *next = st->next;
  *res_out = orc_operator_residual_3d(phi, src, diag, st->cw, st->ce, st->cs, st->cn, st->cb,
                                      st->ct, nx, ny, nz);
  double red1 = 0, red2 = 0;
  ORC_OMP(omp parallel for reduction(+ : red1, red2) schedule(static))
  for (int k = 1; k <= nz; ++k) {
    const int km = k > 1 ? k - 1 : 1, kp = k < nz ? k + 1 : nz;
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        r[c] = 0;
        if (diag[c] > 0) {
          const double vol = st->dx[i - 1] * st->dy[j - 1] * st->dz[k - 1];
          r[c] = vol * (src[c] + st->cw[i - 1] * phi[PH(i - 1, j, k)] +
                        st->ce[i - 1] * phi[PH(i + 1, j, k)] + st->cs[j - 1] * phi[PH(i, j - 1, k)] +
                        st->cn[j - 1] * phi[PH(i, j + 1, k)] + st->cb[k - 1] * phi[PH(i, j, km)] +
                        st->ct[k - 1] * phi[PH(i, j, kp)] - diag[c] * phi[PH(i, j, k)]);
          red1 = red1 + r[c];
          red2 = red2 + 1;
        }
      }
  }
  const double shift = red1 / red2;
  double rz = 0;
  ORC_OMP(omp parallel for reduction(+ : rz) schedule(static))
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        if (diag[c] > 0) r[c] = r[c] - shift;
        if (reset) next[PH(i, j, k)] = rdiag[c] * r[c];
        rz = rz + rdiag[c] * r[c] * r[c];
      }
  *rz_out = rz;
}

/* pcg_iterate, src/fpm_stairstep_mpi_3d.f90:235-293 */
static void pcg_iterate(solve_state *st, double tolerance, int max_sweeps, int check_interval,
                        int *sweeps, double *residual) {
  const int nx = st->nx, ny = st->ny, nz = st->nz;
  double *phi = st->phi, *p = st->next, *r = st->r, *q = st->q;
  const double *diag = st->diag, *rdiag = st->rdiag;
  double rz, rz_new, pq, alpha, beta, res, best;
  int it = 0, restarts = 0, stalled = 0;
  const int ci = check_interval > 1 ? check_interval : 1;
  true_residual(st, 1, &rz, &res);
  best = res;
  while (*sweeps < max_sweeps && res > tolerance) {
    pq = 0;
    ORC_OMP(omp parallel for reduction(+ : pq) schedule(static))
    for (int k = 1; k <= nz; ++k) {
      const int km = k > 1 ? k - 1 : 1, kp = k < nz ? k + 1 : nz;
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t c = CL(i, j, k);
          const double vol = st->dx[i - 1] * st->dy[j - 1] * st->dz[k - 1];
          q[c] = diag[c] > 0
                     ? vol * (diag[c] * p[PH(i, j, k)] - st->cw[i - 1] * p[PH(i - 1, j, k)] -
                              st->ce[i - 1] * p[PH(i + 1, j, k)] - st->cs[j - 1] * p[PH(i, j - 1, k)] -
                              st->cn[j - 1] * p[PH(i, j + 1, k)] - st->cb[k - 1] * p[PH(i, j, km)] -
                              st->ct[k - 1] * p[PH(i, j, kp)])
                     : 0.0;
          pq = pq + p[PH(i, j, k)] * q[c];
        }
    }
    /* The negated tests are also true for NaN. */
    if (!(pq > 0) || !(rz > 0)) {
      restarts = restarts + 1;
      if (restarts > 1) break;
      true_residual(st, 1, &rz, &res);
      continue;
    }
    restarts = 0;
    alpha = rz / pq;
    rz_new = 0;
    ORC_OMP(omp parallel for reduction(+ : rz_new) schedule(static))
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t c = CL(i, j, k);
          phi[PH(i, j, k)] = phi[PH(i, j, k)] + alpha * p[PH(i, j, k)];
          r[c] = r[c] - alpha * q[c];
          rz_new = rz_new + rdiag[c] * r[c] * r[c];
        }
    if (!(rz_new >= 0)) break;
    beta = rz_new / rz;
    rz = rz_new;
    ORC_OMP(omp parallel for schedule(static))
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t c = CL(i, j, k);
          p[PH(i, j, k)] = rdiag[c] * r[c] + beta * p[PH(i, j, k)];
        }
    *sweeps = *sweeps + 1;
    it = it + 1;
    if (it % ci == 0 || *sweeps == max_sweeps) {
      true_residual(st, 0, &rz, &res);
      if (res < 0.99 * best) { best = res; stalled = 0; }
      else stalled = stalled + 1;
      if (stalled >= 50 || !(res == res)) break;
    }
  }
  *residual = res;
}

/* jacobi_iterate, src/fpm_stairstep_mpi_3d.f90:200-225 */
static void jacobi_iterate(solve_state *st, double tolerance, int max_sweeps, int check_interval,
                           int *sweeps, double *residual) {
  const int nx = st->nx, ny = st->ny, nz = st->nz;
  while (*sweeps < max_sweeps && *residual > tolerance) {
    int nsteps = check_interval > 1 ? check_interval : 1;
    if (max_sweeps - *sweeps < nsteps) nsteps = max_sweeps - *sweeps;
    for (int step = 1; step <= nsteps; ++step) {
      if (step % 2 == 1)
        orc_jacobi_sweep_3d(st->phi, st->next, st->src, st->rdiag, st->cw, st->ce, st->cs, st->cn,
                            st->cb, st->ct, nx, ny, nz);
      else
        orc_jacobi_sweep_3d(st->next, st->phi, st->src, st->rdiag, st->cw, st->ce, st->cs, st->cn,
                            st->cb, st->ct, nx, ny, nz);
      *sweeps = *sweeps + 1;
    }
    if (nsteps % 2 == 1)
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) st->phi[PH(i, j, k)] = st->next[PH(i, j, k)];
    *residual = orc_operator_residual_3d(st->phi, st->src, st->diag, st->cw, st->ce, st->cs,
                                         st->cn, st->cb, st->ct, nx, ny, nz);
  }
}

/* solve_stairstep_poisson_3d, src/fpm_stairstep_mpi_3d.f90:106-330 */
int orc_solve_stairstep_poisson_3d(
    double *phi, const double *rhs, const int32_t *solid, const double body_grad[3],
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double tolerance, int max_sweeps, int check_interval,
    int method, int *sweeps, double *residual) {
  if (method != 0 && method != 1) return -1;
  if (nx < 1 || ny < 1 || nz < 2) return -1;
  if (!single_rank(west, east, south, north)) return -4;
  const int use_pcg = method == 0;
  const size_t npad = (size_t)(nx + 2) * (ny + 2) * nz, ncell = (size_t)nx * ny * nz;
  solve_state st;
  st.nx = nx; st.ny = ny; st.nz = nz;
  st.phi = phi;
  st.next = (double *)calloc(npad, sizeof(double));
  st.src = (double *)calloc(ncell, sizeof(double));
  st.diag = (double *)calloc(ncell, sizeof(double));
  st.rdiag = (double *)calloc(ncell, sizeof(double));
  st.r = use_pcg ? (double *)calloc(ncell, sizeof(double)) : NULL;
  st.q = use_pcg ? (double *)calloc(ncell, sizeof(double)) : NULL;
  double *coef = (double *)calloc((size_t)2 * (nx + ny + nz), sizeof(double));
  double *cw = coef, *ce = cw + nx, *cs = ce + nx, *cn = cs + ny, *cb = cn + ny, *ct = cb + nz;
  axis_coefficients(xc, dx, nx, west, east, cw, ce);
  axis_coefficients(yc, dy, ny, south, north, cs, cn);
  z_coefficients(zc, dz, nz, cb, ct);
  st.cw = cw; st.ce = ce; st.cs = cs; st.cn = cn; st.cb = cb; st.ct = ct;
  st.dx = dx; st.dy = dy; st.dz = dz;
  *sweeps = 0;
  *residual = DBL_MAX;
  const int isolated = build_operator(rhs, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, cw, ce,
                                      cs, cn, cb, ct, dx, dy, dz, nx, ny, nz, west, east, south,
                                      north, use_pcg, st.src, st.diag, st.rdiag);
  int status = 0;
  if (isolated > 0) {
    status = -2;
    goto done;
  }
  for (int k = 1; k <= nz; ++k)
    for (int j = 0; j <= ny + 1; ++j)
      for (int i = 0; i <= nx + 1; ++i) {
        if (solid[PH(i, j, k)] != 0 || (i == 0 && is_null(west)) || (i == nx + 1 && is_null(east)) ||
            (j == 0 && is_null(south)) || (j == ny + 1 && is_null(north)))
          phi[PH(i, j, k)] = 0;
        st.next[PH(i, j, k)] = phi[PH(i, j, k)];
      }
  if (use_pcg) pcg_iterate(&st, tolerance, max_sweeps, check_interval, sweeps, residual);
  else jacobi_iterate(&st, tolerance, max_sweeps, check_interval, sweeps, residual);
  {
    double l1 = 0, l2 = 0;
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          if (solid[PH(i, j, k)] == 0) {
            l1 = l1 + phi[PH(i, j, k)] * dx[i - 1] * dy[j - 1] * dz[k - 1];
            l2 = l2 + dx[i - 1] * dy[j - 1] * dz[k - 1];
          }
    const double mean_value = l1 / l2;
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          if (solid[PH(i, j, k)] == 0) phi[PH(i, j, k)] = phi[PH(i, j, k)] - mean_value;
  }
  *residual = orc_operator_residual_3d(phi, st.src, st.diag, cw, ce, cs, cn, cb, ct, nx, ny, nz);
done:
  free(st.next); free(st.src); free(st.diag); free(st.rdiag); free(st.r); free(st.q); free(coef);
  return status;
}

/* face_gradients, src/fpm_stairstep_mpi_3d.f90:472-497 */
static void face_gradients(const double *phi, const int32_t *solid, const double *bg,
                           const double *xlo, const double *xhi, const double *ylo,
                           const double *yhi, const double *zlo, const double *zhi,
                           const double *xc, const double *yc, const double *zc, int nx, int ny,
                           int nz, int west, int east, int south, int north, int i, int j, int k,
                           double g[6]) {
  if (i == 1 && is_null(west)) g[0] = xlo[FXI(j, k)];
  else if (solid[PH(i - 1, j, k)] != 0) g[0] = bg[0];
  else g[0] = (phi[PH(i, j, k)] - phi[PH(i - 1, j, k)]) / (xc[i] - xc[i - 1]);
  if (i == nx && is_null(east)) g[1] = xhi[FXI(j, k)];
  else if (solid[PH(i + 1, j, k)] != 0) g[1] = bg[0];
  else g[1] = (phi[PH(i + 1, j, k)] - phi[PH(i, j, k)]) / (xc[i + 1] - xc[i]);
  if (j == 1 && is_null(south)) g[2] = ylo[FYI(i, k)];
  else if (solid[PH(i, j - 1, k)] != 0) g[2] = bg[1];
  else g[2] = (phi[PH(i, j, k)] - phi[PH(i, j - 1, k)]) / (yc[j] - yc[j - 1]);
  if (j == ny && is_null(north)) g[3] = yhi[FYI(i, k)];
  else if (solid[PH(i, j + 1, k)] != 0) g[3] = bg[1];
  else g[3] = (phi[PH(i, j + 1, k)] - phi[PH(i, j, k)]) / (yc[j + 1] - yc[j]);
  if (k == 1) g[4] = zlo[FZI(i, j)];
  else if (solid[PH(i, j, k - 1)] != 0) g[4] = bg[2];
  else g[4] = (phi[PH(i, j, k)] - phi[PH(i, j, k - 1)]) / (zc[k - 1] - zc[k - 2]);
  if (k == nz) g[5] = zhi[FZI(i, j)];
  else if (solid[PH(i, j, k + 1)] != 0) g[5] = bg[2];
  else g[5] = (phi[PH(i, j, k + 1)] - phi[PH(i, j, k)]) / (zc[k] - zc[k - 1]);
}

/* stairstep_residual_3d, src/fpm_stairstep_mpi_3d.f90:403-419 */
int orc_stairstep_residual_3d(const double *phi, const double *rhs, const int32_t *solid,
                              const double body_grad[3], const double *xlo, const double *xhi,
                              const double *ylo, const double *yhi, const double *zlo,
                              const double *zhi, const double *xc, const double *yc,
                              const double *zc, const double *dx, const double *dy,
                              const double *dz, int nx, int ny, int nz, int west, int east,
                              int south, int north, double *value) {
  if (!single_rank(west, east, south, north)) return -4;
  double local = 0, g[6];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        if (solid[PH(i, j, k)] != 0) continue;
        face_gradients(phi, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, nx, ny, nz,
                       west, east, south, north, i, j, k, g);
        const double v = fabs((g[1] - g[0]) / dx[i - 1] + (g[3] - g[2]) / dy[j - 1] +
                              (g[5] - g[4]) / dz[k - 1] - rhs[CL(i, j, k)]);
        if (v > local) local = v;
      }
  *value = local;
  return 0;
}

/* compute_stairstep_fpm_terms_3d, src/fpm_stairstep_mpi_3d.f90:426-469 */
int orc_compute_stairstep_fpm_terms_3d(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *solid, const double body_grad[3], const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, const double bounds[6], double *gx, double *gy, double *gz,
    double *vortex, double *advection, double totals[2], int nx, int ny, int nz, int west,
    int east, int south, int north) {
  if (!single_rank(west, east, south, north)) return -4;
  double vt = 0, at = 0, g[6];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        face_gradients(phi, solid, body_grad, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, nx, ny, nz,
                       west, east, south, north, i, j, k, g);
        const size_t c = CL(i, j, k);
        const double mask = solid[PH(i, j, k)] == 0 ? 1.0 : 0.0;
        gx[c] = mask * 0.5 * (g[0] + g[1]);
        gy[c] = mask * 0.5 * (g[2] + g[3]);
        gz[c] = mask * 0.5 * (g[4] + g[5]);
        const double cx = w[c] * oy[c] - v[c] * oz[c];
        const double cy = u[c] * oz[c] - w[c] * ox[c];
        const double cz = v[c] * ox[c] - u[c] * oy[c];
        vortex[c] = -(cx * gx[c] + cy * gy[c] + cz * gz[c]);
        advection[c] = -(ax[c] * gx[c] + ay[c] * gy[c] + az[c] * gz[c]);
        const int inside = xc[i] > bounds[0] && xc[i] < bounds[1] && yc[j] > bounds[2] &&
                           yc[j] < bounds[3] && zc[k - 1] > bounds[4] && zc[k - 1] < bounds[5];
        const double volume = inside ? dx[i - 1] * dy[j - 1] * dz[k - 1] : 0.0;
        vt = vt + volume * vortex[c];
        at = at + volume * advection[c];
      }
  totals[0] = vt;
  totals[1] = at;
  return 0;
}

/* ---------------------------------------------------------------------------
 * No-body Neumann solvers, src/fpm_laplace_3d.f90
 * ------------------------------------------------------------------------- */

/* enforce_neumann_compatibility_3d, src/fpm_laplace_3d.f90:9-53 */
int orc_enforce_neumann_compatibility_3d(const double *rhs, double *xlo, double *xhi,
                                         double *ylo, double *yhi, double *zlo, double *zhi,
                                         int nx, int ny, int nz, double dx, double dy, double dz,
                                         double *correction, double *imbalance) {
  double source = 0, flux = 0, area;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) source = source + rhs[CL(i, j, k)] * dx * dy * dz;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) flux = flux + (-xlo[FXI(j, k)] + xhi[FXI(j, k)]) * dy * dz;
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) flux = flux + (-ylo[FYI(i, k)] + yhi[FYI(i, k)]) * dx * dz;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) flux = flux + (-zlo[FZI(i, j)] + zhi[FZI(i, j)]) * dx * dy;
  area = 2.0 * ((double)(ny * nz) * dy * dz + (double)(nx * nz) * dx * dz +
                (double)(nx * ny) * dx * dy);
  const double c = -(flux - source) / area;
  *correction = c;
  for (size_t n = 0; n < (size_t)ny * nz; ++n) { xlo[n] = xlo[n] - c; xhi[n] = xhi[n] + c; }
  for (size_t n = 0; n < (size_t)nx * nz; ++n) { ylo[n] = ylo[n] - c; yhi[n] = yhi[n] + c; }
  for (size_t n = 0; n < (size_t)nx * ny; ++n) { zlo[n] = zlo[n] - c; zhi[n] = zhi[n] + c; }
  flux = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) flux = flux + (-xlo[FXI(j, k)] + xhi[FXI(j, k)]) * dy * dz;
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) flux = flux + (-ylo[FYI(i, k)] + yhi[FYI(i, k)]) * dx * dz;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) flux = flux + (-zlo[FZI(i, j)] + zhi[FZI(i, j)]) * dx * dy;
  *imbalance = flux - source;
  return 0;
}

/* residual_norm, src/fpm_laplace_3d.f90:109-132 */
static double uniform_residual(const double *phi, const double *rhs, const double *xlo,
                               const double *xhi, const double *ylo, const double *yhi,
                               const double *zlo, const double *zhi, int nx, int ny, int nz,
                               double dx, double dy, double dz) {
  double value = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double p = phi[PL(i, j, k)];
        double xm = phi[PL(i > 1 ? i - 1 : 1, j, k)], xp = phi[PL(i < nx ? i + 1 : nx, j, k)];
        double ym = phi[PL(i, j > 1 ? j - 1 : 1, k)], yp = phi[PL(i, j < ny ? j + 1 : ny, k)];
        double zm = phi[PL(i, j, k > 1 ? k - 1 : 1)], zp = phi[PL(i, j, k < nz ? k + 1 : nz)];
        if (i == 1) xm = p - xlo[FXI(j, k)] * dx;
        if (i == nx) xp = p + xhi[FXI(j, k)] * dx;
        if (j == 1) ym = p - ylo[FYI(i, k)] * dy;
        if (j == ny) yp = p + yhi[FYI(i, k)] * dy;
        if (k == 1) zm = p - zlo[FZI(i, j)] * dz;
        if (k == nz) zp = p + zhi[FZI(i, j)] * dz;
        const double res = (xm - 2 * p + xp) / (dx * dx) + (ym - 2 * p + yp) / (dy * dy) +
                           (zm - 2 * p + zp) / (dz * dz) - rhs[CL(i, j, k)];
        if (fabs(res) > value) value = fabs(res);
      }
  return value;
}

/* solve_neumann_poisson_3d, src/fpm_laplace_3d.f90:55-107, set_ghosts 134-148 */
int orc_solve_neumann_poisson_3d(double *phi, const double *rhs, const double *xlo,
                                 const double *xhi, const double *ylo, const double *yhi,
                                 const double *zlo, const double *zhi, int nx, int ny, int nz,
                                 double dx, double dy, double dz, double tolerance,
                                 int max_iterations, int check_interval, int *iterations,
                                 double *final_residual) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const size_t npad = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
  double *next = (double *)calloc(npad, sizeof(double));
  const double idx2 = 1.0 / (dx * dx), idy2 = 1.0 / (dy * dy), idz2 = 1.0 / (dz * dz);
  const double invdiag = 1.0 / (2.0 * (idx2 + idy2 + idz2));
  *iterations = 0;
  *final_residual = DBL_MAX;
  while (*iterations < max_iterations && *final_residual > tolerance) {
    int nsteps = check_interval > 1 ? check_interval : 1;
    if (max_iterations - *iterations < nsteps) nsteps = max_iterations - *iterations;
    for (int step = 1; step <= nsteps; ++step) {
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) {
            const double p = phi[PL(i, j, k)];
            double xm = phi[PL(i > 1 ? i - 1 : 1, j, k)], xp = phi[PL(i < nx ? i + 1 : nx, j, k)];
            double ym = phi[PL(i, j > 1 ? j - 1 : 1, k)], yp = phi[PL(i, j < ny ? j + 1 : ny, k)];
            double zm = phi[PL(i, j, k > 1 ? k - 1 : 1)], zp = phi[PL(i, j, k < nz ? k + 1 : nz)];
            if (i == 1) xm = p - xlo[FXI(j, k)] * dx;
            if (i == nx) xp = p + xhi[FXI(j, k)] * dx;
            if (j == 1) ym = p - ylo[FYI(i, k)] * dy;
            if (j == ny) yp = p + yhi[FYI(i, k)] * dy;
            if (k == 1) zm = p - zlo[FZI(i, j)] * dz;
            if (k == nz) zp = p + zhi[FZI(i, j)] * dz;
            next[PL(i, j, k)] =
                ((xm + xp) * idx2 + (ym + yp) * idy2 + (zm + zp) * idz2 - rhs[CL(i, j, k)]) * invdiag;
          }
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) phi[PL(i, j, k)] = next[PL(i, j, k)];
      *iterations = *iterations + 1;
    }
    *final_residual =
        uniform_residual(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, dx, dy, dz);
  }
  double mean_value = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) mean_value = mean_value + phi[PL(i, j, k)];
  mean_value = mean_value / (double)(nx * ny * nz);
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) phi[PL(i, j, k)] = phi[PL(i, j, k)] - mean_value;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) {
      phi[PL(0, j, k)] = phi[PL(1, j, k)] - xlo[FXI(j, k)] * dx;
      phi[PL(nx + 1, j, k)] = phi[PL(nx, j, k)] + xhi[FXI(j, k)] * dx;
    }
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) {
      phi[PL(i, 0, k)] = phi[PL(i, 1, k)] - ylo[FYI(i, k)] * dy;
      phi[PL(i, ny + 1, k)] = phi[PL(i, ny, k)] + yhi[FYI(i, k)] * dy;
    }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      phi[PL(i, j, 0)] = phi[PL(i, j, 1)] - zlo[FZI(i, j)] * dz;
      phi[PL(i, j, nz + 1)] = phi[PL(i, j, nz)] + zhi[FZI(i, j)] * dz;
    }
  *final_residual = uniform_residual(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, dx, dy, dz);
  free(next);
  return 0;
}

/* enforce_neumann_compatibility_nonuniform_3d, src/fpm_laplace_3d.f90:150-217 */
int orc_enforce_neumann_compatibility_nonuniform_3d(const double *rhs, double *xlo, double *xhi,
                                                    double *ylo, double *yhi, double *zlo,
                                                    double *zhi, const double *dx,
                                                    const double *dy, const double *dz, int nx,
                                                    int ny, int nz, double *correction,
                                                    double *imbalance) {
  double source = 0, flux = 0, area = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        source = source + rhs[CL(i, j, k)] * dx[i - 1] * dy[j - 1] * dz[k - 1];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) {
      flux = flux + (-xlo[FXI(j, k)] + xhi[FXI(j, k)]) * dy[j - 1] * dz[k - 1];
      area = area + 2.0 * dy[j - 1] * dz[k - 1];
    }
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) {
      flux = flux + (-ylo[FYI(i, k)] + yhi[FYI(i, k)]) * dx[i - 1] * dz[k - 1];
      area = area + 2.0 * dx[i - 1] * dz[k - 1];
    }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      flux = flux + (-zlo[FZI(i, j)] + zhi[FZI(i, j)]) * dx[i - 1] * dy[j - 1];
      area = area + 2.0 * dx[i - 1] * dy[j - 1];
    }
  const double c = -(flux - source) / area;
  *correction = c;
  for (size_t n = 0; n < (size_t)ny * nz; ++n) { xlo[n] = xlo[n] - c; xhi[n] = xhi[n] + c; }
  for (size_t n = 0; n < (size_t)nx * nz; ++n) { ylo[n] = ylo[n] - c; yhi[n] = yhi[n] + c; }
  for (size_t n = 0; n < (size_t)nx * ny; ++n) { zlo[n] = zlo[n] - c; zhi[n] = zhi[n] + c; }
  flux = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      flux = flux + (-xlo[FXI(j, k)] + xhi[FXI(j, k)]) * dy[j - 1] * dz[k - 1];
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i)
      flux = flux + (-ylo[FYI(i, k)] + yhi[FYI(i, k)]) * dx[i - 1] * dz[k - 1];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      flux = flux + (-zlo[FZI(i, j)] + zhi[FZI(i, j)]) * dx[i - 1] * dy[j - 1];
  *imbalance = flux - source;
  return 0;
}

/* residual_nonuniform, src/fpm_laplace_3d.f90:295-315 */
static double nonuniform_residual(const double *phi, const double *rhs, const double *xlo,
                                  const double *xhi, const double *ylo, const double *yhi,
                                  const double *zlo, const double *zhi, const double *xc,
                                  const double *yc, const double *zc, const double *dx,
                                  const double *dy, const double *dz, int nx, int ny, int nz) {
  double value = 0, gw, ge, gs, gn, gb, gt;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        if (i == 1) gw = xlo[FXI(j, k)];
        else gw = (phi[c] - phi[CL(i - 1, j, k)]) / (xc[i - 1] - xc[i - 2]);
        if (i == nx) ge = xhi[FXI(j, k)];
        else ge = (phi[CL(i + 1, j, k)] - phi[c]) / (xc[i] - xc[i - 1]);
        if (j == 1) gs = ylo[FYI(i, k)];
        else gs = (phi[c] - phi[CL(i, j - 1, k)]) / (yc[j - 1] - yc[j - 2]);
        if (j == ny) gn = yhi[FYI(i, k)];
        else gn = (phi[CL(i, j + 1, k)] - phi[c]) / (yc[j] - yc[j - 1]);
        if (k == 1) gb = zlo[FZI(i, j)];
        else gb = (phi[c] - phi[CL(i, j, k - 1)]) / (zc[k - 1] - zc[k - 2]);
        if (k == nz) gt = zhi[FZI(i, j)];
        else gt = (phi[CL(i, j, k + 1)] - phi[c]) / (zc[k] - zc[k - 1]);
        const double res =
            (ge - gw) / dx[i - 1] + (gn - gs) / dy[j - 1] + (gt - gb) / dz[k - 1] - rhs[c];
        if (fabs(res) > value) value = fabs(res);
      }
  return value;
}

/* solve_neumann_poisson_nonuniform_3d, src/fpm_laplace_3d.f90:219-293 */
int orc_solve_neumann_poisson_nonuniform_3d(double *phi, const double *rhs, const double *xlo,
                                            const double *xhi, const double *ylo,
                                            const double *yhi, const double *zlo,
                                            const double *zhi, const double *xc,
                                            const double *yc, const double *zc, const double *dx,
                                            const double *dy, const double *dz, int nx, int ny,
                                            int nz, double tolerance, int max_iterations,
                                            int check_interval, int *iterations,
                                            double *final_residual) {
  if (nx < 1 || ny < 1 || nz < 1) return -1;
  const size_t ncell = (size_t)nx * ny * nz;
  double *next = (double *)calloc(ncell, sizeof(double));
  *iterations = 0;
  *final_residual = DBL_MAX;
  while (*iterations < max_iterations && *final_residual > tolerance) {
    int nsteps = check_interval > 1 ? check_interval : 1;
    if (max_iterations - *iterations < nsteps) nsteps = max_iterations - *iterations;
    for (int step = 1; step <= nsteps; ++step) {
      for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
          for (int i = 1; i <= nx; ++i) {
            double aw = 0, ae = 0, as = 0, an = 0, ab = 0, at = 0;
            if (i > 1) aw = 1.0 / (dx[i - 1] * (xc[i - 1] - xc[i - 2]));
            if (i < nx) ae = 1.0 / (dx[i - 1] * (xc[i] - xc[i - 1]));
            if (j > 1) as = 1.0 / (dy[j - 1] * (yc[j - 1] - yc[j - 2]));
            if (j < ny) an = 1.0 / (dy[j - 1] * (yc[j] - yc[j - 1]));
            if (k > 1) ab = 1.0 / (dz[k - 1] * (zc[k - 1] - zc[k - 2]));
            if (k < nz) at = 1.0 / (dz[k - 1] * (zc[k] - zc[k - 1]));
            const double diag = aw + ae + as + an + ab + at;
            const size_t c = CL(i, j, k);
            double numer = -rhs[c];
            if (i > 1) numer = numer + aw * phi[CL(i - 1, j, k)];
            if (i < nx) numer = numer + ae * phi[CL(i + 1, j, k)];
            if (j > 1) numer = numer + as * phi[CL(i, j - 1, k)];
            if (j < ny) numer = numer + an * phi[CL(i, j + 1, k)];
            if (k > 1) numer = numer + ab * phi[CL(i, j, k - 1)];
            if (k < nz) numer = numer + at * phi[CL(i, j, k + 1)];
            if (i == 1) numer = numer - xlo[FXI(j, k)] / dx[i - 1];
            if (i == nx) numer = numer + xhi[FXI(j, k)] / dx[i - 1];
            if (j == 1) numer = numer - ylo[FYI(i, k)] / dy[j - 1];
            if (j == ny) numer = numer + yhi[FYI(i, k)] / dy[j - 1];
            if (k == 1) numer = numer - zlo[FZI(i, j)] / dz[k - 1];
            if (k == nz) numer = numer + zhi[FZI(i, j)] / dz[k - 1];
            next[c] = 0.2 * phi[c] + 0.8 * numer / diag;
          }
      memcpy(phi, next, ncell * sizeof(double));
      *iterations = *iterations + 1;
    }
    *final_residual = nonuniform_residual(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx,
                                          dy, dz, nx, ny, nz);
  }
  double ws = 0, vs = 0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        ws = ws + phi[CL(i, j, k)] * dx[i - 1] * dy[j - 1] * dz[k - 1];
        vs = vs + dx[i - 1] * dy[j - 1] * dz[k - 1];
      }
  const double mean_value = ws / vs;
  for (size_t n = 0; n < ncell; ++n) phi[n] = phi[n] - mean_value;
  *final_residual = nonuniform_residual(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx,
                                        dy, dz, nx, ny, nz);
  free(next);
  return 0;
}

/* compute_fpm_terms_uniform_3d, src/fpm_terms_3d.f90:8-37 */
int orc_compute_fpm_terms_uniform_3d(const double *phi, const double *u, const double *v,
                                     const double *w, const double *ox, const double *oy,
                                     const double *oz, const double *ax, const double *ay,
                                     const double *az, const int32_t *fluid, double *gx,
                                     double *gy, double *gz, double *vortex, double *advection,
                                     double totals[2], int nx, int ny, int nz, double dx,
                                     double dy, double dz) {
  double vt = 0, at = 0;
  const double volume = dx * dy * dz;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        gx[c] = (phi[PL(i + 1, j, k)] - phi[PL(i - 1, j, k)]) / (2.0 * dx);
        gy[c] = (phi[PL(i, j + 1, k)] - phi[PL(i, j - 1, k)]) / (2.0 * dy);
        gz[c] = (phi[PL(i, j, k + 1)] - phi[PL(i, j, k - 1)]) / (2.0 * dz);
        const double cx = w[c] * oy[c] - v[c] * oz[c];
        const double cy = u[c] * oz[c] - w[c] * ox[c];
        const double cz = v[c] * ox[c] - u[c] * oy[c];
        const double mask = (double)fluid[c];
        vortex[c] = -mask * (cx * gx[c] + cy * gy[c] + cz * gz[c]);
        advection[c] = -mask * (ax[c] * gx[c] + ay[c] * gy[c] + az[c] * gz[c]);
        vt = vt + volume * vortex[c];
        at = at + volume * advection[c];
      }
  totals[0] = vt;
  totals[1] = at;
  return 0;
}

/* compute_fpm_terms_nonuniform_3d, src/fpm_terms_3d.f90:39-87 */
int orc_compute_fpm_terms_nonuniform_3d(const double *phi, const double *u, const double *v,
                                        const double *w, const double *ox, const double *oy,
                                        const double *oz, const double *ax, const double *ay,
                                        const double *az, const int32_t *fluid, const double *xc,
                                        const double *yc, const double *zc, const double *dx,
                                        const double *dy, const double *dz, const double *xlo,
                                        const double *xhi, const double *ylo, const double *yhi,
                                        const double *zlo, const double *zhi,
                                        const double bounds[6], double *gx, double *gy,
                                        double *gz, double *vortex, double *advection,
                                        double totals[2], int nx, int ny, int nz) {
  if (nx < 2 || ny < 2 || nz < 2) return -1;
  double vt = 0, at = 0, we, ea, so, no, bo, to;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = CL(i, j, k);
        if (i == 1) we = xlo[FXI(j, k)];
        else we = (phi[c] - phi[CL(i - 1, j, k)]) / (xc[i - 1] - xc[i - 2]);
        if (i == nx) ea = xhi[FXI(j, k)];
        else ea = (phi[CL(i + 1, j, k)] - phi[c]) / (xc[i] - xc[i - 1]);
        if (j == 1) so = ylo[FYI(i, k)];
        else so = (phi[c] - phi[CL(i, j - 1, k)]) / (yc[j - 1] - yc[j - 2]);
        if (j == ny) no = yhi[FYI(i, k)];
        else no = (phi[CL(i, j + 1, k)] - phi[c]) / (yc[j] - yc[j - 1]);
        if (k == 1) bo = zlo[FZI(i, j)];
        else bo = (phi[c] - phi[CL(i, j, k - 1)]) / (zc[k - 1] - zc[k - 2]);
        if (k == nz) to = zhi[FZI(i, j)];
        else to = (phi[CL(i, j, k + 1)] - phi[c]) / (zc[k] - zc[k - 1]);
        gx[c] = 0.5 * (we + ea);
        gy[c] = 0.5 * (so + no);
        gz[c] = 0.5 * (bo + to);
        const double cx = w[c] * oy[c] - v[c] * oz[c];
        const double cy = u[c] * oz[c] - w[c] * ox[c];
        const double cz = v[c] * ox[c] - u[c] * oy[c];
        const double mask = (double)fluid[c];
        vortex[c] = -mask * (cx * gx[c] + cy * gy[c] + cz * gz[c]);
        advection[c] = -mask * (ax[c] * gx[c] + ay[c] * gy[c] + az[c] * gz[c]);
        const double inside = (xc[i - 1] > bounds[0] && xc[i - 1] < bounds[1] &&
                               yc[j - 1] > bounds[2] && yc[j - 1] < bounds[3] &&
                               zc[k - 1] > bounds[4] && zc[k - 1] < bounds[5])
                                  ? 1.0
                                  : 0.0;
        const double volume = inside * dx[i - 1] * dy[j - 1] * dz[k - 1];
        vt = vt + volume * vortex[c];
        at = at + volume * advection[c];
      }
  totals[0] = vt;
  totals[1] = at;
  return 0;
}


/* ------------------------------------------------------------------------
 * fpm_partition_terms_2d, src/fpm_partition_2d.f90:35-154: the force
 * partition of a 2-D stair-step body with rigid in-plane motion (F_KINE,
 * F_VIF, F_VIS, F_BOUND), two serial passes over the cells (j outer, i
 * inner) in the reference's order, every expression left to right.
 * ---------------------------------------------------------------------- */
typedef struct {
  int ii, jj, axis;
  double sgn, hw, len, d, fx, fy;
} P2Face;

/* face(i,j,f), src/fpm_partition_2d.f90:122-140 (1-based i, j; f = 1 east,
 * 2 west, 3 north, 4 south) */
static P2Face p2_face(int i, int j, int f, int nx, int ny, const double *xc, const double *yc,
                      const double *dx, const double *dy) {
  P2Face o;
  switch (f) {
    case 1: o.ii = i + 1; o.jj = j; o.axis = 1; o.sgn = 1; break;
    case 2: o.ii = i - 1; o.jj = j; o.axis = 1; o.sgn = -1; break;
    case 3: o.ii = i; o.jj = j + 1; o.axis = 2; o.sgn = 1; break;
    default: o.ii = i; o.jj = j - 1; o.axis = 2; o.sgn = -1; break;
  }
  if (o.axis == 1) {
    o.hw = 0.5 * dx[i - 1];
    o.len = dy[j - 1];
    o.fx = xc[i - 1] + o.sgn * o.hw;
    o.fy = yc[j - 1];
    o.d = o.hw;
    if (o.ii >= 1 && o.ii <= nx) o.d = fabs(xc[o.ii - 1] - xc[i - 1]);
  } else {
    o.hw = 0.5 * dy[j - 1];
    o.len = dx[i - 1];
    o.fx = xc[i - 1];
    o.fy = yc[j - 1] + o.sgn * o.hw;
    o.d = o.hw;
    if (o.jj >= 1 && o.jj <= ny) o.d = fabs(yc[o.jj - 1] - yc[j - 1]);
  }
  return o;
}

/* ring_index, src/fpm_partition_2d.f90:142-151 (1-based) */
static int p2_ring_index(int i, int j, int f, int nx, int ny) {
  switch (f) {
    case 2: return j;
    case 1: return ny + j;
    case 4: return 2 * ny + i;
    default: return 2 * ny + nx + i;
  }
}

int orc_partition_terms_2d(const double *phi, const double g[2], const double *u,
                           const double *v, const int32_t *solid, const double *xf,
                           const double *yf, const double kin[6], const double *ring, double dt,
                           double nu, int nx, int ny, int slip_body, double terms[4],
                           double *vif_density, double *omega) {
  if (nx < 1 || ny < 1) return -1;
  const size_t n = (size_t)nx * ny;
  const size_t L = 2 * (size_t)nx + 2 * (size_t)ny;
  double *xc = malloc(sizeof(double) * nx), *yc = malloc(sizeof(double) * ny);
  double *dx = malloc(sizeof(double) * nx), *dy = malloc(sizeof(double) * ny);
  double *nxf = calloc(n, sizeof(double)), *nyf = calloc(n, sizeof(double));
  double *gpx = calloc(n, sizeof(double)), *gpy = calloc(n, sizeof(double));
  for (int i = 0; i < nx; ++i) {
    dx[i] = xf[i + 1] - xf[i];
    xc[i] = 0.5 * (xf[i] + xf[i + 1]);
  }
  for (int j = 0; j < ny; ++j) {
    dy[j] = yf[j + 1] - yf[j];
    yc[j] = 0.5 * (yf[j] + yf[j + 1]);
  }
  for (int t = 0; t < 4; ++t) terms[t] = 0;
  for (size_t c = 0; c < n; ++c) vif_density[c] = 0, omega[c] = 0;
#define AT(a, i, j) a[(size_t)((i) - 1) + (size_t)nx * ((j) - 1)]
  /* pass 1 (src/fpm_partition_2d.f90:60-84) */
  for (int j = 1; j <= ny; ++j) {
    for (int i = 1; i <= nx; ++i) {
      if (AT(solid, i, j)) continue;
      double fu[4], fv[4], fp[4];
      for (int f = 1; f <= 4; ++f) {
        const P2Face o = p2_face(i, j, f, nx, ny, xc, yc, dx, dy);
        if (o.ii < 1 || o.ii > nx || o.jj < 1 || o.jj > ny) {
          fu[f - 1] = AT(u, i, j);
          fv[f - 1] = AT(v, i, j);
          fp[f - 1] = 0;
        } else if (AT(solid, o.ii, o.jj)) {
          if (slip_body) {
            fu[f - 1] = AT(u, i, j);
            fv[f - 1] = AT(v, i, j);
          } else {
            fu[f - 1] = kin[0] - kin[2] * o.fy;
            fv[f - 1] = kin[1] + kin[2] * o.fx;
          }
          fp[f - 1] = g[o.axis - 1];
        } else {
          fu[f - 1] = AT(u, i, j) + (AT(u, o.ii, o.jj) - AT(u, i, j)) * o.hw / o.d;
          fv[f - 1] = AT(v, i, j) + (AT(v, o.ii, o.jj) - AT(v, i, j)) * o.hw / o.d;
          fp[f - 1] = o.sgn * (AT(phi, o.ii, o.jj) - AT(phi, i, j)) / o.d;
        }
      }
      AT(nxf, i, j) = AT(u, i, j) * (fu[0] - fu[1]) / dx[i - 1] +
                      AT(v, i, j) * (fu[2] - fu[3]) / dy[j - 1];
      AT(nyf, i, j) = AT(u, i, j) * (fv[0] - fv[1]) / dx[i - 1] +
                      AT(v, i, j) * (fv[2] - fv[3]) / dy[j - 1];
      AT(omega, i, j) = (fv[0] - fv[1]) / dx[i - 1] - (fu[2] - fu[3]) / dy[j - 1];
      AT(gpx, i, j) = 0.5 * (fp[0] + fp[1]);
      AT(gpy, i, j) = 0.5 * (fp[2] + fp[3]);
    }
  }
  /* pass 2 (src/fpm_partition_2d.f90:85-117) */
  for (int j = 1; j <= ny; ++j) {
    for (int i = 1; i <= nx; ++i) {
      if (AT(solid, i, j)) continue;
      double div = 0, lapu = 0, lapv = 0;
      for (int f = 1; f <= 4; ++f) {
        const P2Face o = p2_face(i, j, f, nx, ny, xc, yc, dx, dy);
        if (o.ii < 1 || o.ii > nx || o.jj < 1 || o.jj > ny) {
          const double nn = o.axis == 1 ? AT(nxf, i, j) : AT(nyf, i, j);
          div = div + o.sgn * nn * o.len;
          const int m = p2_ring_index(i, j, f, nx, ny);
          const double dudt = (ring[(m - 1) + L * 2] - ring[(m - 1)]) / (2 * dt);
          const double dvdt = (ring[(m - 1) + L * 3] - ring[(m - 1) + L]) / (2 * dt);
          const double sel = o.axis == 1 ? dudt + AT(nxf, i, j) : dvdt + AT(nyf, i, j);
          terms[3] = terms[3] - o.sgn * sel * AT(phi, i, j) * o.len;
        } else if (AT(solid, o.ii, o.jj)) {
          const double ubx = kin[0] - kin[2] * o.fy, uby = kin[1] + kin[2] * o.fx;
          const double ax = kin[3] - kin[5] * o.fy - kin[2] * kin[2] * o.fx;
          const double ay = kin[4] + kin[5] * o.fx - kin[2] * kin[2] * o.fy;
          const double phif = AT(phi, i, j) + o.sgn * o.hw * g[o.axis - 1];
          terms[0] = terms[0] - o.sgn * (o.axis == 1 ? ax : ay) * phif * o.len;
          if (slip_body) {
            div = div + o.sgn * (o.axis == 1 ? AT(nxf, i, j) : AT(nyf, i, j)) * o.len;
          } else {
            div = div + o.sgn * (o.axis == 1 ? -(kin[2] * uby) : kin[2] * ubx) * o.len;
            lapu = lapu + (ubx - AT(u, i, j)) / o.hw * o.len;
            lapv = lapv + (uby - AT(v, i, j)) / o.hw * o.len;
          }
        } else {
          const double sel =
              o.axis == 1 ? AT(nxf, i, j) + (AT(nxf, o.ii, o.jj) - AT(nxf, i, j)) * o.hw / o.d
                          : AT(nyf, i, j) + (AT(nyf, o.ii, o.jj) - AT(nyf, i, j)) * o.hw / o.d;
          div = div + o.sgn * sel * o.len;
          lapu = lapu + (AT(u, o.ii, o.jj) - AT(u, i, j)) / o.d * o.len;
          lapv = lapv + (AT(v, o.ii, o.jj) - AT(v, i, j)) / o.d * o.len;
        }
      }
      AT(vif_density, i, j) = AT(phi, i, j) * div / (dx[i - 1] * dy[j - 1]);
      terms[1] = terms[1] + AT(phi, i, j) * div;
      terms[2] = terms[2] + nu * (lapu * AT(gpx, i, j) + lapv * AT(gpy, i, j));
    }
  }
#undef AT
  free(xc); free(yc); free(dx); free(dy); free(nxf); free(nyf); free(gpx); free(gpy);
  return 0;
}
