/*
 * CPU oracle for the stair-step influence-potential solve, the vortex-force
 * densities and the no-body Neumann solvers.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C FP64 restatement of the
 * reference's serial (OpenACC-off) arithmetic, kept in the reference's loop
 * order (k outer, j, i inner) and serial summation order so that it can be
 * compared with the reference CPU build. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it, and only as the checker or the
 * timed CPU baseline; the product library never calls it.
 *
 * Array layouts are the reference's Fortran (column-major) layouts:
 *   phi, solid, next : (0:nx+1, 0:ny+1, 1:nz)   x fastest, one x-y halo layer
 *   rhs, fields      : (1:nx, 1:ny, 1:nz)
 *   xlo, xhi         : (ny, nz);  ylo, yhi : (nx, nz);  zlo, zhi : (nx, ny)
 *   xc (0:nx+1), yc (0:ny+1), zc (nz), dx (nx), dy (ny), dz (nz)
 * Neighbour ranks follow MPI: a negative value (MPI_PROC_NULL in every MPI
 * implementation) means a physical boundary. Only one rank is modelled.
 *
 * Every function returns 0 on success or a negative error code:
 *   -1 bad argument, -2 isolated fluid cell, -3 no fluid outer face,
 *   -4 neighbour rank present (multi-rank not modelled).
 */
#ifndef FFPM_ORACLE_H
#define FFPM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* src/fpm_stairstep_mpi_3d.f90:25-36 */
int orc_exchange_solid_halo_3d(int32_t *solid, int nx, int ny, int nz,
                               int west, int east, int south, int north);

/* src/fpm_stairstep_mpi_3d.f90:41-92 */
int orc_enforce_stairstep_compatibility_3d(
    const double *rhs, const int32_t *solid, const double body_grad[3],
    double *xlo, double *xhi, double *ylo, double *yhi, double *zlo, double *zhi,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    int west, int east, int south, int north, double *correction, double *imbalance);

/* src/fpm_stairstep_mpi_3d.f90:106-330. method: 0 = pcg, 1 = jacobi. */
int orc_solve_stairstep_poisson_3d(
    double *phi, const double *rhs, const int32_t *solid, const double body_grad[3],
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz,
    int nx, int ny, int nz, int west, int east, int south, int north,
    double tolerance, int max_sweeps, int check_interval, int method,
    int *sweeps, double *residual);

/* src/fpm_stairstep_mpi_3d.f90:403-419 */
int orc_stairstep_residual_3d(
    const double *phi, const double *rhs, const int32_t *solid, const double body_grad[3],
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz,
    int nx, int ny, int nz, int west, int east, int south, int north, double *value);

/* src/fpm_stairstep_mpi_3d.f90:426-469 */
int orc_compute_stairstep_fpm_terms_3d(
    const double *phi, const double *u, const double *v, const double *w,
    const double *ox, const double *oy, const double *oz,
    const double *ax, const double *ay, const double *az,
    const int32_t *solid, const double body_grad[3],
    const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double bounds[6],
    double *gx, double *gy, double *gz, double *vortex, double *advection,
    double totals[2], int nx, int ny, int nz, int west, int east, int south, int north);

/* Operator pieces exposed for invariant tests (src/fpm_stairstep_mpi_3d.f90:336-379,
 * 499-510, 514-530). src/diag/rdiag are (nx,ny,nz); cw..ct are 1-D. */
int orc_build_operator_3d(
    const double *rhs, const int32_t *solid, const double body_grad[3],
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz,
    int nx, int ny, int nz, int west, int east, int south, int north, int for_pcg,
    double *src, double *diag, double *rdiag, double *cw, double *ce, double *cs,
    double *cn, double *cb, double *ct, int *isolated);
int orc_jacobi_sweep_3d(const double *a, double *b, const double *src, const double *rdiag,
                        const double *cw, const double *ce, const double *cs, const double *cn,
                        const double *cb, const double *ct, int nx, int ny, int nz);
double orc_operator_residual_3d(const double *phi, const double *src, const double *diag,
                                const double *cw, const double *ce, const double *cs,
                                const double *cn, const double *cb, const double *ct,
                                int nx, int ny, int nz);

/* src/fpm_laplace_3d.f90:9-53 */
int orc_enforce_neumann_compatibility_3d(const double *rhs, double *xlo, double *xhi,
    double *ylo, double *yhi, double *zlo, double *zhi, int nx, int ny, int nz,
    double dx, double dy, double dz, double *correction, double *imbalance);
/* src/fpm_laplace_3d.f90:55-148; phi is (0:nx+1,0:ny+1,0:nz+1) */
int orc_solve_neumann_poisson_3d(double *phi, const double *rhs, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, int nx, int ny, int nz, double dx, double dy, double dz,
    double tolerance, int max_iterations, int check_interval, int *iterations,
    double *final_residual);
/* src/fpm_laplace_3d.f90:150-217 */
int orc_enforce_neumann_compatibility_nonuniform_3d(const double *rhs, double *xlo,
    double *xhi, double *ylo, double *yhi, double *zlo, double *zhi, const double *dx,
    const double *dy, const double *dz, int nx, int ny, int nz, double *correction,
    double *imbalance);
/* src/fpm_laplace_3d.f90:219-315; phi is (nx,ny,nz), xc(nx), yc(ny), zc(nz) */
int orc_solve_neumann_poisson_nonuniform_3d(double *phi, const double *rhs,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz,
    int nx, int ny, int nz, double tolerance, int max_iterations, int check_interval,
    int *iterations, double *final_residual);

/* src/fpm_terms_3d.f90:8-37; phi is (0:nx+1,0:ny+1,0:nz+1), fluid (nx,ny,nz) */
int orc_compute_fpm_terms_uniform_3d(const double *phi, const double *u, const double *v,
    const double *w, const double *ox, const double *oy, const double *oz,
    const double *ax, const double *ay, const double *az, const int32_t *fluid,
    double *gx, double *gy, double *gz, double *vortex, double *advection,
    double totals[2], int nx, int ny, int nz, double dx, double dy, double dz);
/* src/fpm_terms_3d.f90:39-87; phi is (nx,ny,nz) */
int orc_compute_fpm_terms_nonuniform_3d(const double *phi, const double *u,
    const double *v, const double *w, const double *ox, const double *oy,
    const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *fluid, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, const double bounds[6], double *gx, double *gy, double *gz,
    double *vortex, double *advection, double totals[2], int nx, int ny, int nz);

/* src/fpm_partition_2d.f90:35-154; phi, u, v, solid (0/1), vif_density, omega
 * are (nx,ny); xf (0:nx), yf (0:ny); kin(6); ring (2nx+2ny, 2, 2);
 * terms = [KINE, VIF, VIS, BOUND] */
int orc_partition_terms_2d(const double *phi, const double g[2], const double *u,
    const double *v, const int32_t *solid, const double *xf, const double *yf,
    const double kin[6], const double *ring, double dt, double nu, int nx, int ny,
    int slip_body, double terms[4], double *vif_density, double *omega);

#ifdef __cplusplus
}
#endif
#endif
