// Kernel declarations and the device-resident solver state.
#pragma once

#include "ffpm_common.hpp"

namespace ffpm {

constexpr int kBreakdown = 1;  // pq <= 0 or rz <= 0 (or NaN): restart from the true residual
constexpr int kNegative = 2;   // rz_new < 0 or NaN: leave the iteration

// Neighbour code of a cell (one byte): which of its six faces lead to a fluid
// neighbour (and so contribute a coefficient to diag), and whether the cell
// itself is fluid.
constexpr unsigned kCodeW = 1u, kCodeE = 2u, kCodeS = 4u, kCodeN = 8u, kCodeB = 16u,
                   kCodeT = 32u, kCodeFluid = 64u;

// direction modes of the fused PCG SpMV
constexpr int kDirUpdate = 0;   // p = M^-1 r + beta p_old
constexpr int kDirReset = 1;    // p = M^-1 r
constexpr int kDirCurrent = 2;  // p_old already holds p

// Scalars the iteration keeps on the device, so that a chunk of iterations
// runs without a host round trip; the host reads them at convergence checks.
struct Scalars {
  double rz, pq, rz_new, alpha, beta;
  double res;         // max-norm residual of the last residual kernel
  double red1, red2;  // sum of the true residual and the fluid-cell count
  double sums[4];     // generic reduction results
  int status;         // 0, kBreakdown or kNegative
  int iters;          // iterations completed since the last chunk start
  double alpha_pend;  // step length of a direction whose x update is deferred
  // The multigrid cycle's level-0 vectors are stored in FP16 scaled by a
  // power of two: sig for the cycle in flight (its first step, iterates and
  // output z carry that factor), sig_next for the next one (from the max of
  // the last output). 0: not set yet (the next FIN_TR1 sets it).
  double sig, sig_next;
  // q = A p of the last update is stored in FP16 times tau (a power of two):
  // |q| <= anorm max |p|, anorm a bound on the infinity norm of A (set with
  // the cycle), pmax a bound on max |p| carried from direction to direction
  double tau, pmax, anorm;
};

// Finish operations of k_finish (what the partials of the preceding
// reduction kernel mean and where their total goes).
enum FinishOp : int {
  FIN_MAX = 0,    // res = max(slot 0)
  FIN_TR1 = 1,    // red1 = sum(slot 0), red2 = sum(slot 1), res = max(slot 2); an unset
                  // sig from max(slot 3) = max |r / V| / diag
  FIN_RZ = 2,     // rz = sum(slot 0)
  FIN_PQ = 3,     // pq = sum(slot 0); breakdown test; alpha = rz / pq
  FIN_RZNEW = 4,  // rz_new = sum(slot 0); rz_new < 0 test; beta; rz; iters += 1
  FIN_SUMS = 5,   // sums[s] = sum(slot s), s < nslots
  FIN_FCG1 = 6,   // rz = sum(slot 0) / sig^2 = z.r; beta = -sum(slot 1) / (sig tau pq) =
                  // -z.(A p_old) / pq (slots of the scaled z and q); sig_next from
                  // max(slot 2); pmax and tau for the next stored q
};

struct Bounds {
  double b[6];
};

__global__ void k_axis_coefficients(int nx, int ny, int nz, const double *xc, const double *yc,
                                    const double *zc, const double *dx, const double *dy,
                                    const double *dz, double *cw, double *ce, double *cs,
                                    double *cn, double *cb, double *ct);
__global__ void k_build_operator(Geom g, const int32_t *solid, const double *rhs,
                                 const double *xlo, const double *xhi, const double *ylo,
                                 const double *yhi, const double *zlo, const double *zhi,
                                 double bg0, double bg1, double bg2, int for_pcg, double *src,
                                 double *diag, double *rdiag, uint8_t *code, int *isolated);
__global__ void k_pack_phi(int nx, int ny, int nz, const int32_t *solid, const double *phi_ref,
                           double *phi);
__global__ void k_unpack_phi(int nx, int ny, int nz, const double *phi, double *phi_ref);
__global__ void k_jacobi(Tiling t, Geom g, const double *a, double *b, const double *src,
                         const double *rdiag);
__global__ void k_residual_max(Tiling t, Geom g, const double *phi, const double *src,
                               const double *diag, double *partials, Scalars *sc);
__global__ void k_true_residual_1(Tiling t, Geom g, const double *phi, const double *src,
                                  const double *diag, double *r, double *partials, Scalars *sc);
__global__ void k_true_residual_2(size_t n, double *r, const double *rdiag, const double *diag,
                                  double *p, int reset, double *z, double *partials, Scalars *sc);
__global__ void k_spmv_dot(Tiling t, Geom g, const double *p, double *q, const double *diag,
                           double *partials, Scalars *sc);
__global__ void k_update_dot(size_t n, double *phi, const double *p, double *r, const double *q,
                             const double *rdiag, double *partials, Scalars *sc);
__global__ void k_direction(size_t n, double *p, const double *r, const double *rdiag,
                            const Scalars *sc);
__global__ void k_finish(int op, int nslots, const double *partials, int nblocks, Scalars *sc);
__global__ void k_chunk_begin(Scalars *sc);
__global__ void k_clear_status(Scalars *sc);
__global__ void k_gauge_sums(Geom g, const int32_t *solid, const double *phi, double *partials,
                             Scalars *sc);
__global__ void k_gauge_apply(int nx, int ny, int nz, const int32_t *solid, double *phi,
                              double *phi_ref, const Scalars *sc);
__global__ void k_compat_sums(Geom g, const int32_t *solid, const double *rhs, const double *xlo,
                              const double *xhi, const double *ylo, const double *yhi,
                              const double *zlo, const double *zhi, double bg0, double bg1,
                              double bg2, double *partials, Scalars *sc);
__global__ void k_shift_faces(int nx, int ny, int nz, double *xlo, double *xhi, double *ylo,
                              double *yhi, double *zlo, double *zhi, double c);
__global__ void k_face_residual(Geom g, const double *phi, const double *rhs,
                                const int32_t *solid, const double *xlo, const double *xhi,
                                const double *ylo, const double *yhi, const double *zlo,
                                const double *zhi, double bg0, double bg1, double bg2,
                                double *partials, Scalars *sc);
__global__ void k_terms(Geom g, const double *phi, const int32_t *solid, const double *u,
                        const double *v, const double *w, const double *ox, const double *oy,
                        const double *oz, const double *ax, const double *ay, const double *az,
                        const double *xlo, const double *xhi, const double *ylo,
                        const double *yhi, const double *zlo, const double *zhi, double bg0,
                        double bg1, double bg2, Bounds bd, double *gx, double *gy, double *gz,
                        double *vortex, double *advection, double *partials, Scalars *sc);

// The per-direction arguments of k_terms_dirs (directions 0..ND-1 used).
struct TermsDirs {
  const double *phi[3];
  const double *xlo[3], *xhi[3], *ylo[3], *yhi[3], *zlo[3], *zhi[3];
  double bg[3][3];
  double *gx[3], *gy[3], *gz[3], *vortex[3], *advection[3];
};

template <int ND>
__global__ void k_terms_dirs(Geom g, TermsDirs a, const int32_t *solid, const double *u,
                             const double *v, const double *w, const double *ox,
                             const double *oy, const double *oz, const double *ax,
                             const double *ay, const double *az, Bounds bd, double *partials);
__global__ void k_terms_dirs_finish(int nd, const double *partials, int nblocks, double *out);

}  // namespace ffpm
