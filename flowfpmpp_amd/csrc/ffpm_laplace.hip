// No-body pure-Neumann solvers (src/fpm_laplace_3d.f90) and their FPM
// densities (src/fpm_terms_3d.f90) on gfx950, FP64, with the C ABI.
//
// The uniform solver sweeps undamped Jacobi with ghost substitution at the
// walls; the non-uniform one sweeps 0.2/0.8-damped Jacobi with its face data
// folded per sweep. Both iterate on dense (nx,ny,nz) ping-pong buffers (the
// reference's copy-back kernel becomes a pointer swap; values are identical)
// and check the max-norm residual every check_interval sweeps.
#include <cfloat>
#include <cstdint>
#include <cstdlib>
#include <initializer_list>
#include <string>

#include "../../include/ffpm.h"
#include "ffpm_pcg_body.hpp"
#include "ffpm_runtime.hpp"

namespace ffpm {

namespace {

constexpr int kThreads = 256;

struct Faces {
  const double *xlo, *xhi, *ylo, *yhi, *zlo, *zhi;
};

__device__ __forceinline__ void cell_of(size_t c, int nx, int ny, int &i, int &j, int &k) {
  i = (int)(c % nx);
  j = (int)((c / nx) % ny);
  k = (int)(c / ((size_t)nx * ny));
}

// Neighbour values with the wall ghost substitution of
// src/fpm_laplace_3d.f90:76-84 (also residual_norm, :119-127).
struct Nb {
  double p, xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Nb uniform_neighbours(const double *a, size_t c, int i, int j, int k,
                                                 int nx, int ny, int nz, const Faces &f,
                                                 double dx, double dy, double dz) {
  Nb n;
  const size_t sy = nx, sz = (size_t)nx * ny;
  n.p = a[c];
  n.xm = a[i > 0 ? c - 1 : c];
  n.xp = a[i < nx - 1 ? c + 1 : c];
  n.ym = a[j > 0 ? c - sy : c];
  n.yp = a[j < ny - 1 ? c + sy : c];
  n.zm = a[k > 0 ? c - sz : c];
  n.zp = a[k < nz - 1 ? c + sz : c];
  if (i == 0) n.xm = n.p - f.xlo[j + (size_t)ny * k] * dx;
  if (i == nx - 1) n.xp = n.p + f.xhi[j + (size_t)ny * k] * dx;
  if (j == 0) n.ym = n.p - f.ylo[i + (size_t)nx * k] * dy;
  if (j == ny - 1) n.yp = n.p + f.yhi[i + (size_t)nx * k] * dy;
  if (k == 0) n.zm = n.p - f.zlo[i + (size_t)nx * j] * dz;
  if (k == nz - 1) n.zp = n.p + f.zhi[i + (size_t)nx * j] * dz;
  return n;
}

struct Uniform {
  int nx, ny, nz;
  double dx, dy, dz, idx2, idy2, idz2, invdiag, rdx2, rdy2, rdz2;
};

// Sweep of src/fpm_laplace_3d.f90:74-87.
__global__ __launch_bounds__(256) void k_lap_uniform_sweep(Uniform u, Faces f,
                                                           const double *__restrict__ a,
                                                           double *__restrict__ b,
                                                           const double *__restrict__ rhs) {
  const size_t n = (size_t)u.nx * u.ny * u.nz;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(c, u.nx, u.ny, i, j, k);
    const Nb nb = uniform_neighbours(a, c, i, j, k, u.nx, u.ny, u.nz, f, u.dx, u.dy, u.dz);
    b[c] = ((nb.xm + nb.xp) * u.idx2 + (nb.ym + nb.yp) * u.idy2 + (nb.zm + nb.zp) * u.idz2 -
            rhs[c]) *
           u.invdiag;
  }
}

// residual_norm, src/fpm_laplace_3d.f90:109-132 -> sc->res.
__global__ __launch_bounds__(256) void k_lap_uniform_residual(Uniform u, Faces f,
                                                              const double *__restrict__ a,
                                                              const double *__restrict__ rhs,
                                                              double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const size_t n = (size_t)u.nx * u.ny * u.nz;
  double m = 0.0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(c, u.nx, u.ny, i, j, k);
    const Nb nb = uniform_neighbours(a, c, i, j, k, u.nx, u.ny, u.nz, f, u.dx, u.dy, u.dz);
    const double res = (nb.xm - 2 * nb.p + nb.xp) / u.rdx2 + (nb.ym - 2 * nb.p + nb.yp) / u.rdy2 +
                       (nb.zm - 2 * nb.p + nb.zp) / u.rdz2 - rhs[c];
    m = fmax(m, fabs(res));
  }
  block_max(m, lds);
  store_partials(partials, 1, &m, gridDim.x);
}

// Sum of phi (uniform) or of phi*V and V (weights given) -> sc->sums[0..1].
__global__ __launch_bounds__(256) void k_lap_mean_sums(int nx, int ny, int nz,
                                                       const double *__restrict__ a,
                                                       const double *__restrict__ dx,
                                                       const double *__restrict__ dy,
                                                       const double *__restrict__ dz,
                                                       double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const size_t n = (size_t)nx * ny * nz;
  double s[2] = {0.0, 0.0};
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    if (dx) {
      int i, j, k;
      cell_of(c, nx, ny, i, j, k);
      s[0] += a[c] * dx[i] * dy[j] * dz[k];
      s[1] += dx[i] * dy[j] * dz[k];
    } else {
      s[0] += a[c];
    }
  }
  block_sum<2>(s, lds);
  store_partials(partials, 2, s, gridDim.x);
}

// mean_value = sums[0] / divisor (divisor < 0: sums[1]); a -= mean_value.
__global__ void k_lap_subtract_mean(size_t n, double *__restrict__ a, double divisor,
                                    const Scalars *sc) {
  const double mean_value = sc->sums[0] / (divisor < 0 ? sc->sums[1] : divisor);
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    a[c] = a[c] - mean_value;
}

// Dense interior <-> the caller's (0:nx+1,0:ny+1,0:nz+1) array.
__global__ void k_lap_pack(int nx, int ny, int nz, const double *__restrict__ ref,
                           double *__restrict__ a) {
  const size_t n = (size_t)nx * ny * nz;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(c, nx, ny, i, j, k);
    a[c] = ref[(size_t)(i + 1) + (size_t)(nx + 2) * ((size_t)(j + 1) + (size_t)(ny + 2) * (k + 1))];
  }
}

// Interior plus set_ghosts (src/fpm_laplace_3d.f90:134-148); edges and
// corners of the caller's array are left as they were.
__global__ void k_lap_unpack_ghosts(int nx, int ny, int nz, Faces f, double dx, double dy,
                                    double dz, const double *__restrict__ a,
                                    double *__restrict__ ref) {
  const size_t sx = (size_t)nx + 2, sy = (size_t)ny + 2;
  const size_t n = sx * sy * ((size_t)nz + 2);
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const int I = (int)(p % sx), J = (int)((p / sx) % sy), K = (int)(p / (sx * sy));
    const bool ix = I >= 1 && I <= nx, jy = J >= 1 && J <= ny, kz = K >= 1 && K <= nz;
    auto A = [&](int i, int j, int k) {
      return a[(size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1))];
    };
    if (ix && jy && kz) {
      ref[p] = A(I, J, K);
    } else if (jy && kz && (I == 0 || I == nx + 1)) {
      ref[p] = I == 0 ? A(1, J, K) - f.xlo[(J - 1) + (size_t)ny * (K - 1)] * dx
                      : A(nx, J, K) + f.xhi[(J - 1) + (size_t)ny * (K - 1)] * dx;
    } else if (ix && kz && (J == 0 || J == ny + 1)) {
      ref[p] = J == 0 ? A(I, 1, K) - f.ylo[(I - 1) + (size_t)nx * (K - 1)] * dy
                      : A(I, ny, K) + f.yhi[(I - 1) + (size_t)nx * (K - 1)] * dy;
    } else if (ix && jy && (K == 0 || K == nz + 1)) {
      ref[p] = K == 0 ? A(I, J, 1) - f.zlo[(I - 1) + (size_t)nx * (J - 1)] * dz
                      : A(I, J, nz) + f.zhi[(I - 1) + (size_t)nx * (J - 1)] * dz;
    }
  }
}

// Per-axis neighbour weights of the non-uniform sweep
// (src/fpm_laplace_3d.f90:241-246): lo[m] = 1/(w_m (c_m - c_{m-1})) for m > 1,
// hi[m] = 1/(w_m (c_{m+1} - c_m)) for m < n, 0 otherwise.
__global__ void k_lap_axis_weights(int n, const double *c, const double *w, double *lo,
                                   double *hi) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  lo[m] = 0.0;
  hi[m] = 0.0;
  if (m > 0) lo[m] = 1.0 / (w[m] * (c[m] - c[m - 1]));
  if (m < n - 1) hi[m] = 1.0 / (w[m] * (c[m + 1] - c[m]));
}

struct NonUniform {
  int nx, ny, nz;
  const double *aw, *ae, *as, *an, *ab, *at;  // per-axis weights
  const double *xc, *yc, *zc, *dx, *dy, *dz;
};

// Damped sweep of src/fpm_laplace_3d.f90:238-264.
__global__ __launch_bounds__(256) void k_lap_nonuniform_sweep(NonUniform g, Faces f,
                                                              const double *__restrict__ a,
                                                              double *__restrict__ b,
                                                              const double *__restrict__ rhs) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz, sy = nx, sz = (size_t)nx * ny;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(c, nx, ny, i, j, k);
    const double aw = g.aw[i], ae = g.ae[i], as = g.as[j], an = g.an[j], ab = g.ab[k],
                 at = g.at[k];
    const double diag = aw + ae + as + an + ab + at;
    double numer = -rhs[c];
    if (i > 0) numer = numer + aw * a[c - 1];
    if (i < nx - 1) numer = numer + ae * a[c + 1];
    if (j > 0) numer = numer + as * a[c - sy];
    if (j < ny - 1) numer = numer + an * a[c + sy];
    if (k > 0) numer = numer + ab * a[c - sz];
    if (k < nz - 1) numer = numer + at * a[c + sz];
    if (i == 0) numer = numer - f.xlo[j + (size_t)ny * k] / g.dx[i];
    if (i == nx - 1) numer = numer + f.xhi[j + (size_t)ny * k] / g.dx[i];
    if (j == 0) numer = numer - f.ylo[i + (size_t)nx * k] / g.dy[j];
    if (j == ny - 1) numer = numer + f.yhi[i + (size_t)nx * k] / g.dy[j];
    if (k == 0) numer = numer - f.zlo[i + (size_t)nx * j] / g.dz[k];
    if (k == nz - 1) numer = numer + f.zhi[i + (size_t)nx * j] / g.dz[k];
    b[c] = 0.2 * a[c] + 0.8 * numer / diag;
  }
}

// FFPM_LAP=classic: the one-cell-per-thread sweeps (the same bits).
bool lap_classic() {
  const char *e = std::getenv("FFPM_LAP");
  return e && std::string(e) == "classic";
}

// Both sweeps, and both max-norm residuals, in the streaming scheme of the
// stair-step kernels: a workgroup owns a tile of 64*CL x rows cells (one
// wavefront per x row, CL adjacent cells per lane in one 16-byte vector for
// CL = 2) and a chunk of z planes, which it walks holding the planes below
// and above in registers. The loads of plane k+1 (south and north rows, the
// edge lanes' west and east cells, rhs, the face data of boundary rows) and
// the own row of plane k+2 are issued before plane k is computed; west and
// east neighbours inside the row come by wave shuffle. Expressions and their
// order are those of k_lap_uniform_sweep / k_lap_uniform_residual
// (NU = false) and k_lap_nonuniform_sweep / k_lap_nonuniform_residual
// (NU = true), so every cell gets the same bits as there, and the residual
// (a maximum, which no order changes) is the same. RES = false writes the
// next iterate to b; RES = true writes this workgroup's largest |residual|
// to partials[blockIdx.x] and nothing else.
template <int CL>
struct LapSet {
  Vec<double, CL> a_next, s, n, r, yl, yh;
  double w, e, xl, xh;
};

template <int CL, bool NU, bool RES>
__global__ __launch_bounds__(512) void k_lap_stream_t(Tiling t, Uniform u, NonUniform g, Faces f,
                                                      const double *__restrict__ a,
                                                      double *__restrict__ b,
                                                      const double *__restrict__ rhs,
                                                      double *partials) {
  const int nx = u.nx, ny = u.ny, nz = u.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz, blockIdx.x);
  double rmax = 0.0;
  if (l.active) {
    using V = Vec<double, CL>;
    const int i = l.i, j = l.j, lane = (int)threadIdx.x;
    const size_t nxy = (size_t)nx * ny;
    const bool hs = j > 0, hn = j < ny - 1;
    const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
    // The edge lanes of a row fetch the cell beyond it; the others re-read
    // their own first cell (a cached load, unused).
    const ptrdiff_t ow = lane == 0 && i > 0 ? -1 : 0;
    const ptrdiff_t oe = lane == kTileX - 1 && i + CL < nx ? CL : 0;
    // Whether this wavefront's row holds the x = 0 or x = nx-1 cell (uniform
    // across the wavefront, so the face loads below are skipped as a whole).
    const int row0 = i - lane * CL;
    const bool fx0 = row0 == 0, fx1 = row0 + kTileX * CL >= nx;
    const size_t base = (size_t)i + (size_t)nx * (size_t)j;
    const int k0 = l.k0, k1 = l.k1, last = min(k1, nz - 1);
    auto at = [&](int k) { return base + nxy * (size_t)k; };
    // Per-lane x and y data: the sweep's weights and the widths, or the
    // residual's centre distances and the widths.
    double cw[CL], ce[CL], rdx[CL];
    double cs = 0.0, cn = 0.0, dyj = 0.0;
    if constexpr (NU) {
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const int ii = i + m;
        if constexpr (RES) {
          cw[m] = ii > 0 ? g.xc[ii] - g.xc[ii - 1] : 1.0;
          ce[m] = ii < nx - 1 ? g.xc[ii + 1] - g.xc[ii] : 1.0;
        } else {
          cw[m] = g.aw[ii];
          ce[m] = g.ae[ii];
        }
        rdx[m] = g.dx[ii];
      }
      if constexpr (RES) {
        cs = hs ? g.yc[j] - g.yc[j - 1] : 1.0;
        cn = hn ? g.yc[j + 1] - g.yc[j] : 1.0;
      } else {
        cs = g.as[j];
        cn = g.an[j];
      }
      dyj = g.dy[j];
    }
    V zl, zh;
    if (k0 == 0) zl = ldv<double, CL>(f.zlo, base);
    if (k1 == nz) zh = ldv<double, CL>(f.zhi, base);
    auto load = [&](LapSet<CL> &ld, int k, int knext) {
      const size_t c = at(k);
      ld.a_next = ldv<double, CL>(a, at(knext));
      ld.s = ldv<double, CL>(a, c + os);
      ld.n = ldv<double, CL>(a, c + on);
      ld.w = a[c + ow];
      ld.e = a[c + oe];
      ld.r = ldv<double, CL>(rhs, c);
      if (fx0) ld.xl = f.xlo[j + (size_t)ny * k];
      if (fx1) ld.xh = f.xhi[j + (size_t)ny * k];
      if (j == 0) ld.yl = ldv<double, CL>(f.ylo, (size_t)i + (size_t)nx * k);
      if (j == ny - 1) ld.yh = ldv<double, CL>(f.yhi, (size_t)i + (size_t)nx * k);
    };
    LapSet<CL> sa, sb;
    V a_cur = ldv<double, CL>(a, at(k0));
    V a_prev = k0 > 0 ? ldv<double, CL>(a, at(k0 - 1)) : a_cur;
    load(sa, k0, min(k0 + 1, last));
    auto stage = [&](int k, const LapSet<CL> &u_, LapSet<CL> &ld) {
      const int kn = min(k + 1, k1 - 1);
      load(ld, kn, min(k + 2, last));
      const double from_w = __shfl_up(a_cur.v[CL - 1], 1, kTileX);
      const double from_e = __shfl_down(a_cur.v[0], 1, kTileX);
      const double vw = lane == 0 ? u_.w : from_w;
      const double ve = lane == kTileX - 1 ? u_.e : from_e;
      double cb = 0.0, ct = 0.0, dzk = 0.0;
      if constexpr (NU) {
        if constexpr (RES) {
          cb = k > 0 ? g.zc[k] - g.zc[k - 1] : 1.0;
          ct = k < nz - 1 ? g.zc[k + 1] - g.zc[k] : 1.0;
        } else {
          cb = g.ab[k];
          ct = g.at[k];
        }
        dzk = g.dz[k];
      }
      V bv;
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const int ii = i + m;
        const double p = a_cur.v[m];
        const double west = m == 0 ? vw : a_cur.v[m - 1];
        const double east = m == CL - 1 ? ve : a_cur.v[m + 1];
        const double below = a_prev.v[m], above = u_.a_next.v[m];
        if constexpr (!NU) {
          double xm = ii > 0 ? west : p, xp = ii < nx - 1 ? east : p;
          double ym = hs ? u_.s.v[m] : p, yp = hn ? u_.n.v[m] : p;
          double zm = k > 0 ? below : p, zp = k < nz - 1 ? above : p;
          if (ii == 0) xm = p - u_.xl * u.dx;
          if (ii == nx - 1) xp = p + u_.xh * u.dx;
          if (j == 0) ym = p - u_.yl.v[m] * u.dy;
          if (j == ny - 1) yp = p + u_.yh.v[m] * u.dy;
          if (k == 0) zm = p - zl.v[m] * u.dz;
          if (k == nz - 1) zp = p + zh.v[m] * u.dz;
          if constexpr (RES) {
            const double res = (xm - 2 * p + xp) / u.rdx2 + (ym - 2 * p + yp) / u.rdy2 +
                               (zm - 2 * p + zp) / u.rdz2 - u_.r.v[m];
            rmax = fmax(rmax, fabs(res));
          } else {
            bv.v[m] = ((xm + xp) * u.idx2 + (ym + yp) * u.idy2 + (zm + zp) * u.idz2 -
                       u_.r.v[m]) *
                      u.invdiag;
          }
        } else if constexpr (RES) {
          const double g0 = ii == 0 ? u_.xl : (p - west) / cw[m];
          const double g1 = ii == nx - 1 ? u_.xh : (east - p) / ce[m];
          const double g2 = j == 0 ? u_.yl.v[m] : (p - u_.s.v[m]) / cs;
          const double g3 = j == ny - 1 ? u_.yh.v[m] : (u_.n.v[m] - p) / cn;
          const double g4 = k == 0 ? zl.v[m] : (p - below) / cb;
          const double g5 = k == nz - 1 ? zh.v[m] : (above - p) / ct;
          const double res = (g1 - g0) / rdx[m] + (g3 - g2) / dyj + (g5 - g4) / dzk - u_.r.v[m];
          rmax = fmax(rmax, fabs(res));
        } else {
          const double diag = cw[m] + ce[m] + cs + cn + cb + ct;
          double numer = -u_.r.v[m];
          if (ii > 0) numer = numer + cw[m] * west;
          if (ii < nx - 1) numer = numer + ce[m] * east;
          if (hs) numer = numer + cs * u_.s.v[m];
          if (hn) numer = numer + cn * u_.n.v[m];
          if (k > 0) numer = numer + cb * below;
          if (k < nz - 1) numer = numer + ct * above;
          if (ii == 0) numer = numer - u_.xl / rdx[m];
          if (ii == nx - 1) numer = numer + u_.xh / rdx[m];
          if (j == 0) numer = numer - u_.yl.v[m] / dyj;
          if (j == ny - 1) numer = numer + u_.yh.v[m] / dyj;
          if (k == 0) numer = numer - zl.v[m] / dzk;
          if (k == nz - 1) numer = numer + zh.v[m] / dzk;
          bv.v[m] = 0.2 * p + 0.8 * numer / diag;
        }
      }
      if constexpr (!RES) stv<double, CL>(b, at(k), bv);
      a_prev = a_cur;
      a_cur = u_.a_next;
    };
    for (int k = k0; k < k1; k += 2) {
      stage(k, sa, sb);
      if (k + 1 < k1) stage(k + 1, sb, sa);
    }
  }
  if constexpr (RES) {
    __shared__ double lds[16];
    block_max(rmax, lds);
    store_partials(partials, 1, &rmax, gridDim.x);
  }
}

// Tiling and cells per lane of the streaming sweeps: CL = 2 where nx is even
// and every array the kernel reads or writes in 16-byte vectors is 16-byte
// aligned, else 1; tiles of 8 rows (FFPM_LAP_ROWS overrides: 1, 2, 4 or 8).
int lap_cells_per_lane(int nx, std::initializer_list<const void *> ptrs) {
  if (nx % 2) return 1;
  for (const void *q : ptrs)
    if (reinterpret_cast<uintptr_t>(q) % 16) return 1;
  return 2;
}

Tiling lap_tiling(int nx, int ny, int nz, int cl) {
  const char *e = std::getenv("FFPM_LAP_ROWS");
  int rows = e ? std::atoi(e) : 8;
  if (rows != 1 && rows != 2 && rows != 4 && rows != 8) rows = 8;
  return make_tiling(nx, ny, nz, 0, rows, kTileX * cl);
}

// Workgroups of the larger of the two streaming launches (partials slots).
size_t lap_max_blocks(int nx, int ny, int nz) {
  const int b1 = lap_tiling(nx, ny, nz, 1).blocks, b2 = lap_tiling(nx, ny, nz, 2).blocks;
  return (size_t)(b1 > b2 ? b1 : b2);
}

template <bool NU, bool RES>
void launch_lap_stream(const Tiling &t, int cl, const Uniform &u, const NonUniform &g,
                       const Faces &f, const double *a, double *b, const double *rhs,
                       double *partials) {
  const dim3 blk(kTileX, t.ty);
  if (cl == 2)
    k_lap_stream_t<2, NU, RES><<<t.blocks, blk, 0, stream()>>>(t, u, g, f, a, b, rhs, partials);
  else
    k_lap_stream_t<1, NU, RES><<<t.blocks, blk, 0, stream()>>>(t, u, g, f, a, b, rhs, partials);
}

// Face gradients of a dense field with prescribed outer-face data
// (src/fpm_laplace_3d.f90:305-310, src/fpm_terms_3d.f90:243-248).
__device__ __forceinline__ void dense_face_gradients(const NonUniform &g, const Faces &f,
                                                     const double *a, size_t c, int i, int j,
                                                     int k, double gr[6]) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t sy = nx, sz = (size_t)nx * ny;
  const double p = a[c];
  gr[0] = i == 0 ? f.xlo[j + (size_t)ny * k] : (p - a[c - 1]) / (g.xc[i] - g.xc[i - 1]);
  gr[1] = i == nx - 1 ? f.xhi[j + (size_t)ny * k] : (a[c + 1] - p) / (g.xc[i + 1] - g.xc[i]);
  gr[2] = j == 0 ? f.ylo[i + (size_t)nx * k] : (p - a[c - sy]) / (g.yc[j] - g.yc[j - 1]);
  gr[3] = j == ny - 1 ? f.yhi[i + (size_t)nx * k] : (a[c + sy] - p) / (g.yc[j + 1] - g.yc[j]);
  gr[4] = k == 0 ? f.zlo[i + (size_t)nx * j] : (p - a[c - sz]) / (g.zc[k] - g.zc[k - 1]);
  gr[5] = k == nz - 1 ? f.zhi[i + (size_t)nx * j] : (a[c + sz] - p) / (g.zc[k + 1] - g.zc[k]);
}

// residual_nonuniform, src/fpm_laplace_3d.f90:295-315 -> sc->res.
__global__ __launch_bounds__(256) void k_lap_nonuniform_residual(NonUniform g, Faces f,
                                                                 const double *__restrict__ a,
                                                                 const double *__restrict__ rhs,
                                                                 double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const size_t n = (size_t)g.nx * g.ny * g.nz;
  double m = 0.0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(c, g.nx, g.ny, i, j, k);
    double gr[6];
    dense_face_gradients(g, f, a, c, i, j, k, gr);
    const double res = (gr[1] - gr[0]) / g.dx[i] + (gr[3] - gr[2]) / g.dy[j] +
                       (gr[5] - gr[4]) / g.dz[k] - rhs[c];
    m = fmax(m, fabs(res));
  }
  block_max(m, lds);
  store_partials(partials, 1, &m, gridDim.x);
}

// Compatibility sums of src/fpm_laplace_3d.f90:18-30 and :160-182:
// source integral, face flux and (non-uniform) face area -> sc->sums[0..2].
// dxv == nullptr selects the uniform spacing (dx, dy, dz).
__global__ __launch_bounds__(256) void k_lap_compat_sums(int nx, int ny, int nz,
                                                         const double *__restrict__ rhs, Faces f,
                                                         const double *dxv, const double *dyv,
                                                         const double *dzv, double dx, double dy,
                                                         double dz, int with_source,
                                                         double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const size_t n = (size_t)nx * ny * nz;
  const size_t nyz = (size_t)ny * nz, nxz = (size_t)nx * nz, nxy = (size_t)nx * ny;
  const size_t total = (with_source ? n : 0) + nyz + nxz + nxy;
  double s[3] = {0.0, 0.0, 0.0};
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total;
       t += (size_t)gridDim.x * blockDim.x) {
    size_t u = t;
    if (with_source) {
      if (u < n) {
        int i, j, k;
        cell_of(u, nx, ny, i, j, k);
        s[0] += dxv ? rhs[u] * dxv[i] * dyv[j] * dzv[k] : rhs[u] * dx * dy * dz;
        continue;
      }
      u -= n;
    }
    if (u < nyz) {
      const int j = (int)(u % ny), k = (int)(u / ny);
      const double a = dxv ? dyv[j] : dy, b = dxv ? dzv[k] : dz;
      s[1] += (-f.xlo[u] + f.xhi[u]) * a * b;
      s[2] += 2.0 * a * b;
    } else if (u < nyz + nxz) {
      u -= nyz;
      const int i = (int)(u % nx), k = (int)(u / nx);
      const double a = dxv ? dxv[i] : dx, b = dxv ? dzv[k] : dz;
      s[1] += (-f.ylo[u] + f.yhi[u]) * a * b;
      s[2] += 2.0 * a * b;
    } else {
      u -= nyz + nxz;
      const int i = (int)(u % nx), j = (int)(u / nx);
      const double a = dxv ? dxv[i] : dx, b = dxv ? dyv[j] : dy;
      s[1] += (-f.zlo[u] + f.zhi[u]) * a * b;
      s[2] += 2.0 * a * b;
    }
  }
  block_sum<3>(s, lds);
  store_partials(partials, 3, s, gridDim.x);
}

// compute_fpm_terms_uniform_3d (src/fpm_terms_3d.f90:8-37): central
// differences through the caller's ghost layers. dense phi with 3-D halo.
// compute_fpm_terms_nonuniform_3d (src/fpm_terms_3d.f90:39-87): face-gradient
// averages with prescribed outer data and bounded integrals.
__global__ __launch_bounds__(256) void k_terms_nobody(
    int uniform, NonUniform g, Faces f, const double *__restrict__ phi, double dx, double dy,
    double dz, const double *__restrict__ u, const double *__restrict__ v,
    const double *__restrict__ w, const double *__restrict__ ox, const double *__restrict__ oy,
    const double *__restrict__ oz, const double *__restrict__ ax, const double *__restrict__ ay,
    const double *__restrict__ az, const int32_t *__restrict__ fluid, Bounds bd,
    double *__restrict__ gx, double *__restrict__ gy, double *__restrict__ gz,
    double *__restrict__ vortex, double *__restrict__ advection, double *partials,
    Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  const double volume_u = dx * dy * dz;
  double s[2] = {0.0, 0.0};
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(c, nx, ny, i, j, k);
    double gxv, gyv, gzv, volume;
    if (uniform) {
      const size_t sx = (size_t)nx + 2, sxy = sx * ((size_t)ny + 2);
      const size_t p = (size_t)(i + 1) + sx * (j + 1) + sxy * (k + 1);
      gxv = (phi[p + 1] - phi[p - 1]) / (2.0 * dx);
      gyv = (phi[p + sx] - phi[p - sx]) / (2.0 * dy);
      gzv = (phi[p + sxy] - phi[p - sxy]) / (2.0 * dz);
      volume = volume_u;
    } else {
      double gr[6];
      dense_face_gradients(g, f, phi, c, i, j, k, gr);
      gxv = 0.5 * (gr[0] + gr[1]);
      gyv = 0.5 * (gr[2] + gr[3]);
      gzv = 0.5 * (gr[4] + gr[5]);
      const bool inside = g.xc[i] > bd.b[0] && g.xc[i] < bd.b[1] && g.yc[j] > bd.b[2] &&
                          g.yc[j] < bd.b[3] && g.zc[k] > bd.b[4] && g.zc[k] < bd.b[5];
      volume = (inside ? 1.0 : 0.0) * g.dx[i] * g.dy[j] * g.dz[k];
    }
    gx[c] = gxv;
    gy[c] = gyv;
    gz[c] = gzv;
    const double cx = w[c] * oy[c] - v[c] * oz[c];
    const double cy = u[c] * oz[c] - w[c] * ox[c];
    const double cz = v[c] * ox[c] - u[c] * oy[c];
    const double mask = (double)fluid[c];
    const double vo = -mask * (cx * gxv + cy * gyv + cz * gzv);
    const double ad = -mask * (ax[c] * gxv + ay[c] * gyv + az[c] * gzv);
    vortex[c] = vo;
    advection[c] = ad;
    s[0] += volume * vo;
    s[1] += volume * ad;
  }
  block_sum<2>(s, lds);
  store_partials(partials, 2, s, gridDim.x);
}

struct Ctx {
  double *partials;
  Scalars *sc;
};

// blocks: the largest launch whose partials the solve stores (beyond the
// elementwise launches).
int ctx_setup(Ctx &x, size_t n, size_t blocks = 0) {
  int st = 0;
  Workspace &ws = workspace();
  const size_t eb = (size_t)elementwise_blocks(n);
  const size_t nb = (blocks > eb ? blocks : eb) + 8192;
  x.partials = (double *)ws.get(S_PARTIALS, sizeof(double) * kReduceSlots * nb, &st);
  if (!x.partials) return st;
  x.sc = (Scalars *)ws.get(S_SCALARS, sizeof(Scalars), &st);
  if (!x.sc) return st;
  FFPM_HIP(hipMemsetAsync(x.sc, 0, sizeof(Scalars), stream()));
  return OK;
}

int finish(const Ctx &x, int op, int nslots, int nblocks) {
  k_finish<<<1, 256, 0, stream()>>>(op, nslots, x.partials, nblocks, x.sc);
  FFPM_LAUNCHED("k_finish");
  return OK;
}

int read_sc(const Ctx &x, Scalars *h) {
  FFPM_HIP(hipMemcpyAsync(h, x.sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  return OK;
}

int bad(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  return OK;
}

}  // namespace

// enforce_neumann_compatibility[_nonuniform]_3d, src/fpm_laplace_3d.f90:9-53,150-217
int lap_compat_dev(const double *rhs, double *xlo, double *xhi, double *ylo, double *yhi,
                   double *zlo, double *zhi, int nx, int ny, int nz, const double *dxv,
                   const double *dyv, const double *dzv, double dx, double dy, double dz,
                   double *correction, double *imbalance) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  Ctx x;
  const size_t n = (size_t)nx * ny * nz;
  const size_t nf = (size_t)ny * nz + (size_t)nx * nz + (size_t)nx * ny;
  if ((st = ctx_setup(x, n + nf))) return st;
  Faces f{xlo, xhi, ylo, yhi, zlo, zhi};
  const int eb = elementwise_blocks(n + nf);
  k_lap_compat_sums<<<eb, kThreads, 0, stream()>>>(nx, ny, nz, rhs, f, dxv, dyv, dzv, dx, dy, dz,
                                                   1, x.partials, x.sc);
  FFPM_LAUNCHED("k_lap_compat_sums");
  finish(x, FIN_SUMS, 3, eb);
  Scalars h;
  if ((st = read_sc(x, &h))) return st;
  const double source = h.sums[0], flux = h.sums[1];
  // uniform: area = 2 (ny nz dy dz + nx nz dx dz + nx ny dx dy) as written
  const double area =
      dxv ? h.sums[2]
          : 2.0 * ((double)(ny * nz) * dy * dz + (double)(nx * nz) * dx * dz +
                   (double)(nx * ny) * dx * dy);
  const double c = -(flux - source) / area;
  k_shift_faces<<<elementwise_blocks(nf), kThreads, 0, stream()>>>(nx, ny, nz, xlo, xhi, ylo,
                                                                   yhi, zlo, zhi, c);
  FFPM_LAUNCHED("k_shift_faces");
  k_lap_compat_sums<<<elementwise_blocks(nf), kThreads, 0, stream()>>>(
      nx, ny, nz, rhs, f, dxv, dyv, dzv, dx, dy, dz, 0, x.partials, x.sc);
  FFPM_LAUNCHED("k_lap_compat_sums");
  finish(x, FIN_SUMS, 3, elementwise_blocks(nf));
  if ((st = read_sc(x, &h))) return st;
  *correction = c;
  *imbalance = h.sums[1] - source;
  return OK;
}

// solve_neumann_poisson_3d, src/fpm_laplace_3d.f90:55-107. phi_ref is the
// caller's (0:nx+1,0:ny+1,0:nz+1) array on the device.
int lap_uniform_solve_dev(double *phi_ref, const double *rhs, const double *xlo,
                          const double *xhi, const double *ylo, const double *yhi,
                          const double *zlo, const double *zhi, int nx, int ny, int nz, double dx,
                          double dy, double dz, double tolerance, int max_iterations,
                          int check_interval, int *iterations, double *final_residual) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  Ctx x;
  const size_t n = (size_t)nx * ny * nz;
  if ((st = ctx_setup(x, n, lap_max_blocks(nx, ny, nz)))) return st;
  double *a = (double *)workspace().get(S_PHI, sizeof(double) * n, &st);
  if (!a) return st;
  double *b = (double *)workspace().get(S_L_NEXT, sizeof(double) * n, &st);
  if (!b) return st;
  Uniform u;
  u.nx = nx; u.ny = ny; u.nz = nz; u.dx = dx; u.dy = dy; u.dz = dz;
  u.idx2 = 1.0 / (dx * dx);
  u.idy2 = 1.0 / (dy * dy);
  u.idz2 = 1.0 / (dz * dz);
  u.invdiag = 1.0 / (2.0 * (u.idx2 + u.idy2 + u.idz2));
  u.rdx2 = dx * dx;
  u.rdy2 = dy * dy;
  u.rdz2 = dz * dz;
  Faces f{xlo, xhi, ylo, yhi, zlo, zhi};
  const int eb = elementwise_blocks(n);
  const bool classic = lap_classic();
  const int cl = lap_cells_per_lane(nx, {a, b, rhs, ylo, yhi, zlo, zhi});
  const Tiling tz = lap_tiling(nx, ny, nz, cl);
  NonUniform g0{};
  // max-norm residual of a -> sc->res
  auto residual = [&](const double *v) {
    if (classic) {
      k_lap_uniform_residual<<<eb, kThreads, 0, stream()>>>(u, f, v, rhs, x.partials, x.sc);
      finish(x, FIN_MAX, 1, eb);
    } else {
      launch_lap_stream<false, true>(tz, cl, u, g0, f, v, nullptr, rhs, x.partials);
      finish(x, FIN_MAX, 1, tz.blocks);
    }
  };
  k_lap_pack<<<eb, kThreads, 0, stream()>>>(nx, ny, nz, phi_ref, a);
  FFPM_LAUNCHED("k_lap_pack");
  *iterations = 0;
  double res = DBL_MAX;
  Scalars h;
  while (*iterations < max_iterations && res > tolerance) {
    int nsteps = check_interval > 1 ? check_interval : 1;
    if (max_iterations - *iterations < nsteps) nsteps = max_iterations - *iterations;
    for (int s = 0; s < nsteps; ++s) {
      {
        TimedLaunch tl("lap_sweep");
        if (classic)
          k_lap_uniform_sweep<<<eb, kThreads, 0, stream()>>>(u, f, a, b, rhs);
        else
          launch_lap_stream<false, false>(tz, cl, u, g0, f, a, b, rhs, nullptr);
      }
      double *t = a;
      a = b;
      b = t;
      *iterations += 1;
    }
    FFPM_LAUNCHED("k_lap_uniform_sweep");
    residual(a);
    FFPM_LAUNCHED("k_lap_uniform_residual");
    if ((st = read_sc(x, &h))) return st;
    res = h.res;
  }
  k_lap_mean_sums<<<eb, kThreads, 0, stream()>>>(nx, ny, nz, a, nullptr, nullptr, nullptr,
                                                 x.partials, x.sc);
  FFPM_LAUNCHED("k_lap_mean_sums");
  finish(x, FIN_SUMS, 2, eb);
  k_lap_subtract_mean<<<eb, kThreads, 0, stream()>>>(n, a, (double)(nx * ny * nz), x.sc);
  FFPM_LAUNCHED("k_lap_subtract_mean");
  const size_t np = ((size_t)nx + 2) * ((size_t)ny + 2) * ((size_t)nz + 2);
  k_lap_unpack_ghosts<<<elementwise_blocks(np), kThreads, 0, stream()>>>(nx, ny, nz, f, dx, dy,
                                                                         dz, a, phi_ref);
  FFPM_LAUNCHED("k_lap_unpack_ghosts");
  residual(a);
  FFPM_LAUNCHED("k_lap_uniform_residual");
  if ((st = read_sc(x, &h))) return st;
  *final_residual = h.res;
  return OK;
}

int nonuniform_setup(NonUniform &g, int nx, int ny, int nz, const double *xc, const double *yc,
                     const double *zc, const double *dx, const double *dy, const double *dz) {
  int st = 0;
  double *w = (double *)workspace().get(S_L_COEF, sizeof(double) * 2 * (nx + ny + nz), &st);
  if (!w) return st;
  double *aw = w, *ae = aw + nx, *as = ae + nx, *an = as + ny, *ab = an + ny, *at = ab + nz;
  k_lap_axis_weights<<<(nx + 255) / 256, 256, 0, stream()>>>(nx, xc, dx, aw, ae);
  k_lap_axis_weights<<<(ny + 255) / 256, 256, 0, stream()>>>(ny, yc, dy, as, an);
  k_lap_axis_weights<<<(nz + 255) / 256, 256, 0, stream()>>>(nz, zc, dz, ab, at);
  FFPM_LAUNCHED("k_lap_axis_weights");
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.aw = aw; g.ae = ae; g.as = as; g.an = an; g.ab = ab; g.at = at;
  g.xc = xc; g.yc = yc; g.zc = zc; g.dx = dx; g.dy = dy; g.dz = dz;
  return OK;
}

// solve_neumann_poisson_nonuniform_3d, src/fpm_laplace_3d.f90:219-293
int lap_nonuniform_solve_dev(double *phi, const double *rhs, const double *xlo,
                             const double *xhi, const double *ylo, const double *yhi,
                             const double *zlo, const double *zhi, const double *xc,
                             const double *yc, const double *zc, const double *dx,
                             const double *dy, const double *dz, int nx, int ny, int nz,
                             double tolerance, int max_iterations, int check_interval,
                             int *iterations, double *final_residual) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  Ctx x;
  const size_t n = (size_t)nx * ny * nz;
  if ((st = ctx_setup(x, n, lap_max_blocks(nx, ny, nz)))) return st;
  NonUniform g;
  if ((st = nonuniform_setup(g, nx, ny, nz, xc, yc, zc, dx, dy, dz))) return st;
  double *b = (double *)workspace().get(S_L_NEXT, sizeof(double) * n, &st);
  if (!b) return st;
  Faces f{xlo, xhi, ylo, yhi, zlo, zhi};
  const int eb = elementwise_blocks(n);
  const bool classic = lap_classic();
  const int cl = lap_cells_per_lane(nx, {phi, b, rhs, ylo, yhi, zlo, zhi});
  const Tiling tz = lap_tiling(nx, ny, nz, cl);
  Uniform u0{};
  u0.nx = nx; u0.ny = ny; u0.nz = nz;
  // max-norm residual of v -> sc->res
  auto residual = [&](const double *v) {
    if (classic) {
      k_lap_nonuniform_residual<<<eb, kThreads, 0, stream()>>>(g, f, v, rhs, x.partials, x.sc);
      finish(x, FIN_MAX, 1, eb);
    } else {
      launch_lap_stream<true, true>(tz, cl, u0, g, f, v, nullptr, rhs, x.partials);
      finish(x, FIN_MAX, 1, tz.blocks);
    }
  };
  double *a = phi;
  *iterations = 0;
  double res = DBL_MAX;
  Scalars h;
  while (*iterations < max_iterations && res > tolerance) {
    int nsteps = check_interval > 1 ? check_interval : 1;
    if (max_iterations - *iterations < nsteps) nsteps = max_iterations - *iterations;
    for (int s = 0; s < nsteps; ++s) {
      {
        TimedLaunch tl("lap_sweep");
        if (classic)
          k_lap_nonuniform_sweep<<<eb, kThreads, 0, stream()>>>(g, f, a, b, rhs);
        else
          launch_lap_stream<true, false>(tz, cl, u0, g, f, a, b, rhs, nullptr);
      }
      double *t = a;
      a = b;
      b = t;
      *iterations += 1;
    }
    FFPM_LAUNCHED("k_lap_nonuniform_sweep");
    residual(a);
    FFPM_LAUNCHED("k_lap_nonuniform_residual");
    if ((st = read_sc(x, &h))) return st;
    res = h.res;
  }
  if (a != phi) {
    FFPM_HIP(hipMemcpyAsync(phi, a, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
    a = phi;
  }
  k_lap_mean_sums<<<eb, kThreads, 0, stream()>>>(nx, ny, nz, a, dx, dy, dz, x.partials, x.sc);
  FFPM_LAUNCHED("k_lap_mean_sums");
  finish(x, FIN_SUMS, 2, eb);
  k_lap_subtract_mean<<<eb, kThreads, 0, stream()>>>(n, a, -1.0, x.sc);
  FFPM_LAUNCHED("k_lap_subtract_mean");
  residual(a);
  FFPM_LAUNCHED("k_lap_nonuniform_residual");
  if ((st = read_sc(x, &h))) return st;
  *final_residual = h.res;
  return OK;
}

int terms_nobody_dev(int uniform, const double *phi, const double *u, const double *v,
                     const double *w, const double *ox, const double *oy, const double *oz,
                     const double *ax, const double *ay, const double *az, const int32_t *fluid,
                     const double *xc, const double *yc, const double *zc, const double *dxv,
                     const double *dyv, const double *dzv, double dx, double dy, double dz,
                     const double *xlo, const double *xhi, const double *ylo, const double *yhi,
                     const double *zlo, const double *zhi, const double *bounds, double *gx,
                     double *gy, double *gz, double *vortex, double *advection, double *totals,
                     int nx, int ny, int nz) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  if (!uniform && (nx < 2 || ny < 2 || nz < 2)) {
    set_error(ERR_SHAPE, "nonuniform 3-D FPM kernel requires at least two cells per direction");
    return ERR_SHAPE;
  }
  Ctx x;
  const size_t n = (size_t)nx * ny * nz;
  if ((st = ctx_setup(x, n))) return st;
  NonUniform g{};
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.xc = xc; g.yc = yc; g.zc = zc; g.dx = dxv; g.dy = dyv; g.dz = dzv;
  Faces f{xlo, xhi, ylo, yhi, zlo, zhi};
  Bounds bd{};
  if (bounds)
    for (int m = 0; m < 6; ++m) bd.b[m] = bounds[m];
  k_terms_nobody<<<elementwise_blocks(n), kThreads, 0, stream()>>>(
      uniform, g, f, phi, dx, dy, dz, u, v, w, ox, oy, oz, ax, ay, az, fluid, bd, gx, gy, gz,
      vortex, advection, x.partials, x.sc);
  FFPM_LAUNCHED("k_terms_nobody");
  finish(x, FIN_SUMS, 2, elementwise_blocks(n));
  Scalars h;
  if ((st = read_sc(x, &h))) return st;
  totals[0] = h.sums[0];
  totals[1] = h.sums[1];
  return OK;
}

namespace {
template <class T>
int up(int slot, const T *host, size_t count, T **dev) {
  int st = 0;
  *dev = (T *)workspace().get(slot, sizeof(T) * count, &st);
  if (!*dev) return st;
  FFPM_HIP(hipMemcpyAsync(*dev, host, sizeof(T) * count, hipMemcpyHostToDevice, stream()));
  return OK;
}
template <class T>
int down(T *host, const T *dev, size_t count) {
  FFPM_HIP(hipMemcpyAsync(host, dev, sizeof(T) * count, hipMemcpyDeviceToHost, stream()));
  return OK;
}
int done() {
  FFPM_HIP(hipStreamSynchronize(stream()));
  return OK;
}
struct F6 {
  double *d[6];
};
int up_faces(int nx, int ny, int nz, const double *const h[6], F6 &f) {
  const size_t cnt[6] = {(size_t)ny * nz, (size_t)ny * nz, (size_t)nx * nz,
                         (size_t)nx * nz, (size_t)nx * ny, (size_t)nx * ny};
  for (int m = 0; m < 6; ++m) {
    int st = up(S_U_XLO + m, h[m], cnt[m], &f.d[m]);
    if (st) return st;
  }
  return OK;
}
int down_faces(int nx, int ny, int nz, double *const h[6], const F6 &f) {
  const size_t cnt[6] = {(size_t)ny * nz, (size_t)ny * nz, (size_t)nx * nz,
                         (size_t)nx * nz, (size_t)nx * ny, (size_t)nx * ny};
  for (int m = 0; m < 6; ++m) {
    int st = down(h[m], f.d[m], cnt[m]);
    if (st) return st;
  }
  return OK;
}
}  // namespace

}  // namespace ffpm

using namespace ffpm;

extern "C" {

int ffpm_enforce_neumann_compatibility_3d_dev(const double *rhs, double *xlo, double *xhi,
                                              double *ylo, double *yhi, double *zlo,
                                              double *zhi, int nx, int ny, int nz, double dx,
                                              double dy, double dz, double *correction,
                                              double *imbalance) {
  int st = lap_compat_dev(rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, nullptr, nullptr,
                          nullptr, dx, dy, dz, correction, imbalance);
  return st ? st : done();
}

int ffpm_enforce_neumann_compatibility_3d(const double *rhs, double *xlo, double *xhi,
                                          double *ylo, double *yhi, double *zlo, double *zhi,
                                          int nx, int ny, int nz, double dx, double dy,
                                          double dz, double *correction, double *imbalance) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  double *d_rhs;
  F6 f;
  const double *hin[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  double *hout[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  if ((st = up(S_U_RHS, rhs, (size_t)nx * ny * nz, &d_rhs))) return st;
  if ((st = up_faces(nx, ny, nz, hin, f))) return st;
  if ((st = lap_compat_dev(d_rhs, f.d[0], f.d[1], f.d[2], f.d[3], f.d[4], f.d[5], nx, ny, nz,
                           nullptr, nullptr, nullptr, dx, dy, dz, correction, imbalance)))
    return st;
  if ((st = down_faces(nx, ny, nz, hout, f))) return st;
  return done();
}

int ffpm_enforce_neumann_compatibility_nonuniform_3d_dev(
    const double *rhs, double *xlo, double *xhi, double *ylo, double *yhi, double *zlo,
    double *zhi, const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double *correction, double *imbalance) {
  int st = lap_compat_dev(rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, dx, dy, dz, 0, 0, 0,
                          correction, imbalance);
  return st ? st : done();
}

int ffpm_enforce_neumann_compatibility_nonuniform_3d(
    const double *rhs, double *xlo, double *xhi, double *ylo, double *yhi, double *zlo,
    double *zhi, const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double *correction, double *imbalance) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  double *d_rhs, *d_dx, *d_dy, *d_dz;
  F6 f;
  const double *hin[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  double *hout[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  if ((st = up(S_U_RHS, rhs, (size_t)nx * ny * nz, &d_rhs))) return st;
  if ((st = up(S_U_DX, dx, (size_t)nx, &d_dx)) || (st = up(S_U_DY, dy, (size_t)ny, &d_dy)) ||
      (st = up(S_U_DZ, dz, (size_t)nz, &d_dz)))
    return st;
  if ((st = up_faces(nx, ny, nz, hin, f))) return st;
  if ((st = lap_compat_dev(d_rhs, f.d[0], f.d[1], f.d[2], f.d[3], f.d[4], f.d[5], nx, ny, nz,
                           d_dx, d_dy, d_dz, 0, 0, 0, correction, imbalance)))
    return st;
  if ((st = down_faces(nx, ny, nz, hout, f))) return st;
  return done();
}

int ffpm_solve_neumann_poisson_3d_dev(double *phi, const double *rhs, const double *xlo,
                                      const double *xhi, const double *ylo, const double *yhi,
                                      const double *zlo, const double *zhi, int nx, int ny,
                                      int nz, double dx, double dy, double dz, double tolerance,
                                      int max_iterations, int check_interval, int *iterations,
                                      double *final_residual) {
  int st = lap_uniform_solve_dev(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, nx, ny, nz, dx, dy, dz,
                                 tolerance, max_iterations, check_interval, iterations,
                                 final_residual);
  return st ? st : done();
}

int ffpm_solve_neumann_poisson_3d(double *phi, const double *rhs, const double *xlo,
                                  const double *xhi, const double *ylo, const double *yhi,
                                  const double *zlo, const double *zhi, int nx, int ny, int nz,
                                  double dx, double dy, double dz, double tolerance,
                                  int max_iterations, int check_interval, int *iterations,
                                  double *final_residual) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  const size_t np = ((size_t)nx + 2) * ((size_t)ny + 2) * ((size_t)nz + 2);
  double *d_phi, *d_rhs;
  F6 f;
  const double *hin[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  if ((st = up(S_U_PHI, phi, np, &d_phi))) return st;
  if ((st = up(S_U_RHS, rhs, (size_t)nx * ny * nz, &d_rhs))) return st;
  if ((st = up_faces(nx, ny, nz, hin, f))) return st;
  if ((st = lap_uniform_solve_dev(d_phi, d_rhs, f.d[0], f.d[1], f.d[2], f.d[3], f.d[4], f.d[5],
                                  nx, ny, nz, dx, dy, dz, tolerance, max_iterations,
                                  check_interval, iterations, final_residual)))
    return st;
  if ((st = down(phi, d_phi, np))) return st;
  return done();
}

int ffpm_solve_neumann_poisson_nonuniform_3d_dev(
    double *phi, const double *rhs, const double *xlo, const double *xhi, const double *ylo,
    const double *yhi, const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, int nx, int ny,
    int nz, double tolerance, int max_iterations, int check_interval, int *iterations,
    double *final_residual) {
  int st = lap_nonuniform_solve_dev(phi, rhs, xlo, xhi, ylo, yhi, zlo, zhi, xc, yc, zc, dx, dy,
                                    dz, nx, ny, nz, tolerance, max_iterations, check_interval,
                                    iterations, final_residual);
  return st ? st : done();
}

int ffpm_solve_neumann_poisson_nonuniform_3d(
    double *phi, const double *rhs, const double *xlo, const double *xhi, const double *ylo,
    const double *yhi, const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, int nx, int ny,
    int nz, double tolerance, int max_iterations, int check_interval, int *iterations,
    double *final_residual) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  const size_t n = (size_t)nx * ny * nz;
  double *d_phi, *d_rhs, *ax[6];
  F6 f;
  const double *hin[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  const double *hax[6] = {xc, yc, zc, dx, dy, dz};
  const size_t nax[6] = {(size_t)nx, (size_t)ny, (size_t)nz, (size_t)nx, (size_t)ny, (size_t)nz};
  if ((st = up(S_U_PHI, phi, n, &d_phi))) return st;
  if ((st = up(S_U_RHS, rhs, n, &d_rhs))) return st;
  if ((st = up_faces(nx, ny, nz, hin, f))) return st;
  for (int m = 0; m < 6; ++m)
    if ((st = up(S_U_XC + m, hax[m], nax[m], &ax[m]))) return st;
  if ((st = lap_nonuniform_solve_dev(d_phi, d_rhs, f.d[0], f.d[1], f.d[2], f.d[3], f.d[4],
                                     f.d[5], ax[0], ax[1], ax[2], ax[3], ax[4], ax[5], nx, ny,
                                     nz, tolerance, max_iterations, check_interval, iterations,
                                     final_residual)))
    return st;
  if ((st = down(phi, d_phi, n))) return st;
  return done();
}

int ffpm_compute_fpm_terms_uniform_3d_dev(const double *phi, const double *u, const double *v,
                                          const double *w, const double *ox, const double *oy,
                                          const double *oz, const double *ax, const double *ay,
                                          const double *az, const int32_t *fluid, double *gx,
                                          double *gy, double *gz, double *vortex,
                                          double *advection, double *totals, int nx, int ny,
                                          int nz, double dx, double dy, double dz) {
  int st = terms_nobody_dev(1, phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, nullptr, nullptr,
                            nullptr, nullptr, nullptr, nullptr, dx, dy, dz, nullptr, nullptr,
                            nullptr, nullptr, nullptr, nullptr, nullptr, gx, gy, gz, vortex,
                            advection, totals, nx, ny, nz);
  return st ? st : done();
}

static int terms_nobody_host(int uniform, const double *phi, const double *const fin[9],
                             const int32_t *fluid, const double *const axes[6], double dx,
                             double dy, double dz, const double *const faces[6],
                             const double *bounds, double *const outs[5], double *totals, int nx,
                             int ny, int nz) {
  int st;
  if ((st = bad(nx, ny, nz))) return st;
  const size_t n = (size_t)nx * ny * nz;
  const size_t nphi = uniform ? ((size_t)nx + 2) * ((size_t)ny + 2) * ((size_t)nz + 2) : n;
  double *d_phi, *f[9], *o[5], *a[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int32_t *d_fluid;
  F6 fc{};
  if ((st = up(S_U_PHI, phi, nphi, &d_phi))) return st;
  if ((st = up(S_U_SOLID, fluid, n, &d_fluid))) return st;
  for (int m = 0; m < 9; ++m)
    if ((st = up(S_U_F0 + m, fin[m], n, &f[m]))) return st;
  for (int m = 0; m < 5; ++m) {
    o[m] = (double *)workspace().get(S_U_O0 + m, sizeof(double) * n, &st);
    if (!o[m]) return st;
  }
  if (!uniform) {
    const size_t nax[6] = {(size_t)nx, (size_t)ny, (size_t)nz, (size_t)nx, (size_t)ny, (size_t)nz};
    for (int m = 0; m < 6; ++m)
      if ((st = up(S_U_XC + m, axes[m], nax[m], &a[m]))) return st;
    if ((st = up_faces(nx, ny, nz, faces, fc))) return st;
  }
  if ((st = terms_nobody_dev(uniform, d_phi, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8],
                             d_fluid, a[0], a[1], a[2], a[3], a[4], a[5], dx, dy, dz, fc.d[0],
                             fc.d[1], fc.d[2], fc.d[3], fc.d[4], fc.d[5], bounds, o[0], o[1],
                             o[2], o[3], o[4], totals, nx, ny, nz)))
    return st;
  for (int m = 0; m < 5; ++m)
    if ((st = down(outs[m], o[m], n))) return st;
  return done();
}

int ffpm_compute_fpm_terms_uniform_3d(const double *phi, const double *u, const double *v,
                                      const double *w, const double *ox, const double *oy,
                                      const double *oz, const double *ax, const double *ay,
                                      const double *az, const int32_t *fluid, double *gx,
                                      double *gy, double *gz, double *vortex, double *advection,
                                      double *totals, int nx, int ny, int nz, double dx,
                                      double dy, double dz) {
  const double *fin[9] = {u, v, w, ox, oy, oz, ax, ay, az};
  double *outs[5] = {gx, gy, gz, vortex, advection};
  const double *axes[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  return terms_nobody_host(1, phi, fin, fluid, axes, dx, dy, dz, axes, nullptr, outs, totals, nx,
                           ny, nz);
}

int ffpm_compute_fpm_terms_nonuniform_3d_dev(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *fluid, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, const double *xlo, const double *xhi,
    const double *ylo, const double *yhi, const double *zlo, const double *zhi,
    const double *bounds, double *gx, double *gy, double *gz, double *vortex, double *advection,
    double *totals, int nx, int ny, int nz) {
  int st = terms_nobody_dev(0, phi, u, v, w, ox, oy, oz, ax, ay, az, fluid, xc, yc, zc, dx, dy,
                            dz, 0, 0, 0, xlo, xhi, ylo, yhi, zlo, zhi, bounds, gx, gy, gz, vortex,
                            advection, totals, nx, ny, nz);
  return st ? st : done();
}

int ffpm_compute_fpm_terms_nonuniform_3d(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *fluid, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, const double *xlo, const double *xhi,
    const double *ylo, const double *yhi, const double *zlo, const double *zhi,
    const double *bounds, double *gx, double *gy, double *gz, double *vortex, double *advection,
    double *totals, int nx, int ny, int nz) {
  if (nx < 2 || ny < 2 || nz < 2) {
    set_error(ERR_SHAPE, "nonuniform 3-D FPM kernel requires at least two cells per direction");
    return ERR_SHAPE;
  }
  const double *fin[9] = {u, v, w, ox, oy, oz, ax, ay, az};
  double *outs[5] = {gx, gy, gz, vortex, advection};
  const double *axes[6] = {xc, yc, zc, dx, dy, dz};
  const double *faces[6] = {xlo, xhi, ylo, yhi, zlo, zhi};
  return terms_nobody_host(0, phi, fin, fluid, axes, 0, 0, 0, faces, bounds, outs, totals, nx, ny,
                           nz);
}

}  // extern "C"
