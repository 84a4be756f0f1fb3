// gfx950 kernels of the stair-step influence-potential solve and the FPM
// force densities (src/fpm_stairstep_mpi_3d.f90). FP64 throughout; built with
// -ffp-contract=off so every expression keeps the reference's rounding.
//
// Stencil kernels stream along z: a workgroup owns a 64x4 x-y tile and a
// chunk of z-planes, keeps the (k-1, k, k+1) values of its own column in
// registers and reads the four x-y neighbours of plane k through the cache
// (each row is read from HBM once; the halo rows come from the neighbour
// tile, which the XCD-aware tile order keeps in the same L2). The operator is
// mask-free: solid cells hold zero in every iterate and zero src/rdiag/diag,
// and the zero physical halo is substituted instead of stored.
#include "ffpm_bodies.hpp"
#include "ffpm_frame_body.hpp"

namespace ffpm {

// axis_coefficients (src/fpm_stairstep_mpi_3d.f90:499-510) and the z
// coefficients (src/fpm_stairstep_mpi_3d.f90:138-142), one thread per entry.
__global__ void k_axis_coefficients(int nx, int ny, int nz, const double *xc, const double *yc,
                                    const double *zc, const double *dx, const double *dy,
                                    const double *dz, double *cw, double *ce, double *cs,
                                    double *cn, double *cb, double *ct) {
  axis_coefficients_body(blockIdx.x * blockDim.x + threadIdx.x, nx, ny, nz, xc, yc, zc, dx, dy,
                         dz, cw, ce, cs, cn, cb, ct);
}

// build_operator (src/fpm_stairstep_mpi_3d.f90:336-379); body in
// ffpm_frame_body.hpp.
__global__ __launch_bounds__(512) void k_build_operator(
    Geom g, const int32_t *__restrict__ solid, const double *__restrict__ rhs,
    const double *__restrict__ xlo, const double *__restrict__ xhi,
    const double *__restrict__ ylo, const double *__restrict__ yhi,
    const double *__restrict__ zlo, const double *__restrict__ zhi, double bg0, double bg1,
    double bg2, int for_pcg, double *__restrict__ src, double *__restrict__ diag,
    double *__restrict__ rdiag, uint8_t *__restrict__ code, int *__restrict__ isolated) {
  build_operator_body(blockIdx.x, gridDim.x, g, solid, rhs, xlo, xhi, ylo, yhi, zlo, zhi, bg0,
                      bg1, bg2, for_pcg, src, diag, rdiag, code, isolated, true);
}

// Init loop of src/fpm_stairstep_mpi_3d.f90:154-160 on the dense copy.
__global__ void k_pack_phi(int nx, int ny, int nz, const int32_t *__restrict__ solid,
                           const double *__restrict__ phi_ref, double *__restrict__ phi) {
  pack_phi_body(blockIdx.x, gridDim.x, nx, ny, nz, solid, phi_ref, phi);
}

// Writes the dense field back to the caller's halo layout with the physical
// halo at zero (what the reference leaves there after its init loop).
__global__ void k_unpack_phi(int nx, int ny, int nz, const double *__restrict__ phi,
                             double *__restrict__ phi_ref) {
  const size_t n = (size_t)(nx + 2) * (ny + 2) * nz;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n;
       p += (size_t)gridDim.x * blockDim.x) {
    const int I = (int)(p % (nx + 2)), J = (int)((p / (nx + 2)) % (ny + 2));
    const int k = (int)(p / ((size_t)(nx + 2) * (ny + 2)));
    double v = 0.0;
    if (I >= 1 && I <= nx && J >= 1 && J <= ny)
      v = phi[(size_t)(I - 1) + (size_t)nx * ((size_t)(J - 1) + (size_t)ny * k)];
    phi_ref[p] = v;
  }
}

// jacobi_sweep (src/fpm_stairstep_mpi_3d.f90:384-399), z-streaming.
__global__ __launch_bounds__(512) void k_jacobi(Tiling t, Geom g, const double *__restrict__ a,
                                                double *__restrict__ b,
                                                const double *__restrict__ src,
                                                const double *__restrict__ rdiag) {
  const TileCoord tc = tile_coord(t, g.nx, g.ny, g.nz);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (!tc.valid || tc.i >= nx || tc.j >= ny) return;
  const int i = tc.i, j = tc.j;
  const size_t nxy = (size_t)nx * ny;
  size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * tc.k0);
  const double cw = g.cw[i], ce = g.ce[i], cs = g.cs[j], cn = g.cn[j];
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
  double a_cur = a[c];
  double a_prev = tc.k0 > 0 ? a[c - nxy] : a_cur;
  for (int k = tc.k0; k < tc.k1; ++k) {
    const double a_next = (k + 1 < nz) ? a[c + nxy] : a_cur;
    const double aw = hw ? a[c - 1] : 0.0;
    const double ae = he ? a[c + 1] : 0.0;
    const double as = hs ? a[c - nx] : 0.0;
    const double an = hn ? a[c + nx] : 0.0;
    b[c] = 0.2 * a_cur + rdiag[c] * (src[c] + cw * aw + ce * ae + cs * as + cn * an +
                                     g.cb[k] * a_prev + g.ct[k] * a_next);
    a_prev = a_cur;
    a_cur = a_next;
    c += nxy;
  }
}

// operator_residual (src/fpm_stairstep_mpi_3d.f90:514-530); the partials'
// maximum goes to sc->res (k_finish FIN_MAX).
__global__ __launch_bounds__(512) void k_residual_max(Tiling t, Geom g,
                                                      const double *__restrict__ phi,
                                                      const double *__restrict__ src,
                                                      const double *__restrict__ diag,
                                                      double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  residual_max_body(t, blockIdx.x, gridDim.x, g, phi, src, diag, partials, lds);
}

// true_residual, first pass (src/fpm_stairstep_mpi_3d.f90:303-319); body in
// ffpm_bodies.hpp.
__global__ __launch_bounds__(512) void k_true_residual_1(Tiling t, Geom g,
                                                         const double *__restrict__ phi,
                                                         const double *__restrict__ src,
                                                         const double *__restrict__ diag,
                                                         double *__restrict__ r,
                                                         double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  true_residual_1_body(t, blockIdx.x, gridDim.x, g, phi, src, diag, r, partials, lds);
}

// true_residual, second pass (src/fpm_stairstep_mpi_3d.f90:319-328).
__global__ __launch_bounds__(512) void k_true_residual_2(size_t n, double *__restrict__ r,
                                                         const double *__restrict__ rdiag,
                                                         const double *__restrict__ diag,
                                                         double *__restrict__ p, int reset,
                                                         double *__restrict__ z,
                                                         double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  true_residual_2_body(n, blockIdx.x, gridDim.x, r, rdiag, diag, p, reset, z, partials, sc, lds);
}

// PCG SpMV + dot (src/fpm_stairstep_mpi_3d.f90:246-265): q = A p with
// A = V (diag - offdiag), 0 in solids; pq = p.q; then the breakdown test and
// alpha = rz / pq on the device.
__global__ __launch_bounds__(512) void k_spmv_dot(Tiling t, Geom g, const double *__restrict__ p,
                                                  double *__restrict__ q,
                                                  const double *__restrict__ diag,
                                                  double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  if (sc->status != 0) return;
  const TileCoord tc = tile_coord(t, g.nx, g.ny, g.nz);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  double s[1] = {0.0};
  if (tc.valid && tc.i < nx && tc.j < ny) {
    const int i = tc.i, j = tc.j;
    const size_t nxy = (size_t)nx * ny;
    size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * tc.k0);
    const double cw = g.cw[i], ce = g.ce[i], cs = g.cs[j], cn = g.cn[j];
    const double dxdy = g.dx[i] * g.dy[j];
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    double p_cur = p[c];
    double p_prev = tc.k0 > 0 ? p[c - nxy] : p_cur;
    for (int k = tc.k0; k < tc.k1; ++k) {
      const double p_next = (k + 1 < nz) ? p[c + nxy] : p_cur;
      const double d = diag[c];
      double qv = 0.0;
      if (d > 0) {
        const double pw = hw ? p[c - 1] : 0.0;
        const double pe = he ? p[c + 1] : 0.0;
        const double ps = hs ? p[c - nx] : 0.0;
        const double pn = hn ? p[c + nx] : 0.0;
        qv = dxdy * g.dz[k] *
             (d * p_cur - cw * pw - ce * pe - cs * ps - cn * pn - g.cb[k] * p_prev -
              g.ct[k] * p_next);
      }
      q[c] = qv;
      s[0] += p_cur * qv;
      p_prev = p_cur;
      p_cur = p_next;
      c += nxy;
    }
  }
  block_sum<1>(s, lds);
  store_partials(partials, 1, s, gridDim.x);
}

// PCG update (src/fpm_stairstep_mpi_3d.f90:267-278): phi += alpha p,
// r -= alpha q, rz_new = r.M^-1 r; beta on the device; counts the iteration.
__global__ __launch_bounds__(512) void k_update_dot(size_t n, double *__restrict__ phi,
                                                    const double *__restrict__ p,
                                                    double *__restrict__ r,
                                                    const double *__restrict__ q,
                                                    const double *__restrict__ rdiag,
                                                    double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  if (sc->status != 0) return;
  const double alpha = sc->alpha;
  double s[1] = {0.0};
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    phi[c] = phi[c] + alpha * p[c];
    const double rv = r[c] - alpha * q[c];
    r[c] = rv;
    s[0] += rdiag[c] * rv * rv;
  }
  block_sum<1>(s, lds);
  store_partials(partials, 1, s, gridDim.x);
}

// PCG direction (src/fpm_stairstep_mpi_3d.f90:279-283): p = M^-1 r + beta p.
__global__ __launch_bounds__(512) void k_direction(size_t n, double *__restrict__ p,
                                                   const double *__restrict__ r,
                                                   const double *__restrict__ rdiag,
                                                   const Scalars *sc) {
  if (sc->status != 0) return;
  const double beta = sc->beta;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    p[c] = rdiag[c] * r[c] + beta * p[c];
}

// One workgroup: totals the partials of the reduction kernel launched just
// before on the same stream and applies the PCG scalar logic on the device.
__global__ __launch_bounds__(512) void k_finish(int op, int nslots, const double *partials,
                                                int nblocks, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  finish_body(op, nslots, partials, nblocks, sc, lds);
}

__global__ void k_chunk_begin(Scalars *sc) {
  if (threadIdx.x == 0) sc->iters = 0;
}

// After a breakdown: the multigrid cycle scale also starts afresh (a
// non-finite cycle output ends in a breakdown).
__global__ void k_clear_status(Scalars *sc) {
  if (threadIdx.x == 0) {
    sc->status = 0;
    sc->iters = 0;
    sc->sig = 0.0;
  }
}

// Gauge sums (src/fpm_stairstep_mpi_3d.f90:170-179) -> sc->sums[0..1] after
// k_finish FIN_SUMS.
__global__ __launch_bounds__(512) void k_gauge_sums(Geom g, const int32_t *__restrict__ solid,
                                                   const double *__restrict__ phi,
                                                   double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  gauge_sums_body(blockIdx.x, gridDim.x, g, solid, phi, partials, lds);
}

// Gauge shift (src/fpm_stairstep_mpi_3d.f90:180-185) with the caller's
// halo-layout copy written in the same pass.
__global__ void k_gauge_apply(int nx, int ny, int nz, const int32_t *__restrict__ solid,
                              double *__restrict__ phi, double *__restrict__ phi_ref,
                              const Scalars *sc) {
  gauge_apply_body(blockIdx.x, gridDim.x, nx, ny, nz, solid, phi, phi_ref, sc);
}

// Flux balance of src/fpm_stairstep_mpi_3d.f90:49-52,69-91: optional fluid
// source integral, net outward flux through fluid outer faces and body faces,
// and fluid outer-face area -> sc->sums[0..2] (source, flux, area). rhs null:
// the source is not formed (sums[0] = 0) and rhs is not read, as in the
// recomputation of the flux after the correction.
__global__ __launch_bounds__(512) void k_compat_sums(
    Geom g, const int32_t *__restrict__ solid, const double *__restrict__ rhs,
    const double *__restrict__ xlo, const double *__restrict__ xhi,
    const double *__restrict__ ylo, const double *__restrict__ yhi,
    const double *__restrict__ zlo, const double *__restrict__ zhi, double bg0, double bg1,
    double bg2, double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  double s[3] = {0.0, 0.0, 0.0};
  const size_t c0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x,
               stride = (size_t)gridDim.x * blockDim.x;
  Walk3 w(c0, stride, nx, ny);
  for (size_t c = c0; c < n; c += stride, w.next()) {
    const int i = (int)w.i, j = (int)w.j, k = (int)w.k;
    const int I = i + 1, J = j + 1;
    // operands loaded before the branches, as in build_operator_body
    const size_t pc = ph(I, J, k, nx, ny), pxy = (size_t)(nx + 2) * (ny + 2);
    const int32_t s0 = solid[pc], sw = solid[pc - 1], se = solid[pc + 1],
                  ss = solid[pc - (nx + 2)], sn = solid[pc + (nx + 2)],
                  sb = k > 0 ? solid[pc - pxy] : 0, st = k < nz - 1 ? solid[pc + pxy] : 0;
    const double rv = rhs ? rhs[c] : 0.0;
    const double dxi = g.dx[i], dyj = g.dy[j], dzk = g.dz[k];
    if (s0 != 0) continue;
    if (rhs) s[0] += rv * dxi * dyj * dzk;
    double f = 0.0, ar = 0.0, a;
    a = dyj * dzk;
    if (I == 1) { f = f - xlo[j + (size_t)ny * k] * a; ar = ar + a; }
    else if (sw != 0) { f = f - bg0 * a; }
    if (I == nx) { f = f + xhi[j + (size_t)ny * k] * a; ar = ar + a; }
    else if (se != 0) { f = f + bg0 * a; }
    a = dxi * dzk;
    if (J == 1) { f = f - ylo[i + (size_t)nx * k] * a; ar = ar + a; }
    else if (ss != 0) { f = f - bg1 * a; }
    if (J == ny) { f = f + yhi[i + (size_t)nx * k] * a; ar = ar + a; }
    else if (sn != 0) { f = f + bg1 * a; }
    a = dxi * dyj;
    if (k == 0) { f = f - zlo[i + (size_t)nx * j] * a; ar = ar + a; }
    else if (sb != 0) { f = f - bg2 * a; }
    if (k == nz - 1) { f = f + zhi[i + (size_t)nx * j] * a; ar = ar + a; }
    else if (st != 0) { f = f + bg2 * a; }
    s[1] += f;
    s[2] += ar;
  }
  block_sum<3>(s, lds);
  store_partials(partials, 3, s, gridDim.x);
}

// Uniform correction of the outer-face data (src/fpm_stairstep_mpi_3d.f90:58-62).
__global__ void k_shift_faces(int nx, int ny, int nz, double *xlo, double *xhi, double *ylo,
                              double *yhi, double *zlo, double *zhi, double c) {
  const size_t nyz = (size_t)ny * nz, nxz = (size_t)nx * nz, nxy = (size_t)nx * ny;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < nyz + nxz + nxy;
       t += (size_t)gridDim.x * blockDim.x) {
    if (t < nyz) {
      xlo[t] = xlo[t] - c;
      xhi[t] = xhi[t] + c;
    } else if (t < nyz + nxz) {
      const size_t u = t - nyz;
      ylo[u] = ylo[u] - c;
      yhi[u] = yhi[u] + c;
    } else {
      const size_t u = t - nyz - nxz;
      zlo[u] = zlo[u] - c;
      zhi[u] = zhi[u] + c;
    }
  }
}

// face_gradients (src/fpm_stairstep_mpi_3d.f90:472-497) of cell (I,J,k),
// I,J 1-based, k 0-based; phi in the halo layout.
__device__ __forceinline__ void face_gradients(const Geom &g, const double *__restrict__ phi,
                                               const int32_t *__restrict__ solid,
                                               const double *__restrict__ xlo,
                                               const double *__restrict__ xhi,
                                               const double *__restrict__ ylo,
                                               const double *__restrict__ yhi,
                                               const double *__restrict__ zlo,
                                               const double *__restrict__ zhi, double bg0,
                                               double bg1, double bg2, int I, int J, int k,
                                               double gr[6]) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int i = I - 1, j = J - 1;
  const size_t pc = ph(I, J, k, nx, ny);
  const size_t sy = (size_t)(nx + 2), sz = (size_t)(nx + 2) * (ny + 2);
  const bool kb = k > 0, kt = k < nz - 1;
  // every operand is loaded before the branches (the halo layout holds the
  // x and y neighbours of every cell; z neighbours only inside); the
  // arithmetic is the reference's
  const double p0 = phi[pc], pw = phi[pc - 1], pe = phi[pc + 1], ps = phi[pc - sy],
               pn = phi[pc + sy], pb = kb ? phi[pc - sz] : 0.0, pt = kt ? phi[pc + sz] : 0.0;
  const int32_t sw = solid[pc - 1], se = solid[pc + 1], ss = solid[pc - sy],
                sn = solid[pc + sy], sb = kb ? solid[pc - sz] : 0, st = kt ? solid[pc + sz] : 0;
  const double xm = g.xc[I - 1], x0 = g.xc[I], xp = g.xc[I + 1];
  const double ym = g.yc[J - 1], y0 = g.yc[J], yp = g.yc[J + 1];
  const double zm = kb ? g.zc[k - 1] : 0.0, z0 = g.zc[k], zp = kt ? g.zc[k + 1] : 0.0;
  if (I == 1) gr[0] = xlo[j + (size_t)ny * k];
  else if (sw != 0) gr[0] = bg0;
  else gr[0] = (p0 - pw) / (x0 - xm);
  if (I == nx) gr[1] = xhi[j + (size_t)ny * k];
  else if (se != 0) gr[1] = bg0;
  else gr[1] = (pe - p0) / (xp - x0);
  if (J == 1) gr[2] = ylo[i + (size_t)nx * k];
  else if (ss != 0) gr[2] = bg1;
  else gr[2] = (p0 - ps) / (y0 - ym);
  if (J == ny) gr[3] = yhi[i + (size_t)nx * k];
  else if (sn != 0) gr[3] = bg1;
  else gr[3] = (pn - p0) / (yp - y0);
  if (k == 0) gr[4] = zlo[i + (size_t)nx * j];
  else if (sb != 0) gr[4] = bg2;
  else gr[4] = (p0 - pb) / (z0 - zm);
  if (k == nz - 1) gr[5] = zhi[i + (size_t)nx * j];
  else if (st != 0) gr[5] = bg2;
  else gr[5] = (pt - p0) / (zp - z0);
}

// stairstep_residual_3d (src/fpm_stairstep_mpi_3d.f90:403-419) from face
// gradients -> sc->res.
__global__ __launch_bounds__(512) void k_face_residual(
    Geom g, const double *__restrict__ phi, const double *__restrict__ rhs,
    const int32_t *__restrict__ solid, const double *__restrict__ xlo,
    const double *__restrict__ xhi, const double *__restrict__ ylo,
    const double *__restrict__ yhi, const double *__restrict__ zlo,
    const double *__restrict__ zhi, double bg0, double bg1, double bg2, double *partials,
    Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  double m = 0.0;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((size_t)nx * ny));
    if (solid[ph(i + 1, j + 1, k, nx, ny)] != 0) continue;
    double gr[6];
    face_gradients(g, phi, solid, xlo, xhi, ylo, yhi, zlo, zhi, bg0, bg1, bg2, i + 1, j + 1, k,
                   gr);
    m = fmax(m, fabs((gr[1] - gr[0]) / g.dx[i] + (gr[3] - gr[2]) / g.dy[j] +
                     (gr[5] - gr[4]) / g.dz[k] - rhs[c]));
  }
  block_max(m, lds);
  store_partials(partials, 1, &m, gridDim.x);
}

// compute_stairstep_fpm_terms_3d (src/fpm_stairstep_mpi_3d.f90:426-469):
// face gradients fused with the densities and their bounded integrals
// -> sc->sums[0..1].
__global__ __launch_bounds__(512) void k_terms(
    Geom g, const double *__restrict__ phi, const int32_t *__restrict__ solid,
    const double *__restrict__ u, const double *__restrict__ v, const double *__restrict__ w,
    const double *__restrict__ ox, const double *__restrict__ oy,
    const double *__restrict__ oz, const double *__restrict__ ax,
    const double *__restrict__ ay, const double *__restrict__ az,
    const double *__restrict__ xlo, const double *__restrict__ xhi,
    const double *__restrict__ ylo, const double *__restrict__ yhi,
    const double *__restrict__ zlo, const double *__restrict__ zhi, double bg0, double bg1,
    double bg2, Bounds bd, double *__restrict__ gx, double *__restrict__ gy,
    double *__restrict__ gz, double *__restrict__ vortex, double *__restrict__ advection,
    double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  double s[2] = {0.0, 0.0};
  const size_t c0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x,
               stride = (size_t)gridDim.x * blockDim.x;
  Walk3 wk(c0, stride, nx, ny);
  for (size_t c = c0; c < n; c += stride, wk.next()) {
    const int i = (int)wk.i, j = (int)wk.j, k = (int)wk.k;
    // the flow fields are loaded with the gradients' operands; they and the
    // outputs are touched once, so they bypass the caches that hold phi
    const double uu = __builtin_nontemporal_load(u + c), vv = __builtin_nontemporal_load(v + c),
                 ww = __builtin_nontemporal_load(w + c),
                 ox_ = __builtin_nontemporal_load(ox + c),
                 oy_ = __builtin_nontemporal_load(oy + c),
                 oz_ = __builtin_nontemporal_load(oz + c);
    const double axc = __builtin_nontemporal_load(ax + c),
                 ayc = __builtin_nontemporal_load(ay + c),
                 azc = __builtin_nontemporal_load(az + c);
    double gr[6];
    face_gradients(g, phi, solid, xlo, xhi, ylo, yhi, zlo, zhi, bg0, bg1, bg2, i + 1, j + 1, k,
                   gr);
    const double mask = solid[ph(i + 1, j + 1, k, nx, ny)] == 0 ? 1.0 : 0.0;
    const double gxv = mask * 0.5 * (gr[0] + gr[1]);
    const double gyv = mask * 0.5 * (gr[2] + gr[3]);
    const double gzv = mask * 0.5 * (gr[4] + gr[5]);
    __builtin_nontemporal_store(gxv, gx + c);
    __builtin_nontemporal_store(gyv, gy + c);
    __builtin_nontemporal_store(gzv, gz + c);
    const double cx = ww * oy_ - vv * oz_;
    const double cy = uu * oz_ - ww * ox_;
    const double cz = vv * ox_ - uu * oy_;
    const double vo = -(cx * gxv + cy * gyv + cz * gzv);
    const double ad = -(axc * gxv + ayc * gyv + azc * gzv);
    __builtin_nontemporal_store(vo, vortex + c);
    __builtin_nontemporal_store(ad, advection + c);
    const double X = g.xc[i + 1], Y = g.yc[j + 1], Z = g.zc[k];
    const bool inside = X > bd.b[0] && X < bd.b[1] && Y > bd.b[2] && Y < bd.b[3] &&
                        Z > bd.b[4] && Z < bd.b[5];
    const double volume = inside ? g.dx[i] * g.dy[j] * g.dz[k] : 0.0;
    s[0] += volume * vo;
    s[1] += volume * ad;
  }
  block_sum<2>(s, lds);
  store_partials(partials, 2, s, gridDim.x);
}

// compute_stairstep_fpm_terms_3d (src/fpm_stairstep_mpi_3d.f90:426-469) for
// ND directions in one pass: the nine flow fields of a cell are loaded once
// and each direction's face gradients, densities and bounded integrals are
// formed as k_terms forms them (the same cells per thread in the same order,
// one block_sum per direction), so every direction's outputs and partial
// sums equal those of its own k_terms launch bit for bit. Partials:
// direction d's two slots at partials + 2 d gridDim.x.
template <int ND>
__global__ __launch_bounds__(512) void k_terms_dirs(
    Geom g, TermsDirs a, const int32_t *__restrict__ solid, const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, const double *__restrict__ ox,
    const double *__restrict__ oy, const double *__restrict__ oz, const double *__restrict__ ax,
    const double *__restrict__ ay, const double *__restrict__ az, Bounds bd, double *partials) {
  __shared__ double lds[16 * kReduceSlots];
  const int nx = g.nx, ny = g.ny;
  const size_t n = (size_t)nx * ny * g.nz;
  double s[ND][2];
#pragma unroll
  for (int d = 0; d < ND; ++d) s[d][0] = s[d][1] = 0.0;
  const size_t c0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x,
               stride = (size_t)gridDim.x * blockDim.x;
  Walk3 wk(c0, stride, nx, ny);
  for (size_t c = c0; c < n; c += stride, wk.next()) {
    const int i = (int)wk.i, j = (int)wk.j, k = (int)wk.k;
    const double uu = __builtin_nontemporal_load(u + c), vv = __builtin_nontemporal_load(v + c),
                 ww = __builtin_nontemporal_load(w + c),
                 ox_ = __builtin_nontemporal_load(ox + c),
                 oy_ = __builtin_nontemporal_load(oy + c),
                 oz_ = __builtin_nontemporal_load(oz + c);
    const double axc = __builtin_nontemporal_load(ax + c),
                 ayc = __builtin_nontemporal_load(ay + c),
                 azc = __builtin_nontemporal_load(az + c);
    const double mask = solid[ph(i + 1, j + 1, k, nx, ny)] == 0 ? 1.0 : 0.0;
    const double cx = ww * oy_ - vv * oz_;
    const double cy = uu * oz_ - ww * ox_;
    const double cz = vv * ox_ - uu * oy_;
    const double X = g.xc[i + 1], Y = g.yc[j + 1], Z = g.zc[k];
    const bool inside = X > bd.b[0] && X < bd.b[1] && Y > bd.b[2] && Y < bd.b[3] &&
                        Z > bd.b[4] && Z < bd.b[5];
    const double volume = inside ? g.dx[i] * g.dy[j] * g.dz[k] : 0.0;
    // every direction's gradients before any store: the outputs may alias
    // the potentials as far as the compiler knows, so a store between them
    // would hold the next direction's loads back
    double gr[ND][6];
#pragma unroll
    for (int d = 0; d < ND; ++d)
      face_gradients(g, a.phi[d], solid, a.xlo[d], a.xhi[d], a.ylo[d], a.yhi[d], a.zlo[d],
                     a.zhi[d], a.bg[d][0], a.bg[d][1], a.bg[d][2], i + 1, j + 1, k, gr[d]);
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const double gxv = mask * 0.5 * (gr[d][0] + gr[d][1]);
      const double gyv = mask * 0.5 * (gr[d][2] + gr[d][3]);
      const double gzv = mask * 0.5 * (gr[d][4] + gr[d][5]);
      __builtin_nontemporal_store(gxv, a.gx[d] + c);
      __builtin_nontemporal_store(gyv, a.gy[d] + c);
      __builtin_nontemporal_store(gzv, a.gz[d] + c);
      const double vo = -(cx * gxv + cy * gyv + cz * gzv);
      const double ad = -(axc * gxv + ayc * gyv + azc * gzv);
      __builtin_nontemporal_store(vo, a.vortex[d] + c);
      __builtin_nontemporal_store(ad, a.advection[d] + c);
      s[d][0] += volume * vo;
      s[d][1] += volume * ad;
    }
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    block_sum<2>(s[d], lds);
    store_partials(partials + (size_t)2 * d * gridDim.x, 2, s[d], gridDim.x);
  }
}

template __global__ void k_terms_dirs<1>(Geom, TermsDirs, const int32_t *, const double *,
                                         const double *, const double *, const double *,
                                         const double *, const double *, const double *,
                                         const double *, const double *, Bounds, double *);
template __global__ void k_terms_dirs<2>(Geom, TermsDirs, const int32_t *, const double *,
                                         const double *, const double *, const double *,
                                         const double *, const double *, const double *,
                                         const double *, const double *, Bounds, double *);
template __global__ void k_terms_dirs<3>(Geom, TermsDirs, const int32_t *, const double *,
                                         const double *, const double *, const double *,
                                         const double *, const double *, const double *,
                                         const double *, const double *, Bounds, double *);

// Totals of k_terms_dirs: direction d's two sums into out[2d], out[2d + 1],
// summed as k_finish sums FIN_SUMS partials (one workgroup of the same size).
__global__ void k_terms_dirs_finish(int nd, const double *partials, int nblocks, double *out) {
  __shared__ double lds[16 * kReduceSlots];
  for (int d = 0; d < nd; ++d) {
    double v[2];
    grid_finish_sum<2>(partials + (size_t)2 * d * nblocks, nblocks, v, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      out[2 * d] = v[0];
      out[2 * d + 1] = v[1];
    }
  }
}

}  // namespace ffpm
