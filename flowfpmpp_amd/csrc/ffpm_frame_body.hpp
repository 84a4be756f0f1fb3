// Device bodies of a solve's set-up and finishing kernels (1-D coefficients,
// operator build, dense copy of phi, gauge, final residual), shared by the
// single-solve kernels (ffpm_stairstep_kernels.hip) and the batched ones
// (ffpm_batch.hip). Every body takes its workgroup's index within its own
// solve's launch, so a batched solve computes what its standalone solve
// computes, reductions included. Elementwise bodies are grid-stride loops
// over workgroups of blockDim.x threads.
#pragma once

#include "ffpm_bodies.hpp"

namespace ffpm {

// (0:nx+1, 0:ny+1, 1:nz) halo layout, 0-based z
__device__ __forceinline__ size_t ph(int i, int j, int k, int nx, int ny) {
  return (size_t)i + (size_t)(nx + 2) * ((size_t)j + (size_t)(ny + 2) * (size_t)k);
}

// (i, j, k) of the cells a grid-stride loop visits in an nx x ny x (any)
// box: the first from one division, every later one by adding the stride's
// own (i, j, k) with a carry per axis (no 64-bit division per cell).
struct Walk3 {
  unsigned i, j, k, si, sj, sk, nx, ny;
  __device__ __forceinline__ Walk3(size_t c, size_t stride, int nx_, int ny_)
      : nx((unsigned)nx_), ny((unsigned)ny_) {
    const size_t nxy = (size_t)nx_ * ny_;
    i = (unsigned)(c % nx_);
    j = (unsigned)((c / nx_) % ny_);
    k = (unsigned)(c / nxy);
    si = (unsigned)(stride % nx_);
    sj = (unsigned)((stride / nx_) % ny_);
    sk = (unsigned)(stride / nxy);
  }
  __device__ __forceinline__ void next() {
    i += si;
    j += sj;
    k += sk;
    if (i >= nx) {
      i -= nx;
      ++j;
    }
    if (j >= ny) {
      j -= ny;
      ++k;
    }
  }
};

// axis_coefficients (src/fpm_stairstep_mpi_3d.f90:499-510) and the z
// coefficients (src/fpm_stairstep_mpi_3d.f90:138-142): entry t of the
// nx + ny + nz.
__device__ __forceinline__ void axis_coefficients_body(int t, int nx, int ny, int nz,
                                                       const double *xc, const double *yc,
                                                       const double *zc, const double *dx,
                                                       const double *dy, const double *dz,
                                                       double *cw, double *ce, double *cs,
                                                       double *cn, double *cb, double *ct) {
  if (t < nx) {
    const int m = t + 1;
    cw[t] = (m == 1) ? 0.0 : 1 / (dx[t] * (xc[m] - xc[m - 1]));
    ce[t] = (m == nx) ? 0.0 : 1 / (dx[t] * (xc[m + 1] - xc[m]));
  } else if (t < nx + ny) {
    const int u = t - nx, m = u + 1;
    cs[u] = (m == 1) ? 0.0 : 1 / (dy[u] * (yc[m] - yc[m - 1]));
    cn[u] = (m == ny) ? 0.0 : 1 / (dy[u] * (yc[m + 1] - yc[m]));
  } else if (t < nx + ny + nz) {
    const int k = t - nx - ny;  // 0-based; zc is (1:nz)
    cb[k] = (k > 0) ? 1 / (dz[k] * (zc[k] - zc[k - 1])) : 0.0;
    ct[k] = (k < nz - 1) ? 1 / (dz[k] * (zc[k + 1] - zc[k])) : 0.0;
  }
}

// build_operator (src/fpm_stairstep_mpi_3d.f90:336-379): folds the outer and
// body face gradients into src, sums fluid-neighbour coefficients into diag,
// and sets rdiag = 1/(V diag) (PCG) or 0.8/diag (Jacobi); 0 in solids
// (rdiag null: not formed, for solvers that rebuild what they need from the
// neighbour code). With write_operator false only src is stored (another
// right-hand side of an operator that is already built).
__device__ __forceinline__ void build_operator_body(
    int block, int nblocks, const Geom &g, const int32_t *__restrict__ solid,
    const double *__restrict__ rhs, const double *__restrict__ xlo,
    const double *__restrict__ xhi, const double *__restrict__ ylo,
    const double *__restrict__ yhi, const double *__restrict__ zlo,
    const double *__restrict__ zhi, double bg0, double bg1, double bg2, int for_pcg,
    double *__restrict__ src, double *__restrict__ diag, double *__restrict__ rdiag,
    uint8_t *__restrict__ code, int *__restrict__ isolated, bool write_operator) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  int local_isolated = 0;
  const size_t c0 = block * (size_t)blockDim.x + threadIdx.x, stride = (size_t)nblocks * blockDim.x;
  Walk3 w(c0, stride, nx, ny);
  for (size_t c = c0; c < n; c += stride, w.next()) {
    const int i = (int)w.i, j = (int)w.j, k = (int)w.k;
    const int I = i + 1, J = j + 1;
    // every operand is loaded before the branches (a load under each branch
    // waits for the previous one); the arithmetic keeps the reference's order
    const size_t pc = ph(I, J, k, nx, ny), pxy = (size_t)(nx + 2) * (ny + 2);
    const int32_t s0 = solid[pc], sw = solid[pc - 1], se = solid[pc + 1],
                  ss = solid[pc - (nx + 2)], sn = solid[pc + (nx + 2)],
                  sb = k > 0 ? solid[pc - pxy] : 0, st = k < nz - 1 ? solid[pc + pxy] : 0;
    const double rv = rhs[c];
    const double dxi = g.dx[i], dyj = g.dy[j], dzk = g.dz[k];
    const double cwi = g.cw[i], cei = g.ce[i], csj = g.cs[j], cnj = g.cn[j], cbk = g.cb[k],
                 ctk = g.ct[k];
    double s = 0, d = 0;
    unsigned cb = 0;
    if (s0 == 0) {
      cb = kCodeFluid;
      s = -rv;
      if (I == 1) s = s - xlo[j + (size_t)ny * k] / dxi;
      else if (sw != 0) s = s - bg0 / dxi;
      else { d = d + cwi; cb |= kCodeW; }
      if (I == nx) s = s + xhi[j + (size_t)ny * k] / dxi;
      else if (se != 0) s = s + bg0 / dxi;
      else { d = d + cei; cb |= kCodeE; }
      if (J == 1) s = s - ylo[i + (size_t)nx * k] / dyj;
      else if (ss != 0) s = s - bg1 / dyj;
      else { d = d + csj; cb |= kCodeS; }
      if (J == ny) s = s + yhi[i + (size_t)nx * k] / dyj;
      else if (sn != 0) s = s + bg1 / dyj;
      else { d = d + cnj; cb |= kCodeN; }
      if (k == 0) s = s - zlo[i + (size_t)nx * j] / dzk;
      else if (sb != 0) s = s - bg2 / dzk;
      else { d = d + cbk; cb |= kCodeB; }
      if (k == nz - 1) s = s + zhi[i + (size_t)nx * j] / dzk;
      else if (st != 0) s = s + bg2 / dzk;
      else { d = d + ctk; cb |= kCodeT; }
      if (d <= 0) local_isolated += 1;
    }
    src[c] = s;
    if (write_operator) {
      diag[c] = d;
      code[c] = (uint8_t)cb;
      if (rdiag) {
        double rd;
        if (d <= 0) rd = 0;
        else if (for_pcg) rd = 1 / (dxi * dyj * dzk * d);
        else rd = 0.8 / d;
        rdiag[c] = rd;
      }
    }
  }
  if (local_isolated) atomicAdd(isolated, local_isolated);
}

// Init loop of src/fpm_stairstep_mpi_3d.f90:154-160 on the dense copy: phi in
// solid cells becomes 0 (the halo is substituted, not stored).
__device__ __forceinline__ void pack_phi_body(int block, int nblocks, int nx, int ny, int nz,
                                              const int32_t *__restrict__ solid,
                                              const double *__restrict__ phi_ref,
                                              double *__restrict__ phi) {
  const size_t n = (size_t)nx * ny * nz;
  const size_t c0 = block * (size_t)blockDim.x + threadIdx.x, stride = (size_t)nblocks * blockDim.x;
  Walk3 w(c0, stride, nx, ny);
  for (size_t c = c0; c < n; c += stride, w.next()) {
    const size_t p = ph((int)w.i + 1, (int)w.j + 1, (int)w.k, nx, ny);
    phi[c] = solid[p] != 0 ? 0.0 : phi_ref[p];
  }
}

// operator_residual (src/fpm_stairstep_mpi_3d.f90:514-530): max-norm of
// src + sum(c phi_nb) - diag phi over diag > 0; one partial per workgroup.
__device__ __forceinline__ void residual_max_body(const Tiling &t, int block, int nblocks,
                                                  const Geom &g, const double *__restrict__ phi,
                                                  const double *__restrict__ src,
                                                  const double *__restrict__ diag,
                                                  double *partials, double *lds) {
  const TileCoord tc = tile_coord(t, g.nx, g.ny, g.nz, block);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  double m = 0.0;
  if (tc.valid && tc.i < nx && tc.j < ny) {
    const int i = tc.i, j = tc.j;
    const size_t nxy = (size_t)nx * ny;
    size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * tc.k0);
    const double cw = g.cw[i], ce = g.ce[i], cs = g.cs[j], cn = g.cn[j];
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    double p_cur = phi[c];
    double p_prev = tc.k0 > 0 ? phi[c - nxy] : p_cur;
    for (int k = tc.k0; k < tc.k1; ++k) {
      const double p_next = (k + 1 < nz) ? phi[c + nxy] : p_cur;
      const double d = diag[c];
      if (d > 0) {
        const double pw = hw ? phi[c - 1] : 0.0;
        const double pe = he ? phi[c + 1] : 0.0;
        const double ps = hs ? phi[c - nx] : 0.0;
        const double pn = hn ? phi[c + nx] : 0.0;
        const double v = fabs(src[c] + cw * pw + ce * pe + cs * ps + cn * pn + g.cb[k] * p_prev +
                              g.ct[k] * p_next - d * p_cur);
        m = fmax(m, v);
      }
      p_prev = p_cur;
      p_cur = p_next;
      c += nxy;
    }
  }
  block_max(m, lds);
  store_partials(partials, 1, &m, nblocks, block);
}

// Gauge sums (src/fpm_stairstep_mpi_3d.f90:170-179): fluid-volume integral
// of phi and the fluid volume, two partial slots per workgroup.
__device__ __forceinline__ void gauge_sums_body(int block, int nblocks, const Geom &g,
                                                const int32_t *__restrict__ solid,
                                                const double *__restrict__ phi,
                                                double *partials, double *lds) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t n = (size_t)nx * ny * nz;
  double s[2] = {0.0, 0.0};
  const size_t c0 = block * (size_t)blockDim.x + threadIdx.x, stride = (size_t)nblocks * blockDim.x;
  Walk3 w(c0, stride, nx, ny);
  for (size_t c = c0; c < n; c += stride, w.next()) {
    const int i = (int)w.i, j = (int)w.j, k = (int)w.k;
    if (solid[ph(i + 1, j + 1, k, nx, ny)] == 0) {
      s[0] += phi[c] * g.dx[i] * g.dy[j] * g.dz[k];
      s[1] += g.dx[i] * g.dy[j] * g.dz[k];
    }
  }
  block_sum<2>(s, lds);
  store_partials(partials, 2, s, nblocks, block);
}

// Gauge shift (src/fpm_stairstep_mpi_3d.f90:180-185) on the dense field, with
// the caller's halo-layout copy written in the same pass (mean = sums[0] /
// sums[1] of sc).
__device__ __forceinline__ void gauge_apply_body(int block, int nblocks, int nx, int ny, int nz,
                                                 const int32_t *__restrict__ solid,
                                                 double *__restrict__ phi,
                                                 double *__restrict__ phi_ref,
                                                 const Scalars *sc) {
  const double mean_value = sc->sums[0] / sc->sums[1];
  const size_t n = (size_t)(nx + 2) * (ny + 2) * nz;
  const size_t p0 = block * (size_t)blockDim.x + threadIdx.x, stride = (size_t)nblocks * blockDim.x;
  Walk3 w(p0, stride, nx + 2, ny + 2);
  for (size_t p = p0; p < n; p += stride, w.next()) {
    const int I = (int)w.i, J = (int)w.j, k = (int)w.k;
    double v = 0.0;
    if (I >= 1 && I <= nx && J >= 1 && J <= ny) {
      const size_t c = (size_t)(I - 1) + (size_t)nx * ((size_t)(J - 1) + (size_t)ny * k);
      v = phi[c];
      if (solid[p] == 0) {
        v = v - mean_value;
        phi[c] = v;
      }
    }
    phi_ref[p] = v;
  }
}

}  // namespace ffpm
