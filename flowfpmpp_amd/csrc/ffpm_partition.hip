// Force partition of a 2-D stair-step body with rigid in-plane motion
// (src/fpm_partition_2d.f90:35-154): F_KINE, F_VIF, F_VIS and F_BOUND per unit
// span and density, the F_VIF density phi div[(u.grad)u] and the vorticity.
// Two passes, one thread per cell: pass 1 stores the convective acceleration
// (u.grad)u, the vorticity and grad(phi) of every fluid cell; pass 2 forms the
// face fluxes, the densities and each workgroup's share of the four terms,
// totalled in a fixed order by k_finish. Every per-cell expression keeps the
// reference's order, so the densities and vorticity are bit-identical to the
// serial reference; the terms agree to summation-order roundoff.
#include <vector>

#include "../../include/ffpm.h"
#include "ffpm_bodies.hpp"
#include "ffpm_runtime.hpp"

namespace ffpm {

namespace {

constexpr int kThreads = 256;

struct Kin {
  double g[2];
  double k[6];
  double dt, nu;
  int slip;
};

struct Face {
  int ii, jj, axis;
  double sgn, hw, len, d, fx, fy;
};

// face(i,j,f) (src/fpm_partition_2d.f90:122-140), 0-based cells; f = 1 east,
// 2 west, 3 north, 4 south.
__device__ inline Face face_of(int i, int j, int f, int nx, int ny, const double *xc,
                               const double *yc, const double *dx, const double *dy) {
  Face o;
  if (f == 1) { o.ii = i + 1; o.jj = j; o.axis = 1; o.sgn = 1; }
  else if (f == 2) { o.ii = i - 1; o.jj = j; o.axis = 1; o.sgn = -1; }
  else if (f == 3) { o.ii = i; o.jj = j + 1; o.axis = 2; o.sgn = 1; }
  else { o.ii = i; o.jj = j - 1; o.axis = 2; o.sgn = -1; }
  if (o.axis == 1) {
    o.hw = 0.5 * dx[i];
    o.len = dy[j];
    o.fx = xc[i] + o.sgn * o.hw;
    o.fy = yc[j];
    o.d = o.hw;
    if (o.ii >= 0 && o.ii < nx) o.d = fabs(xc[o.ii] - xc[i]);
  } else {
    o.hw = 0.5 * dy[j];
    o.len = dx[i];
    o.fx = xc[i];
    o.fy = yc[j] + o.sgn * o.hw;
    o.d = o.hw;
    if (o.jj >= 0 && o.jj < ny) o.d = fabs(yc[o.jj] - yc[j]);
  }
  return o;
}

// ring_index (src/fpm_partition_2d.f90:142-151), 0-based.
__device__ inline int ring_index(int i, int j, int f, int nx, int ny) {
  if (f == 2) return j;
  if (f == 1) return ny + j;
  if (f == 4) return 2 * ny + i;
  return 2 * ny + nx + i;
}

__global__ void k_part_axes(int nx, int ny, const double *xf, const double *yf, double *xc,
                            double *yc, double *dx, double *dy) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nx) {
    dx[t] = xf[t + 1] - xf[t];
    xc[t] = 0.5 * (xf[t] + xf[t + 1]);
  } else if (t < nx + ny) {
    const int j = t - nx;
    dy[j] = yf[j + 1] - yf[j];
    yc[j] = 0.5 * (yf[j] + yf[j + 1]);
  }
}

// Pass 1 (src/fpm_partition_2d.f90:60-84).
__global__ __launch_bounds__(256) void k_part_pass1(
    int nx, int ny, Kin kn, const double *__restrict__ phi, const double *__restrict__ u,
    const double *__restrict__ v, const int32_t *__restrict__ solid,
    const double *__restrict__ xc, const double *__restrict__ yc,
    const double *__restrict__ dx, const double *__restrict__ dy, double *__restrict__ nxf,
    double *__restrict__ nyf, double *__restrict__ gpx, double *__restrict__ gpy,
    double *__restrict__ omega) {
  const size_t n = (size_t)nx * ny;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx), j = (int)(c / nx);
    if (solid[c]) {
      nxf[c] = 0;
      nyf[c] = 0;
      gpx[c] = 0;
      gpy[c] = 0;
      omega[c] = 0;
      continue;
    }
    double fu[4], fv[4], fp[4];
    for (int f = 1; f <= 4; ++f) {
      const Face o = face_of(i, j, f, nx, ny, xc, yc, dx, dy);
      const size_t cn = (size_t)o.ii + (size_t)nx * o.jj;
      if (o.ii < 0 || o.ii >= nx || o.jj < 0 || o.jj >= ny) {
        fu[f - 1] = u[c];
        fv[f - 1] = v[c];
        fp[f - 1] = 0;
      } else if (solid[cn]) {
        if (kn.slip) {
          fu[f - 1] = u[c];
          fv[f - 1] = v[c];
        } else {
          fu[f - 1] = kn.k[0] - kn.k[2] * o.fy;
          fv[f - 1] = kn.k[1] + kn.k[2] * o.fx;
        }
        fp[f - 1] = kn.g[o.axis - 1];
      } else {
        fu[f - 1] = u[c] + (u[cn] - u[c]) * o.hw / o.d;
        fv[f - 1] = v[c] + (v[cn] - v[c]) * o.hw / o.d;
        fp[f - 1] = o.sgn * (phi[cn] - phi[c]) / o.d;
      }
    }
    nxf[c] = u[c] * (fu[0] - fu[1]) / dx[i] + v[c] * (fu[2] - fu[3]) / dy[j];
    nyf[c] = u[c] * (fv[0] - fv[1]) / dx[i] + v[c] * (fv[2] - fv[3]) / dy[j];
    omega[c] = (fv[0] - fv[1]) / dx[i] - (fu[2] - fu[3]) / dy[j];
    gpx[c] = 0.5 * (fp[0] + fp[1]);
    gpy[c] = 0.5 * (fp[2] + fp[3]);
  }
}

// Pass 2 (src/fpm_partition_2d.f90:85-117): densities and the workgroup's
// partial terms (slots: KINE, VIF, VIS, BOUND).
__global__ __launch_bounds__(256) void k_part_pass2(
    int nx, int ny, Kin kn, const double *__restrict__ phi, const double *__restrict__ u,
    const double *__restrict__ v, const int32_t *__restrict__ solid,
    const double *__restrict__ xc, const double *__restrict__ yc,
    const double *__restrict__ dx, const double *__restrict__ dy,
    const double *__restrict__ ring, const double *__restrict__ nxf,
    const double *__restrict__ nyf, const double *__restrict__ gpx,
    const double *__restrict__ gpy, double *__restrict__ vif, double *partials) {
  __shared__ double lds[16 * kReduceSlots];
  const size_t n = (size_t)nx * ny;
  const size_t L = 2 * (size_t)nx + 2 * (size_t)ny;
  double t[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx), j = (int)(c / nx);
    if (solid[c]) {
      vif[c] = 0;
      continue;
    }
    double div = 0, lapu = 0, lapv = 0;
    for (int f = 1; f <= 4; ++f) {
      const Face o = face_of(i, j, f, nx, ny, xc, yc, dx, dy);
      const size_t cn = (size_t)o.ii + (size_t)nx * o.jj;
      if (o.ii < 0 || o.ii >= nx || o.jj < 0 || o.jj >= ny) {
        const double nn = o.axis == 1 ? nxf[c] : nyf[c];
        div = div + o.sgn * nn * o.len;
        const size_t m = (size_t)ring_index(i, j, f, nx, ny);
        const double dudt = (ring[m + L * 2] - ring[m]) / (2 * kn.dt);
        const double dvdt = (ring[m + L * 3] - ring[m + L]) / (2 * kn.dt);
        const double sel = o.axis == 1 ? dudt + nxf[c] : dvdt + nyf[c];
        t[3] = t[3] - o.sgn * sel * phi[c] * o.len;
      } else if (solid[cn]) {
        const double ubx = kn.k[0] - kn.k[2] * o.fy, uby = kn.k[1] + kn.k[2] * o.fx;
        const double ax = kn.k[3] - kn.k[5] * o.fy - kn.k[2] * kn.k[2] * o.fx;
        const double ay = kn.k[4] + kn.k[5] * o.fx - kn.k[2] * kn.k[2] * o.fy;
        const double phif = phi[c] + o.sgn * o.hw * kn.g[o.axis - 1];
        t[0] = t[0] - o.sgn * (o.axis == 1 ? ax : ay) * phif * o.len;
        if (kn.slip) {
          div = div + o.sgn * (o.axis == 1 ? nxf[c] : nyf[c]) * o.len;
        } else {
          div = div + o.sgn * (o.axis == 1 ? -(kn.k[2] * uby) : kn.k[2] * ubx) * o.len;
          lapu = lapu + (ubx - u[c]) / o.hw * o.len;
          lapv = lapv + (uby - v[c]) / o.hw * o.len;
        }
      } else {
        const double sel = o.axis == 1 ? nxf[c] + (nxf[cn] - nxf[c]) * o.hw / o.d
                                       : nyf[c] + (nyf[cn] - nyf[c]) * o.hw / o.d;
        div = div + o.sgn * sel * o.len;
        lapu = lapu + (u[cn] - u[c]) / o.d * o.len;
        lapv = lapv + (v[cn] - v[c]) / o.d * o.len;
      }
    }
    vif[c] = phi[c] * div / (dx[i] * dy[j]);
    t[1] = t[1] + phi[c] * div;
    t[2] = t[2] + kn.nu * (lapu * gpx[c] + lapv * gpy[c]);
  }
  block_sum<4>(t, lds);
  store_partials(partials, 4, t, gridDim.x);
}

int partition_dev(const double *phi, const double *g, const double *u, const double *v,
                  const int32_t *solid, const double *xf, const double *yf, const double *kin,
                  const double *ring, double dt, double nu, int nx, int ny, int slip,
                  double *terms, double *vif, double *omega) {
  if (nx < 1 || ny < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  if (!phi || !g || !u || !v || !solid || !xf || !yf || !kin || !ring || !terms || !vif ||
      !omega) {
    set_error(ERR_ARGUMENT, "null array argument");
    return ERR_ARGUMENT;
  }
  int st = 0;
  Workspace &ws = workspace();
  const size_t n = (size_t)nx * ny;
  const int eb = elementwise_blocks(n);
  double *axes = (double *)ws.get(S_P2_AXES, sizeof(double) * 2 * (nx + ny), &st);
  double *work = axes ? (double *)ws.get(S_P2_WORK, sizeof(double) * 4 * n, &st) : nullptr;
  double *partials = work ? (double *)ws.get(S_P2_PART, sizeof(double) * 4 * eb, &st) : nullptr;
  Scalars *sc = partials ? (Scalars *)ws.get(S_P2_SC, sizeof(Scalars), &st) : nullptr;
  if (!sc) return st;
  double *xc = axes, *yc = axes + nx, *dx = axes + nx + ny, *dy = axes + 2 * nx + ny;
  double *nxf = work, *nyf = work + n, *gpx = work + 2 * n, *gpy = work + 3 * n;
  Kin kn{};
  kn.g[0] = g[0];
  kn.g[1] = g[1];
  for (int q = 0; q < 6; ++q) kn.k[q] = kin[q];
  kn.dt = dt;
  kn.nu = nu;
  kn.slip = slip ? 1 : 0;
  k_part_axes<<<(nx + ny + 255) / 256, 256, 0, stream()>>>(nx, ny, xf, yf, xc, yc, dx, dy);
  FFPM_LAUNCHED("k_part_axes");
  k_part_pass1<<<eb, kThreads, 0, stream()>>>(nx, ny, kn, phi, u, v, solid, xc, yc, dx, dy, nxf,
                                              nyf, gpx, gpy, omega);
  FFPM_LAUNCHED("k_part_pass1");
  k_part_pass2<<<eb, kThreads, 0, stream()>>>(nx, ny, kn, phi, u, v, solid, xc, yc, dx, dy, ring,
                                              nxf, nyf, gpx, gpy, vif, partials);
  FFPM_LAUNCHED("k_part_pass2");
  k_finish<<<1, 256, 0, stream()>>>(FIN_SUMS, 4, partials, eb, sc);
  FFPM_LAUNCHED("k_finish");
  Scalars h;
  FFPM_HIP(hipMemcpyAsync(&h, sc, sizeof(Scalars), hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  for (int q = 0; q < 4; ++q) terms[q] = h.sums[q];
  return OK;
}

}  // namespace

}  // namespace ffpm

using namespace ffpm;

extern "C" {

int ffpm_partition_terms_2d_dev(const double *phi, const double *g, const double *u,
                                const double *v, const int32_t *solid, const double *xf,
                                const double *yf, const double *kin, const double *ring,
                                double dt, double nu, int nx, int ny, int slip_body,
                                double *terms, double *vif_density, double *omega) {
  return partition_dev(phi, g, u, v, solid, xf, yf, kin, ring, dt, nu, nx, ny, slip_body, terms,
                       vif_density, omega);
}

int ffpm_partition_terms_2d(const double *phi, const double *g, const double *u,
                            const double *v, const int32_t *solid, const double *xf,
                            const double *yf, const double *kin, const double *ring, double dt,
                            double nu, int nx, int ny, int slip_body, double *terms,
                            double *vif_density, double *omega) {
  if (nx < 1 || ny < 1) {
    set_error(ERR_SHAPE, "empty grid");
    return ERR_SHAPE;
  }
  if (!phi || !g || !u || !v || !solid || !xf || !yf || !kin || !ring || !terms ||
      !vif_density || !omega) {
    set_error(ERR_ARGUMENT, "null array argument");
    return ERR_ARGUMENT;
  }
  const size_t n = (size_t)nx * ny;
  const size_t nr = 4 * (2 * (size_t)nx + 2 * (size_t)ny);
  int st = 0;
  Workspace &ws = workspace();
  double *d = (double *)ws.get(S_P2_HOST, sizeof(double) * (5 * n + nx + ny + 2 + nr), &st);
  int32_t *ds = d ? (int32_t *)ws.get(S_P2_SOLID, sizeof(int32_t) * n, &st) : nullptr;
  if (!ds) return st;
  double *dphi = d, *du = d + n, *dv = d + 2 * n, *dvif = d + 3 * n, *dom = d + 4 * n;
  double *dxf = d + 5 * n, *dyf = dxf + nx + 1, *dring = dyf + ny + 1;
  FFPM_HIP(hipMemcpyAsync(dphi, phi, sizeof(double) * n, hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(du, u, sizeof(double) * n, hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(dv, v, sizeof(double) * n, hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(ds, solid, sizeof(int32_t) * n, hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(dxf, xf, sizeof(double) * (nx + 1), hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(dyf, yf, sizeof(double) * (ny + 1), hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipMemcpyAsync(dring, ring, sizeof(double) * nr, hipMemcpyHostToDevice, stream()));
  if ((st = partition_dev(dphi, g, du, dv, ds, dxf, dyf, kin, dring, dt, nu, nx, ny, slip_body,
                          terms, dvif, dom)))
    return st;
  FFPM_HIP(hipMemcpyAsync(vif_density, dvif, sizeof(double) * n, hipMemcpyDeviceToHost,
                          stream()));
  FFPM_HIP(hipMemcpyAsync(omega, dom, sizeof(double) * n, hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  return FFPM_OK;
}

}  // extern "C"
