// Geometric multigrid V-cycle on the stair-step operator (preconditioner of
// the flexible CG in ffpm_stairstep_host.hip). Layout and cycle as described
// in ffpm_mg.hpp. Cell kernels run 64 x 4 workgroups over (x, y) tiles with
// one z plane per grid row, so no index divisions; all loads are
// unconditional with clamped indices and masked coefficients.
#include <cstdlib>

#include "ffpm_mg.hpp"
#include "ffpm_pcg_body.hpp"
#include "ffpm_runtime.hpp"

namespace ffpm {

namespace {

constexpr int kMx = 64, kMy = 4;

inline dim3 cell_grid(int nx, int ny, int nz) {
  return dim3((unsigned)((nx + kMx - 1) / kMx), (unsigned)((ny + kMy - 1) / kMy), (unsigned)nz);
}

struct Idx {
  int i, j, k;
  size_t c;
  bool in;
};

__device__ __forceinline__ Idx cell_idx(int nx, int ny) {
  Idx t;
  t.i = blockIdx.x * kMx + threadIdx.x;
  t.j = blockIdx.y * kMy + threadIdx.y;
  t.k = blockIdx.z;
  t.in = t.i < nx && t.j < ny;
  t.c = (size_t)t.i + (size_t)nx * ((size_t)t.j + (size_t)ny * t.k);
  return t;
}

// ---- level 0: operator from the neighbour code and FP32 1-D coefficients ----
struct Fine {
  float d, V, nb;  // diag, volume, sum of coefficient * neighbour
};

__device__ __forceinline__ Fine fine_row(const GeomT<float> &g, unsigned cd,
                                         const float *__restrict__ x, size_t c, int i, int j,
                                         int k) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t sx = (size_t)nx, sxy = (size_t)nx * ny;
  const float cw = g.cw[i], ce = g.ce[i], cs = g.cs[j], cn = g.cn[j], cb = g.cb[k], ct = g.ct[k];
  const float xw = x[i > 0 ? c - 1 : c], xe = x[i < nx - 1 ? c + 1 : c];
  const float xs = x[j > 0 ? c - sx : c], xn = x[j < ny - 1 ? c + sx : c];
  const float xb = x[k > 0 ? c - sxy : c], xt = x[k < nz - 1 ? c + sxy : c];
  Fine f;
  f.d = diag_code<float>(cd, cw, ce, cs, cn, cb, ct);
  f.V = g.dx[i] * g.dy[j] * g.dz[k];
  float s = (cd & kCodeW) ? cw * xw : 0.f;
  s += (cd & kCodeE) ? ce * xe : 0.f;
  s += (cd & kCodeS) ? cs * xs : 0.f;
  s += (cd & kCodeN) ? cn * xn : 0.f;
  s += (cd & kCodeB) ? cb * xb : 0.f;
  s += (cd & kCodeT) ? ct * xt : 0.f;
  f.nb = s;
  return f;
}

// x = omega D^-1 r (the first sweep from zero); D = V diag.
__global__ __launch_bounds__(256) void k_mgf_jac0(GeomT<float> g, const uint8_t *__restrict__ code,
                                                  const double *__restrict__ r,
                                                  float *__restrict__ x, float omega) {
  const Idx t = cell_idx(g.nx, g.ny);
  if (!t.in) return;
  const unsigned cd = code[t.c];
  const float d = diag_code<float>(cd, g.cw[t.i], g.ce[t.i], g.cs[t.j], g.cn[t.j], g.cb[t.k],
                                   g.ct[t.k]);
  const float V = g.dx[t.i] * g.dy[t.j] * g.dz[t.k];
  x[t.c] = d > 0.f ? omega * ((float)r[t.c] / V) / d : 0.f;
}

// One damped-Jacobi sweep on A x = r: out = in + omega D^-1 (r - A in).
__global__ __launch_bounds__(256) void k_mgf_sweep(GeomT<float> g,
                                                   const uint8_t *__restrict__ code,
                                                   const double *__restrict__ r,
                                                   const float *__restrict__ in,
                                                   float *__restrict__ out, float omega) {
  const Idx t = cell_idx(g.nx, g.ny);
  if (!t.in) return;
  const unsigned cd = code[t.c];
  const Fine f = fine_row(g, cd, in, t.c, t.i, t.j, t.k);
  const float xv = in[t.c];
  out[t.c] = f.d > 0.f ? xv + omega * ((float)r[t.c] / f.V + f.nb - f.d * xv) / f.d : 0.f;
}

// Coarse right-hand side: the level-0 residual r - A x summed over the
// children of each level-1 cell (one thread per level-1 cell).
__global__ __launch_bounds__(256) void k_mgf_restrict(GeomT<float> g,
                                                      const uint8_t *__restrict__ code,
                                                      const double *__restrict__ r,
                                                      const float *__restrict__ x, int cnx,
                                                      int cny, float *__restrict__ bc) {
  const Idx t = cell_idx(cnx, cny);
  if (!t.in) return;
  float s = 0.f;
  for (int dk = 0; dk < 2; ++dk) {
    const int k = 2 * t.k + dk;
    if (k >= g.nz) break;
    for (int dj = 0; dj < 2; ++dj) {
      const int j = 2 * t.j + dj;
      if (j >= g.ny) break;
      for (int di = 0; di < 2; ++di) {
        const int i = 2 * t.i + di;
        if (i >= g.nx) break;
        const size_t c = (size_t)i + (size_t)g.nx * ((size_t)j + (size_t)g.ny * k);
        const unsigned cd = code[c];
        if (!(cd & kCodeFluid)) continue;
        const Fine f = fine_row(g, cd, x, c, i, j, k);
        s += (float)r[c] - f.V * (f.d * x[c] - f.nb);
      }
    }
  }
  bc[t.c] = s;
}

// ---- coarse levels: per-cell conductances ------------------------------------
struct Coarse {
  float d, nb;
};

__device__ __forceinline__ Coarse coarse_row(int nx, int ny, int nz, const float *__restrict__ gE,
                                             const float *__restrict__ gN,
                                             const float *__restrict__ gT,
                                             const float *__restrict__ x, size_t c, int i,
                                             int j, int k) {
  const size_t sx = (size_t)nx, sxy = (size_t)nx * ny;
  const size_t cw = i > 0 ? c - 1 : c, cs = j > 0 ? c - sx : c, cb = k > 0 ? c - sxy : c;
  const size_t ce = i < nx - 1 ? c + 1 : c, cn = j < ny - 1 ? c + sx : c,
               ct = k < nz - 1 ? c + sxy : c;
  const float ge = gE[c], gn = gN[c], gt = gT[c];
  const float gw = i > 0 ? gE[cw] : 0.f, gs = j > 0 ? gN[cs] : 0.f, gb = k > 0 ? gT[cb] : 0.f;
  Coarse o;
  o.d = gw + ge + gs + gn + gb + gt;
  o.nb = gw * x[cw] + ge * x[ce] + gs * x[cs] + gn * x[cn] + gb * x[cb] + gt * x[ct];
  return o;
}

__device__ __forceinline__ float coarse_diag(int nx, int ny, const float *__restrict__ gE,
                                             const float *__restrict__ gN,
                                             const float *__restrict__ gT, size_t c, int i,
                                             int j, int k) {
  const size_t sx = (size_t)nx, sxy = (size_t)nx * ny;
  const float gw = i > 0 ? gE[c - 1] : 0.f, gs = j > 0 ? gN[c - sx] : 0.f,
              gb = k > 0 ? gT[c - sxy] : 0.f;
  return gw + gE[c] + gs + gN[c] + gb + gT[c];
}

__global__ __launch_bounds__(256) void k_mgc_jac0(int nx, int ny, const float *__restrict__ gE,
                                                  const float *__restrict__ gN,
                                                  const float *__restrict__ gT,
                                                  const float *__restrict__ b,
                                                  float *__restrict__ x, float omega) {
  const Idx t = cell_idx(nx, ny);
  if (!t.in) return;
  const float d = coarse_diag(nx, ny, gE, gN, gT, t.c, t.i, t.j, t.k);
  x[t.c] = d > 0.f ? omega * b[t.c] / d : 0.f;
}

__global__ __launch_bounds__(256) void k_mgc_sweep(int nx, int ny, int nz,
                                                   const float *__restrict__ gE,
                                                   const float *__restrict__ gN,
                                                   const float *__restrict__ gT,
                                                   const float *__restrict__ b,
                                                   const float *__restrict__ in,
                                                   float *__restrict__ out, float omega) {
  const Idx t = cell_idx(nx, ny);
  if (!t.in) return;
  const Coarse o = coarse_row(nx, ny, nz, gE, gN, gT, in, t.c, t.i, t.j, t.k);
  const float xv = in[t.c];
  out[t.c] = o.d > 0.f ? xv + omega * (b[t.c] + o.nb - o.d * xv) / o.d : 0.f;
}

__global__ __launch_bounds__(256) void k_mgc_restrict(int nx, int ny, int nz,
                                                      const float *__restrict__ gE,
                                                      const float *__restrict__ gN,
                                                      const float *__restrict__ gT,
                                                      const float *__restrict__ b,
                                                      const float *__restrict__ x, int cnx,
                                                      int cny, float *__restrict__ bc) {
  const Idx t = cell_idx(cnx, cny);
  if (!t.in) return;
  float s = 0.f;
  for (int dk = 0; dk < 2; ++dk) {
    const int k = 2 * t.k + dk;
    if (k >= nz) break;
    for (int dj = 0; dj < 2; ++dj) {
      const int j = 2 * t.j + dj;
      if (j >= ny) break;
      for (int di = 0; di < 2; ++di) {
        const int i = 2 * t.i + di;
        if (i >= nx) break;
        const size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
        const Coarse o = coarse_row(nx, ny, nz, gE, gN, gT, x, c, i, j, k);
        if (o.d > 0.f) s += b[c] - (o.d * x[c] - o.nb);
      }
    }
  }
  bc[t.c] = s;
}

// Trilinear cell-centred prolongation from the coarse level (cnx, cny, cnz;
// values xc, unknown mask mc): fine cell 2I+a takes 3/4 of coarse cell I and
// 1/4 of its neighbour on the side of a, per axis; weights are renormalised
// over coarse cells with mc = 1. Only fine cells that carry unknowns change.
__device__ __forceinline__ float prolong_value(int i, int j, int k, int cnx, int cny, int cnz,
                                               const float *__restrict__ xc,
                                               const float *__restrict__ mc) {
  const int I = i >> 1, J = j >> 1, K = k >> 1;
  const int I2 = (i & 1) ? I + 1 : I - 1, J2 = (j & 1) ? J + 1 : J - 1,
            K2 = (k & 1) ? K + 1 : K - 1;
  const bool vi = I2 >= 0 && I2 < cnx, vj = J2 >= 0 && J2 < cny, vk = K2 >= 0 && K2 < cnz;
  const int ia[2] = {I, vi ? I2 : I}, ja[2] = {J, vj ? J2 : J}, ka[2] = {K, vk ? K2 : K};
  const float wi[2] = {0.75f, vi ? 0.25f : 0.f}, wj[2] = {0.75f, vj ? 0.25f : 0.f},
              wk[2] = {0.75f, vk ? 0.25f : 0.f};
  float num = 0.f, den = 0.f;
#pragma unroll
  for (int c3 = 0; c3 < 2; ++c3)
#pragma unroll
    for (int b3 = 0; b3 < 2; ++b3)
#pragma unroll
      for (int a3 = 0; a3 < 2; ++a3) {
        const size_t cc = (size_t)ia[a3] + (size_t)cnx * ((size_t)ja[b3] + (size_t)cny * ka[c3]);
        const float w = wi[a3] * wj[b3] * wk[c3];
        num += w * xc[cc];
        den += w * mc[cc];
      }
  return den > 0.f ? num / den : 0.f;
}

__global__ __launch_bounds__(256) void k_mgf_prolong(int nx, int ny,
                                                     const uint8_t *__restrict__ code,
                                                     float *__restrict__ x, int cnx, int cny,
                                                     int cnz, const float *__restrict__ xc,
                                                     const float *__restrict__ mc) {
  const Idx t = cell_idx(nx, ny);
  if (!t.in) return;
  if (!(code[t.c] & kCodeFluid)) return;
  x[t.c] += prolong_value(t.i, t.j, t.k, cnx, cny, cnz, xc, mc);
}

__global__ __launch_bounds__(256) void k_mgc_prolong(int nx, int ny, const float *__restrict__ m,
                                                     float *__restrict__ x, int cnx, int cny,
                                                     int cnz, const float *__restrict__ xc,
                                                     const float *__restrict__ mc) {
  const Idx t = cell_idx(nx, ny);
  if (!t.in) return;
  if (!(m[t.c] > 0.f)) return;
  x[t.c] += prolong_value(t.i, t.j, t.k, cnx, cny, cnz, xc, mc);
}

// Coarsest level (n <= 1024 cells, one workgroup): `sweeps` damped-Jacobi
// sweeps from zero with the iterate in LDS.
__global__ __launch_bounds__(1024) void k_mgc_coarsest(int nx, int ny, int nz,
                                                       const float *__restrict__ gE,
                                                       const float *__restrict__ gN,
                                                       const float *__restrict__ gT,
                                                       const float *__restrict__ b,
                                                       float *__restrict__ x, float omega,
                                                       int sweeps) {
  __shared__ float xs[kMgCoarsestCells];
  const int t = threadIdx.x;
  const int n = nx * ny * nz;
  const bool on = t < n;
  int i = 0, j = 0, k = 0;
  float gw = 0.f, ge = 0.f, gs = 0.f, gn = 0.f, gb = 0.f, gt = 0.f, bb = 0.f;
  if (on) {
    i = t % nx;
    j = (t / nx) % ny;
    k = t / (nx * ny);
    ge = gE[t];
    gn = gN[t];
    gt = gT[t];
    gw = i > 0 ? gE[t - 1] : 0.f;
    gs = j > 0 ? gN[t - nx] : 0.f;
    gb = k > 0 ? gT[t - nx * ny] : 0.f;
    bb = b[t];
  }
  const float d = gw + ge + gs + gn + gb + gt;
  float xv = on && d > 0.f ? omega * bb / d : 0.f;
  const int tw = i > 0 ? t - 1 : t, te = i < nx - 1 ? t + 1 : t;
  const int ts = j > 0 ? t - nx : t, tn = j < ny - 1 ? t + nx : t;
  const int tb = k > 0 ? t - nx * ny : t, tt = k < nz - 1 ? t + nx * ny : t;
  for (int s = 1; s < sweeps; ++s) {
    if (on) xs[t] = xv;
    __syncthreads();
    if (on && d > 0.f) {
      const float nb = gw * xs[tw] + ge * xs[te] + gs * xs[ts] + gn * xs[tn] + gb * xs[tb] +
                       gt * xs[tt];
      xv = xv + omega * (bb + nb - d * xv) / d;
    }
    __syncthreads();
  }
  if (on) x[t] = xv;
}

// Last post-smoothing sweep of level 0, written as z in FP64, with the partial
// sums z.r and z.(A p_old) (A in FP64 from the code and the FP64
// coefficients, the operator of the outer iteration).
__global__ __launch_bounds__(256) void k_mgf_last(GeomT<float> g, GeomT<double> gd,
                                                  const uint8_t *__restrict__ code,
                                                  const double *__restrict__ r,
                                                  const float *__restrict__ in,
                                                  const double *__restrict__ pold,
                                                  double *__restrict__ z, double *partials,
                                                  float omega, int nblocks) {
  __shared__ double lds[16 * kReduceSlots];
  const int i = blockIdx.x * kMx + threadIdx.x, j = blockIdx.y * kMy + threadIdx.y;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  double s[2] = {0.0, 0.0};
  if (i < nx && j < ny) {
    const size_t sx = (size_t)nx, sxy = (size_t)nx * ny;
    for (int k = blockIdx.z; k < nz; k += gridDim.z) {
      const size_t c = (size_t)i + sx * ((size_t)j + (size_t)ny * k);
      const unsigned cd = code[c];
      const Fine f = fine_row(g, cd, in, c, i, j, k);
      const float xv = in[c];
      const double rv = r[c];
      const float zf = f.d > 0.f ? xv + omega * ((float)rv / f.V + f.nb - f.d * xv) / f.d : 0.f;
      const double zv = (double)zf;
      z[c] = zv;
      s[0] += zv * rv;
      if (pold && (cd & kCodeFluid)) {
        const double cw = gd.cw[i], ce = gd.ce[i], cs = gd.cs[j], cn = gd.cn[j], cb = gd.cb[k],
                     ct = gd.ct[k];
        const double pw = pold[i > 0 ? c - 1 : c], pe = pold[i < nx - 1 ? c + 1 : c];
        const double ps = pold[j > 0 ? c - sx : c], pn = pold[j < ny - 1 ? c + sx : c];
        const double pb = pold[k > 0 ? c - sxy : c], pt = pold[k < nz - 1 ? c + sxy : c];
        double nb = (cd & kCodeW) ? cw * pw : 0.0;
        nb += (cd & kCodeE) ? ce * pe : 0.0;
        nb += (cd & kCodeS) ? cs * ps : 0.0;
        nb += (cd & kCodeN) ? cn * pn : 0.0;
        nb += (cd & kCodeB) ? cb * pb : 0.0;
        nb += (cd & kCodeT) ? ct * pt : 0.0;
        const double d = diag_code<double>(cd, cw, ce, cs, cn, cb, ct);
        const double V = gd.dx[i] * gd.dy[j] * gd.dz[k];
        s[1] += zv * (V * (d * pold[c] - nb));
      }
    }
  }
  block_sum<2>(s, lds);
  const int block = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  store_partials(partials, 2, s, nblocks, block);
}

// ---- setup --------------------------------------------------------------------
// Level-1 conductances from level 0 (FP64 sums, stored FP32): half the sum of
// V c over the fluid-fluid level-0 faces on each aggregate face.
__global__ __launch_bounds__(256) void k_mg_coarsen0(GeomT<double> g,
                                                     const uint8_t *__restrict__ code, int cnx,
                                                     int cny, float scale,
                                                     float *__restrict__ gE,
                                                     float *__restrict__ gN,
                                                     float *__restrict__ gT) {
  const Idx t = cell_idx(cnx, cny);
  if (!t.in) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  auto at = [&](int i, int j, int k) {
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
  };
  double se = 0.0, sn = 0.0, st = 0.0;
  const int i1 = 2 * t.i + 1, j1 = 2 * t.j + 1, k1 = 2 * t.k + 1;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      {  // +x face: level-0 cells (i1, j, k), j = 2J+a, k = 2K+b
        const int j = 2 * t.j + a, k = 2 * t.k + b;
        if (i1 < nx && j < ny && k < nz && (code[at(i1, j, k)] & kCodeE))
          se += g.dx[i1] * g.dy[j] * g.dz[k] * g.ce[i1];
      }
      {  // +y face
        const int i = 2 * t.i + a, k = 2 * t.k + b;
        if (j1 < ny && i < nx && k < nz && (code[at(i, j1, k)] & kCodeN))
          sn += g.dx[i] * g.dy[j1] * g.dz[k] * g.cn[j1];
      }
      {  // +z face
        const int i = 2 * t.i + a, j = 2 * t.j + b;
        if (k1 < nz && i < nx && j < ny && (code[at(i, j, k1)] & kCodeT))
          st += g.dx[i] * g.dy[j] * g.dz[k1] * g.ct[k1];
      }
    }
  gE[t.c] = (float)(scale * se);
  gN[t.c] = (float)(scale * sn);
  gT[t.c] = (float)(scale * st);
}

__global__ __launch_bounds__(256) void k_mg_coarsen(int nx, int ny, int nz,
                                                    const float *__restrict__ fE,
                                                    const float *__restrict__ fN,
                                                    const float *__restrict__ fT, int cnx,
                                                    int cny, float scale,
                                                    float *__restrict__ gE,
                                                    float *__restrict__ gN,
                                                    float *__restrict__ gT) {
  const Idx t = cell_idx(cnx, cny);
  if (!t.in) return;
  auto at = [&](int i, int j, int k) {
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
  };
  float se = 0.f, sn = 0.f, st = 0.f;
  const int i1 = 2 * t.i + 1, j1 = 2 * t.j + 1, k1 = 2 * t.k + 1;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      {
        const int j = 2 * t.j + a, k = 2 * t.k + b;
        if (i1 < nx && j < ny && k < nz) se += fE[at(i1, j, k)];
      }
      {
        const int i = 2 * t.i + a, k = 2 * t.k + b;
        if (j1 < ny && i < nx && k < nz) sn += fN[at(i, j1, k)];
      }
      {
        const int i = 2 * t.i + a, j = 2 * t.j + b;
        if (k1 < nz && i < nx && j < ny) st += fT[at(i, j, k1)];
      }
    }
  gE[t.c] = scale * se;
  gN[t.c] = scale * sn;
  gT[t.c] = scale * st;
}

__global__ __launch_bounds__(256) void k_mg_mask(int nx, int ny, const float *__restrict__ gE,
                                                 const float *__restrict__ gN,
                                                 const float *__restrict__ gT,
                                                 float *__restrict__ m) {
  const Idx t = cell_idx(nx, ny);
  if (!t.in) return;
  m[t.c] = coarse_diag(nx, ny, gE, gN, gT, t.c, t.i, t.j, t.k) > 0.f ? 1.f : 0.f;
}

float env_float(const char *name, float fallback) {
  const char *e = std::getenv(name);
  return e ? (float)std::atof(e) : fallback;
}

}  // namespace

int mg_setup(Multigrid &mg, const GeomT<double> &g64, const GeomT<float> &g32,
             const uint8_t *code) {
  mg.g64 = g64;
  mg.g32 = g32;
  mg.code = code;
  mg.omega = env_float("FFPM_MG_OMEGA", 0.8f);
  mg.scale = env_float("FFPM_MG_SCALE", 0.5f);
  mg.coarse_sweeps = (int)env_float("FFPM_MG_COARSE_SWEEPS", 30.f);
  if (mg.coarse_sweeps < 1) mg.coarse_sweeps = 1;
  // level extents: halve (ceil) until the level fits one workgroup
  int L = 0;
  int nx = g64.nx, ny = g64.ny, nz = g64.nz;
  mg.lv[0].nx = nx; mg.lv[0].ny = ny; mg.lv[0].nz = nz;
  mg.lv[0].n = (size_t)nx * ny * nz;
  while ((size_t)nx * ny * nz > (size_t)kMgCoarsestCells && L + 1 < kMgMaxLevels) {
    nx = (nx + 1) / 2;
    ny = (ny + 1) / 2;
    nz = (nz + 1) / 2;
    ++L;
    mg.lv[L].nx = nx; mg.lv[L].ny = ny; mg.lv[L].nz = nz;
    mg.lv[L].n = (size_t)nx * ny * nz;
  }
  if (L == 0) {  // a grid of at most 1024 cells still gets one coarse level
    ++L;
    mg.lv[1].nx = (nx + 1) / 2; mg.lv[1].ny = (ny + 1) / 2; mg.lv[1].nz = (nz + 1) / 2;
    mg.lv[1].n = (size_t)mg.lv[1].nx * mg.lv[1].ny * mg.lv[1].nz;
  }
  if ((size_t)mg.lv[L].n > (size_t)kMgCoarsestCells) {
    set_error(ERR_SHAPE, "multigrid hierarchy does not reach a one-workgroup level");
    return ERR_SHAPE;
  }
  mg.levels = L + 1;
  // last kernel's launch: z planes strided over at most 16 grid rows
  mg.kz = g64.nz < 16 ? g64.nz : 16;
  const dim3 lg = cell_grid(g64.nx, g64.ny, mg.kz);
  mg.last_blocks = (int)(lg.x * lg.y * lg.z);
  // one arena: level-0 x, y; per coarse level gE gN gT m x y b; partials
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t bytes = 2 * up(sizeof(float) * mg.lv[0].n);
  for (int l = 1; l < mg.levels; ++l) bytes += 7 * up(sizeof(float) * mg.lv[l].n);
  bytes += up(sizeof(double) * 2 * (size_t)mg.last_blocks);
  int st = 0;
  char *base = (char *)workspace().get(S_MG, bytes, &st);
  if (!base) return st;
  auto take = [&](size_t b) {
    char *p = base;
    base += up(b);
    return p;
  };
  mg.lv[0].x = (float *)take(sizeof(float) * mg.lv[0].n);
  mg.lv[0].y = (float *)take(sizeof(float) * mg.lv[0].n);
  for (int l = 1; l < mg.levels; ++l) {
    MgLevel &v = mg.lv[l];
    const size_t b = sizeof(float) * v.n;
    v.gE = (float *)take(b); v.gN = (float *)take(b); v.gT = (float *)take(b);
    v.m = (float *)take(b); v.x = (float *)take(b); v.y = (float *)take(b);
    v.b = (float *)take(b);
  }
  mg.partials = (double *)take(sizeof(double) * 2 * (size_t)mg.last_blocks);
  // conductances and masks of the coarse levels
  {
    const MgLevel &c = mg.lv[1];
    k_mg_coarsen0<<<cell_grid(c.nx, c.ny, c.nz), dim3(kMx, kMy), 0, stream()>>>(
        g64, code, c.nx, c.ny, mg.scale, c.gE, c.gN, c.gT);
    FFPM_LAUNCHED("k_mg_coarsen0");
  }
  for (int l = 2; l < mg.levels; ++l) {
    const MgLevel &f = mg.lv[l - 1], &c = mg.lv[l];
    k_mg_coarsen<<<cell_grid(c.nx, c.ny, c.nz), dim3(kMx, kMy), 0, stream()>>>(
        f.nx, f.ny, f.nz, f.gE, f.gN, f.gT, c.nx, c.ny, mg.scale, c.gE, c.gN, c.gT);
    FFPM_LAUNCHED("k_mg_coarsen");
  }
  for (int l = 1; l < mg.levels; ++l) {
    const MgLevel &c = mg.lv[l];
    k_mg_mask<<<cell_grid(c.nx, c.ny, c.nz), dim3(kMx, kMy), 0, stream()>>>(c.nx, c.ny, c.gE,
                                                                          c.gN, c.gT, c.m);
    FFPM_LAUNCHED("k_mg_mask");
  }
  return OK;
}

int mg_vcycle(const Multigrid &mg, const double *r, const double *p_old, double *z) {
  const dim3 blk(kMx, kMy);
  const MgLevel &f0 = mg.lv[0];
  const GeomT<float> &g = mg.g32;
  const float w = mg.omega;
  const int L = mg.levels - 1;  // coarsest index
  {
    TimedLaunch tl("mg_fine_pre");
    k_mgf_jac0<<<cell_grid(f0.nx, f0.ny, f0.nz), blk, 0, stream()>>>(g, mg.code, r, f0.x, w);
    k_mgf_sweep<<<cell_grid(f0.nx, f0.ny, f0.nz), blk, 0, stream()>>>(g, mg.code, r, f0.x, f0.y,
                                                                      w);
  }
  FFPM_LAUNCHED("k_mgf_sweep");
  {
    const MgLevel &c = mg.lv[1];
    TimedLaunch tl("mg_fine_restrict");
    k_mgf_restrict<<<cell_grid(c.nx, c.ny, c.nz), blk, 0, stream()>>>(g, mg.code, r, f0.y, c.nx,
                                                                       c.ny, c.b);
  }
  FFPM_LAUNCHED("k_mgf_restrict");
  // down: pre-smoothing (result in y) and restriction
  for (int l = 1; l < L; ++l) {
    const MgLevel &v = mg.lv[l], &c = mg.lv[l + 1];
    const dim3 gr = cell_grid(v.nx, v.ny, v.nz);
    k_mgc_jac0<<<gr, blk, 0, stream()>>>(v.nx, v.ny, v.gE, v.gN, v.gT, v.b, v.x, w);
    k_mgc_sweep<<<gr, blk, 0, stream()>>>(v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, v.x, v.y, w);
    k_mgc_restrict<<<cell_grid(c.nx, c.ny, c.nz), blk, 0, stream()>>>(
        v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, v.y, c.nx, c.ny, c.b);
    FFPM_LAUNCHED("k_mgc_restrict");
  }
  {
    const MgLevel &c = mg.lv[L];
    k_mgc_coarsest<<<1, kMgCoarsestCells, 0, stream()>>>(c.nx, c.ny, c.nz, c.gE, c.gN, c.gT, c.b,
                                                         c.y, w, mg.coarse_sweeps);
    FFPM_LAUNCHED("k_mgc_coarsest");
  }
  // up: prolongation into y, two sweeps y -> x -> y
  for (int l = L - 1; l >= 1; --l) {
    const MgLevel &v = mg.lv[l], &c = mg.lv[l + 1];
    const dim3 gr = cell_grid(v.nx, v.ny, v.nz);
    k_mgc_prolong<<<gr, blk, 0, stream()>>>(v.nx, v.ny, v.m, v.y, c.nx, c.ny, c.nz, c.y, c.m);
    k_mgc_sweep<<<gr, blk, 0, stream()>>>(v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, v.y, v.x, w);
    k_mgc_sweep<<<gr, blk, 0, stream()>>>(v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, v.x, v.y, w);
    FFPM_LAUNCHED("k_mgc_sweep");
  }
  {
    const MgLevel &c = mg.lv[1];
    TimedLaunch tl("mg_fine_post");
    k_mgf_prolong<<<cell_grid(f0.nx, f0.ny, f0.nz), blk, 0, stream()>>>(
        f0.nx, f0.ny, mg.code, f0.y, c.nx, c.ny, c.nz, c.y, c.m);
    k_mgf_sweep<<<cell_grid(f0.nx, f0.ny, f0.nz), blk, 0, stream()>>>(g, mg.code, r, f0.y, f0.x,
                                                                      w);
  }
  FFPM_LAUNCHED("k_mgf_prolong");
  {
    TimedLaunch tl("mg_fine_last");
    k_mgf_last<<<cell_grid(f0.nx, f0.ny, mg.kz), blk, 0, stream()>>>(
        g, mg.g64, mg.code, r, f0.x, p_old, z, mg.partials, w, mg.last_blocks);
  }
  FFPM_LAUNCHED("k_mgf_last");
  return OK;
}

}  // namespace ffpm
