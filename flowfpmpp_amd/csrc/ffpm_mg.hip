// Geometric multigrid V-cycle on the stair-step operator (preconditioner of
// the flexible CG in ffpm_stairstep_host.hip). Layout and cycle as described
// in ffpm_mg.hpp; the kernel bodies are in ffpm_mg_body.hpp.
#include <cstdlib>
#include <utility>

#include "ffpm_mg.hpp"
#include "ffpm_mg_body.hpp"
#include "ffpm_runtime.hpp"
#include "ffpm_unet.hpp"

namespace ffpm {

namespace {

using namespace mgk;

// ---- standalone kernels: one solve per launch --------------------------------
__global__ __launch_bounds__(256) void k_mgf_jac0(GeomT<float> g, const uint8_t *__restrict__ code,
                                                  double *__restrict__ r,
                                                  MgStore *__restrict__ x, float omega,
                                                  const Scalars *sc) {
  mgf_jac0_body(blk3(), g, code, r, x, omega, sc);
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_mgf_restrict_z(GeomT<float> g,
                                                        const uint8_t *__restrict__ code,
                                                        const MgStore *__restrict__ src,
                                                        const MgStore *__restrict__ x, int cnx,
                                                        int cny, int cnz, int kchunk,
                                                        CoarseOut out, float sscale) {
  mgf_restrict_z_body<VEC>(blk3(), g, code, src, x, cnx, cny, cnz, kchunk, out, sscale);
}

template <int CL, bool NT>
__global__ __launch_bounds__(512) void k_mg_restrict_t(Tiling t, GeomT<float> g,
                                                       const MgStore *__restrict__ x,
                                                       const MgStore *__restrict__ src,
                                                       const uint8_t *__restrict__ code, int cnx,
                                                       int cny, CoarseOut out, float sscale) {
  extern __shared__ float mg_rs_lds[];  // mg_restrict_lds(t.ty, CL) bytes
  mg_restrict_body<CL, NT>(blockIdx.x, mg_rs_lds, t, g, x, src, code, cnx, cny, out, sscale);
}

template <bool VEC>
__global__ __launch_bounds__(256) void k_mgf_prolong_z(int nx, int ny, int nz,
                                                       const uint8_t *__restrict__ code,
                                                       MgStore *__restrict__ x, int cnx, int cny,
                                                       int cnz, const MgStore *__restrict__ xc,
                                                       const MgStore *__restrict__ mc,
                                                       int kchunk) {
  mgf_prolong_z_body<VEC>(blk3(), nx, ny, nz, code, x, cnx, cny, cnz, xc, mc, kchunk);
}

__global__ __launch_bounds__(256) void k_mgc_sweep(int nx, int ny, int nz,
                                                   const MgStore *__restrict__ gE,
                                                   const MgStore *__restrict__ gN,
                                                   const MgStore *__restrict__ gT,
                                                   const MgStore *__restrict__ b,
                                                   const MgStore *__restrict__ in,
                                                   const MgStore *__restrict__ prev,
                                                   MgStore *__restrict__ out, float c1, float c2) {
  mgc_sweep_body(blk3(), nx, ny, nz, gE, gN, gT, b, in, prev, out, c1, c2);
}

template <int CL>
__global__ __launch_bounds__(256) void k_mgc_sweep_t(Tiling t, int nx, int ny, int nz,
                                                     const MgStore *__restrict__ gE,
                                                     const MgStore *__restrict__ gN,
                                                     const MgStore *__restrict__ gT,
                                                     const MgStore *__restrict__ b,
                                                     const MgStore *__restrict__ in,
                                                     const MgStore *__restrict__ prev,
                                                     MgStore *__restrict__ out, float c1,
                                                     float c2) {
  mgc_sweep_t_body<CL>(blockIdx.x, t, nx, ny, nz, gE, gN, gT, b, in, prev, out, c1, c2);
}

__global__ __launch_bounds__(256) void k_mgc_restrict(int nx, int ny, int nz,
                                                      const MgStore *__restrict__ gE,
                                                      const MgStore *__restrict__ gN,
                                                      const MgStore *__restrict__ gT,
                                                      const MgStore *__restrict__ b,
                                                      const MgStore *__restrict__ x, int cnx,
                                                      int cny, CoarseOut out) {
  mgc_restrict_body(blk3(), nx, ny, nz, gE, gN, gT, b, x, cnx, cny, out);
}

__global__ __launch_bounds__(256) void k_mgc_prolong(int nx, int ny, const MgStore *__restrict__ m,
                                                     MgStore *__restrict__ x, int cnx, int cny,
                                                     int cnz, const MgStore *__restrict__ xc,
                                                     const MgStore *__restrict__ mc) {
  mgc_prolong_body(blk3(), nx, ny, m, x, cnx, cny, cnz, xc, mc);
}

__global__ __launch_bounds__(1024) void k_mgc_coarsest(int nx, int ny, int nz,
                                                       const MgStore *__restrict__ gE,
                                                       const MgStore *__restrict__ gN,
                                                       const MgStore *__restrict__ gT,
                                                       const MgStore *__restrict__ b,
                                                       const MgStore *__restrict__ rd,
                                                       MgStore *__restrict__ x, float omega,
                                                       int sweeps) {
  __shared__ float xs[kMgCoarsestCells];
  mgc_coarsest_body(xs, nx, ny, nz, gE, gN, gT, b, rd, x, omega, sweeps);
}

template <int CL, bool LAST, bool NT, bool CHEB>
__global__ __launch_bounds__(1024) void k_mg_sweep_t(Tiling t, GeomT<float> g,
                                                     const MgStore *__restrict__ a,
                                                     const MgStore *__restrict__ r,
                                                     const uint8_t *__restrict__ code,
                                                     MgStore *__restrict__ b, double *partials,
                                                     float c2, const MgStore *__restrict__ q,
                                                     const MgStore *__restrict__ prev, float c1,
                                                     float sscale) {
  __shared__ double lds[16 * kReduceSlots];
  mg_sweep_body<CL, LAST, NT, CHEB>(blockIdx.x, gridDim.x, lds, t, g, a, r, code, b, partials,
                                    c2, q, prev, c1, sscale);
}

template <int CL, int XM>
__global__ __launch_bounds__(1024) void k_mg_upd_t(Tiling t, GeomT<double> g,
                                                   const float *__restrict__ p,
                                                   const uint8_t *__restrict__ code,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   MgStore *__restrict__ x1, Scalars *sc,
                                                   double omega, double *rmax,
                                                   MgStore *__restrict__ q16,
                                                   const float *__restrict__ ppend) {
  __shared__ double lds[16 * kReduceSlots];
  mg_upd_body<CL, XM>(blockIdx.x, gridDim.x, lds, t, g, p, code, x, r, x1, sc, omega, rmax, q16,
                      ppend);
}

// The flexible CG's direction step on the cycle's output (the PCG direction
// body with the FP16 z widened and unscaled on load, p formed and stored in
// FP32, q = A p in FP32 and p.q summed in FP64: the step length alpha only
// needs p.q to FP32 accuracy, while the update recomputes A p in FP64 for r).
template <int CL, int MODE>
__global__ __launch_bounds__(1024) void k_mg_dir_t(Tiling t, GeomT<float> g,
                                                   const MgStore *__restrict__ z,
                                                   const uint8_t *__restrict__ code,
                                                   const float *__restrict__ pin,
                                                   float *__restrict__ pout, double *partials,
                                                   Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  pcg_dir_body<float, CL, MODE, MgStore, float>(t, blockIdx.x, gridDim.x, g, z, code, pin,
                                                pout, partials, sc, lds);
}

// ---- the learned coarse correction -----------------------------------------------
// Cell volumes of level L: a level-L cell aggregates level-0 cells
// [I 2^L, (I + 1) 2^L) per axis (clipped to the grid: ceil halving).
__global__ __launch_bounds__(256) void k_lrn_volume(int L, int nx, int ny, int nz,
                                                    const double *__restrict__ dx,
                                                    const double *__restrict__ dy,
                                                    const double *__restrict__ dz, int cnx,
                                                    int cny, float *__restrict__ vol) {
  const Idx t = cell_idx(cnx, cny, blk3());
  if (!t.in) return;
  const int s = 1 << L;
  auto width = [&](const double *d, int n, int I) {
    double w = 0.0;
    for (int i = I * s; i < (I + 1) * s && i < n; ++i) w += d[i];
    return w;
  };
  vol[t.c] = (float)(width(dx, nx, t.i) * width(dy, ny, t.j) * width(dz, nz, t.k));
}

// res = b - A x where the level carries an unknown (else 0), and one partial
// max |res| / V per workgroup.
__global__ __launch_bounds__(256) void k_lrn_residual(int nx, int ny, int nz,
                                                      const MgStore *__restrict__ gE,
                                                      const MgStore *__restrict__ gN,
                                                      const MgStore *__restrict__ gT,
                                                      const MgStore *__restrict__ b,
                                                      const MgStore *__restrict__ x,
                                                      const float *__restrict__ vol,
                                                      float *__restrict__ res,
                                                      double *__restrict__ part) {
  __shared__ double lds[16];
  const Idx t = cell_idx(nx, ny, blk3());
  double mx = 0.0;
  if (t.in) {
    const Coarse o = coarse_row(nx, ny, nz, gE, gN, gT, x, t.c, t.i, t.j, t.k);
    const float r = o.d > 0.f ? b[t.c] - (o.d * x[t.c] - o.nb) : 0.f;
    res[t.c] = r;
    mx = fabs((double)r) / (double)vol[t.c];
  }
  block_max(mx, lds);
  if (threadIdx.x == 0) part[blockIdx.x + (size_t)gridDim.x * blockIdx.z] = mx;
}

// sc[0] = sigma = the largest partial; sc[1] = alpha from the c.res and c.A c
// partials (slots 1, 2): c.res / c.A c, 0 unless c.A c > 0.
__global__ __launch_bounds__(256) void k_lrn_finish(int what, const double *__restrict__ part,
                                                    int blocks, double *__restrict__ sc) {
  __shared__ double lds[32];
  if (what == 0) {
    const double m = grid_finish_max(part, 0, blocks, lds);
    if (threadIdx.x == 0) sc[0] = m;
  } else {
    double v[2];
    grid_finish_sum<2>(part + blocks, blocks, v, lds);
    if (threadIdx.x == 0) sc[1] = v[1] > 0.0 && v[0] == v[0] ? v[0] / v[1] : 0.0;
  }
}

// The network's source density s = res / (V sigma) (0 when sigma is 0).
__global__ __launch_bounds__(256) void k_lrn_source(size_t n, const float *__restrict__ res,
                                                    const float *__restrict__ vol,
                                                    const double *__restrict__ sc,
                                                    float *__restrict__ src) {
  const double sig = sc[0];
  const float inv = sig > 0.0 ? (float)(1.0 / sig) : 0.f;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    src[c] = res[c] / vol[c] * inv;
}

// q = A c where the level carries an unknown; partials of c.res and c.q.
__global__ __launch_bounds__(256) void k_lrn_dots(int nx, int ny, int nz,
                                                  const MgStore *__restrict__ gE,
                                                  const MgStore *__restrict__ gN,
                                                  const MgStore *__restrict__ gT,
                                                  const float *__restrict__ c,
                                                  const float *__restrict__ res,
                                                  double *__restrict__ part) {
  __shared__ double lds[32];
  const Idx t = cell_idx(nx, ny, blk3());
  double v[2] = {0.0, 0.0};
  if (t.in) {
    const Coarse o = coarse_row(nx, ny, nz, gE, gN, gT, c, t.c, t.i, t.j, t.k);
    if (o.d > 0.f) {
      const float q = o.d * c[t.c] - o.nb;
      v[0] = (double)c[t.c] * (double)res[t.c];
      v[1] = (double)c[t.c] * (double)q;
    }
  }
  block_sum<2>(v, lds);
  const int blocks = gridDim.x * gridDim.z, b = blockIdx.x + gridDim.x * blockIdx.z;
  if (threadIdx.x == 0) {
    part[(size_t)blocks + b] = v[0];
    part[(size_t)2 * blocks + b] = v[1];
  }
}

// x += alpha c where the level carries an unknown.
__global__ __launch_bounds__(256) void k_lrn_add(size_t n, const MgStore *__restrict__ m,
                                                 const float *__restrict__ c,
                                                 const double *__restrict__ sc,
                                                 MgStore *__restrict__ x) {
  const float a = (float)sc[1];
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    if (m[i] > 0.f) x[i] += a * c[i];
}

// ---- setup ----------------------------------------------------------------------
// sc->anorm (mg_anorm_body), one workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_mg_anorm(GeomT<double> g, Scalars *sc) {
  __shared__ double lds[16];
  mg_anorm_body(g, sc, lds);
}

__global__ __launch_bounds__(256) void k_mg_coarsen0(GeomT<double> g,
                                                     const uint8_t *__restrict__ code, int cnx,
                                                     int cny, float scale,
                                                     MgStore *__restrict__ gE,
                                                     MgStore *__restrict__ gN,
                                                     MgStore *__restrict__ gT) {
  mg_coarsen0_body(blk3(), g, code, cnx, cny, scale, gE, gN, gT);
}

__global__ __launch_bounds__(256) void k_mg_coarsen(int nx, int ny, int nz,
                                                    const MgStore *__restrict__ fE,
                                                    const MgStore *__restrict__ fN,
                                                    const MgStore *__restrict__ fT, int cnx,
                                                    int cny, float scale,
                                                    MgStore *__restrict__ gE,
                                                    MgStore *__restrict__ gN,
                                                    MgStore *__restrict__ gT) {
  mg_coarsen_body(blk3(), nx, ny, nz, fE, fN, fT, cnx, cny, scale, gE, gN, gT);
}

__global__ __launch_bounds__(256) void k_mg_mask(int nx, int ny, const MgStore *__restrict__ gE,
                                                 const MgStore *__restrict__ gN,
                                                 const MgStore *__restrict__ gT, float first,
                                                 MgStore *__restrict__ m, MgStore *__restrict__ rd) {
  mg_mask_body(blk3(), nx, ny, gE, gN, gT, first, m, rd);
}

float env_float(const char *name, float fallback) {
  const char *e = std::getenv(name);
  return e ? (float)std::atof(e) : fallback;
}

// The steps of one level (MgSchedule) for nu smoothing steps before and after
// the coarse correction; `first` receives the scale of the first step from
// zero. Chebyshev (Saad's three-term recurrence on D^-1 A over [lo, hi], hi
// = 2.1 above the bound 2 of the spectrum, lo = 2 / ratio): step 0 is
// x1 = x0 + (1 / theta) D^-1 (b - A x0); step k >= 1 adds
// rho_k rho_(k-1) (x_k - x_(k-1)) + (2 rho_k / delta) D^-1 (b - A x_k), with
// theta and delta the interval's centre and half-width, rho_0 = delta / theta
// and rho_k = 1 / (2 theta / delta - rho_(k-1)). Damped Jacobi: every step
// x + omega D^-1 (b - A x). Buffers rotate over 0..2 (the step writes the
// one holding neither its iterate nor its previous iterate); with to_z
// (level 0) over 1, 2 and 4, and the last step writes buffer 3.
MgSchedule make_schedule(int nu, bool cheb, float omega, double ratio, bool to_z,
                         float *first) {
  const double hi = 2.1, lo = 2.0 / (ratio > 1.0 ? ratio : 25.0);
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
  float c1[8] = {}, c2[8] = {};
  double rho = 1.0 / sigma;
  c2[0] = cheb ? (float)(1.0 / theta) : omega;
  for (int k = 1; k < nu; ++k) {
    if (cheb) {
      const double rn = 1.0 / (2.0 * sigma - rho);
      c1[k] = (float)(rn * rho);
      c2[k] = (float)(2.0 * rn / delta);
      rho = rn;
    } else {
      c2[k] = omega;
    }
  }
  *first = c2[0];
  // Level 0 (to_z) keeps the first step in buffer 0 for the whole cycle (it
  // also stands for the source) and rotates 1, 2 and 4; the coarse levels
  // rotate 0, 1 and 2.
  auto other = [to_z](int a, int p) {
    const int set[3] = {to_z ? 1 : 0, to_z ? 2 : 1, to_z ? 4 : 2};
    for (int i = 0; i < 3; ++i)
      if (set[i] != a && set[i] != p) return set[i];
    return set[2];
  };
  MgSchedule s;
  int cur = 0, prev = -1;  // the first step from zero leaves buffer 0
  for (int k = 1; k < nu; ++k) {
    const int o = other(cur, prev);
    s.pre[s.npre++] = MgStep{cur, prev, o, c1[k], c2[k]};
    prev = cur;
    cur = o;
  }
  s.mid = cur;
  prev = -1;
  for (int k = 0; k < nu; ++k) {
    const int o = (to_z && k == nu - 1) ? 3 : other(cur, prev);
    s.post[s.npost++] = MgStep{cur, k == 0 ? -1 : prev, o, c1[k], c2[k]};
    prev = cur;
    cur = o;
  }
  s.out = cur;
  return s;
}

}  // namespace

int mg_check_interval(int check_interval) {
  const char *e = std::getenv("FFPM_MG_CHECK");
  int c = e ? std::atoi(e) : 2;
  if (check_interval > 0 && check_interval < c) c = check_interval;
  return c < 1 ? 1 : c;
}

int mg_cells_per_lane(int nx) {
  const int want = (int)env_float("FFPM_MG_CL", 4.f);
  return cells_per_lane(nx, want == 1 || want == 2 || want == 4 || want == 8 ? want : 4);
}

Tiling mg_cycle_tiling(int nx, int ny, int nz, int rows, int cl) {
  return make_tiling(nx, ny, nz, (int)env_float("FFPM_MG_ZCHUNK", 0.f), rows, kTileX * cl);
}

int mg_rows() {
  const char *e = std::getenv("FFPM_MG_ROWS");
  const int r = e ? std::atoi(e) : 4;
  return r == 1 || r == 2 || r == 4 || r == 8 || r == 16 ? r : 4;
}

int mg_plan(Multigrid &mg, const GeomT<double> &g64, const GeomT<float> &g32,
            const uint8_t *code, const Tiling &tl, int cl, double *partials) {
  mg.g64 = g64;
  mg.g32 = g32;
  mg.code = code;
  mg.tld = tl;
  mg.cld = cl;
  mg.cl = mg_cells_per_lane(g64.nx);
  // the caller's tiling unless another shape is asked for
  const bool own = mg.cl != cl || env_float("FFPM_MG_ZCHUNK", 0.f) > 0.f;
  mg.tl = own ? mg_cycle_tiling(g64.nx, g64.ny, g64.nz, tl.ty, mg.cl) : tl;
  mg.partials = partials;
  {
    const Tiling &tc = mg.tl;
    const int rows = tc.ty < 2 ? 2 : tc.ty > 8 ? 8 : tc.ty;
    int zc = tc.zchunk + (tc.zchunk & 1);
    if (zc > kMgRestrictPlanes) zc = kMgRestrictPlanes;
    mg.tlr = make_tiling(g64.nx, g64.ny, g64.nz, zc, rows, kTileX * mg.cl);
    const char *e = std::getenv("FFPM_MG_RESTRICT");
    mg.stream_restrict = !(e && std::strcmp(e, "column") == 0);
  }
  mg.omega = env_float("FFPM_MG_OMEGA", 0.9f);
  mg.omega_coarse = env_float("FFPM_MG_OMEGA_COARSE", mg.omega);
  mg.scale = env_float("FFPM_MG_SCALE", 0.5f);
  mg.coarse_sweeps = (int)env_float("FFPM_MG_COARSE_SWEEPS", 30.f);
  if (mg.coarse_sweeps < 1) mg.coarse_sweeps = 1;
  mg.nontemporal = env_float("FFPM_MG_NT", 1.f) != 0.f;
  mg.cheb = env_float("FFPM_MG_CHEB", 1.f) != 0.f;
  mg.nu = (int)env_float("FFPM_MG_NU", 3.f);
  mg.nu_coarse = (int)env_float("FFPM_MG_NU_COARSE", mg.cheb ? 3.f : 4.f);
  if (mg.nu < 2) mg.nu = 2;  // the update forms the first sweep, the last one writes z
  if (mg.nu > 8) mg.nu = 8;
  if (mg.nu_coarse < 1) mg.nu_coarse = 1;
  if (mg.nu_coarse > 8) mg.nu_coarse = 8;
  // lo = 2/30: 20/20/20 iterations on the 512x384x256 frame against 22/20/20
  // with 2/25, the same or fewer on the other frames measured
  const double ratio = env_float("FFPM_MG_CHEB_RATIO", 30.f);
  mg.s0 = make_schedule(mg.nu, mg.cheb, mg.omega, ratio, true, &mg.first);
  mg.sscale = 1.f / mg.first;
  mg.sc = make_schedule(mg.nu_coarse, mg.cheb, mg.omega_coarse, ratio, false,
                        &mg.first_coarse);
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
  // the streaming sweep on the coarse levels that fill the chip (4-row tiles
  // of 64 lanes; FFPM_MGC_STREAM=0: the one-cell sweep everywhere)
  const bool cstream = env_float("FFPM_MGC_STREAM", 1.f) != 0.f;
  for (int l = 1; l <= L; ++l) {
    MgLevel &v = mg.lv[l];
    v.cl = 0;
    if (cstream && l < L && v.n >= kMgStreamCells) {
      v.cl = cells_per_lane(v.nx, 4);
      v.tl = make_tiling(v.nx, v.ny, v.nz, 0, 4, kTileX * v.cl);
    }
  }
  if ((size_t)mg.lv[L].n > (size_t)kMgCoarsestCells) {
    set_error(ERR_SHAPE, "multigrid hierarchy does not reach a one-workgroup level");
    return ERR_SHAPE;
  }
  mg.levels = L + 1;
  return OK;
}

namespace {

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

template <typename T>
T *carve(char *&base, size_t n) {
  T *p = reinterpret_cast<T *>(base);
  base += up256(sizeof(T) * n);
  return p;
}

}  // namespace

size_t mg_operator_bytes(const Multigrid &mg) {
  size_t bytes = 0;
  for (int l = 1; l < mg.levels; ++l) bytes += 5 * up256(sizeof(MgStore) * mg.lv[l].n);
  return bytes;
}

size_t mg_state_bytes(const Multigrid &mg) {
  size_t bytes = 6 * up256(sizeof(MgStore) * mg.lv[0].n);
  for (int l = 1; l < mg.levels; ++l) bytes += 4 * up256(sizeof(MgStore) * mg.lv[l].n);
  return bytes;
}

void mg_bind_operator(Multigrid &mg, char *&base) {
  for (int l = 1; l < mg.levels; ++l) {
    MgLevel &v = mg.lv[l];
    v.gE = carve<MgStore>(base, v.n);
    v.gN = carve<MgStore>(base, v.n);
    v.gT = carve<MgStore>(base, v.n);
    v.m = carve<MgStore>(base, v.n);
    v.rd = carve<MgStore>(base, v.n);
  }
}

void mg_bind_state(Multigrid &mg, char *&base) {
  for (int i = 0; i < 5; ++i) mg.f0[i] = carve<MgStore>(base, mg.lv[0].n);
  mg.q16 = carve<MgStore>(base, mg.lv[0].n);
  for (int l = 1; l < mg.levels; ++l) {
    MgLevel &v = mg.lv[l];
    v.x = carve<MgStore>(base, v.n);
    v.y = carve<MgStore>(base, v.n);
    v.w = carve<MgStore>(base, v.n);
    v.b = carve<MgStore>(base, v.n);
  }
}

int mg_build(const Multigrid &mg) {
  {
    const MgLevel &c = mg.lv[1];
    k_mg_coarsen0<<<cell_grid(c.nx, c.ny, c.nz), dim3(kMt), 0, stream()>>>(
        mg.g64, mg.code, c.nx, c.ny, mg.scale, c.gE, c.gN, c.gT);
    FFPM_LAUNCHED("k_mg_coarsen0");
  }
  for (int l = 2; l < mg.levels; ++l) {
    const MgLevel &f = mg.lv[l - 1], &c = mg.lv[l];
    k_mg_coarsen<<<cell_grid(c.nx, c.ny, c.nz), dim3(kMt), 0, stream()>>>(
        f.nx, f.ny, f.nz, f.gE, f.gN, f.gT, c.nx, c.ny, mg.scale, c.gE, c.gN, c.gT);
    FFPM_LAUNCHED("k_mg_coarsen");
  }
  for (int l = 1; l < mg.levels; ++l) {
    const MgLevel &c = mg.lv[l];
    k_mg_mask<<<cell_grid(c.nx, c.ny, c.nz), dim3(kMt), 0, stream()>>>(
        c.nx, c.ny, c.gE, c.gN, c.gT, mg.first_coarse, c.m, c.rd);
    FFPM_LAUNCHED("k_mg_mask");
  }
  return OK;
}

int mg_setup(Multigrid &mg, const GeomT<double> &g64, const GeomT<float> &g32,
             const uint8_t *code, const Tiling &tl, int cl, double *partials) {
  int st = mg_plan(mg, g64, g32, code, tl, cl, partials);
  if (st) return st;
  char *base =
      (char *)workspace().get(S_MG, mg_operator_bytes(mg) + mg_state_bytes(mg), &st);
  if (!base) return st;
  mg_bind_state(mg, base);
  mg_bind_operator(mg, base);
  return mg_build(mg);
}

int mg_learned_setup(Multigrid &mg, UNet *net, const float *features, int level, int mode) {
  if (mode != 0 && mode != 1) {
    set_error(ERR_ARGUMENT, "learned correction mode must be 0 (replace) or 1 (augment)");
    return ERR_ARGUMENT;
  }
  if (!net || !features) {
    set_error(ERR_ARGUMENT, "learned coarse correction needs a network and its features");
    return ERR_ARGUMENT;
  }
  if (level < 1 || level >= mg.levels - 1) {
    set_error(ERR_ARGUMENT, "learned level must lie between level 1 and the coarsest level");
    return ERR_ARGUMENT;
  }
  const MgLevel &v = mg.lv[level];
  if (v.nx < 16 || v.nx % 4 || v.ny % 4 || v.nz % 4) {
    set_error(ERR_SHAPE, "the learned level needs nx >= 16 and every extent a multiple of 4");
    return ERR_SHAPE;
  }
  MgLearned &L = mg.lrn;
  L.net = net;
  L.features = features;
  L.level = level;
  L.mode = mode;
  const dim3 grid = cell_grid(v.nx, v.ny, v.nz);
  L.blocks = (int)(grid.x * grid.z);
  const size_t fb = (v.n * sizeof(float) + 255) / 256 * 256;
  const size_t pb = ((size_t)3 * L.blocks * sizeof(double) + 255) / 256 * 256;
  int st = 0;
  char *base = (char *)workspace().get(S_MG_LRN, 4 * fb + pb + 256, &st);
  if (!base) return st;
  L.res = (float *)base;
  L.src = (float *)(base + fb);
  L.q = (float *)(base + 2 * fb);
  L.vol = (float *)(base + 3 * fb);
  L.part = (double *)(base + 4 * fb);
  L.sc = (double *)(base + 4 * fb + pb);
  k_lrn_volume<<<grid, dim3(kMt), 0, stream()>>>(level, mg.g64.nx, mg.g64.ny, mg.g64.nz,
                                                 mg.g64.dx, mg.g64.dy, mg.g64.dz, v.nx, v.ny,
                                                 L.vol);
  FFPM_LAUNCHED("k_lrn_volume");
  return OK;
}

namespace {

// Level-0 streaming sweep launch (one step of a schedule) for the
// multigrid's cells-per-lane choice.
template <bool LAST, bool NT, bool CHEB>
void launch_step0(const Multigrid &mg, const MgStep &st, const MgStore *q) {
  const dim3 blk(kTileX, mg.tl.ty);
  const MgStore *a = mg.buf0(st.a), *p = st.p >= 0 ? mg.buf0(st.p) : nullptr;
  MgStore *o = mg.buf0(st.o);
  auto go = [&](auto kern) {
    kern<<<mg.tl.blocks, blk, 0, stream()>>>(mg.tl, mg.g32, a, mg.buf0(0), mg.code, o,
                                             mg.partials, st.c2, q, p, st.c1, mg.sscale);
  };
  if (mg.cl == 8) go(k_mg_sweep_t<8, LAST, NT, CHEB>);
  else if (mg.cl == 4) go(k_mg_sweep_t<4, LAST, NT, CHEB>);
  else if (mg.cl == 2) go(k_mg_sweep_t<2, LAST, NT, CHEB>);
  else go(k_mg_sweep_t<1, LAST, NT, CHEB>);
}

template <bool LAST>
void launch_step(const Multigrid &mg, const MgStep &st, const MgStore *q = nullptr) {
  if (mg.cheb) {
    if (mg.nontemporal) launch_step0<LAST, true, true>(mg, st, q);
    else launch_step0<LAST, false, true>(mg, st, q);
  } else {
    if (mg.nontemporal) launch_step0<LAST, true, false>(mg, st, q);
    else launch_step0<LAST, false, false>(mg, st, q);
  }
}

}  // namespace

int mg_anorm(const Multigrid &mg, Scalars *sc) {
  k_mg_anorm<<<1, 256, 0, stream()>>>(mg.g64, sc);
  FFPM_LAUNCHED("k_mg_anorm");
  return OK;
}

int mg_first(const Multigrid &mg, double *r, const Scalars *sc) {
  const MgLevel &f0 = mg.lv[0];
  k_mgf_jac0<<<cell_grid(f0.nx, f0.ny, f0.nz), dim3(kMt), 0, stream()>>>(
      mg.g32, mg.code, r, mg.buf0(0), mg.first, sc);
  FFPM_LAUNCHED("k_mgf_jac0");
  return OK;
}

int mg_update(const Multigrid &mg, const Tiling &t, int cl, const float *p, double *x,
              double *r, Scalars *sc, double *rmax, int xmode, const float *ppend) {
  const dim3 blk(kTileX, t.ty);
  const double w = (double)mg.first;
  MgStore *x1 = mg.buf0(0);
  if (xmode == XM_TWO && !ppend) {
    set_error(ERR_ARGUMENT, "multigrid update: the deferred step needs its direction");
    return ERR_ARGUMENT;
  }
  TimedLaunch tl("mg_update");
  auto go = [&](auto kern) {
    kern<<<t.blocks, blk, 0, stream()>>>(t, mg.g64, p, mg.code, x, r, x1, sc, w, rmax, mg.q16,
                                         ppend);
  };
  if (cl == 2) {
    if (xmode == XM_DEFER) go(k_mg_upd_t<2, XM_DEFER>);
    else if (xmode == XM_TWO) go(k_mg_upd_t<2, XM_TWO>);
    else go(k_mg_upd_t<2, XM_NOW>);
  } else {
    if (xmode == XM_DEFER) go(k_mg_upd_t<1, XM_DEFER>);
    else if (xmode == XM_TWO) go(k_mg_upd_t<1, XM_TWO>);
    else go(k_mg_upd_t<1, XM_NOW>);
  }
  FFPM_LAUNCHED("k_mg_upd_t");
  return OK;
}

namespace {

CoarseOut coarse_out(const Multigrid &mg, int l) {
  const MgLevel &v = mg.lv[l];
  CoarseOut o;
  o.b = v.b;
  o.x = v.x;
  o.rd = v.rd;
  return o;
}

}  // namespace

int mg_vcycle(const Multigrid &mg, bool with_q) {
  const dim3 blk(kMt);
  const MgStore *r = mg.buf0(0);  // the source, read as x1 d sscale
  const MgLevel &f0 = mg.lv[0];
  const GeomT<float> &g = mg.g32;
  const MgSchedule &s0 = mg.s0, &sc = mg.sc;
  const int L = mg.levels - 1;  // coarsest index
  for (int k = 0; k < s0.npre; ++k) {
    TimedLaunch tl("mg_fine_sweep");
    launch_step<false>(mg, s0.pre[k]);
  }
  FFPM_LAUNCHED("k_mg_sweep_t");
  if (mg.stream_restrict) {
    const MgLevel &c = mg.lv[1];
    TimedLaunch tl("mg_fine_restrict");
    const dim3 rb(kTileX, mg.tlr.ty);
    const MgStore *xm = mg.buf0(s0.mid);
    const size_t lb = mg_restrict_lds(mg.tlr.ty, mg.cl);
    auto go = [&](auto kern) {
      kern<<<mg.tlr.blocks, rb, lb, stream()>>>(mg.tlr, g, xm, r, mg.code, c.nx, c.ny,
                                                coarse_out(mg, 1), mg.sscale);
    };
    if (mg.nontemporal) {
      if (mg.cl == 8) go(k_mg_restrict_t<8, true>);
      else if (mg.cl == 4) go(k_mg_restrict_t<4, true>);
      else if (mg.cl == 2) go(k_mg_restrict_t<2, true>);
      else go(k_mg_restrict_t<1, true>);
    } else {
      if (mg.cl == 8) go(k_mg_restrict_t<8, false>);
      else if (mg.cl == 4) go(k_mg_restrict_t<4, false>);
      else if (mg.cl == 2) go(k_mg_restrict_t<2, false>);
      else go(k_mg_restrict_t<1, false>);
    }
  } else {
    const MgLevel &c = mg.lv[1];
    TimedLaunch tl("mg_fine_restrict");
    const int kc = column_chunk(c.nx, c.ny, c.nz);
    if (f0.nx % 2 == 0)
      k_mgf_restrict_z<true><<<column_grid(c.nx, c.ny, c.nz, kc), blk, 0, stream()>>>(
          g, mg.code, r, mg.buf0(s0.mid), c.nx, c.ny, c.nz, kc, coarse_out(mg, 1), mg.sscale);
    else
      k_mgf_restrict_z<false><<<column_grid(c.nx, c.ny, c.nz, kc), blk, 0, stream()>>>(
          g, mg.code, r, mg.buf0(s0.mid), c.nx, c.ny, c.nz, kc, coarse_out(mg, 1), mg.sscale);
  }
  FFPM_LAUNCHED("k_mgf_restrict_z");
  auto csweep = [&](const MgLevel &v, const MgStep &st) {
    const MgStore *a = Multigrid::bufc(v, st.a),
                  *p = st.p >= 0 ? Multigrid::bufc(v, st.p) : nullptr;
    MgStore *o = Multigrid::bufc(v, st.o);
    if (v.cl) {
      auto go = [&](auto kern) {
        kern<<<v.tl.blocks, dim3(kTileX, v.tl.ty), 0, stream()>>>(
            v.tl, v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, a, p, o, st.c1, st.c2);
      };
      if (v.cl == 4) go(k_mgc_sweep_t<4>);
      else if (v.cl == 2) go(k_mgc_sweep_t<2>);
      else go(k_mgc_sweep_t<1>);
    } else {
      k_mgc_sweep<<<cell_grid(v.nx, v.ny, v.nz), blk, 0, stream()>>>(
          v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, a, p, o, st.c1, st.c2);
    }
  };
  // The learned correction either replaces everything below its level (the
  // cycle stops descending there) or, in augment mode, adds its A-optimal
  // correction of the level's remaining residual after the recursive coarse
  // correction, before the level's post-smoothing.
  const bool replace = mg.lrn.net && mg.lrn.mode == 0;
  const bool augment = mg.lrn.net && mg.lrn.mode == 1;
  const int Ln = replace ? mg.lrn.level : L;
  auto learned_correct = [&](const MgLevel &v, MgStore *xm) -> int {
    const MgLearned &N = mg.lrn;
    const dim3 grid = cell_grid(v.nx, v.ny, v.nz);
    k_lrn_residual<<<grid, blk, 0, stream()>>>(v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, xm,
                                               N.vol, N.res, N.part);
    k_lrn_finish<<<1, 256, 0, stream()>>>(0, N.part, N.blocks, N.sc);
    const int eb = elementwise_blocks(v.n);
    k_lrn_source<<<eb, 256, 0, stream()>>>(v.n, N.res, N.vol, N.sc, N.src);
    FFPM_LAUNCHED("k_lrn_source");
    const float *c = nullptr;
    int st = unet_apply(*N.net, v.nx, v.ny, v.nz, N.features, N.src, &c);
    if (st) return st;
    k_lrn_dots<<<grid, blk, 0, stream()>>>(v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, c, N.res, N.part);
    k_lrn_finish<<<1, 256, 0, stream()>>>(1, N.part, N.blocks, N.sc);
    k_lrn_add<<<eb, 256, 0, stream()>>>(v.n, v.m, c, N.sc, xm);
    FFPM_LAUNCHED("k_lrn_add");
    return OK;
  };
  // down: pre-smoothing (after the first step, formed by the restriction
  // into the level) and restriction
  for (int l = 1; l < Ln; ++l) {
    const MgLevel &v = mg.lv[l], &c = mg.lv[l + 1];
    for (int k = 0; k < sc.npre; ++k) csweep(v, sc.pre[k]);
    k_mgc_restrict<<<cell_grid(c.nx, c.ny, c.nz), blk, 0, stream()>>>(
        v.nx, v.ny, v.nz, v.gE, v.gN, v.gT, v.b, Multigrid::bufc(v, sc.mid), c.nx, c.ny,
        coarse_out(mg, l + 1));
    FFPM_LAUNCHED("k_mgc_restrict");
  }
  if (replace) {
    // the learned level: pre-smoothing, the network's correction of the
    // smoothed iterate, post-smoothing
    const MgLevel &v = mg.lv[Ln];
    TimedLaunch tl("mg_learned");
    for (int k = 0; k < sc.npre; ++k) csweep(v, sc.pre[k]);
    MgStore *xm = Multigrid::bufc(v, sc.mid);
    int st = learned_correct(v, xm);
    if (st) return st;
    for (int k = 0; k < sc.npost; ++k) csweep(v, sc.post[k]);
    FFPM_LAUNCHED("k_mgc_sweep");
  } else {
    // the coarsest level leaves its solution where the other coarse levels
    // leave theirs
    const MgLevel &c = mg.lv[L];
    k_mgc_coarsest<<<1, kMgCoarsestCells, 0, stream()>>>(c.nx, c.ny, c.nz, c.gE, c.gN, c.gT, c.b,
                                                         c.rd, Multigrid::bufc(c, sc.out),
                                                         mg.omega_coarse, mg.coarse_sweeps);
    FFPM_LAUNCHED("k_mgc_coarsest");
  }
  // up: prolongation into the level's iterate, then post-smoothing
  for (int l = Ln - 1; l >= 1; --l) {
    const MgLevel &v = mg.lv[l], &c = mg.lv[l + 1];
    k_mgc_prolong<<<cell_grid(v.nx, v.ny, v.nz), blk, 0, stream()>>>(
        v.nx, v.ny, v.m, Multigrid::bufc(v, sc.mid), c.nx, c.ny, c.nz,
        Multigrid::bufc(c, sc.out), c.m);
    if (augment && l == mg.lrn.level) {
      TimedLaunch tl("mg_learned");
      int st = learned_correct(v, Multigrid::bufc(v, sc.mid));
      if (st) return st;
    }
    for (int k = 0; k < sc.npost; ++k) csweep(v, sc.post[k]);
    FFPM_LAUNCHED("k_mgc_sweep");
  }
  {
    const MgLevel &c = mg.lv[1];
    TimedLaunch tl("mg_fine_prolong");
    const int kc = column_chunk(c.nx, c.ny, c.nz);
    MgStore *x0 = mg.buf0(s0.mid);
    const MgStore *xc = Multigrid::bufc(c, sc.out);
    if (f0.nx % 2 == 0)
      k_mgf_prolong_z<true><<<column_grid(c.nx, c.ny, c.nz, kc), blk, 0, stream()>>>(
          f0.nx, f0.ny, f0.nz, mg.code, x0, c.nx, c.ny, c.nz, xc, c.m, kc);
    else
      k_mgf_prolong_z<false><<<column_grid(c.nx, c.ny, c.nz, kc), blk, 0, stream()>>>(
          f0.nx, f0.ny, f0.nz, mg.code, x0, c.nx, c.ny, c.nz, xc, c.m, kc);
  }
  FFPM_LAUNCHED("k_mgf_prolong_z");
  for (int k = 0; k + 1 < s0.npost; ++k) {
    TimedLaunch tl("mg_fine_sweep");
    launch_step<false>(mg, s0.post[k]);
  }
  FFPM_LAUNCHED("k_mg_sweep_t");
  {
    TimedLaunch tl("mg_fine_last");
    launch_step<true>(mg, s0.post[s0.npost - 1], with_q ? mg.q16 : nullptr);
  }
  FFPM_LAUNCHED("k_mg_sweep_t");
  return OK;
}

int mg_direction(const Multigrid &mg, const Tiling &t, int cl, const float *pin, float *pout,
                 int mode, Scalars *sc) {
  if (mode != kDirUpdate && mode != kDirReset) {
    set_error(ERR_ARGUMENT, "multigrid direction mode");
    return ERR_ARGUMENT;
  }
  const dim3 blk(kTileX, t.ty);
  TimedLaunch tl("mg_direction");
  auto go = [&](auto kern) {
    kern<<<t.blocks, blk, 0, stream()>>>(t, mg.g32, mg.buf0(3), mg.code, pin, pout, mg.partials,
                                         sc);
  };
  if (cl == 4) {
    if (mode == kDirUpdate) go(k_mg_dir_t<4, kDirUpdate>);
    else go(k_mg_dir_t<4, kDirReset>);
  } else if (cl == 2) {
    if (mode == kDirUpdate) go(k_mg_dir_t<2, kDirUpdate>);
    else go(k_mg_dir_t<2, kDirReset>);
  } else {
    if (mode == kDirUpdate) go(k_mg_dir_t<1, kDirUpdate>);
    else go(k_mg_dir_t<1, kDirReset>);
  }
  FFPM_LAUNCHED("k_mg_dir_t");
  return OK;
}

}  // namespace ffpm
