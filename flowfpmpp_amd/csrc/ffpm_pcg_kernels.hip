// Fused PCG iteration kernels, templated on the value type T (double for the
// FP64 solve, float for the inner solve of mixed precision) and on CL, the
// x-adjacent cells one lane owns (16-byte loads: CL = 2 for double, 4 for
// float; CL = 1 when nx is not a multiple of the vector width).
//
// Iteration (src/fpm_stairstep_mpi_3d.f90:246-283, formulas unchanged):
//   k_pcg_dir_t : p = z + beta p_old (the previous iteration's direction
//                 update, deferred), q = A p, p.q (FP64 sum)
//   k_pcg_upd_t : q = A p recomputed bit-identically, x += alpha p,
//                 r -= alpha q, z = M^-1 r, z.r (FP64 sum)
// A = V (diag - offdiag) with diag and rdiag = 1/(V diag) rebuilt in
// registers from the one-byte neighbour code and the 1-D coefficients, so the
// kernels read no diagonal arrays. Each lane rebuilds its neighbours' new
// direction from z and p_old through the cache: no LDS exchange, no barrier.
// Kernel time is bound by wavefront memory instructions rather than bytes, so
// every value moves in 16-byte vectors where the layout allows.
#include "ffpm_pcg.hpp"

namespace ffpm {

namespace {

template <typename T, int CL>
struct alignas(CL * sizeof(T)) Vec {
  T v[CL];
};

template <int CL>
struct alignas(CL) Codes {
  uint8_t b[CL];
};

template <typename T, int CL>
__device__ __forceinline__ Vec<T, CL> ldv(const T *__restrict__ a, size_t c) {
  return *reinterpret_cast<const Vec<T, CL> *>(a + c);
}

template <typename T, int CL>
__device__ __forceinline__ void stv(T *__restrict__ a, size_t c, const Vec<T, CL> &x) {
  *reinterpret_cast<Vec<T, CL> *>(a + c) = x;
}

// Direction value (see kDir* modes): z + beta p_old, z, or p_old.
template <typename T>
__device__ __forceinline__ T dir1(int mode, T beta, const T *__restrict__ z,
                                  const T *__restrict__ pin, size_t c) {
  if (mode == kDirCurrent) return pin[c];
  if (mode == kDirReset) return z[c];
  return z[c] + beta * pin[c];
}

template <typename T, int CL>
__device__ __forceinline__ Vec<T, CL> dirv(int mode, T beta, const T *__restrict__ z,
                                           const T *__restrict__ pin, size_t c) {
  if (mode == kDirCurrent) return ldv<T, CL>(pin, c);
  Vec<T, CL> zz = ldv<T, CL>(z, c);
  if (mode == kDirReset) return zz;
  const Vec<T, CL> pp = ldv<T, CL>(pin, c);
#pragma unroll
  for (int m = 0; m < CL; ++m) zz.v[m] = zz.v[m] + beta * pp.v[m];
  return zz;
}

// diag from the neighbour code: the fluid-neighbour coefficients summed in
// build_operator's order (src/fpm_stairstep_mpi_3d.f90:353-370), so for
// T = double the value is bit-identical to the stored diag.
template <typename T>
__device__ __forceinline__ T diag_code(unsigned code, const GeomT<T> &g, int i, int j, int k) {
  T d = 0;
  if (code & kCodeW) d = d + g.cw[i];
  if (code & kCodeE) d = d + g.ce[i];
  if (code & kCodeS) d = d + g.cs[j];
  if (code & kCodeN) d = d + g.cn[j];
  if (code & kCodeB) d = d + g.cb[k];
  if (code & kCodeT) d = d + g.ct[k];
  return d;
}

template <typename T>
__device__ __forceinline__ T rdiag_code(T d, T dxdy, T dzk) {
  return d > 0 ? T(1) / (dxdy * dzk * d) : T(0);
}

struct Lane {
  int i, j, k0, k1;
  bool active;
};

template <int CL>
__device__ __forceinline__ Lane lane_of(const Tiling &t, int nx, int ny, int nz) {
  const TileCoord tc = tile_coord(t, nx, ny, nz);
  Lane l;
  l.i = (tc.i - (int)threadIdx.x) * CL + CL * (int)threadIdx.x;
  l.j = tc.j;
  l.k0 = tc.k0;
  l.k1 = tc.k1;
  l.active = tc.valid && l.i < nx && l.j < ny;
  return l;
}

}  // namespace

template <typename T, int CL>
__global__ __launch_bounds__(512) void k_pcg_dir_t(Tiling t, GeomT<T> g, const T *__restrict__ z,
                                                   const uint8_t *__restrict__ code,
                                                   const T *__restrict__ pin, T *__restrict__ pout,
                                                   int mode, double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  if (sc->status != 0) return;
  const T beta = (T)sc->beta;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz);
  double s = 0.0;
  if (l.active) {
    const int i = l.i, j = l.j;
    const size_t nxy = (size_t)nx * ny;
    T cw[CL], ce[CL], dxdy[CL];
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      cw[m] = g.cw[i + m];
      ce[m] = g.ce[i + m];
      dxdy[m] = g.dx[i + m] * g.dy[j];
    }
    const T cs = g.cs[j], cn = g.cn[j];
    const bool hw = i > 0, he = i + CL < nx, hs = j > 0, hn = j < ny - 1;
    const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
    const ptrdiff_t ow = hw ? -1 : 0, oe = he ? CL : CL - 1;
    size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * l.k0);
    Vec<T, CL> p_cur = dirv<T, CL>(mode, beta, z, pin, c);
    Vec<T, CL> p_prev = l.k0 > 0 ? dirv<T, CL>(mode, beta, z, pin, c - nxy) : p_cur;
    for (int k = l.k0; k < l.k1; ++k) {
      const size_t cn_ = (k + 1 < nz) ? c + nxy : c;
      const Vec<T, CL> p_next = dirv<T, CL>(mode, beta, z, pin, cn_);
      const Vec<T, CL> vs = dirv<T, CL>(mode, beta, z, pin, c + os);
      const Vec<T, CL> vn = dirv<T, CL>(mode, beta, z, pin, c + on);
      const T vw = dir1<T>(mode, beta, z, pin, c + ow);
      const T ve = dir1<T>(mode, beta, z, pin, c + oe);
      const Codes<CL> cc = *reinterpret_cast<const Codes<CL> *>(code + c);
      const T dzk = g.dz[k], cb = g.cb[k], ct = g.ct[k];
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const T west = m == 0 ? (hw ? vw : T(0)) : p_cur.v[m - 1];
        const T east = m == CL - 1 ? (he ? ve : T(0)) : p_cur.v[m + 1];
        const T south = hs ? vs.v[m] : T(0), north = hn ? vn.v[m] : T(0);
        const T d = diag_code<T>(cc.b[m], g, i + m, j, k);
        const T q = dxdy[m] * dzk *
                    (d * p_cur.v[m] - cw[m] * west - ce[m] * east - cs * south - cn * north -
                     cb * p_prev.v[m] - ct * p_next.v[m]);
        s += (double)p_cur.v[m] * (double)(d > 0 ? q : T(0));
      }
      if (mode != kDirCurrent) stv<T, CL>(pout, c, p_cur);
      p_prev = p_cur;
      p_cur = p_next;
      c += nxy;
    }
  }
  double v[1] = {s};
  block_sum<1>(v, lds);
  store_partials(partials, 1, v, gridDim.x);
}

template <typename T, int CL>
__global__ __launch_bounds__(512) void k_pcg_upd_t(Tiling t, GeomT<T> g, const T *__restrict__ p,
                                                   const uint8_t *__restrict__ code,
                                                   T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ z, double *partials,
                                                   Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  if (sc->status != 0) return;
  const T alpha = (T)sc->alpha;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz);
  double s = 0.0;
  if (l.active) {
    const int i = l.i, j = l.j;
    const size_t nxy = (size_t)nx * ny;
    T cw[CL], ce[CL], dxdy[CL];
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      cw[m] = g.cw[i + m];
      ce[m] = g.ce[i + m];
      dxdy[m] = g.dx[i + m] * g.dy[j];
    }
    const T cs = g.cs[j], cn = g.cn[j];
    const bool hw = i > 0, he = i + CL < nx, hs = j > 0, hn = j < ny - 1;
    const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
    const ptrdiff_t ow = hw ? -1 : 0, oe = he ? CL : CL - 1;
    size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * l.k0);
    Vec<T, CL> p_cur = ldv<T, CL>(p, c);
    Vec<T, CL> p_prev = l.k0 > 0 ? ldv<T, CL>(p, c - nxy) : p_cur;
    for (int k = l.k0; k < l.k1; ++k) {
      const size_t cn_ = (k + 1 < nz) ? c + nxy : c;
      const Vec<T, CL> p_next = ldv<T, CL>(p, cn_);
      const Vec<T, CL> vs = ldv<T, CL>(p, c + os), vn = ldv<T, CL>(p, c + on);
      const T vw = p[c + ow], ve = p[c + oe];
      const Codes<CL> cc = *reinterpret_cast<const Codes<CL> *>(code + c);
      Vec<T, CL> xv = ldv<T, CL>(x, c), rv = ldv<T, CL>(r, c), zv;
      const T dzk = g.dz[k], cb = g.cb[k], ct = g.ct[k];
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const T west = m == 0 ? (hw ? vw : T(0)) : p_cur.v[m - 1];
        const T east = m == CL - 1 ? (he ? ve : T(0)) : p_cur.v[m + 1];
        const T south = hs ? vs.v[m] : T(0), north = hn ? vn.v[m] : T(0);
        const T d = diag_code<T>(cc.b[m], g, i + m, j, k);
        const T qr = dxdy[m] * dzk *
                     (d * p_cur.v[m] - cw[m] * west - ce[m] * east - cs * south - cn * north -
                      cb * p_prev.v[m] - ct * p_next.v[m]);
        const T q = d > 0 ? qr : T(0);
        xv.v[m] = xv.v[m] + alpha * p_cur.v[m];
        rv.v[m] = rv.v[m] - alpha * q;
        zv.v[m] = rdiag_code<T>(d, dxdy[m], dzk) * rv.v[m];
        s += (double)zv.v[m] * (double)rv.v[m];
      }
      stv<T, CL>(x, c, xv);
      stv<T, CL>(r, c, rv);
      stv<T, CL>(z, c, zv);
      p_prev = p_cur;
      p_cur = p_next;
      c += nxy;
    }
  }
  double v[1] = {s};
  block_sum<1>(v, lds);
  store_partials(partials, 1, v, gridDim.x);
}

// LDS-tiled direction + SpMV: the same results as k_pcg_dir_t with fewer
// global memory instructions. A workgroup of ty wavefronts owns ty rows of
// 64*CL cells; for every plane the new direction of its rows, the two halo
// rows and the two edge columns is built once, exchanged through LDS, and each
// lane reads its x-y neighbours from there (z-neighbours stay in registers).
// Two LDS planes alternate, so one barrier per plane suffices: plane k+1 is
// built while plane k is consumed.
template <typename T, int CL, int TY>
__global__ __launch_bounds__(64 * TY) void k_pcg_dir_lds_t(
    Tiling t, GeomT<T> g, const T *__restrict__ z, const uint8_t *__restrict__ code,
    const T *__restrict__ pin, T *__restrict__ pout, int mode, double *partials, Scalars *sc) {
  constexpr int W = 64 * CL + 2;  // LDS row: west edge, 64*CL cells, east edge
  __shared__ T tile[2][TY + 2][W];
  __shared__ double lds[16 * kReduceSlots];
  if (sc->status != 0) return;
  const T beta = (T)sc->beta;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const TileCoord tc = tile_coord(t, nx, ny, nz);
  const int lane = threadIdx.x, w = threadIdx.y;
  const int i0 = (tc.i - lane) * CL;  // first cell of the tile row
  const int j0 = tc.j - w;
  const int i = i0 + lane * CL, j = tc.j;
  const bool active = tc.valid && i < nx && j < ny;
  const size_t nxy = (size_t)nx * ny;
  // The halo rows are built by the first and last wavefront, the edge
  // columns by lanes of the first wavefront.
  const int hrow = (w == 0) ? j0 - 1 : (w == TY - 1 ? j0 + TY : -2);
  const int hslot = (w == 0) ? 0 : TY + 1;
  const bool hrow_in = tc.valid && hrow >= 0 && hrow < ny && i < nx;
  // Only the owned rows need edge columns (south/north reads stay inside).
  int erow = -1, ecol = -1, eslot_r = 0, eslot_c = 0;
  if (w == 0 && lane < TY) {
    erow = j0 + lane; ecol = i0 - 1; eslot_r = lane + 1; eslot_c = 0;
  } else if (w == 0 && lane >= 32 && lane < 32 + TY) {
    erow = j0 + lane - 32; ecol = i0 + 64 * CL; eslot_r = lane - 31; eslot_c = W - 1;
  }
  const bool edge_in = tc.valid && erow >= 0 && erow < ny && ecol >= 0 && ecol < nx;
  const bool edge_lane = erow != -1;
  auto plane_base = [&](int jj, int kk) -> size_t {
    return (size_t)i + (size_t)nx * ((size_t)jj + (size_t)ny * kk);
  };
  auto put_row = [&](int b, int slot, const Vec<T, CL> &v) {
#pragma unroll
    for (int m = 0; m < CL; ++m) tile[b][slot][1 + lane * CL + m] = v.v[m];
  };
  Vec<T, CL> zero;
#pragma unroll
  for (int m = 0; m < CL; ++m) zero.v[m] = T(0);
  // Builds plane kk of buffer b from this thread's share of the loads.
  auto build = [&](int b, int kk, const Vec<T, CL> &own) {
    put_row(b, w + 1, active ? own : zero);
    if (hrow != -2)
      put_row(b, hslot, hrow_in ? dirv<T, CL>(mode, beta, z, pin, plane_base(hrow, kk)) : zero);
    if (edge_lane)
      tile[b][eslot_r][eslot_c] =
          edge_in ? dir1<T>(mode, beta, z, pin,
                            (size_t)ecol + (size_t)nx * ((size_t)erow + (size_t)ny * kk))
                  : T(0);
  };
  double s = 0.0;
  const int k0 = tc.k0, k1 = tc.valid ? tc.k1 : tc.k0;
  T cw[CL], ce[CL], dxdy[CL];
  T cs = 0, cn = 0;
  if (active) {
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      cw[m] = g.cw[i + m];
      ce[m] = g.ce[i + m];
      dxdy[m] = g.dx[i + m] * g.dy[j];
    }
    cs = g.cs[j];
    cn = g.cn[j];
  }
  const bool hw = i > 0, he = i + CL < nx, hs = j > 0, hn = j < ny - 1;
  Vec<T, CL> p_cur = zero, p_prev = zero;
  if (k0 < k1) {
    if (active) {
      const size_t c = plane_base(j, k0);
      p_cur = dirv<T, CL>(mode, beta, z, pin, c);
      p_prev = k0 > 0 ? dirv<T, CL>(mode, beta, z, pin, c - nxy) : p_cur;
    }
    build(0, k0, p_cur);
  }
  __syncthreads();
  for (int k = k0; k < k1; ++k) {
    const int b = (k - k0) & 1;
    Vec<T, CL> p_next = p_cur;
    if (active && k + 1 < nz) p_next = dirv<T, CL>(mode, beta, z, pin, plane_base(j, k + 1));
    if (k + 1 < k1) build(b ^ 1, k + 1, p_next);
    if (active) {
      const size_t c = plane_base(j, k);
      const Codes<CL> cc = *reinterpret_cast<const Codes<CL> *>(code + c);
      const T dzk = g.dz[k], cb = g.cb[k], ct = g.ct[k];
      const int col = 1 + lane * CL;
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const T west = m > 0 ? p_cur.v[m - 1] : (hw ? tile[b][w + 1][col - 1] : T(0));
        const T east = m < CL - 1 ? p_cur.v[m + 1] : (he ? tile[b][w + 1][col + CL] : T(0));
        const T south = hs ? tile[b][w][col + m] : T(0);
        const T north = hn ? tile[b][w + 2][col + m] : T(0);
        const T d = diag_code<T>(cc.b[m], g, i + m, j, k);
        const T q = dxdy[m] * dzk *
                    (d * p_cur.v[m] - cw[m] * west - ce[m] * east - cs * south - cn * north -
                     cb * p_prev.v[m] - ct * p_next.v[m]);
        s += (double)p_cur.v[m] * (double)(d > 0 ? q : T(0));
      }
      if (mode != kDirCurrent) stv<T, CL>(pout, c, p_cur);
    }
    p_prev = p_cur;
    p_cur = p_next;
    __syncthreads();
  }
  double v[1] = {s};
  block_sum<1>(v, lds);
  store_partials(partials, 1, v, gridDim.x);
}

// Deferred direction update p = z + beta p (before a convergence check).
template <typename T>
__global__ __launch_bounds__(256) void k_dir_t(size_t n, T *__restrict__ p,
                                               const T *__restrict__ z, const Scalars *sc) {
  if (sc->status != 0) return;
  const T beta = (T)sc->beta;
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    p[c] = z[c] + beta * p[c];
}

// Inner right-hand side of a defect correction: r32 = r64 (the projected FP64
// true residual), z32 = M32^-1 r32, e32 = 0, and rz = z32.r32 summed in FP64.
__global__ __launch_bounds__(256) void k_mixed_begin(GeomT<float> g, const double *__restrict__ r64,
                                                     const uint8_t *__restrict__ code,
                                                     float *__restrict__ r32,
                                                     float *__restrict__ z32,
                                                     float *__restrict__ e32, double *partials) {
  __shared__ double lds[16 * kReduceSlots];
  const int nx = g.nx, ny = g.ny;
  const size_t n = (size_t)nx * ny * g.nz;
  double s[1] = {0.0};
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((size_t)nx * ny));
    const float d = diag_code<float>(code[c], g, i, j, k);
    const float rv = (float)r64[c];
    const float zv = rdiag_code<float>(d, g.dx[i] * g.dy[j], g.dz[k]) * rv;
    r32[c] = rv;
    z32[c] = zv;
    e32[c] = 0.0f;
    s[0] += (double)zv * (double)rv;
  }
  block_sum<1>(s, lds);
  store_partials(partials, 1, s, gridDim.x);
}

// phi += e (FP64 accumulation of the FP32 correction).
__global__ __launch_bounds__(256) void k_mixed_accumulate(size_t n, double *__restrict__ phi,
                                                          const float *__restrict__ e32) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n;
       c += (size_t)gridDim.x * blockDim.x)
    phi[c] = phi[c] + (double)e32[c];
}

// FP32 copies of the 1-D coefficients and widths.
__global__ void k_to_float(int n, const double *__restrict__ a, float *__restrict__ b) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

#define FFPM_INSTANTIATE(T, CL)                                                                 \
  template __global__ void k_pcg_dir_t<T, CL>(Tiling, GeomT<T>, const T *, const uint8_t *,    \
                                              const T *, T *, int, double *, Scalars *);       \
  template __global__ void k_pcg_upd_t<T, CL>(Tiling, GeomT<T>, const T *, const uint8_t *,    \
                                              T *, T *, T *, double *, Scalars *);
FFPM_INSTANTIATE(double, 1)
FFPM_INSTANTIATE(double, 2)
FFPM_INSTANTIATE(float, 1)
FFPM_INSTANTIATE(float, 2)
FFPM_INSTANTIATE(float, 4)
#undef FFPM_INSTANTIATE
#define FFPM_INSTANTIATE_LDS(T, CL, TY)                                                        \
  template __global__ void k_pcg_dir_lds_t<T, CL, TY>(Tiling, GeomT<T>, const T *,               \
                                                      const uint8_t *, const T *, T *, int,      \
                                                      double *, Scalars *);
FFPM_INSTANTIATE_LDS(double, 1, 4)
FFPM_INSTANTIATE_LDS(double, 1, 8)
FFPM_INSTANTIATE_LDS(double, 2, 4)
FFPM_INSTANTIATE_LDS(double, 2, 8)
FFPM_INSTANTIATE_LDS(float, 1, 4)
FFPM_INSTANTIATE_LDS(float, 1, 8)
FFPM_INSTANTIATE_LDS(float, 2, 4)
FFPM_INSTANTIATE_LDS(float, 2, 8)
FFPM_INSTANTIATE_LDS(float, 4, 4)
FFPM_INSTANTIATE_LDS(float, 4, 8)
#undef FFPM_INSTANTIATE_LDS
template __global__ void k_dir_t<double>(size_t, double *, const double *, const Scalars *);
template __global__ void k_dir_t<float>(size_t, float *, const float *, const Scalars *);

}  // namespace ffpm
