// Fused PCG iteration kernels (single solve) and the mixed-precision helper
// kernels. The iteration bodies live in ffpm_pcg_body.hpp.
#include "ffpm_bodies.hpp"
#include "ffpm_pcg_body.hpp"

namespace ffpm {

template <typename T, int CL, int MODE>
__global__ __launch_bounds__(1024) void k_pcg_dir_t(Tiling t, GeomT<T> g, const T *__restrict__ z,
                                                   const uint8_t *__restrict__ code,
                                                   const T *__restrict__ pin, T *__restrict__ pout,
                                                   double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  pcg_dir_body<T, CL, MODE>(t, blockIdx.x, gridDim.x, g, z, code, pin, pout, partials, sc, lds);
}

template <typename T, int CL>
__global__ __launch_bounds__(1024) void k_pcg_upd_t(Tiling t, GeomT<T> g, const T *__restrict__ p,
                                                   const uint8_t *__restrict__ code,
                                                   T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ z, double *partials,
                                                   Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  pcg_upd_body<T, CL>(t, blockIdx.x, gridDim.x, g, p, code, x, r, z, partials, sc, lds);
}

// Damped Jacobi sweep (src/fpm_stairstep_mpi_3d.f90:384-399): b = 0.2 a +
// rdiag (src + sum c a_nb), rdiag = 0.8 / diag in fluid cells and 0 in solids,
// rebuilt from the neighbour code (the same division as build_operator's, so
// the sweep is bit-identical to one with the stored rdiag). Same streaming
// scheme as the PCG kernels: the loads of plane k+1 (neighbours, src) and of
// plane k+2 (own row) are issued before plane k is computed; west/east
// neighbours by wave shuffle.
template <int CL>
__global__ __launch_bounds__(1024) void k_jacobi_t(Tiling t, GeomT<double> g,
                                                  const double *__restrict__ a,
                                                  double *__restrict__ b,
                                                  const double *__restrict__ src,
                                                  const uint8_t *__restrict__ code) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz, blockIdx.x);
  if (!l.active) return;
  using V = Vec<double, CL>;
  const int i = l.i, j = l.j;
  const size_t nxy = (size_t)nx * ny;
  double cw[CL], ce[CL];
#pragma unroll
  for (int m = 0; m < CL; ++m) {
    cw[m] = g.cw[i + m];
    ce[m] = g.ce[i + m];
  }
  const double cs = g.cs[j], cn = g.cn[j];
  const bool hw = i > 0, he = i + CL < nx, hs = j > 0, hn = j < ny - 1;
  const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
  const int lane = (int)threadIdx.x;
  const ptrdiff_t row0 = -(ptrdiff_t)(lane * CL);
  const ptrdiff_t ow = lane == 0 ? (hw ? -1 : 0) : row0;
  const ptrdiff_t oe = lane == kTileX - 1 ? (he ? CL : CL - 1) : row0;
  const size_t base = (size_t)i + (size_t)nx * (size_t)j;
  const int k0 = l.k0, k1 = l.k1, last = min(k1, nz - 1);
  auto at = [&](int k) { return base + nxy * (size_t)k; };
  struct Set {
    V a_next, s;
    Plane<kDirCurrent, double, CL> pl;
  };
  Set sa, sb;
  V a_cur = ldv<double, CL>(a, at(k0));
  V a_prev = k0 > 0 ? ldv<double, CL>(a, at(k0 - 1)) : a_cur;
  sa.a_next = ldv<double, CL>(a, at(min(k0 + 1, last)));
  sa.pl.load(a, a, code, at(k0), os, on, ow, oe);
  sa.s = ldv<double, CL>(src, at(k0));
  auto stage = [&](int k, const Set &u, Set &ld) {
    const int kn = min(k + 1, k1 - 1);
    ld.a_next = ldv<double, CL>(a, at(min(k + 2, last)));
    ld.pl.load(a, a, code, at(kn), os, on, ow, oe);
    ld.s = ldv<double, CL>(src, at(kn));
    const double from_w = __shfl_up(a_cur.v[CL - 1], 1, kTileX);
    const double from_e = __shfl_down(a_cur.v[0], 1, kTileX);
    const double vw = lane == 0 ? u.pl.w.a : from_w;
    const double ve = lane == kTileX - 1 ? u.pl.e.a : from_e;
    const double cb = g.cb[k], ct = g.ct[k];
    V bv;
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      const double west = m == 0 ? (hw ? vw : 0.0) : a_cur.v[m - 1];
      const double east = m == CL - 1 ? (he ? ve : 0.0) : a_cur.v[m + 1];
      const double south = hs ? u.pl.s.a.v[m] : 0.0, north = hn ? u.pl.n.a.v[m] : 0.0;
      const double d = diag_code<double>(u.pl.code.b[m], cw[m], ce[m], cs, cn, cb, ct);
      const double rd = d > 0 ? 0.8 / d : 0.0;
      bv.v[m] = 0.2 * a_cur.v[m] + rd * (u.s.v[m] + cw[m] * west + ce[m] * east + cs * south +
                                         cn * north + cb * a_prev.v[m] + ct * u.a_next.v[m]);
    }
    stv<double, CL>(b, at(k), bv);
    a_prev = a_cur;
    a_cur = u.a_next;
  };
  for (int k = k0; k < k1; k += 2) {
    stage(k, sa, sb);
    if (k + 1 < k1) stage(k + 1, sb, sa);
  }
}

template __global__ void k_jacobi_t<1>(Tiling, GeomT<double>, const double *, double *,
                                       const double *, const uint8_t *);
template __global__ void k_jacobi_t<2>(Tiling, GeomT<double>, const double *, double *,
                                       const double *, const uint8_t *);

// Deferred direction update p = z + beta p (before a convergence check).
template <typename T>
__global__ __launch_bounds__(256) void k_dir_t(size_t n, T *__restrict__ p,
                                               const T *__restrict__ z, const Scalars *sc) {
  dir_update_body<T>(n, blockIdx.x, gridDim.x, p, z, sc);
}

// Inner right-hand side of a defect correction (body in ffpm_pcg_body.hpp).
__global__ __launch_bounds__(256) void k_mixed_begin(GeomT<float> g, const double *__restrict__ r64,
                                                     const uint8_t *__restrict__ code,
                                                     float *__restrict__ r32,
                                                     float *__restrict__ z32,
                                                     float *__restrict__ e32, double *partials) {
  __shared__ double lds[16 * kReduceSlots];
  mixed_begin_body(g, r64, code, r32, z32, e32, partials, blockIdx.x, gridDim.x, lds);
}

// phi += e (FP64 accumulation of the FP32 correction).
__global__ __launch_bounds__(256) void k_mixed_accumulate(size_t n, double *__restrict__ phi,
                                                          const float *__restrict__ e32) {
  mixed_accumulate_body(n, phi, e32, blockIdx.x, gridDim.x);
}

// FP32 copies of the 1-D coefficients and widths.
__global__ void k_to_float(int n, const double *__restrict__ a, float *__restrict__ b) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

#define FFPM_INSTANTIATE_DIR(T, CL, MODE)                                                       \
  template __global__ void k_pcg_dir_t<T, CL, MODE>(Tiling, GeomT<T>, const T *,               \
                                                    const uint8_t *, const T *, T *, double *,  \
                                                    Scalars *);
#define FFPM_INSTANTIATE(T, CL)                                                                 \
  FFPM_INSTANTIATE_DIR(T, CL, kDirUpdate)                                                       \
  FFPM_INSTANTIATE_DIR(T, CL, kDirReset)                                                        \
  FFPM_INSTANTIATE_DIR(T, CL, kDirCurrent)                                                      \
  template __global__ void k_pcg_upd_t<T, CL>(Tiling, GeomT<T>, const T *, const uint8_t *,    \
                                              T *, T *, T *, double *, Scalars *);
FFPM_INSTANTIATE(double, 1)
FFPM_INSTANTIATE(double, 2)
FFPM_INSTANTIATE(float, 1)
FFPM_INSTANTIATE(float, 2)
FFPM_INSTANTIATE(float, 4)
#undef FFPM_INSTANTIATE
#undef FFPM_INSTANTIATE_DIR
template __global__ void k_dir_t<double>(size_t, double *, const double *, const Scalars *);
template __global__ void k_dir_t<float>(size_t, float *, const float *, const Scalars *);

}  // namespace ffpm
