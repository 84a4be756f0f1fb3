// Templated fused PCG kernels (value type T, CL x-adjacent cells per lane).
#pragma once

#include "ffpm_kernels.hpp"

namespace ffpm {

// Geometry in the solver's value type: the 1-D stencil coefficients and cell
// widths (FP64 arrays for the FP64 solve, FP32 copies for the mixed inner).
template <typename T>
struct GeomT {
  int nx, ny, nz;
  const T *cw, *ce, *cs, *cn, *cb, *ct;
  const T *dx, *dy, *dz;
};

inline GeomT<double> geom_t(const Geom &g) {
  return GeomT<double>{g.nx, g.ny, g.nz, g.cw, g.ce, g.cs, g.cn, g.cb, g.ct, g.dx, g.dy, g.dz};
}

// How an update kernel treats the iterate x: XM_NOW, x += alpha p; XM_DEFER,
// x is neither read nor written (its step alpha p is applied with the next
// direction's); XM_TWO, x = (x + alpha_pend p_pend) + alpha p, the deferred
// step and this one in the order two updates take them (so x is
// bit-identical to updating it at every iteration).
constexpr int XM_NOW = 0, XM_DEFER = 1, XM_TWO = 2;

// MODE is kDirUpdate, kDirReset or kDirCurrent.
template <typename T, int CL, int MODE>
__global__ void k_pcg_dir_t(Tiling t, GeomT<T> g, const T *z, const uint8_t *code, const T *pin,
                            T *pout, double *partials, Scalars *sc);
template <typename T, int CL>
__global__ void k_pcg_upd_t(Tiling t, GeomT<T> g, const T *p, const uint8_t *code, T *x, T *r,
                            T *z, double *partials, Scalars *sc);
template <typename T>
__global__ void k_dir_t(size_t n, T *p, const T *z, const Scalars *sc);
template <int CL>
__global__ void k_jacobi_t(Tiling t, GeomT<double> g, const double *a, double *b,
                           const double *src, const uint8_t *code);
__global__ void k_mixed_begin(GeomT<float> g, const double *r64, const uint8_t *code,
                              float *r32, float *z32, float *e32, double *partials);
__global__ void k_mixed_accumulate(size_t n, double *phi, const float *e32);
__global__ void k_to_float(int n, const double *a, float *b);

}  // namespace ffpm
