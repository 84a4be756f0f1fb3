// Fused PCG iteration kernels (single solve) and the mixed-precision helper
// kernels. The iteration bodies live in ffpm_pcg_body.hpp.
#include "ffpm_bodies.hpp"
#include "ffpm_pcg_body.hpp"

namespace ffpm {

template <typename T, int CL, int MODE>
__global__ __launch_bounds__(512) void k_pcg_dir_t(Tiling t, GeomT<T> g, const T *__restrict__ z,
                                                   const uint8_t *__restrict__ code,
                                                   const T *__restrict__ pin, T *__restrict__ pout,
                                                   double *partials, Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  pcg_dir_body<T, CL, MODE>(t, blockIdx.x, gridDim.x, g, z, code, pin, pout, partials, sc, lds);
}

template <typename T, int CL>
__global__ __launch_bounds__(512) void k_pcg_upd_t(Tiling t, GeomT<T> g, const T *__restrict__ p,
                                                   const uint8_t *__restrict__ code,
                                                   T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ z, double *partials,
                                                   Scalars *sc) {
  __shared__ double lds[16 * kReduceSlots];
  pcg_upd_body<T, CL>(t, blockIdx.x, gridDim.x, g, p, code, x, r, z, partials, sc, lds);
}

// Deferred direction update p = z + beta p (before a convergence check).
template <typename T>
__global__ __launch_bounds__(256) void k_dir_t(size_t n, T *__restrict__ p,
                                               const T *__restrict__ z, const Scalars *sc) {
  dir_update_body<T>(n, blockIdx.x, gridDim.x, p, z, sc);
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
