// Shared device/host helpers of the stair-step solver library (gfx950).
//
// Internal field layout ("dense"): cell (i,j,k), 0-based, lives at
// i + nx*(j + ny*k) -- the Fortran layout of an (nx,ny,nz) array, so rhs and
// the flow fields are used in place. phi and solid keep the caller's
// one-layer x-y halo layout (0:nx+1, 0:ny+1, nz) only at the interface; the
// iterations run on dense fields and substitute the zero physical halo.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

namespace ffpm {

// ---------------------------------------------------------------- errors --
void set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define FFPM_HIP(call)                                                   \
  do {                                                                   \
    hipError_t _e = (call);                                              \
    if (_e != hipSuccess) return ::ffpm::hip_fail(_e, #call, __FILE__, __LINE__); \
  } while (0)

#define FFPM_LAUNCHED(what)                                              \
  do {                                                                   \
    hipError_t _e = hipGetLastError();                                   \
    if (_e != hipSuccess) return ::ffpm::hip_fail(_e, what, __FILE__, __LINE__); \
  } while (0)

enum Status : int {
  OK = 0,
  ERR_ARGUMENT = 1,
  ERR_ISOLATED = 2,
  ERR_NO_OUTER_FACE = 3,
  ERR_METHOD = 4,
  ERR_SHAPE = 5,
  ERR_DECOMPOSITION = 6,
  ERR_DEVICE = 7,
};

// ------------------------------------------------------------ launch shape --
constexpr int kTileX = 64;   // one wavefront along x
constexpr int kTileY = 4;    // four wavefronts per workgroup
constexpr int kBlock = kTileX * kTileY;
constexpr int kXcds = 8;     // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kReduceSlots = 4;

// z-streaming tiles: (tiles_x, tiles_y, z chunks) flattened x-fastest. A
// workgroup is 64 x ty threads (one wavefront per row of the tile). With
// xcd_map the tiles are dealt to the XCDs in contiguous ranges, so that
// y-neighbour tiles (whose halo rows a tile reads) share an L2.
struct Tiling {
  int tiles_x, tiles_y, chunks_z, zchunk, ty;
  int total;      // tiles
  int per_xcd;    // ceil(total / 8)
  int blocks;     // launched workgroups
  int xcd_map;
};

// cells_x: cells one wavefront covers along x (64, or 128 for the paired-cell
// kernels).
Tiling make_tiling(int nx, int ny, int nz, int zchunk, int ty = 0, int cells_x = kTileX);

struct TileCoord {
  int i, j, k0, k1;
  bool valid;  // workgroup owns a tile
};

// block: the workgroup's index within its own solve's launch (blockIdx.x,
// except in the batched kernels, which run many solves in one grid).
__device__ inline TileCoord tile_coord(const Tiling &t, int nx, int ny, int nz, int block) {
  TileCoord c;
  const int L = block;
  const int tile = t.xcd_map ? (L % kXcds) * t.per_xcd + L / kXcds : L;
  c.valid = tile < t.total;
  const int tx = tile % t.tiles_x;
  const int rest = tile / t.tiles_x;
  const int ty = rest % t.tiles_y;
  const int tz = rest / t.tiles_y;
  c.i = tx * kTileX + threadIdx.x;
  c.j = ty * t.ty + threadIdx.y;
  c.k0 = tz * t.zchunk;
  c.k1 = min(nz, c.k0 + t.zchunk);
  return c;
}

__device__ inline TileCoord tile_coord(const Tiling &t, int nx, int ny, int nz) {
  return tile_coord(t, nx, ny, nz, blockIdx.x);
}

// ---------------------------------------------------------------- grids ----
// Geometry of one frame on the device: 1-D stencil coefficients and widths.
struct Geom {
  int nx, ny, nz;
  const double *cw, *ce, *cs, *cn, *cb, *ct;  // src/fpm_stairstep_mpi_3d.f90:499-510,138-142
  const double *dx, *dy, *dz;
  const double *xc, *yc, *zc;  // xc(0:nx+1), yc(0:ny+1), zc(nz)
};

// ------------------------------------------------------------ reductions ---
// Block-level sum of up to kReduceSlots doubles (wave shuffles, then LDS), and
// a grid-level finish in the last workgroup to arrive. The final sum runs over
// the per-workgroup partials in a fixed order, so a reduction is reproducible
// run to run for a fixed launch shape.
template <int N>
__device__ inline void block_sum(double (&v)[N], double *lds /* [N * 16] */) {
  __syncthreads();
#pragma unroll
  for (int s = 0; s < N; ++s) {
    double x = v[s];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[s] = x;
  }
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, wave = tid >> 6;
  const int nwaves = (blockDim.x * blockDim.y + 63) >> 6;
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < N; ++s) lds[s * 16 + wave] = v[s];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int s = 0; s < N; ++s) {
      double x = 0.0;
      for (int w = 0; w < nwaves; ++w) x += lds[s * 16 + w];
      v[s] = x;
    }
  }
}

__device__ inline void block_max(double &v, double *lds) {
  __syncthreads();
  double x = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off, 64));
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, wave = tid >> 6;
  const int nwaves = (blockDim.x * blockDim.y + 63) >> 6;
  if (lane == 0) lds[wave] = x;
  __syncthreads();
  if (tid == 0) {
    double m = 0.0;
    for (int w = 0; w < nwaves; ++w) m = fmax(m, lds[w]);
    v = m;
  }
}

// Grid reductions: each workgroup reduces its values and thread 0 stores them
// as this workgroup's partials (plain stores); a one-workgroup finish kernel
// launched next on the same stream sums the partials in a fixed order, so a
// reduction is reproducible run to run for a fixed launch shape. The kernel
// boundary makes the partials visible without a per-workgroup device fence.
__device__ inline void store_partials(double *partials, int nslots, const double *vals,
                                      int nblocks, int block) {
  if (threadIdx.x == 0 && threadIdx.y == 0)
    for (int s = 0; s < nslots; ++s) partials[(size_t)s * nblocks + block] = vals[s];
}

__device__ inline void store_partials(double *partials, int nslots, const double *vals,
                                      int nblocks) {
  store_partials(partials, nslots, vals, nblocks, blockIdx.x);
}

// Called by every thread of the finishing workgroup; thread 0 gets the sums.
// Each thread adds partials tid, tid + nthreads, ... in that order; the loads
// of four consecutive terms are issued together (the finish kernels would
// otherwise wait one memory latency per term).
template <int N>
__device__ inline void grid_finish_sum(const double *partials, int nblocks, double (&v)[N],
                                       double *lds) {
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nthreads = blockDim.x * blockDim.y;
#pragma unroll
  for (int s = 0; s < N; ++s) {
    const double *p = partials + (size_t)s * nblocks;
    double x = 0.0;
    int b = tid;
    for (; b + 3 * nthreads < nblocks; b += 4 * nthreads) {
      const double t0 = p[b], t1 = p[b + nthreads], t2 = p[b + 2 * nthreads],
                   t3 = p[b + 3 * nthreads];
      x += t0;
      x += t1;
      x += t2;
      x += t3;
    }
    for (; b < nblocks; b += nthreads) x += p[b];
    v[s] = x;
  }
  block_sum<N>(v, lds);
}

__device__ inline double grid_finish_max(const double *partials, int slot, int nblocks,
                                         double *lds) {
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int nthreads = blockDim.x * blockDim.y;
  const double *p = partials + (size_t)slot * nblocks;
  double x = 0.0;
  int b = tid;
  for (; b + 3 * nthreads < nblocks; b += 4 * nthreads) {
    const double t0 = p[b], t1 = p[b + nthreads], t2 = p[b + 2 * nthreads],
                 t3 = p[b + 3 * nthreads];
    x = fmax(x, t0);
    x = fmax(x, t1);
    x = fmax(x, t2);
    x = fmax(x, t3);
  }
  for (; b < nblocks; b += nthreads) x = fmax(x, p[b]);
  block_max(x, lds);
  return x;
}

// Cells per lane of the streaming kernels (a 64-lane row of a tile covers
// 64 x cells cells of an x row), at most vmax (the 16-byte vector) and
// dividing nx. FFPM_LANE_FIT=1: the width that leaves the fewest lanes of a
// row idle (ties to the wider), else the widest; FFPM_X2=0: one cell.
inline int cells_per_lane(int nx, int vmax) {
  const char *e = std::getenv("FFPM_X2");
  if (e && std::strcmp(e, "0") == 0) return 1;
  const char *f = std::getenv("FFPM_LANE_FIT");
  const bool fit = f && std::strcmp(f, "1") == 0;
  int best = 1;
  long best_idle = -1;
  for (int c = 1; c <= vmax; c *= 2) {
    if (nx % c) continue;
    const long row = 64L * c, idle = (nx + row - 1) / row * row - nx;
    if (!fit || best_idle < 0 || idle <= best_idle) best = c, best_idle = idle;
  }
  return best;
}

// Elementwise launch shape: grid-stride over n cells.
inline int elementwise_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace ffpm
