// Device bodies of the multigrid cycle's kernels (layout and cycle as
// described in ffpm_mg.hpp), shared by the standalone kernels (ffpm_mg.hip,
// one solve per launch) and the batched ones (ffpm_batch.hip, many solves per
// launch): every body takes its workgroup's coordinates within its own
// solve's launch, so a batched solve computes exactly what its standalone
// solve computes. Cell and column bodies run 256-thread workgroups over the
// cells (or coarse columns) of one z plane taken in storage order, so every
// lane has work on narrow levels too; all loads are unconditional with
// clamped indices and masked coefficients.
#pragma once

#include <cstdlib>

#include "ffpm_mg.hpp"
#include "ffpm_pcg_body.hpp"

namespace ffpm {
namespace mgk {

constexpr int kMt = 256;  // threads of a cell or column workgroup

// Workgroups per z plane of an nx x ny level (the grid is that x 1 x planes).
__host__ __device__ inline int plane_blocks(int nx, int ny) { return (nx * ny + kMt - 1) / kMt; }

inline dim3 cell_grid(int nx, int ny, int nz) {
  return dim3((unsigned)plane_blocks(nx, ny), 1u, (unsigned)nz);
}

// Wavefronts the column launches aim for: FFPM_MG_COLUMN_WAVES, default 16384.
inline long column_waves() {
  static const long target = [] {
    const char *e = std::getenv("FFPM_MG_COLUMN_WAVES");
    const long v = e ? std::atol(e) : 16384;
    return v < 1 ? 1L : v;
  }();
  return target;
}

// Column kernels (one thread per coarse (I, J) walking K): K chunk so that
// the launch has at least `target` wavefronts (column_waves() for one solve
// per launch; a batch shares the target among its solves).
inline int column_chunk(int cnx, int cny, int cnz, long target = column_waves()) {
  const long waves_per_plane = (long)plane_blocks(cnx, cny) * (kMt / 64);
  long chunks = (target + waves_per_plane - 1) / waves_per_plane;
  if (chunks < 1) chunks = 1;
  if (chunks > cnz) chunks = cnz;
  return (int)((cnz + chunks - 1) / chunks);
}

inline dim3 column_grid(int cnx, int cny, int cnz, int kchunk) {
  return dim3((unsigned)plane_blocks(cnx, cny), 1u, (unsigned)((cnz + kchunk - 1) / kchunk));
}

// Workgroup coordinates of a cell or column launch: blockIdx in the
// standalone kernels, decoded from the workgroup's index within its own
// solve's launch in the batched ones.
struct Blk {
  int x, y, z;
};

__device__ __forceinline__ Blk blk3() {
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z};
}

// Local workgroup b of a launch of gx x gy x (any) workgroups, x fastest.
__device__ __forceinline__ Blk blk_of(int b, int gx, int gy) {
  const int q = b / gx;
  return Blk{b - q * gx, q % gy, q / gy};
}

struct Idx {
  int i, j, k;
  size_t c;
  bool in;
};

__device__ __forceinline__ Idx cell_idx(int nx, int ny, Blk bk) {
  Idx t;
  const int p = bk.x * kMt + (int)threadIdx.x;  // index in the plane
  t.j = p / nx;
  t.i = p - t.j * nx;
  t.k = bk.z;
  t.in = t.j < ny;
  t.c = (size_t)t.i + (size_t)nx * ((size_t)t.j + (size_t)ny * t.k);
  return t;
}

// Diagonal of a coarse level: the sum of the six face conductances.
__device__ __forceinline__ float coarse_diag(int nx, int ny, const MgStore *__restrict__ gE,
                                             const MgStore *__restrict__ gN,
                                             const MgStore *__restrict__ gT, size_t c, int i,
                                             int j, int k) {
  const size_t sx = (size_t)nx, sxy = (size_t)nx * ny;
  const float gw = i > 0 ? gE[c - 1] : 0.f, gs = j > 0 ? gN[c - sx] : 0.f,
              gb = k > 0 ? gT[c - sxy] : 0.f;
  return gw + gE[c] + gs + gN[c] + gb + gT[c];
}

// The coarse level's first sweep from zero, x = omega b / diag (0 where the
// diagonal is 0), formed by the restriction that produces b.
struct CoarseOut {
  MgStore *b, *x;
  const MgStore *rd;  // omega / diag, 0 where the diagonal is 0
  __device__ __forceinline__ void store(size_t c, float bv) const {
    const MgStore bh = (MgStore)bv;
    b[c] = bh;
    x[c] = (MgStore)((float)rd[c] * (float)bh);
  }
};

// ---- level 0: operator from the neighbour code and FP32 1-D coefficients ----
// The level-0 diagonal of a lane's CL cells in one plane. Almost every cell
// is fluid with all its neighbours inside the domain fluid: its code is the
// one the domain walls alone leave (exy[m] | ez: a lane constant and a plane
// constant), and its diagonal is dxy[m] + cz, the lane's in-plane
// coefficients towards the neighbours inside the domain plus the plane's
// (one add). A wave with any other cell (next to a body) sums the
// coefficients of its codes.
template <int CL>
__device__ __forceinline__ void level0_diag(float (&d)[CL], const Codes<CL> &code,
                                            const unsigned (&exy)[CL], unsigned ez,
                                            const float (&cw)[CL], const float (&ce)[CL],
                                            const float (&dxy)[CL], float cs, float cn, float cb,
                                            float ct, float cz) {
  bool full = true;
#pragma unroll
  for (int m = 0; m < CL; ++m) full = full && code.b[m] == (exy[m] | ez);
  if (__all(full)) {
#pragma unroll
    for (int m = 0; m < CL; ++m) d[m] = dxy[m] + cz;
  } else {
#pragma unroll
    for (int m = 0; m < CL; ++m) d[m] = diag_code<float>(code.b[m], cw[m], ce[m], cs, cn, cb, ct);
  }
}

// The code of a fluid cell whose neighbours inside the domain are all fluid:
// the in-plane part for the lane's cells (walls at i = 0, i = nx - 1 on the
// lane's edge cells, j = 0, j = ny - 1) and the plane's part.
template <int CL>
__device__ __forceinline__ void wall_codes(unsigned (&exy)[CL], bool hw, bool he, bool hs,
                                           bool hn) {
#pragma unroll
  for (int m = 0; m < CL; ++m)
    exy[m] = kCodeFluid | (m > 0 || hw ? kCodeW : 0u) | (m < CL - 1 || he ? kCodeE : 0u) |
             (hs ? kCodeS : 0u) | (hn ? kCodeN : 0u);
}

__device__ __forceinline__ unsigned wall_code_z(int k, int nz) {
  return (k > 0 ? kCodeB : 0u) | (k < nz - 1 ? kCodeT : 0u);
}

// The first sweep from zero, x = sig omega D^-1 r = sig omega s / d (D = V
// diag, s = r / V rounded to FP32, 0 where d = 0; sig = sc->sig, the cycle
// scale), stored in FP16, which the rest of the cycle also reads as its
// (scaled) source s = x d / omega. The true residual's projection
// (src/fpm_stairstep_mpi_3d.f90:319-322) is applied first: r -= red1 / red2
// (from sc) in the cells with a fluid neighbour (those whose diag is
// positive), and the projected r is stored.
__device__ __forceinline__ void mgf_jac0_body(Blk bk, GeomT<float> g,
                                              const uint8_t *__restrict__ code,
                                              double *__restrict__ r, MgStore *__restrict__ x,
                                              float omega, const Scalars *sc) {
  const Idx t = cell_idx(g.nx, g.ny, bk);
  if (!t.in) return;
  const unsigned cd = code[t.c];
  double rv = r[t.c];
  if (cd & (kCodeW | kCodeE | kCodeS | kCodeN | kCodeB | kCodeT)) {
    rv = rv - sc->red1 / sc->red2;
    r[t.c] = rv;
  }
  const float sig = (float)sc->sig;
  const float d = diag_code<float>(cd, g.cw[t.i], g.ce[t.i], g.cs[t.j], g.cn[t.j], g.cb[t.k],
                                   g.ct[t.k]);
  const float V = g.dx[t.i] * g.dy[t.j] * g.dz[t.k];
  const float sv = d > 0.f ? (float)rv / V : 0.f;
  x[t.c] = d > 0.f ? (MgStore)(omega * sv / d * sig) : (MgStore)0.f;
}

// Coarse right-hand side: the level-0 residual r - A x = V (s - (d x - sum c
// x_nb)) summed over the children of each level-1 cell. One thread per
// level-1 (I, J) column walks a range of K: each fine plane's 2x2 block of x
// is loaded once and kept for
// the planes above and below; only the in-plane halo comes from other
// threads' cells (cache hits). VEC (even nx): the two x-adjacent children of
// a row move as one 8-byte (x), 16-byte (r) and 2-byte (code) access.
template <bool VEC>
__device__ __forceinline__ void mgf_restrict_z_body(Blk bk, GeomT<float> g,
                                                    const uint8_t *__restrict__ code,
                                                    const MgStore *__restrict__ src,
                                                    const MgStore *__restrict__ x, int cnx, int cny,
                                                    int cnz, int kchunk, CoarseOut out,
                                                    float sscale) {
  const int pc = bk.x * kMt + (int)threadIdx.x;  // coarse column in the plane
  const int J = pc / cnx, I = pc - J * cnx;
  if (J >= cny) return;
  const int K0 = bk.z * kchunk, K1 = min(cnz, K0 + kchunk);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const size_t sx = (size_t)nx;
  const int i0 = 2 * I, j0 = 2 * J;
  const bool va[2] = {true, i0 + 1 < nx}, vb[2] = {true, j0 + 1 < ny};
  const int ia[2] = {i0, va[1] ? i0 + 1 : i0}, jb[2] = {j0, vb[1] ? j0 + 1 : j0};
  float cw[2], ce[2], dxa[2], cs[2], cn[2], dyb[2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    cw[a] = g.cw[ia[a]];
    ce[a] = g.ce[ia[a]];
    dxa[a] = g.dx[ia[a]];
    cs[a] = g.cs[jb[a]];
    cn[a] = g.cn[jb[a]];
    dyb[a] = g.dy[jb[a]];
  }
  const bool hw = i0 > 0, he = i0 + 2 < nx, hs = j0 > 0, hn = j0 + 2 < ny;
  auto at = [&](int i, int j, int k) { return (size_t)i + sx * ((size_t)j + (size_t)ny * k); };
  auto load = [&](int k, float (&v)[2][2]) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const bool ok = k >= 0 && k < nz && vb[b];
      if constexpr (VEC) {
        const Vec<MgStore, 2> t =
            ok ? ldv<MgStore, 2>(x, at(i0, jb[b], k)) : Vec<MgStore, 2>{};
        v[b][0] = (float)t.v[0];
        v[b][1] = (float)t.v[1];
      } else {
#pragma unroll
        for (int a = 0; a < 2; ++a) v[b][a] = ok && va[a] ? (float)x[at(ia[a], jb[b], k)] : 0.f;
      }
    }
  };
  auto plane = [&](int k, const float (&dn)[2][2], const float (&c)[2][2],
                   const float (&up)[2][2]) {
    float s = 0.f;
    if (k >= nz) return s;
    const float cb = g.cb[k], ct = g.ct[k], dzk = g.dz[k];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      if (!vb[b]) continue;
      const size_t c0 = at(i0, jb[b], k);
      unsigned cds[2];
      float ss[2];
      if constexpr (VEC) {
        const uchar2 cd2 = *reinterpret_cast<const uchar2 *>(code + c0);
        const Vec<MgStore, 2> s2 = ldv<MgStore, 2>(src, c0);
        cds[0] = cd2.x;
        cds[1] = cd2.y;
        ss[0] = (float)s2.v[0];
        ss[1] = (float)s2.v[1];
      } else {
        cds[0] = code[c0];
        ss[0] = (float)src[c0];
        cds[1] = va[1] ? code[c0 + 1] : 0u;
        ss[1] = va[1] ? (float)src[c0 + 1] : 0.f;
      }
      const float xw = hw ? (float)x[c0 - 1] : 0.f, xe = he ? (float)x[c0 + 2] : 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const unsigned cd = cds[a];
        if (!(cd & kCodeFluid)) continue;
        const size_t cc = c0 + a;
        const float west = a == 1 ? c[b][0] : xw;
        const float east = a == 0 ? c[b][1] : xe;
        const float south = b == 1 ? c[0][a] : (hs ? (float)x[cc - sx] : 0.f);
        const float north = b == 0 ? c[1][a] : (hn ? (float)x[cc + sx] : 0.f);
        float nb = (cd & kCodeW) ? cw[a] * west : 0.f;
        nb += (cd & kCodeE) ? ce[a] * east : 0.f;
        nb += (cd & kCodeS) ? cs[b] * south : 0.f;
        nb += (cd & kCodeN) ? cn[b] * north : 0.f;
        nb += (cd & kCodeB) ? cb * dn[b][a] : 0.f;
        nb += (cd & kCodeT) ? ct * up[b][a] : 0.f;
        const float d = diag_code<float>(cd, cw[a], ce[a], cs[b], cn[b], cb, ct);
        s += dxa[a] * dyb[b] * dzk * (ss[a] * d * sscale - (d * c[b][a] - nb));
      }
    }
    return s;
  };
  float below[2][2], c0[2][2], c1[2][2], above[2][2];
  load(2 * K0 - 1, below);
  load(2 * K0, c0);
  load(2 * K0 + 1, c1);
  for (int K = K0; K < K1; ++K) {
    load(2 * K + 2, above);
    const float s = plane(2 * K, below, c0, c1) + plane(2 * K + 1, c0, c1, above);
    out.store((size_t)I + (size_t)cnx * ((size_t)J + (size_t)cny * K), s);
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        below[b][a] = c1[b][a];
        c0[b][a] = above[b][a];
      }
    load(2 * K + 3, c1);
  }
}

// Coarse right-hand side in the streaming scheme of the level-0 sweeps: the
// level-0 residual V (s - (d x - sum c x_nb)) of each fluid cell, summed over
// the 2x2x2 children of each level-1 cell. The workgroups walk the tiles of
// tlr (mg_plan: an even number of rows, z chunks of an even number of planes,
// so every child block lies in one workgroup); each lane forms its cells'
// residuals plane by plane with the loads of the next plane in flight, adds
// x-adjacent pairs (in its registers, or from the next lane when it owns one
// cell), then the two planes of a pair, and parks the sum in LDS; after the
// chunk (at most kMgRestrictPlanes planes) one barrier, then the even rows
// add the odd rows' sums and store. The sums are taken in that fixed order
// whatever the tile shape and cells per lane, so standalone and batched
// solves get the same bits. lds: mg_restrict_lds(t.ty, CL) bytes. (A
// barrier per plane pair would wait for the next plane's loads in flight.)
template <int CL, bool NT>
__device__ __forceinline__ void mg_restrict_body(int block, float *lds, const Tiling &t,
                                                 const GeomT<float> &g,
                                                 const MgStore *__restrict__ a,
                                                 const MgStore *__restrict__ r,
                                                 const uint8_t *__restrict__ code, int cnx,
                                                 int cny, CoarseOut out, float sscale) {
  constexpr int RC = CL >= 2 ? CL / 2 : 1;  // coarse cells per lane
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const TileCoord tc = tile_coord(t, nx, ny, nz, block);
  if (!tc.valid) return;  // the whole workgroup
  using V = Vec<float, CL>;
  using H = Vec<MgStore, CL>;  // loads stay FP16 until used (in flight meanwhile)
  const int lane = (int)threadIdx.x, row = (int)threadIdx.y;
  const int i = (tc.i - lane) * CL + CL * lane, j = tc.j;
  const bool act = i < nx && j < ny;
  // Lanes past the grid load from a clamped cell (in bounds, unused) and add
  // nothing, so every wave of the workgroup reaches every barrier.
  const int ic = i < nx ? i : nx - CL, jc = j < ny ? j : ny - 1;
  const size_t nxy = (size_t)nx * ny;
  float cw[CL], ce[CL], dxv[CL], dxy[CL], cwn[CL], cen[CL];
  const float cs = g.cs[jc], cn = g.cn[jc], dyj = g.dy[jc];
  const bool hw = ic > 0, he = ic + CL < nx, hs = jc > 0, hn = jc < ny - 1;
  const float csn = hs ? cs : 0.f, cnn = hn ? cn : 0.f;
  unsigned exy[CL];
  wall_codes<CL>(exy, hw, he, hs, hn);
#pragma unroll
  for (int m = 0; m < CL; ++m) {
    cw[m] = g.cw[ic + m];
    ce[m] = g.ce[ic + m];
    dxv[m] = g.dx[ic + m];
    // neighbour terms past the domain walls drop out with their coefficient
    cwn[m] = m == 0 && !hw ? 0.f : cw[m];
    cen[m] = m == CL - 1 && !he ? 0.f : ce[m];
    dxy[m] = ((cwn[m] + cen[m]) + csn) + cnn;
  }
  const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
  // only the edge lanes' west / east loads leave their own cells (a clamped
  // lane's row start could lie before the array)
  const ptrdiff_t ow = lane == 0 ? (hw ? -1 : 0) : 0;
  const ptrdiff_t oe = lane == kTileX - 1 ? (he ? CL : CL - 1) : 0;
  const size_t base = (size_t)ic + (size_t)nx * (size_t)jc;
  const int k0 = tc.k0, k1 = tc.k1, last = min(k1, nz - 1);
  auto at = [&](int k) { return base + nxy * (size_t)k; };
  struct Set {
    H a_next, r;
    Plane<kDirCurrent, float, CL, MgStore, MgStore> pl;
  };
  Set sa, sb;
  V a_cur = convert<float>(ldv<MgStore, CL>(a, at(k0)));
  V a_prev = k0 > 0 ? convert<float>(ldv<MgStore, CL>(a, at(k0 - 1))) : a_cur;
  sa.a_next = ldv<MgStore, CL>(a, at(min(k0 + 1, last)));
  sa.pl.load(a, a, code, at(k0), os, on, ow, oe);
  sa.r = ldv_s<NT, MgStore, CL>(r, at(k0));
  float acc[RC];
  const int stride = kTileX * RC, ty = (int)blockDim.y;
  auto stage = [&](int k, const Set &u, Set &ld) {
    const int kn = min(k + 1, k1 - 1);
    ld.a_next = ldv<MgStore, CL>(a, at(min(k + 2, last)));
    ld.pl.load(a, a, code, at(kn), os, on, ow, oe);
    ld.r = ldv_s<NT, MgStore, CL>(r, at(kn));
    const float from_w = __shfl_up(a_cur.v[CL - 1], 1, kTileX);
    const float from_e = __shfl_down(a_cur.v[0], 1, kTileX);
    const float vw = lane == 0 ? (float)u.pl.w.a : from_w;
    const float ve = lane == kTileX - 1 ? (float)u.pl.e.a : from_e;
    const float cb = g.cb[k], ct = g.ct[k], dzk = g.dz[k];
    const float cbk = k > 0 ? cb : 0.f, ctk = k < nz - 1 ? ct : 0.f;
    float d[CL], con[CL];
    level0_diag<CL>(d, u.pl.code, exy, wall_code_z(k, nz), cw, ce, dxy, cs, cn, cb, ct,
                    cbk + ctk);
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      // the iterate is 0 outside the unknowns, so every neighbour term enters
      const float west = m == 0 ? vw : a_cur.v[m - 1];
      const float east = m == CL - 1 ? ve : a_cur.v[m + 1];
      const float nb = cwn[m] * west + cen[m] * east + csn * (float)u.pl.s.a.v[m] +
                       cnn * (float)u.pl.n.a.v[m] + cbk * a_prev.v[m] +
                       ctk * (float)u.a_next.v[m];
      const float src = (float)u.r.v[m] * d[m] * sscale;
      const bool fl = act && (u.pl.code.b[m] & kCodeFluid);
      con[m] = fl ? dxv[m] * dyj * dzk * (src - (d[m] * a_cur.v[m] - nb)) : 0.f;
    }
    float px[RC];
    if constexpr (CL >= 2) {
#pragma unroll
      for (int q = 0; q < RC; ++q) px[q] = con[2 * q] + con[2 * q + 1];
    } else {
      // the odd lane's cell joins the even lane's (whose sum is the one kept)
      px[0] = con[0] + __shfl_down(con[0], 1, kTileX);
    }
#pragma unroll
    for (int q = 0; q < RC; ++q) acc[q] = (k & 1) ? acc[q] + px[q] : px[q];
    if ((k & 1) || k == nz - 1) {
      float *slot = lds + (size_t)(((k - k0) >> 1) * ty + row) * stride + lane * RC;
#pragma unroll
      for (int q = 0; q < RC; ++q) slot[q] = acc[q];
    }
    a_prev = a_cur;
    a_cur = convert<float>(u.a_next);
  };
  for (int k = k0; k < k1; k += 2) {
    stage(k, sa, sb);
    if (k + 1 < k1) stage(k + 1, sb, sa);
  }
  __syncthreads();
  if (!(row & 1) && act && (CL >= 2 || !(lane & 1))) {
    const int J = j >> 1;
    for (int kc = 0; kc < (k1 - k0 + 1) >> 1; ++kc) {
      const float *e = lds + (size_t)(kc * ty + row) * stride + lane * RC;
      const int K = (k0 >> 1) + kc;
#pragma unroll
      for (int q = 0; q < RC; ++q) {
        const int I = (i >> 1) + q;
        if (I < cnx && J < cny)
          out.store((size_t)I + (size_t)cnx * ((size_t)J + (size_t)cny * K), e[q] + e[stride + q]);
      }
    }
  }
}

// Planes per z chunk of the streaming restriction (its LDS holds the sums of
// half as many coarse planes) and its LDS bytes for tiles of ty rows.
constexpr int kMgRestrictPlanes = 16;
__host__ __device__ inline size_t mg_restrict_lds(int ty, int cl) {
  return sizeof(float) * (size_t)(kMgRestrictPlanes / 2) * ty * kTileX * (cl >= 2 ? cl / 2 : 1);
}

// Prolongation onto level 0 with one thread per level-1 (I, J) column walking
// a range of K: fine plane 2K + c takes 3/4 of coarse plane K and 1/4 of
// plane K - 1 (c = 0) or K + 1 (c = 1). Each fine cell takes the trilinear
// weights of prolong_value (only fine cells that carry unknowns change; the
// sum is formed in FP32 and stored in FP16). The in-plane part of the
// interpolation (for the four fine columns of the coarse column: the
// weighted values and unknown masks of the 3x3 coarse patch) depends on the
// coarse plane alone, and each coarse plane serves four fine planes (two as
// the 3/4 plane, two as the 1/4 plane): it is formed once per coarse plane,
// when the patch is first used, so the patch loads of the next plane are in
// flight while a fine plane is computed. VEC (even nx): each fine row pair
// moves as one 4-byte x and 2-byte code access.
template <bool VEC>
__device__ __forceinline__ void mgf_prolong_z_body(Blk bk, int nx, int ny, int nz,
                                                   const uint8_t *__restrict__ code,
                                                   MgStore *__restrict__ x, int cnx, int cny, int cnz,
                                                   const MgStore *__restrict__ xc,
                                                   const MgStore *__restrict__ mc, int kchunk) {
  const int pc = bk.x * kMt + (int)threadIdx.x;  // coarse column in the plane
  const int J = pc / cnx, I = pc - J * cnx;
  if (J >= cny) return;
  const int K0 = bk.z * kchunk, K1 = min(cnz, K0 + kchunk);
  const bool vi[3] = {I > 0, true, I + 1 < cnx}, vj[3] = {J > 0, true, J + 1 < cny};
  const int ii[3] = {vi[0] ? I - 1 : I, I, vi[2] ? I + 1 : I};
  const int jj[3] = {vj[0] ? J - 1 : J, J, vj[2] ? J + 1 : J};
  const size_t fsx = (size_t)nx, cplane = (size_t)cnx * cny;
  const bool has_i1 = 2 * I + 1 < nx;
  const int jo[3] = {jj[0] * cnx, jj[1] * cnx, jj[2] * cnx};
  // a coarse plane's raw 3x3 patch; every load in bounds (indices clamped
  // into the level), the values of absent neighbours replaced by 0 when the
  // patch is used: no branch around a load
  struct Patch {
    MgStore v[3][3], m[3][3];
    bool kok;
  };
  auto cload = [&](int K, Patch &pt) {
    pt.kok = K >= 0 && K < cnz;
    const size_t kp = (size_t)(K < 0 ? 0 : K < cnz ? K : cnz - 1) * cplane;
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        pt.v[b][a] = xc[kp + jo[b] + ii[a]];
        pt.m[b][a] = mc[kp + jo[b] + ii[a]];
      }
  };
  // the in-plane interpolation of a patch for fine column (a, b): the sums
  // w00 x[1][1] + w10 x[1][ai] + w01 x[bj][1] + w11 x[bj][ai] of values and
  // of masks, ai = 0 or 2, bj = 0 or 2 by the side of a and b
  struct Interp {
    float num[2][2], den[2][2];
  };
  auto interp = [&](const Patch &pt, Interp &ip) {
    float v[3][3], m[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const bool ok = pt.kok && vi[a] && vj[b];
        v[b][a] = ok ? (float)pt.v[b][a] : 0.f;
        m[b][a] = ok ? (float)pt.m[b][a] : 0.f;
      }
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int bj = b ? 2 : 0;
      const float wj = vj[bj] ? 0.25f : 0.f;
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int ai = a ? 2 : 0;
        const float wi = vi[ai] ? 0.25f : 0.f;
        const float w00 = 0.75f * 0.75f, w10 = wi * 0.75f, w01 = 0.75f * wj, w11 = wi * wj;
        ip.num[b][a] = w00 * v[1][1] + w10 * v[1][ai] + w01 * v[bj][1] + w11 * v[bj][ai];
        ip.den[b][a] = w00 * m[1][1] + w10 * m[1][ai] + w01 * m[bj][1] + w11 * m[bj][ai];
      }
    }
  };
  // The fine rows of plane k (x and code of the two fine rows of the
  // column), loaded one fine plane ahead of their use (rows past the grid
  // are loaded from its last row and plane and never stored).
  struct Fine {
    Vec<MgStore, 2> x[2];
    uchar2 cd[2];
  };
  auto fload = [&](int k, Fine &f) {
    const int kc = k < nz ? k : nz - 1;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int j = 2 * J + b;
      const size_t c0 = (size_t)(2 * I) + fsx * ((size_t)(j < ny ? j : ny - 1) + (size_t)ny * kc);
      if constexpr (VEC) {
        f.x[b] = ldv<MgStore, 2>(x, c0);
        f.cd[b] = *reinterpret_cast<const uchar2 *>(code + c0);
      } else {
        f.x[b].v[0] = x[c0];
        f.cd[b].x = code[c0];
        f.x[b].v[1] = x[has_i1 ? c0 + 1 : c0];
        f.cd[b].y = has_i1 ? code[c0 + 1] : 0;
      }
    }
  };
  // fine plane k: 3/4 of ic (its own coarse plane) and wk of iz (the z
  // neighbour's, 0 past the level)
  auto fine_plane = [&](int k, float wk, const Interp &ic, const Interp &iz, const Fine &f) {
    if (k >= nz) return;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      float xv[2] = {(float)f.x[b].v[0], (float)f.x[b].v[1]};
      const unsigned cd[2] = {f.cd[b].x, f.cd[b].y};
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const float num = 0.75f * ic.num[b][a] + wk * iz.num[b][a];
        const float den = 0.75f * ic.den[b][a] + wk * iz.den[b][a];
        const float add = den > 0.f ? num * __builtin_amdgcn_rcpf(den) : 0.f;
        xv[a] = (cd[a] & kCodeFluid) ? xv[a] + add : xv[a];
      }
      const int j = 2 * J + b;
      if (j >= ny) continue;
      const size_t c0 = (size_t)(2 * I) + fsx * ((size_t)j + (size_t)ny * k);
      if constexpr (VEC) {
        Vec<MgStore, 2> t;
        t.v[0] = (MgStore)xv[0];
        t.v[1] = (MgStore)xv[1];
        stv<MgStore, 2>(x, c0, t);
      } else {
        x[c0] = (MgStore)xv[0];
        if (has_i1) x[c0 + 1] = (MgStore)xv[1];
      }
    }
  };
  // ibelow, icur: the interpolations of planes K - 1 and K; the patch of
  // plane K + 1 loads while fine plane 2K is computed
  Patch pt;
  Interp ibelow, icur, iabove;
  cload(K0 - 1, pt);
  interp(pt, ibelow);
  cload(K0, pt);
  interp(pt, icur);
  Fine fa, fb;
  fload(2 * K0, fa);
  for (int K = K0; K < K1; ++K) {
    cload(K + 1, pt);
    fload(2 * K + 1, fb);
    fine_plane(2 * K, K > 0 ? 0.25f : 0.f, icur, ibelow, fa);
    interp(pt, iabove);
    if (K + 1 < K1) fload(2 * K + 2, fa);
    fine_plane(2 * K + 1, K + 1 < cnz ? 0.25f : 0.f, icur, iabove, fb);
    ibelow = icur;
    icur = iabove;
  }
}

// ---- coarse levels: per-cell conductances ------------------------------------
struct Coarse {
  float d, nb;
};

// x: the level's FP16 iterate, or an FP32 vector on the level (the learned
// correction's).
template <typename TX>
__device__ __forceinline__ Coarse coarse_row(int nx, int ny, int nz, const MgStore *__restrict__ gE,
                                             const MgStore *__restrict__ gN,
                                             const MgStore *__restrict__ gT,
                                             const TX *__restrict__ x, size_t c, int i,
                                             int j, int k) {
  const size_t sx = (size_t)nx, sxy = (size_t)nx * ny;
  const size_t cw = i > 0 ? c - 1 : c, cs = j > 0 ? c - sx : c, cb = k > 0 ? c - sxy : c;
  const size_t ce = i < nx - 1 ? c + 1 : c, cn = j < ny - 1 ? c + sx : c,
               ct = k < nz - 1 ? c + sxy : c;
  const float ge = gE[c], gn = gN[c], gt = gT[c];
  const float gw = i > 0 ? gE[cw] : 0.f, gs = j > 0 ? gN[cs] : 0.f, gb = k > 0 ? gT[cb] : 0.f;
  Coarse o;
  o.d = gw + ge + gs + gn + gb + gt;
  o.nb = gw * (float)x[cw] + ge * (float)x[ce] + gs * (float)x[cs] + gn * (float)x[cn] +
         gb * (float)x[cb] + gt * (float)x[ct];
  return o;
}

// One smoothing step of a coarse level: out = x + c2 (b - A x) / d, plus
// c1 (x - prev) for a Chebyshev step after the first (prev null: a zero
// previous iterate); damped Jacobi is c2 = omega, c1 = 0.
__device__ __forceinline__ void mgc_sweep_body(Blk bk, int nx, int ny, int nz,
                                               const MgStore *__restrict__ gE,
                                               const MgStore *__restrict__ gN,
                                               const MgStore *__restrict__ gT,
                                               const MgStore *__restrict__ b,
                                               const MgStore *__restrict__ in,
                                               const MgStore *__restrict__ prev,
                                               MgStore *__restrict__ out, float c1, float c2) {
  const Idx t = cell_idx(nx, ny, bk);
  if (!t.in) return;
  const Coarse o = coarse_row(nx, ny, nz, gE, gN, gT, in, t.c, t.i, t.j, t.k);
  const float xv = in[t.c];
  float v = o.d > 0.f ? xv + c2 * (b[t.c] + o.nb - o.d * xv) / o.d : 0.f;
  if (c1 != 0.f && o.d > 0.f) v += c1 * (xv - (prev ? prev[t.c] : 0.f));
  out[t.c] = v;
}

// The coarse-level smoothing step of mgc_sweep_body in the streaming scheme
// of the level-0 kernels, for the levels large enough to fill the chip
// (MgLevel::tl): a lane owns CL x-adjacent cells of a row (16-byte loads for
// CL = 4) and walks its z range with the planes below and above in registers
// (x, and gT of the plane below); the next plane's loads are issued before
// the current one is computed; west/east values by wave shuffle. The
// neighbour values, clamped at the walls, and the order of every sum are
// those of coarse_row and mgc_sweep_body, so both give the same bits.
template <int CL>
__device__ __forceinline__ void mgc_sweep_t_body(int block, const Tiling &t, int nx, int ny,
                                                 int nz, const MgStore *__restrict__ gE,
                                                 const MgStore *__restrict__ gN,
                                                 const MgStore *__restrict__ gT,
                                                 const MgStore *__restrict__ b,
                                                 const MgStore *__restrict__ in,
                                                 const MgStore *__restrict__ prev,
                                                 MgStore *__restrict__ out, float c1, float c2) {
  const Lane l = lane_of<CL>(t, nx, ny, nz, block);
  if (!l.active) return;
  using V = Vec<float, CL>;
  using H = Vec<MgStore, CL>;  // loads stay FP16 until used (in flight meanwhile)
  const int i = l.i, j = l.j, lane = (int)threadIdx.x;
  const size_t nxy = (size_t)nx * ny;
  const bool hw = i > 0, he = i + CL < nx, hs = j > 0, hn = j < ny - 1;
  const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
  const ptrdiff_t ow = lane == 0 && hw ? -1 : 0;
  const ptrdiff_t oe = lane == kTileX - 1 ? (he ? CL : CL - 1) : 0;
  const size_t base = (size_t)i + (size_t)nx * (size_t)j;
  const int k0 = l.k0, k1 = l.k1, last = min(k1, nz - 1);
  auto at = [&](int k) { return base + nxy * (size_t)k; };
  struct Set {
    H x_next, xs, xn, bb, pp, ge, gn, gt, gs;
    MgStore xw, xe, gw;
  };
  auto load = [&](Set &s, int k, int knext) {
    const size_t c = at(k);
    s.x_next = ldv<MgStore, CL>(in, at(knext));
    s.xs = ldv<MgStore, CL>(in, c + os);
    s.xn = ldv<MgStore, CL>(in, c + on);
    s.xw = in[c + ow];
    s.xe = in[c + oe];
    s.bb = ldv<MgStore, CL>(b, c);
    if (prev) s.pp = ldv<MgStore, CL>(prev, c);
    s.ge = ldv<MgStore, CL>(gE, c);
    s.gn = ldv<MgStore, CL>(gN, c);
    s.gt = ldv<MgStore, CL>(gT, c);
    s.gs = ldv<MgStore, CL>(gN, c + os);
    s.gw = gE[c + ow];
  };
  Set sa, sb;
  V x_cur = convert<float>(ldv<MgStore, CL>(in, at(k0)));
  V x_prev = k0 > 0 ? convert<float>(ldv<MgStore, CL>(in, at(k0 - 1))) : x_cur;
  V gt_prev = k0 > 0 ? convert<float>(ldv<MgStore, CL>(gT, at(k0 - 1))) : x_cur;
  load(sa, k0, min(k0 + 1, last));
  auto stage = [&](int k, const Set &u, Set &ld) {
    load(ld, min(k + 1, k1 - 1), min(k + 2, last));
    const float from_w = __shfl_up(x_cur.v[CL - 1], 1, kTileX);
    const float from_e = __shfl_down(x_cur.v[0], 1, kTileX);
    const float gfrom_w = __shfl_up((float)u.ge.v[CL - 1], 1, kTileX);
    H o;
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      const float xv = x_cur.v[m];
      const float xw = m > 0 ? x_cur.v[m - 1] : hw ? (lane == 0 ? (float)u.xw : from_w) : xv;
      const float xe = m < CL - 1 ? x_cur.v[m + 1]
                                  : he ? (lane == kTileX - 1 ? (float)u.xe : from_e) : xv;
      const float xs = hs ? (float)u.xs.v[m] : xv, xn = hn ? (float)u.xn.v[m] : xv;
      const float xb = k > 0 ? x_prev.v[m] : xv, xt = k < nz - 1 ? (float)u.x_next.v[m] : xv;
      const float ge = u.ge.v[m], gn = u.gn.v[m], gt = u.gt.v[m];
      const float gw = m > 0 ? (float)u.ge.v[m - 1] : hw ? (lane == 0 ? (float)u.gw : gfrom_w)
                                                         : 0.f;
      const float gs = hs ? (float)u.gs.v[m] : 0.f, gb = k > 0 ? gt_prev.v[m] : 0.f;
      const float d = gw + ge + gs + gn + gb + gt;
      const float nb = gw * xw + ge * xe + gs * xs + gn * xn + gb * xb + gt * xt;
      float v = d > 0.f ? xv + c2 * ((float)u.bb.v[m] + nb - d * xv) / d : 0.f;
      if (c1 != 0.f && d > 0.f) v += c1 * (xv - (prev ? (float)u.pp.v[m] : 0.f));
      o.v[m] = (MgStore)v;
    }
    stv<MgStore, CL>(out, at(k), o);
    x_prev = x_cur;
    x_cur = convert<float>(u.x_next);
    gt_prev = convert<float>(u.gt);
  };
  for (int k = k0; k < k1; k += 2) {
    stage(k, sa, sb);
    if (k + 1 < k1) stage(k + 1, sb, sa);
  }
}

__device__ __forceinline__ void mgc_restrict_body(Blk bk, int nx, int ny, int nz,
                                                  const MgStore *__restrict__ gE,
                                                  const MgStore *__restrict__ gN,
                                                  const MgStore *__restrict__ gT,
                                                  const MgStore *__restrict__ b,
                                                  const MgStore *__restrict__ x, int cnx, int cny,
                                                  CoarseOut out) {
  const Idx t = cell_idx(cnx, cny, bk);
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
  out.store(t.c, s);
}

// Trilinear cell-centred prolongation from the coarse level (cnx, cny, cnz;
// values xc, unknown mask mc): fine cell 2I+a takes 3/4 of coarse cell I and
// 1/4 of its neighbour on the side of a, per axis; weights are renormalised
// over coarse cells with mc = 1. Only fine cells that carry unknowns change.
__device__ __forceinline__ float prolong_value(int i, int j, int k, int cnx, int cny, int cnz,
                                               const MgStore *__restrict__ xc,
                                               const MgStore *__restrict__ mc) {
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

__device__ __forceinline__ void mgc_prolong_body(Blk bk, int nx, int ny,
                                                 const MgStore *__restrict__ m, MgStore *__restrict__ x,
                                                 int cnx, int cny, int cnz,
                                                 const MgStore *__restrict__ xc,
                                                 const MgStore *__restrict__ mc) {
  const Idx t = cell_idx(nx, ny, bk);
  if (!t.in) return;
  if (!(m[t.c] > 0.f)) return;
  x[t.c] += prolong_value(t.i, t.j, t.k, cnx, cny, cnz, xc, mc);
}

// Coarsest level (n <= 1024 cells, one workgroup): `sweeps` damped-Jacobi
// sweeps from zero with the iterate in LDS; the first one is rd b, as the
// restriction into a level forms it.
__device__ __forceinline__ void mgc_coarsest_body(float *xs, int nx, int ny, int nz,
                                                  const MgStore *__restrict__ gE,
                                                  const MgStore *__restrict__ gN,
                                                  const MgStore *__restrict__ gT,
                                                  const MgStore *__restrict__ b,
                                                  const MgStore *__restrict__ rd,
                                                  MgStore *__restrict__ x, float omega, int sweeps) {
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
  float xv = on ? rd[t] * bb : 0.f;
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

// Streaming damped-Jacobi sweep of level 0 in FP32 arithmetic on FP16
// vectors (the structure of the reference-Jacobi kernel k_jacobi_t: each lane
// owns CL x-adjacent cells of one row and walks its z range; plane k+1's
// loads are issued before plane k is computed and stay FP16 until used;
// west/east neighbours by wave shuffle):
//   out = (1 - w) a + (w / d) (s + sum c a_nb), 0 where d = 0, s = r / V
// (read from the first step, see below). Every vector carries the cycle
// scale sig.
// LAST: out is the cycle's output z (times sig), and the partials of the
// stored (rounded) values go to three slots: z.r = z.(s V) and z.q (q = A
// p_old from the previous update, FP16 times sc->tau, or z.q = 0 when q is
// null), and max |z|.
//
// CHEB: a Chebyshev step, out = a + c2 (s - A a) / d + c1 (a - prev) (prev
// null: a zero previous iterate); otherwise damped Jacobi with omega = c2.
// The source is read from the cycle's first step x1 = first D^-1 r (r holds
// x1): s = x1 d sscale, sscale = 1 / first. When a is x1 itself (the step
// after the first) the one load serves both.
template <int CL, bool LAST, bool NT, bool CHEB = false>
__device__ __forceinline__ void mg_sweep_body(int block, int nblocks, double *lds, Tiling t,
                                              GeomT<float> g, const MgStore *__restrict__ a,
                                              const MgStore *__restrict__ r,
                                              const uint8_t *__restrict__ code,
                                              MgStore *__restrict__ b, double *partials,
                                              float omega, const MgStore *__restrict__ q,
                                              const MgStore *__restrict__ prev, float c1,
                                              float sscale) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz, block);
  double s = 0.0, sq = 0.0, zmx = 0.0;
  if (l.active) {
    using V = Vec<float, CL>;
    using H = Vec<MgStore, CL>;
    const int i = l.i, j = l.j;
    const size_t nxy = (size_t)nx * ny;
    float cw[CL], ce[CL], dxdy[CL], dxy[CL];
    const float cs = g.cs[j], cn = g.cn[j];
    const float keep = 1.f - omega;
    const bool hw = i > 0, he = i + CL < nx, hs = j > 0, hn = j < ny - 1;
    // neighbour terms past the domain walls drop out with their coefficient
    // (the loads there are clamped to cells of the grid)
    float cwn[CL], cen[CL];
    const float csn = hs ? cs : 0.f, cnn = hn ? cn : 0.f;
    unsigned exy[CL];
    wall_codes<CL>(exy, hw, he, hs, hn);
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      cw[m] = g.cw[i + m];
      ce[m] = g.ce[i + m];
      dxdy[m] = g.dx[i + m] * g.dy[j];
      cwn[m] = m == 0 && !hw ? 0.f : cw[m];
      cen[m] = m == CL - 1 && !he ? 0.f : ce[m];
      dxy[m] = ((cwn[m] + cen[m]) + csn) + cnn;
    }
    const ptrdiff_t os = hs ? -(ptrdiff_t)nx : 0, on = hn ? (ptrdiff_t)nx : 0;
    const int lane = (int)threadIdx.x;
    const ptrdiff_t row0 = -(ptrdiff_t)(lane * CL);
    const ptrdiff_t ow = lane == 0 ? (hw ? -1 : 0) : row0;
    const ptrdiff_t oe = lane == kTileX - 1 ? (he ? CL : CL - 1) : row0;
    const size_t base = (size_t)i + (size_t)nx * (size_t)j;
    const int k0 = l.k0, k1 = l.k1, last = min(k1, nz - 1);
    auto at = [&](int k) { return base + nxy * (size_t)k; };
    struct Set {
      H a_next, r, p;
      Plane<kDirCurrent, float, CL, MgStore, MgStore> pl;
    };
    Set sa, sb;
    const bool same = r == a;  // uniform
    V a_cur = convert<float>(ldv<MgStore, CL>(a, at(k0)));
    V a_prev = k0 > 0 ? convert<float>(ldv<MgStore, CL>(a, at(k0 - 1))) : a_cur;
    sa.a_next = ldv<MgStore, CL>(a, at(min(k0 + 1, last)));
    sa.pl.load(a, a, code, at(k0), os, on, ow, oe);
    if (!same) sa.r = ldv_s<NT, MgStore, CL>(r, at(k0));
    if constexpr (CHEB) {
      if (prev) sa.p = ldv_s<NT, MgStore, CL>(prev, at(k0));
    }
    auto stage = [&](int k, const Set &u, Set &ld) {
      const int kn = min(k + 1, k1 - 1);
      ld.a_next = ldv<MgStore, CL>(a, at(min(k + 2, last)));
      ld.pl.load(a, a, code, at(kn), os, on, ow, oe);
      if (!same) ld.r = ldv_s<NT, MgStore, CL>(r, at(kn));
      if constexpr (CHEB) {
        if (prev) ld.p = ldv_s<NT, MgStore, CL>(prev, at(kn));
      }
      const float from_w = __shfl_up(a_cur.v[CL - 1], 1, kTileX);
      const float from_e = __shfl_down(a_cur.v[0], 1, kTileX);
      const float vw = lane == 0 ? (float)u.pl.w.a : from_w;
      const float ve = lane == kTileX - 1 ? (float)u.pl.e.a : from_e;
      const float cb = g.cb[k], ct = g.ct[k], dzk = g.dz[k];
      const float cbk = k > 0 ? cb : 0.f, ctk = k < nz - 1 ? ct : 0.f;
      H bv, qz;
      if constexpr (LAST) {
        if (q) qz = ldv_s<NT, MgStore, CL>(q, at(k));
      }
      float d[CL];
      level0_diag<CL>(d, u.pl.code, exy, wall_code_z(k, nz), cw, ce, dxy, cs, cn, cb, ct,
                      cbk + ctk);
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const float west = m == 0 ? vw : a_cur.v[m - 1];
        const float east = m == CL - 1 ? ve : a_cur.v[m + 1];
        // the source s = r d sscale enters divided by d: rs = r sscale
        const float rs = (same ? a_cur.v[m] : (float)u.r.v[m]) * sscale;
        const float nb = cwn[m] * west + cen[m] * east + csn * (float)u.pl.s.a.v[m] +
                         cnn * (float)u.pl.n.a.v[m] + cbk * a_prev.v[m] +
                         ctk * (float)u.a_next.v[m];
        const float rd = __builtin_amdgcn_rcpf(d[m]);
        const float t = nb * rd + rs;  // (s + sum c a_nb) / d
        float o;
        if constexpr (CHEB) {
          const float av = a_cur.v[m];
          o = av + omega * (t - av);
          if (c1 != 0.f) o += c1 * (av - (prev ? (float)u.p.v[m] : 0.f));
        } else {
          o = keep * a_cur.v[m] + omega * t;
        }
        o = d[m] > 0.f ? o : 0.f;
        bv.v[m] = (MgStore)o;
        if constexpr (LAST) {
          // the sums take z as stored
          const double zd = (double)(float)bv.v[m];
          const float src = rs * d[m];
          s += zd * ((double)src * (double)(dxdy[m] * dzk));
          if (q) sq += zd * (double)(float)qz.v[m];
          zmx = fmax(zmx, fabs(zd));
        }
      }
      stv_s<NT, MgStore, CL>(b, at(k), bv);
      a_prev = a_cur;
      a_cur = convert<float>(u.a_next);
    };
    for (int k = k0; k < k1; k += 2) {
      stage(k, sa, sb);
      if (k + 1 < k1) stage(k + 1, sb, sa);
    }
  }
  if constexpr (LAST) {
    double v[2] = {s, sq};
    block_sum<2>(v, lds);
    block_max(zmx, lds + 32);
    const double vals[3] = {v[0], v[1], zmx};
    store_partials(partials, 3, vals, nblocks, block);
  }
}

// x += alpha p, r -= alpha A p (the fused PCG update body, p in FP32), the
// first smoothing step of the next cycle, x1 = sig w D^-1 r (FP16 with the
// next cycle's scale sig = sc->sig_next, which becomes sc->sig; x1 also
// carries the cycle's source), and q = A p in FP16 times sc->tau; with rmax
// the per-workgroup max |r / V|; counts the iteration.
// XM (XM_NOW, XM_DEFER, XM_TWO) and ppend: the treatment of x (see
// pcg_upd_body); a deferring update keeps its alpha in sc->alpha_pend for the
// next one, which applies it with p_pend.
template <int CL, int XM = XM_NOW>
__device__ __forceinline__ void mg_upd_body(int block, int nblocks, double *lds, Tiling t,
                                            GeomT<double> g, const float *__restrict__ p,
                                            const uint8_t *__restrict__ code,
                                            double *__restrict__ x, double *__restrict__ r,
                                            MgStore *__restrict__ x1, Scalars *sc, double omega,
                                            double *rmax, MgStore *__restrict__ q16,
                                            const float *__restrict__ ppend = nullptr) {
  const bool broke = sc->status != 0;
  const double sig = sc->sig_next;  // no workgroup of this kernel reads sc->sig
  pcg_upd_body<double, CL, float, false, float, XM, MgStore, MgStore>(
      t, block, nblocks, g, p, code, x, r, x1, nullptr, sc, lds, omega * sig, rmax, q16,
      nullptr, ppend, sc->tau);
  if (block == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    // a deferring iteration that broke down leaves no step behind
    if constexpr (XM == XM_DEFER) sc->alpha_pend = broke ? 0.0 : sc->alpha;
    if (!broke) sc->iters += 1;
    sc->sig = sig;
  }
}

// ---- setup --------------------------------------------------------------------
// sc->anorm: a bound on the infinity norm of the level-0 operator A = V (diag
// - offdiag), whose rows sum |entries| to at most 2 V (cw + ce + cs + cn + cb
// + ct) <= 2 (max dx (cw + ce) max dy max dz + max dx max dy (cs + cn) max dz
// + max dx max dy max dz (cb + ct)). One workgroup; lds: 16 doubles.
__device__ __forceinline__ void mg_anorm_body(const GeomT<double> &g, Scalars *sc, double *lds) {
  double m[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
    m[0] = fmax(m[0], g.dx[i] * (g.cw[i] + g.ce[i]));
    m[1] = fmax(m[1], g.dx[i]);
  }
  for (int j = threadIdx.x; j < g.ny; j += blockDim.x) {
    m[2] = fmax(m[2], g.dy[j] * (g.cs[j] + g.cn[j]));
    m[3] = fmax(m[3], g.dy[j]);
  }
  for (int k = threadIdx.x; k < g.nz; k += blockDim.x) {
    m[4] = fmax(m[4], g.dz[k] * (g.cb[k] + g.ct[k]));
    m[5] = fmax(m[5], g.dz[k]);
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) block_max(m[q], lds);
  if (threadIdx.x == 0)
    sc->anorm = 2.0 * (m[0] * m[3] * m[5] + m[1] * m[2] * m[5] + m[1] * m[3] * m[4]);
}

// Level-1 conductances from level 0 (FP64 sums, stored FP16): half the sum of
// V c over the fluid-fluid level-0 faces on each aggregate face.
__device__ __forceinline__ void mg_coarsen0_body(Blk bk, const GeomT<double> &g,
                                                 const uint8_t *__restrict__ code, int cnx,
                                                 int cny, float scale, MgStore *__restrict__ gE,
                                                 MgStore *__restrict__ gN,
                                                 MgStore *__restrict__ gT) {
  const Idx t = cell_idx(cnx, cny, bk);
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

// Level l + 1 conductances from level l: the sum over the aggregate face.
__device__ __forceinline__ void mg_coarsen_body(Blk bk, int nx, int ny, int nz,
                                                const MgStore *__restrict__ fE,
                                                const MgStore *__restrict__ fN,
                                                const MgStore *__restrict__ fT, int cnx, int cny,
                                                float scale, MgStore *__restrict__ gE,
                                                MgStore *__restrict__ gN,
                                                MgStore *__restrict__ gT) {
  const Idx t = cell_idx(cnx, cny, bk);
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

// Unknown mask (diag > 0) and the first-step scale over diag of a level.
__device__ __forceinline__ void mg_mask_body(Blk bk, int nx, int ny,
                                             const MgStore *__restrict__ gE,
                                             const MgStore *__restrict__ gN,
                                             const MgStore *__restrict__ gT, float first,
                                             MgStore *__restrict__ m, MgStore *__restrict__ rd) {
  const Idx t = cell_idx(nx, ny, bk);
  if (!t.in) return;
  const float d = coarse_diag(nx, ny, gE, gN, gT, t.c, t.i, t.j, t.k);
  m[t.c] = d > 0.f ? 1.f : 0.f;
  rd[t.c] = d > 0.f ? first / d : 0.f;
}

}  // namespace mgk
}  // namespace ffpm
