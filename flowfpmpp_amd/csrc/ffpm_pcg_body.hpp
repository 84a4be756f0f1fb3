// Device bodies of the fused PCG iteration, templated on the value type T
// (double for the FP64 solve, float for the inner solve of mixed precision)
// and on CL, the x-adjacent cells one lane owns (16-byte loads: CL = 2 for
// double, 4 for float; CL = 1 when nx is not a multiple of the vector width).
// Used by the single-solve kernels and by the batched kernels.
//
// Iteration (src/fpm_stairstep_mpi_3d.f90:246-283, formulas unchanged):
//   pcg_dir_body : p = z + beta p_old (the previous iteration's direction
//                  update, deferred), q = A p, p.q (FP64 sum)
//   pcg_upd_body : q = A p recomputed bit-identically, x += alpha p,
//                  r -= alpha q, z = M^-1 r, z.r (FP64 sum)
// A = V (diag - offdiag) with diag and rdiag = 1/(V diag) rebuilt in
// registers from the one-byte neighbour code and the 1-D coefficients, so the
// kernels read no diagonal arrays. Each workgroup streams its tile along z;
// the loads of the next plane are issued before the current plane is
// computed, and every value moves in 16-byte vectors where the layout allows.
#pragma once

#include <type_traits>

#include "ffpm_pcg.hpp"

namespace ffpm {


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

// Element-wise conversion of a vector (exact where the values are already
// representable in U, as the rounded directions are).
template <typename U, typename T, int CL>
__device__ __forceinline__ Vec<U, CL> convert(const Vec<T, CL> &x) {
  if constexpr (std::is_same<U, T>::value) {
    return x;
  } else {
    Vec<U, CL> o;
#pragma unroll
    for (int m = 0; m < CL; ++m) o.v[m] = (U)x.v[m];
    return o;
  }
}

// Streaming accesses of values used once (the source, the outputs): the
// nontemporal form marks the cache lines for early eviction, which keeps
// the L2 for the neighbour rows the stencil re-reads. NT = false keeps the
// plain access (FFPM_MG_NT=0 for the multigrid kernels).
template <typename T, int CL>
struct NtVec;
template <> struct NtVec<float, 1> { typedef float type; };
template <> struct NtVec<float, 2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<float, 4> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct NtVec<double, 1> { typedef double type; };
template <> struct NtVec<double, 2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<double, 4> { typedef double type __attribute__((ext_vector_type(4))); };
template <> struct NtVec<_Float16, 1> { typedef _Float16 type; };
template <> struct NtVec<_Float16, 2> { typedef _Float16 type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<_Float16, 4> { typedef _Float16 type __attribute__((ext_vector_type(4))); };
template <> struct NtVec<_Float16, 8> { typedef _Float16 type __attribute__((ext_vector_type(8))); };
template <> struct NtVec<float, 8> { typedef float type __attribute__((ext_vector_type(8))); };

template <bool NT, typename T, int CL>
__device__ __forceinline__ Vec<T, CL> ldv_s(const T *__restrict__ a, size_t c) {
  if constexpr (NT) {
    typedef typename NtVec<T, CL>::type V;
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(a + c));
    Vec<T, CL> o;
    if constexpr (CL == 1) {
      o.v[0] = v;
    } else {
#pragma unroll
      for (int m = 0; m < CL; ++m) o.v[m] = v[m];
    }
    return o;
  } else {
    return ldv<T, CL>(a, c);
  }
}

template <bool NT, typename T, int CL>
__device__ __forceinline__ void stv_s(T *__restrict__ a, size_t c, const Vec<T, CL> &x) {
  if constexpr (NT) {
    typedef typename NtVec<T, CL>::type V;
    V v;
    if constexpr (CL == 1) {
      v = x.v[0];
    } else {
#pragma unroll
      for (int m = 0; m < CL; ++m) v[m] = x.v[m];
    }
    __builtin_nontemporal_store(v, reinterpret_cast<V *>(a + c));
  } else {
    stv<T, CL>(a, c, x);
  }
}

// Raw loads behind one direction vector (z and p_old for kDirUpdate, z for
// kDirReset, p_old for kDirCurrent), combined only when used, so the loads of
// a whole plane can be issued one plane ahead. The mode is a template
// argument: with a run-time mode the compiler puts every load pair in its own
// branch and waits for it there, which serialises the loads of a plane. z may
// be stored in a narrower type ZT (the multigrid cycle's FP32 output), widened
// exactly on use (kDirCurrent, which loads p_old into a, needs ZT = T). The
// directions may be stored in a narrower type PT (the multigrid CG's FP32
// directions): p = z + beta p_old is formed in T and rounded to PT, and every
// use of p (its own value, its neighbours in A p) takes the rounded value,
// so the kernels that read the stored p later see exactly what was used here.
// z stored in FP16 is the multigrid cycle's output scaled by a power of two
// (Scalars::sig): value() multiplies it by zinv = 1 / sig (exact). In a
// kDirCurrent plane (a stencil operand a read alone) a and the array share
// one type, which may be narrower than T.
template <typename T, typename PT>
__device__ __forceinline__ T round_to(T v) {
  if constexpr (std::is_same<T, PT>::value) return v;
  else return (T)(PT)v;
}

template <int MODE, typename T, int CL, typename ZT = T, typename PT = T>
struct RawV {
  static_assert(MODE != kDirCurrent || std::is_same<ZT, PT>::value, "a is read from pin");
  Vec<ZT, CL> a;  // z (p_old for kDirCurrent)
  Vec<PT, CL> b;  // p_old for kDirUpdate
  __device__ __forceinline__ void load(const ZT *__restrict__ z, const PT *__restrict__ pin,
                                       size_t c) {
    if constexpr (MODE == kDirCurrent) {
      a = ldv<ZT, CL>(pin, c);
    } else if constexpr (MODE == kDirReset) {
      a = ldv<ZT, CL>(z, c);
    } else {
      a = ldv<ZT, CL>(z, c);
      b = ldv<PT, CL>(pin, c);
    }
  }
  __device__ __forceinline__ Vec<T, CL> value(T beta, T zinv) const {
    Vec<T, CL> v;
#pragma unroll
    for (int m = 0; m < CL; ++m) {
      T zv = (T)a.v[m];
      if constexpr (std::is_same<ZT, _Float16>::value && MODE != kDirCurrent) zv = zv * zinv;
      v.v[m] = round_to<T, PT>(MODE == kDirUpdate ? zv + beta * (T)b.v[m] : zv);
    }
    return v;
  }
};

template <int MODE, typename T, typename ZT = T, typename PT = T>
struct Raw1 {
  static_assert(MODE != kDirCurrent || std::is_same<ZT, PT>::value, "a is read from pin");
  ZT a;  // z (p_old for kDirCurrent)
  PT b;  // p_old for kDirUpdate
  __device__ __forceinline__ void load(const ZT *__restrict__ z, const PT *__restrict__ pin,
                                       size_t c) {
    if constexpr (MODE == kDirCurrent) {
      a = pin[c];
    } else if constexpr (MODE == kDirReset) {
      a = z[c];
    } else {
      a = z[c];
      b = pin[c];
    }
  }
  __device__ __forceinline__ T value(T beta, T zinv) const {
    T zv = (T)a;
    if constexpr (std::is_same<ZT, _Float16>::value && MODE != kDirCurrent) zv = zv * zinv;
    if constexpr (MODE != kDirUpdate) return round_to<T, PT>(zv);
    else return round_to<T, PT>(zv + beta * (T)b);
  }
};

// The in-plane neighbours of a lane's cells (south and north rows, the west
// and east cells) and their neighbour codes, for one plane.
template <int MODE, typename T, int CL, typename ZT = T, typename PT = T>
struct Plane {
  RawV<MODE, T, CL, ZT, PT> s, n;
  Raw1<MODE, T, ZT, PT> w, e;
  Codes<CL> code;
  __device__ __forceinline__ void load(const ZT *__restrict__ z, const PT *__restrict__ pin,
                                       const uint8_t *__restrict__ cd, size_t c, ptrdiff_t os,
                                       ptrdiff_t on, ptrdiff_t ow, ptrdiff_t oe) {
    s.load(z, pin, c + os);
    n.load(z, pin, c + on);
    w.load(z, pin, c + ow);
    e.load(z, pin, c + oe);
    code = *reinterpret_cast<const Codes<CL> *>(cd + c);
  }
};

// diag from the neighbour code: the fluid-neighbour coefficients summed in
// build_operator's order (src/fpm_stairstep_mpi_3d.f90:353-370), so for
// T = double the value is bit-identical to the stored diag.
template <typename T>
__device__ __forceinline__ T diag_code(unsigned code, T cw, T ce, T cs, T cn, T cb, T ct) {
  T d = 0;
  if (code & kCodeW) d = d + cw;
  if (code & kCodeE) d = d + ce;
  if (code & kCodeS) d = d + cs;
  if (code & kCodeN) d = d + cn;
  if (code & kCodeB) d = d + cb;
  if (code & kCodeT) d = d + ct;
  return d;
}

// The same from the coefficient arrays (for kernels that hold no copies).
template <typename T>
__device__ __forceinline__ T diag_code(unsigned code, const GeomT<T> &g, int i, int j, int k) {
  return diag_code<T>(code, g.cw[i], g.ce[i], g.cs[j], g.cn[j], g.cb[k], g.ct[k]);
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
__device__ __forceinline__ Lane lane_of(const Tiling &t, int nx, int ny, int nz, int block) {
  const TileCoord tc = tile_coord(t, nx, ny, nz, block);
  Lane l;
  l.i = (tc.i - (int)threadIdx.x) * CL + CL * (int)threadIdx.x;
  l.j = tc.j;
  l.k0 = tc.k0;
  l.k1 = tc.k1;
  l.active = tc.valid && l.i < nx && l.j < ny;
  return l;
}


// p = z + beta p_old (or z, or p_old, by MODE), q = A p and p.q for the
// workgroup `block` of a launch of nblocks workgroups; z in T or in a
// narrower ZT.
template <typename T, int CL, int MODE, typename ZT = T, typename PT = T>
__device__ __forceinline__ void pcg_dir_body(const Tiling &t, int block, int nblocks,
                                             const GeomT<T> &g, const ZT *__restrict__ z,
                                             const uint8_t *__restrict__ code,
                                             const PT *__restrict__ pin, PT *__restrict__ pout,
                                             double *partials, const Scalars *sc, double *lds) {
  if (sc->status != 0) return;
  const T beta = (T)sc->beta;
  // an FP16 z carries the cycle scale sig (a power of two)
  const T zinv = std::is_same<ZT, _Float16>::value ? (T)(1.0 / sc->sig) : T(1);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz, block);
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
    // West and east neighbours come from the adjacent lanes' registers by a
    // wave shuffle; only lanes 0 and 63 load theirs (from the adjacent tile).
    // The other lanes point their edge loads at the row start, so those load
    // instructions touch one cache line.
    const int lane = (int)threadIdx.x;
    const ptrdiff_t row0 = -(ptrdiff_t)(lane * CL);
    const ptrdiff_t ow = lane == 0 ? (hw ? -1 : 0) : row0;
    const ptrdiff_t oe = lane == kTileX - 1 ? (he ? CL : CL - 1) : row0;
    // Software pipeline: the loads of plane k+1's neighbours and of plane
    // k+2's own row are issued before plane k is computed. Two register sets
    // alternate (the loop steps two planes), so no loaded value is copied
    // before it arrives. Planes past the chunk are clamped to planes already
    // loaded (their values are unused); past the top, p_next is p_cur's own
    // plane as in the unpipelined form.
    const size_t base = (size_t)i + (size_t)nx * (size_t)j;
    const int k0 = l.k0, k1 = l.k1, last = min(k1, nz - 1);
    auto at = [&](int k) { return base + nxy * (size_t)k; };
    RawV<MODE, T, CL, ZT, PT> own_a, own_b;
    Plane<MODE, T, CL, ZT, PT> pl_a, pl_b;
    own_a.load(z, pin, at(k0));
    Vec<T, CL> p_cur = own_a.value(beta, zinv);
    Vec<T, CL> p_prev = p_cur;
    if (k0 > 0) {
      own_a.load(z, pin, at(k0 - 1));
      p_prev = own_a.value(beta, zinv);
    }
    own_a.load(z, pin, at(min(k0 + 1, last)));
    pl_a.load(z, pin, code, at(k0), os, on, ow, oe);
    auto stage = [&](int k, const RawV<MODE, T, CL, ZT, PT> &own_use,
                     const Plane<MODE, T, CL, ZT, PT> &pl, RawV<MODE, T, CL, ZT, PT> &own_load,
                     Plane<MODE, T, CL, ZT, PT> &pl_load) {
      own_load.load(z, pin, at(min(k + 2, last)));
      pl_load.load(z, pin, code, at(min(k + 1, k1 - 1)), os, on, ow, oe);
      const Vec<T, CL> p_next = own_use.value(beta, zinv);
      const Vec<T, CL> vs = pl.s.value(beta, zinv), vn = pl.n.value(beta, zinv);
      // The shuffles run on every active lane (a source lane that is masked
      // off reads as 0), then the edge lanes take their loaded values.
      const T from_w = __shfl_up(p_cur.v[CL - 1], 1, kTileX);
      const T from_e = __shfl_down(p_cur.v[0], 1, kTileX);
      const T vw = lane == 0 ? pl.w.value(beta, zinv) : from_w;
      const T ve = lane == kTileX - 1 ? pl.e.value(beta, zinv) : from_e;
      const T dzk = g.dz[k], cb = g.cb[k], ct = g.ct[k];
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const T west = m == 0 ? (hw ? vw : T(0)) : p_cur.v[m - 1];
        const T east = m == CL - 1 ? (he ? ve : T(0)) : p_cur.v[m + 1];
        const T south = hs ? vs.v[m] : T(0), north = hn ? vn.v[m] : T(0);
        const T d = diag_code<T>(pl.code.b[m], cw[m], ce[m], cs, cn, cb, ct);
        const T q = dxdy[m] * dzk *
                    (d * p_cur.v[m] - cw[m] * west - ce[m] * east - cs * south - cn * north -
                     cb * p_prev.v[m] - ct * p_next.v[m]);
        s += (double)p_cur.v[m] * (double)(d > 0 ? q : T(0));
      }
      if constexpr (MODE != kDirCurrent) stv<PT, CL>(pout, at(k), convert<PT>(p_cur));
      p_prev = p_cur;
      p_cur = p_next;
    };
    for (int k = k0; k < k1; k += 2) {
      stage(k, own_a, pl_a, own_b, pl_b);
      if (k + 1 < k1) stage(k + 1, own_b, pl_b, own_a, pl_a);
    }
  }
  double v[1] = {s};
  block_sum<1>(v, lds);
  store_partials(partials, 1, v, nblocks, block);
}

// q = A p (recomputed), x += alpha p, r -= alpha q, z = M^-1 r and z.r.
// With DOT false (the multigrid-preconditioned CG) z = zscale M^-1 r is stored
// in ZT (the multigrid's first smoothing step from zero) and no dot is formed;
// with rmax non-null each workgroup also stores max |r / V| over its fluid
// cells there (an estimate of the max-norm residual), with qout non-null
// q = A p is stored in ZT (for the flexible-CG coefficient of the next step),
// and with sout non-null r / V (0 outside fluid cells) is stored in ZT. XT:
// the type z is stored in (ZT unless given); QT: the type q is stored in, as
// qscale q (ZT and 1 unless given).
template <typename T, int CL, typename ZT = T, bool DOT = true, typename PT = T,
          int XM = XM_NOW, typename XT = ZT, typename QT = ZT>
__device__ __forceinline__ void pcg_upd_body(const Tiling &t, int block, int nblocks,
                                             const GeomT<T> &g, const PT *__restrict__ p,
                                             const uint8_t *__restrict__ code,
                                             T *__restrict__ x, T *__restrict__ r,
                                             XT *__restrict__ z, double *partials,
                                             const Scalars *sc, double *lds,
                                             T zscale = T(1), double *rmax = nullptr,
                                             QT *__restrict__ qout = nullptr,
                                             ZT *__restrict__ sout = nullptr,
                                             const PT *__restrict__ ppend = nullptr,
                                             T qscale = T(1)) {
  if (sc->status != 0) {
    if constexpr (XM == XM_TWO) {
      // the iteration broke down: only the deferred step is applied (none
      // when the deferring iteration broke down too: alpha_pend = 0)
      const T ap = (T)sc->alpha_pend;
      const Lane l = lane_of<CL>(t, g.nx, g.ny, g.nz, block);
      if (l.active && ap != T(0)) {
        const size_t nxy = (size_t)g.nx * g.ny;
        for (int k = l.k0; k < l.k1; ++k) {
          const size_t c = (size_t)l.i + (size_t)g.nx * (size_t)l.j + nxy * (size_t)k;
          Vec<T, CL> xv = ldv<T, CL>(x, c);
          const Vec<T, CL> pp = convert<T>(ldv<PT, CL>(ppend, c));
#pragma unroll
          for (int m = 0; m < CL; ++m) xv.v[m] = xv.v[m] + ap * pp.v[m];
          stv<T, CL>(x, c, xv);
        }
      }
    }
    return;
  }
  const T alpha = (T)sc->alpha;
  const T alpha_pend = XM == XM_TWO ? (T)sc->alpha_pend : T(0);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const Lane l = lane_of<CL>(t, nx, ny, nz, block);
  double s = 0.0, rmx = 0.0;
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
    // West and east neighbours come from the adjacent lanes' registers by a
    // wave shuffle; only lanes 0 and 63 load theirs (from the adjacent tile).
    // The other lanes point their edge loads at the row start, so those load
    // instructions touch one cache line.
    const int lane = (int)threadIdx.x;
    const ptrdiff_t row0 = -(ptrdiff_t)(lane * CL);
    const ptrdiff_t ow = lane == 0 ? (hw ? -1 : 0) : row0;
    const ptrdiff_t oe = lane == kTileX - 1 ? (he ? CL : CL - 1) : row0;
    // Software pipeline as in k_pcg_dir_t: plane k+1's neighbours, phi and r
    // and plane k+2's own direction row load while plane k is computed.
    const size_t base = (size_t)i + (size_t)nx * (size_t)j;
    const int k0 = l.k0, k1 = l.k1, last = min(k1, nz - 1);
    auto at = [&](int k) { return base + nxy * (size_t)k; };
    // p may be stored narrower (PT): it is widened where it is used, so the
    // loads stay in flight across the computation of a plane.
    struct Set {
      Vec<PT, CL> p_next, pp;
      Vec<T, CL> x, r;
      Plane<kDirCurrent, PT, CL> pl;
    };
    Set sa, sb;
    Vec<T, CL> p_cur = convert<T>(ldv<PT, CL>(p, at(k0)));
    Vec<T, CL> p_prev = k0 > 0 ? convert<T>(ldv<PT, CL>(p, at(k0 - 1))) : p_cur;
    sa.p_next = ldv<PT, CL>(p, at(min(k0 + 1, last)));
    sa.pl.load(p, p, code, at(k0), os, on, ow, oe);
    if constexpr (XM != XM_DEFER) sa.x = ldv<T, CL>(x, at(k0));
    if constexpr (XM == XM_TWO) sa.pp = ldv<PT, CL>(ppend, at(k0));
    sa.r = ldv<T, CL>(r, at(k0));
    auto stage = [&](int k, const Set &u, Set &ld) {
      const int kn = min(k + 1, k1 - 1);
      ld.p_next = ldv<PT, CL>(p, at(min(k + 2, last)));
      ld.pl.load(p, p, code, at(kn), os, on, ow, oe);
      if constexpr (XM != XM_DEFER) ld.x = ldv<T, CL>(x, at(kn));
      if constexpr (XM == XM_TWO) ld.pp = ldv<PT, CL>(ppend, at(kn));
      ld.r = ldv<T, CL>(r, at(kn));
      const T dzk = g.dz[k], cb = g.cb[k], ct = g.ct[k];
      Vec<T, CL> xv = u.x, rv = u.r;
      Vec<XT, CL> zv;
      Vec<QT, CL> qv;
      Vec<ZT, CL> sv;
      const Vec<T, CL> p_next = convert<T>(u.p_next);
      const T from_w = __shfl_up(p_cur.v[CL - 1], 1, kTileX);
      const T from_e = __shfl_down(p_cur.v[0], 1, kTileX);
      const T vw = lane == 0 ? (T)u.pl.w.a : from_w;
      const T ve = lane == kTileX - 1 ? (T)u.pl.e.a : from_e;
#pragma unroll
      for (int m = 0; m < CL; ++m) {
        const T west = m == 0 ? (hw ? vw : T(0)) : p_cur.v[m - 1];
        const T east = m == CL - 1 ? (he ? ve : T(0)) : p_cur.v[m + 1];
        const T south = hs ? (T)u.pl.s.a.v[m] : T(0), north = hn ? (T)u.pl.n.a.v[m] : T(0);
        const T d = diag_code<T>(u.pl.code.b[m], cw[m], ce[m], cs, cn, cb, ct);
        const T qr = dxdy[m] * dzk *
                     (d * p_cur.v[m] - cw[m] * west - ce[m] * east - cs * south - cn * north -
                      cb * p_prev.v[m] - ct * p_next.v[m]);
        const T q = d > 0 ? qr : T(0);
        if constexpr (XM == XM_TWO) xv.v[m] = xv.v[m] + alpha_pend * (T)u.pp.v[m];
        xv.v[m] = xv.v[m] + alpha * p_cur.v[m];
        rv.v[m] = rv.v[m] - alpha * q;
        if constexpr (DOT) {
          zv.v[m] = rdiag_code<T>(d, dxdy[m], dzk) * rv.v[m];
          s += (double)zv.v[m] * (double)rv.v[m];
        } else {
          zv.v[m] = (XT)(zscale * rdiag_code<T>(d, dxdy[m], dzk) * rv.v[m]);
          qv.v[m] = (QT)(qscale * q);
          sv.v[m] = d > 0 ? (ZT)(rv.v[m] / (dxdy[m] * dzk)) : ZT(0);
          if (rmax && d > 0) rmx = fmax(rmx, fabs((double)rv.v[m]) / (double)(dxdy[m] * dzk));
        }
      }
      // the multigrid variant writes its outputs with nontemporal stores:
      // none is read again by this kernel
      if constexpr (XM != XM_DEFER) stv_s<!DOT, T, CL>(x, at(k), xv);
      stv_s<!DOT, T, CL>(r, at(k), rv);
      stv_s<!DOT, XT, CL>(z, at(k), zv);
      if constexpr (!DOT) {
        if (qout) stv_s<true, QT, CL>(qout, at(k), qv);
        if (sout) stv_s<true, ZT, CL>(sout, at(k), sv);
      }
      p_prev = p_cur;
      p_cur = p_next;
    };
    for (int k = k0; k < k1; k += 2) {
      stage(k, sa, sb);
      if (k + 1 < k1) stage(k + 1, sb, sa);
    }
  }
  if constexpr (DOT) {
    double v[1] = {s};
    block_sum<1>(v, lds);
    store_partials(partials, 1, v, nblocks, block);
  } else {
    if (rmax) {
      block_max(rmx, lds);
      store_partials(rmax, 1, &rmx, nblocks, block);
    }
  }
}

// Inner right-hand side of a defect correction: r32 = r64 (the projected FP64
// true residual), z32 = M32^-1 r32, e32 = 0, and rz = z32.r32 summed in FP64;
// grid-stride over the cells with nblocks workgroups.
__device__ inline void mixed_begin_body(const GeomT<float> &g, const double *__restrict__ r64,
                                        const uint8_t *__restrict__ code,
                                        float *__restrict__ r32, float *__restrict__ z32,
                                        float *__restrict__ e32, double *partials, int block,
                                        int nblocks, double *lds) {
  const int nx = g.nx, ny = g.ny;
  const size_t n = (size_t)nx * ny * g.nz;
  const size_t nt = (size_t)blockDim.x * blockDim.y;
  const size_t tid = threadIdx.x + (size_t)blockDim.x * threadIdx.y;
  double s[1] = {0.0};
  for (size_t c = block * nt + tid; c < n; c += (size_t)nblocks * nt) {
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
  store_partials(partials, 1, s, nblocks, block);
}

// phi += e (FP64 accumulation of the FP32 correction).
__device__ inline void mixed_accumulate_body(size_t n, double *__restrict__ phi,
                                             const float *__restrict__ e32, int block,
                                             int nblocks) {
  const size_t nt = (size_t)blockDim.x * blockDim.y;
  const size_t tid = threadIdx.x + (size_t)blockDim.x * threadIdx.y;
  for (size_t c = block * nt + tid; c < n; c += (size_t)nblocks * nt) phi[c] = phi[c] + (double)e32[c];
}

}  // namespace ffpm
