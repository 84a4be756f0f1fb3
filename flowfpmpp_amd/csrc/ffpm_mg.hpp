// Geometric multigrid V-cycle used as the preconditioner of flexible CG on
// the stair-step operator A = V (diag - offdiag) (src/fpm_stairstep_mpi_3d.f90:
// 336-379, 247-256). There is no counterpart in the reference.
//
// Hierarchy: level 0 is the solve's own grid, whose operator is rebuilt from
// the per-cell neighbour code and FP32 copies of the 1-D coefficients; each
// coarser level aggregates 2x2x2 cells (ceil for odd extents) and stores, per
// cell, the conductance to its +x, +y and +z neighbour: half the sum of the
// finer conductances across the shared face (the Galerkin sum of piecewise-
// constant aggregation, scaled by 1/2 so that on a smooth grid it equals the
// rediscretised operator). Rows sum to zero on every level, so the constant
// null space of the Neumann problem is kept. A cell carries an unknown when
// its diagonal is positive.
//
// Cycle (FP32 arithmetic, FP64 input r): nu smoothing steps
// from zero (the first one fused into the outer update or the restriction
// that forms the level's right-hand side), restriction of the residual by
// summation over the children, recursion, trilinear cell-centred
// prolongation with weights renormalised over the coarse cells that carry
// unknowns, nu smoothing steps. The smoother is Chebyshev of degree nu on
// D^-1 A over [2.1 / ratio, 2.1] (rows sum to zero, so the spectrum of D^-1 A
// lies in [0, 2]; ratio 25), or damped Jacobi (weight 0.9); nu = 3 on every
// level. The coarsest level (at most 1024 cells) runs a fixed number of
// damped-Jacobi sweeps in one workgroup. Every step is a fixed linear map, so the
// CG around it sees the same preconditioner at every iteration.
//
// Storage: every vector of the cycle (on level 0 its first step, which also
// carries the source, the iterates and the output z) is FP16, scaled by a
// power of two per cycle (Scalars::sig, from the previous cycle's largest
// output) so that their values sit near 2^10: the cycle is linear, so the
// scale divides out exactly where z leaves it (the direction step, the z.r
// and z.q sums). That halves the level-0 bytes of the cycle; with the
// rounding (11 bits) the 512x384x256 frame takes the same 20 iterations per
// direction as with FP32 vectors (a CPU model of the cycle: one more of 18
// in one direction of a 256x192x128 frame). The coarse levels' vectors
// carry the same scale; their conductances are FP16 too. Arithmetic is FP32
// throughout.
#pragma once

#include "ffpm_pcg.hpp"

namespace ffpm {

// storage of the cycle's level-0 vectors
using MgStore = _Float16;

constexpr int kMgMaxLevels = 20;
constexpr int kMgCoarsestCells = 1024;

struct MgLevel {
  int nx = 0, ny = 0, nz = 0;
  size_t n = 0;
  // coarse levels only, all FP16: the conductances, the unknown mask (1
  // where diag > 0), the first-step scale over diag, iterates and right-hand
  // side
  MgStore *gE = nullptr, *gN = nullptr, *gT = nullptr;
  MgStore *m = nullptr;
  MgStore *rd = nullptr;
  MgStore *x = nullptr, *y = nullptr, *w = nullptr;
  MgStore *b = nullptr;
  // coarse levels of at least kMgStreamCells cells: the streaming sweep's
  // launch shape and cells per lane (cl = 0: the one-cell sweep)
  Tiling tl{};
  int cl = 0;
};

constexpr size_t kMgStreamCells = 1u << 18;

// One smoothing step of a level: iterate buffer a (and the previous iterate
// p, -1 for none) into buffer o, out = a + c2 (b - A a) / d + c1 (a - p).
// Level-0 buffers (Multigrid::f0): 0 the first step x1 (kept through the
// cycle: it also stands for the source), 1 and 2 iterates, 3 the cycle's
// output z, 4 a fifth iterate; coarse: 0 x, 1 y, 2 w.
struct MgStep {
  int a, p, o;
  float c1, c2;
};

// The steps of one level: `pre` after the first step (which leaves buffer 0),
// then the restriction from buffer `mid`, and after the prolongation into
// `mid` the `post` steps, ending in buffer `out`. The same for every level
// of a kind and every solve, so batched solves share it.
struct MgSchedule {
  int npre = 0, npost = 0, mid = 0, out = 0;
  MgStep pre[8], post[8];
};

struct UNet;

// The learned coarse correction (no counterpart in the reference) on level
// `level` of the cycle: the level's residual res = b - A x is scaled to a
// source density s = res / (V sigma), sigma = max |res / V|; the network maps
// s (with the level's geometry features) to a correction c, whose length is
// the A-optimal one, alpha = c.res / c.A c; x += alpha c, then the level's
// usual post-smoothing. Mode 0 (replace): the network takes the place of
// everything coarser, applied after the level's pre-smoothing. Mode 1
// (augment): the recursion runs as usual and the network corrects what is
// left after the prolongation from below. Buffers are the level's cells.
struct MgLearned {
  UNet *net = nullptr;               // null: the plain cycle
  const float *features = nullptr;   // 8 FP32 per level cell, channels-last
  int level = 0;
  int mode = 0;
  float *res = nullptr, *src = nullptr, *q = nullptr, *vol = nullptr;
  double *part = nullptr;            // per-workgroup partials, 3 slots of `blocks`
  double *sc = nullptr;              // sigma, alpha
  int blocks = 0;                    // workgroups of the level's cell launch
};

struct Multigrid {
  int levels = 0;
  MgLevel lv[kMgMaxLevels];
  GeomT<float> g32{};
  GeomT<double> g64{};
  const uint8_t *code = nullptr;
  Tiling tl{};              // the cycle's level-0 streaming launch shape (cl cells per lane)
  Tiling tlr{};             // the restriction's: tl with an even number of rows (2-8)
                            // and z chunks of an even number of planes
  bool stream_restrict = true;  // streaming restriction (FFPM_MG_RESTRICT=column: the
                                // column kernel)
  int cl = 1;
  Tiling tld{};             // the direction step's (FP32 p: the caller's tiling, cld cells
  int cld = 1;              // per lane)
  float omega = 0.9f;        // damped-Jacobi weight on level 0
  float omega_coarse = 0.9f; // and on the coarse levels
  float scale = 0.5f;
  int coarse_sweeps = 30;
  int nu = 3;                // level-0 sweeps before and after the coarse correction
  int nu_coarse = 4;         // coarse-level sweeps before and after
  bool nontemporal = true;   // level-0 sweeps: nontemporal source loads and stores
  bool cheb = true;          // Chebyshev smoothing (else damped Jacobi)
  float first = 0.9f;        // scale of the first step from zero, x = first b / d
  float first_coarse = 0.9f; // the same on the coarse levels
  MgSchedule s0, sc;         // level 0 (its last step writes z), coarse levels
  MgLearned lrn;               // the learned coarse correction, when net is set
  double *partials = nullptr;  // z.r and z.q partials of the last sweep (tl.blocks)
  MgStore *q16 = nullptr;      // q = A p of the last update, FP16 times Scalars::tau
  float sscale = 1.f;          // 1 / first: the cycle's (scaled) source r / V is
                               // read as f0[0] d sscale (f0[0] = sig first D^-1 r)
  // the level-0 vectors, FP16 times Scalars::sig (index as in MgStep; 3 is
  // the cycle's output z = B r)
  MgStore *f0[5] = {};
  __host__ __device__ MgStore *buf0(int i) const { return f0[i]; }
  __host__ __device__ static MgStore *bufc(const MgLevel &v, int i) {
    return i == 0 ? v.x : i == 1 ? v.y : v.w;
  }
};

// Builds the hierarchy for the frame with FP64 geometry g64, its FP32 copy
// g32 and the neighbour code (workspace slot S_MG); tl and cl are the
// level-0 streaming launch shape for FP32 values.
int mg_setup(Multigrid &mg, const GeomT<double> &g64, const GeomT<float> &g32,
             const uint8_t *code, const Tiling &tl, int cl, double *partials);

// The steps of mg_setup for callers that place the arrays themselves (the
// batched solves): parameters and level extents (no allocation); the bytes of
// the level operators (gE gN gT m rd of the coarse levels, shared by solves
// of one operator) and of one solve's cycle state (the five FP16 level-0
// vectors and q16, coarse x y w b), every array 256-byte aligned; carving
// them from base (advanced past them); and the launches that fill the level
// operators.
int mg_plan(Multigrid &mg, const GeomT<double> &g64, const GeomT<float> &g32,
            const uint8_t *code, const Tiling &tl, int cl, double *partials);
size_t mg_operator_bytes(const Multigrid &mg);
size_t mg_state_bytes(const Multigrid &mg);
void mg_bind_operator(Multigrid &mg, char *&base);
void mg_bind_state(Multigrid &mg, char *&base);
int mg_build(const Multigrid &mg);

// Puts the network on `features` at `level` (1 <= level < levels - 1, whose
// extents are multiples of 4 with nx >= 16) in mode 0 (replace) or 1
// (augment); allocates the level's work buffers (workspace slot S_MG_LRN) and
// computes its cell volumes.
int mg_learned_setup(Multigrid &mg, UNet *net, const float *features, int level, int mode);

// Cells per lane of the cycle's level-0 kernels (FFPM_MG_CL: 1, 2, 4 or 8,
// default 4; fewer where nx is not a multiple) and their launch shape for
// `rows` rows (z chunk FFPM_MG_ZCHUNK, default chosen by make_tiling).
int mg_cells_per_lane(int nx);
Tiling mg_cycle_tiling(int nx, int ny, int nz, int rows, int cl);

// Rows per workgroup of the cycle's level-0 kernels and the direction step:
// FFPM_MG_ROWS (1, 2, 4, 8, 16), default 4, measured faster than 2 or 8 rows
// at 512x384x256 (FP16 sweeps 98 us against 99 and 110).
int mg_rows();

// Convergence-check granularity of the multigrid CG: FFPM_MG_CHECK
// iterations (default 2, at most the caller's check interval); the FP64 true
// residual itself is formed only near convergence.
int mg_check_interval(int check_interval);

// The first smoothing step from zero after a true residual's first pass,
// f0[0] = sig w D^-1 r (which also carries the cycle's source r / V =
// f0[0] d / (w sig)): r is first projected off the constants in place, r -=
// red1 / red2 where diag > 0, which completes that true residual (sc: the
// scalars of that first pass, with the cycle scale sig).
int mg_first(const Multigrid &mg, double *r, const Scalars *sc);

// sc->anorm, the bound on the level-0 operator's infinity norm that scales
// the stored q = A p (Scalars::tau); once per solve, before its first update.
int mg_anorm(const Multigrid &mg, Scalars *sc);

// x += alpha p, r -= alpha A p, f0[0] = sig w D^-1 r (sig = sc->sig_next,
// which becomes sc->sig) and q16 = tau A p (alpha in sc; no-op after a
// breakdown); counts the iteration. t, cl: FP64 launch
// shape. The direction p is stored in FP32 (mg_direction forms it) and
// widened exactly; x, r and the arithmetic stay FP64. With rmax non-null, one
// partial max |r / V| per workgroup (t.blocks) goes there. xmode XM_DEFER
// leaves x alone and keeps alpha in sc->alpha_pend; the next update, XM_TWO
// with ppend the deferred direction, applies both steps in order (after a
// breakdown only the deferred one), so x ends bit-identical to XM_NOW at
// every iteration while the skipped iteration neither reads nor writes it.
int mg_update(const Multigrid &mg, const Tiling &t, int cl, const float *p, double *x,
              double *r, Scalars *sc, double *rmax = nullptr, int xmode = XM_NOW,
              const float *ppend = nullptr);

// f0[3] = sig B r (the rest of the cycle after the first step, which must be
// in f0[0]). Leaves three slots of partials (tl.blocks each) of the scaled
// output: z.r and z.q16 (0 without with_q) times sig^2 and sig tau, and max |z|
// times sig (k_finish FIN_FCG1 divides the scale out).
int mg_vcycle(const Multigrid &mg, bool with_q);

// The flexible CG's direction step with the cycle's output: p = z (MODE
// kDirReset) or z + beta p_old (kDirUpdate), z = f0[3] / sig, formed and
// stored in FP32, q = A p in FP32 and the p.q partials summed in FP64
// (t.blocks; launch shape t with cl cells per lane: the FP32 shape tld, cld).
int mg_direction(const Multigrid &mg, const Tiling &t, int cl, const float *pin, float *pout,
                 int mode, Scalars *sc);

}  // namespace ffpm
