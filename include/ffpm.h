/*
 * ffpm.h -- C ABI of the MI355X stair-step influence-potential solver and the
 * FPM force densities (library: flowfpmpp_amd/lib/libffpm.so).
 *
 * Every entry point replaces one Fortran module procedure of the reference
 * (cited per declaration) with the same argument order and meaning. Arrays
 * are plain pointers to column-major (Fortran-order) data, x fastest:
 *
 *   phi, solid : (0:nx+1, 0:ny+1, nz)  one x-y halo layer
 *   rhs, fields, outputs : (nx, ny, nz)
 *   xlo, xhi : (ny, nz)   ylo, yhi : (nx, nz)   zlo, zhi : (nx, ny)
 *   xc : (0:nx+1)  yc : (0:ny+1)  zc : (nz)  dx : (nx)  dy : (ny)  dz : (nz)
 *   body_grad : (3)  bounds : (6)  totals : (2)
 *
 * Entry points without a suffix take HOST pointers (the drop-in for callers
 * holding host arrays): they stage the arrays in device memory, run on the
 * GPU and copy the results back. Entry points ending in _dev take DEVICE
 * pointers for every array argument (body_grad and bounds stay on the host;
 * scalar outputs are host pointers); their inputs stay resident in HBM.
 *
 * Neighbour ranks follow MPI: any negative value (MPI_PROC_NULL in MPICH,
 * Open MPI and Intel MPI) marks a physical boundary. One GPU holds the whole
 * domain, so every neighbour must be a physical boundary; otherwise the call
 * returns FFPM_ERR_DECOMPOSITION. The comm argument of the reference is not
 * needed and is not passed.
 *
 * Every function returns FFPM_OK (0) or an error status; ffpm_last_error()
 * returns the message of the last failure on the calling thread. Messages of
 * the reference's own checks are reproduced verbatim.
 *
 * Calls are blocking and run on the calling thread's HIP stream
 * (ffpm_set_stream, per thread; the null stream until set). The library keeps
 * a grow-only device workspace per stream between calls, so calls from
 * different threads on different streams may overlap on one device (the
 * three force directions of a frame solved at once, for example); calls that
 * overlap on one stream must be serialised by the caller.
 */
#ifndef FFPM_H
#define FFPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FFPM_OK = 0,
  FFPM_ERR_ARGUMENT = 1,
  FFPM_ERR_ISOLATED = 2,      /* 'stair-step fluid cell has no fluid neighbour' */
  FFPM_ERR_NO_OUTER_FACE = 3, /* 'stair-step compatibility needs at least one fluid outer face' */
  FFPM_ERR_METHOD = 4,        /* 'stair-step method must be pcg or jacobi' */
  FFPM_ERR_SHAPE = 5,         /* 'stair-step solve needs nz>=2 and nonempty ranks' */
  FFPM_ERR_DECOMPOSITION = 6, /* a neighbour rank is present */
  FFPM_ERR_DEVICE = 7         /* HIP failure: no device, launch or memory error */
};

/* Iteration selector (src/fpm_stairstep_mpi_3d.f90:125-129). PCG_MIXED and
 * FCG_MG have no counterpart in the reference. PCG_MIXED: FP64 defect
 * correction around an FP32 PCG (FP32 stencil and preconditioner, FP64 dots),
 * stopping on the same FP64 max-norm residual; sweeps then counts inner
 * iterations. FCG_MG: FP64 flexible CG preconditioned by an FP32 geometric
 * multigrid V-cycle on the same operator, stopping on the same FP64 max-norm
 * residual (tested every 2 iterations, or every check_interval if smaller,
 * once a recursive estimate is near the tolerance); sweeps counts iterations. */
enum { FFPM_METHOD_PCG = 0, FFPM_METHOD_JACOBI = 1, FFPM_METHOD_PCG_MIXED = 2,
       FFPM_METHOD_FCG = 3 /* only through ffpm_solve_stairstep_fcg_dev */,
       FFPM_METHOD_FCG_MG = 4 };

const char *ffpm_last_error(void);
/* Sets the HIP stream (hipStream_t) every later call of the calling thread
 * launches on; 0 = null stream. No counterpart in the reference. */
int ffpm_set_stream(void *stream);
/* A new non-blocking HIP stream of default priority for ffpm_set_stream, and
 * its release (after its work has finished). No counterpart in the
 * reference: for callers that run independent solves concurrently. */
int ffpm_stream_create(void **stream);
int ffpm_stream_destroy(void *stream);
/* Frees the library's device workspaces (every stream's; no call may be in
 * flight). */
int ffpm_release(void);

/* Device memory for callers that hold their arrays in HBM (no counterpart in
 * the reference, whose OpenACC data regions played this role). Copies are
 * synchronous except device-to-device and the zero fill, which are ordered
 * on the library stream. */
int ffpm_device_count(int *count);
int ffpm_malloc(void **ptr, size_t bytes);
int ffpm_free(void *ptr);
int ffpm_memcpy_htod(void *dst, const void *src, size_t bytes);
int ffpm_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int ffpm_memcpy_dtod(void *dst, const void *src, size_t bytes);
int ffpm_memset_zero(void *ptr, size_t bytes);
int ffpm_synchronize(void);

/* trace_push / trace_pop of the reference's trace module (src/fpm_trace.f90:
 * 7-13, NVTX version src/fpm_trace_nvtx.f90:8-15): a named range on the
 * rocprofv3 timeline (roctx; shown with --marker-trace). The solver emits
 * the reference's own four ranges itself: 'stair-step setup', 'stair-step
 * iterations', 'stair-step gauge' and 'stair-step FPM terms'
 * (src/fpm_stairstep_mpi_3d.f90:131,162,169,441). name is NUL-terminated. */
int ffpm_trace_push(const char *name);
int ffpm_trace_pop(void);

/* Frames: host arrays kept resident in HBM across calls, the role the
 * caller's OpenACC data region plays for the reference (`!$acc data copy(phi)
 * copyin(...)`, src/fpm_stairstep_mpi_3d.f90:145-148; the benchmark's frame
 * region, tests/benchmark_stairstep_3d.f90:151). ffpm_frame_bind uploads a
 * host array once; until ffpm_frame_destroy every host-pointer entry point
 * uses that array's device copy in place instead of staging it (inputs are
 * not uploaded again, outputs are not downloaded). ffpm_frame_destroy copies
 * the arrays bound with copy_out != 0 back to the host and frees the device
 * copies. ffpm_frame_upload / ffpm_frame_download refresh one bound array in
 * either direction (`!$acc update device / self`). An array may be bound to
 * one frame at a time; an argument that overlaps a bound array without lying
 * inside it is an error. No counterpart in the reference's own API. */
typedef struct ffpm_frame ffpm_frame;
int ffpm_frame_create(ffpm_frame **frame);
int ffpm_frame_bind(ffpm_frame *frame, void *host, size_t bytes, int copy_out);
int ffpm_frame_upload(ffpm_frame *frame, const void *host);
int ffpm_frame_download(ffpm_frame *frame, void *host);
int ffpm_frame_destroy(ffpm_frame *frame);
/* Host-to-device and device-to-host copies made by the library since the
 * process started: counts[0] uploads, counts[1] bytes uploaded, counts[2]
 * downloads, counts[3] bytes downloaded. */
int ffpm_transfer_counts(long long counts[4]);

/* ------------------------------------------------------------------------ *
 * Stair-step solve, src/fpm_stairstep_mpi_3d.f90
 * ------------------------------------------------------------------------ */

/* exchange_solid_halo_3d, src/fpm_stairstep_mpi_3d.f90:25-36 */
int ffpm_exchange_solid_halo_3d(int32_t *solid, int nx, int ny, int nz, int west, int east,
                                int south, int north);
int ffpm_exchange_solid_halo_3d_dev(int32_t *solid, int nx, int ny, int nz, int west, int east,
                                    int south, int north);

/* enforce_stairstep_compatibility_3d, src/fpm_stairstep_mpi_3d.f90:41-92 */
int ffpm_enforce_stairstep_compatibility_3d(
    const double *rhs, const int32_t *solid, const double *body_grad, double *xlo, double *xhi,
    double *ylo, double *yhi, double *zlo, double *zhi, const double *dx, const double *dy,
    const double *dz, int nx, int ny, int nz, int west, int east, int south, int north,
    double *correction, double *imbalance);
int ffpm_enforce_stairstep_compatibility_3d_dev(
    const double *rhs, const int32_t *solid, const double *body_grad, double *xlo, double *xhi,
    double *ylo, double *yhi, double *zlo, double *zhi, const double *dx, const double *dy,
    const double *dz, int nx, int ny, int nz, int west, int east, int south, int north,
    double *correction, double *imbalance);

/* solve_stairstep_poisson_3d, src/fpm_stairstep_mpi_3d.f90:106-330.
 * method: FFPM_METHOD_PCG or FFPM_METHOD_JACOBI; tolerance < 0 runs exactly
 * max_sweeps iterations. phi is the initial guess on entry; on return it
 * holds the gauged solution, with 0 in solid cells and in the halo. */
int ffpm_solve_stairstep_poisson_3d(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double tolerance, int max_sweeps, int check_interval,
    int method, int *sweeps, double *residual);
int ffpm_solve_stairstep_poisson_3d_dev(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double tolerance, int max_sweeps, int check_interval,
    int method, int *sweeps, double *residual);

/* Outer defect corrections and inner FP32 iterations of the last
 * FFPM_METHOD_PCG_MIXED solve on the calling thread (no reference counterpart). */
int ffpm_last_mixed_solve(int *outer, int *inner);
/* Multigrid levels and convergence checks of the last FFPM_METHOD_FCG_MG
 * solve on the calling thread (no reference counterpart). */
int ffpm_last_mg_solve(int *levels, int *checks);

/* Batched solves (no counterpart in the reference, which solves one frame per
 * call of solve_stairstep_poisson_3d, src/fpm_stairstep_mpi_3d.f90:106-330).
 * Each problem is one call's worth of device arrays in the same layouts; the
 * problems of one batch advance together (one grid per kernel, per-problem
 * convergence masks, one HIP graph per check interval) and each ends exactly
 * as its own FP64 PCG solve would: same iterations, same phi. Problems with
 * the same solid and axis pointers (e.g. the three force directions of a
 * frame) share one operator. method is FFPM_METHOD_PCG, _PCG_MIXED (each
 * problem then ends as its standalone mixed solve, sweeps = FP32 inner
 * iterations) or _FCG_MG (each problem ends as its standalone multigrid CG;
 * a frame's problems also share the multigrid level operators). rows/zchunk
 * fix the workgroup height (0, 4 or 8) and z chunk of the streaming kernels
 * (0 = the standalone defaults, with which each problem is bit-identical to
 * its standalone solve; other shapes change only the reduction order);
 * solve_ms is the time from the batch start until the problem converged. */
typedef struct ffpm_stairstep_problem {
  int nx, ny, nz;
  double *phi;             /* (0:nx+1,0:ny+1,nz), in/out */
  const double *rhs;       /* (nx,ny,nz) */
  const int32_t *solid;    /* (0:nx+1,0:ny+1,nz) */
  double body_grad[3];
  const double *xlo, *xhi, *ylo, *yhi, *zlo, *zhi;
  const double *xc, *yc, *zc, *dx, *dy, *dz;
  int sweeps;              /* out */
  double residual;         /* out */
  double solve_ms;         /* out */
} ffpm_stairstep_problem;
int ffpm_solve_stairstep_batch_dev(ffpm_stairstep_problem *problems, int count,
                                   double tolerance, int max_sweeps, int check_interval,
                                   int method, int rows, int zchunk, int use_graph);
/* Times `iterations` batched iterations (FP64 PCG, FP32 inner iterations for
 * PCG_MIXED, multigrid CG iterations for FCG_MG) for each candidate launch
 * shape (rows 4, 8 x z chunks 8, 16, 32) and returns the fastest; phi is not
 * written. */
int ffpm_autotune_stairstep_batch_dev(ffpm_stairstep_problem *problems, int count, int method,
                                      int iterations, int *rows, int *zchunk,
                                      double *ms_per_iteration);

/* Learned preconditioner (no counterpart in the reference): a 3-D U-Net
 * (3 levels, 16/32/64 channels, 3x3x3 convolutions, ~0.36M parameters), BF16
 * weights on MFMA with FP32 accumulation, trained by its own forward, backward
 * and Adam kernels. Device arrays: features (nx,ny,nz,8) FP32 channels-last
 * (fluid mask, signed distance, dx, dy, dz, direction one-hot), source and
 * out (nx,ny,nz) FP32; every extent a multiple of 4 and nx >= 16. */
typedef struct ffpm_unet ffpm_unet;
int ffpm_unet_create(ffpm_unet **net, uint64_t seed);
int ffpm_unet_destroy(ffpm_unet *net);
int ffpm_unet_param_count(ffpm_unet *net, size_t *count);
/* FP32 master parameters (setting them also resets the Adam moments). */
int ffpm_unet_set_params(ffpm_unet *net, const float *host, size_t count);
int ffpm_unet_get_params(ffpm_unet *net, float *host, size_t count);
/* Gradients of the last training step (before its Adam update). */
int ffpm_unet_get_grads(ffpm_unet *net, float *host, size_t count);
int ffpm_unet_forward_dev(ffpm_unet *net, int nx, int ny, int nz, const float *features,
                          const float *source, float *out);
/* One step on the masked mean-square loss against target: forward, backward,
 * and an Adam update with learning rate lr (lr <= 0: gradients only). */
int ffpm_unet_train_step_dev(ffpm_unet *net, int nx, int ny, int nz, const float *features,
                             const float *source, const float *target, const float *mask,
                             float lr, float *loss);

/* y = A x for the PCG operator of the frame (A = V (diag - offdiag), 0 in
 * solid cells; src/fpm_stairstep_mpi_3d.f90:246-256), x and y dense
 * (nx,ny,nz) device arrays. No counterpart in the reference; used to form
 * training pairs for the network. */
int ffpm_stairstep_apply_operator_dev(const int32_t *solid, const double *xc, const double *yc,
                                      const double *zc, const double *dx, const double *dy,
                                      const double *dz, int nx, int ny, int nz, const double *x,
                                      double *y);

/* Flexible CG preconditioned by the network (no counterpart in the
 * reference): arguments as ffpm_solve_stairstep_poisson_3d_dev (one rank),
 * plus the network and its features for this frame and direction. The first
 * preconditioned step from the given phi is the network's initial guess;
 * three convergence checks without a 1% gain switch to diagonal PCG from the
 * current phi (fallback_at = FCG iterations before the switch, -1 if none).
 * sweeps counts FCG plus PCG iterations; residual is the FP64 max-norm
 * residual of src/fpm_stairstep_mpi_3d.f90:514-530. */
int ffpm_solve_stairstep_fcg_dev(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double tolerance, int max_sweeps, int check_interval, ffpm_unet *net, const float *features,
    int *sweeps, double *residual, int *fallback_at);

/* The multigrid-preconditioned CG (FFPM_METHOD_FCG_MG) with a learned coarse
 * correction (no counterpart in the reference) on multigrid level `level`
 * (1 <= level < the coarsest; extents multiples of 4, nx >= 16): the network
 * maps the level's scaled residual density to a correction, added with the
 * A-optimal length before the level's post-smoothing. mode 0: it replaces
 * the recursive coarse solve below the level (applied after the level's
 * pre-smoothing); mode 1: the recursion runs as usual and the network
 * corrects the residual left after the prolongation from below. features: device
 * FP32, 8 per cell of that level (channels-last; fluid fraction, signed
 * distance, three cell widths over 8/128, the direction one-hot). Stops on
 * the reference's FP64 max-norm residual; after three convergence checks
 * without a 1% gain it falls back to diagonal PCG from the current phi
 * (*fallback_at: iterations before the switch, -1 if none). */
int ffpm_solve_stairstep_fcg_mg_learned_dev(
    double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double tolerance, int max_sweeps, int check_interval, ffpm_unet *net, int level, int mode,
    const float *features, int *sweeps, double *residual, int *fallback_at);

/* fpm_partition_terms_2d, src/fpm_partition_2d.f90:35-154: force partition of
 * a 2-D stair-step body with rigid in-plane motion, per unit span and density.
 * phi, u, v (nx,ny); solid (nx,ny) int32, nonzero = solid (the reference's
 * logical array); xf(0:nx), yf(0:ny) cell faces; g(2) the body-face gradient
 * of phi's direction; kin(6) body motion; ring(2nx+2ny,2,2) outer-boundary
 * velocities at t-dt and t+dt; slip_body as the optional argument (0 =
 * absent/.false.). Outputs terms(4) = [KINE, VIF, VIS, BOUND] (host),
 * vif_density and omega (nx,ny). The _dev variant takes device arrays (g,
 * kin and terms stay host arrays). */
int ffpm_partition_terms_2d(const double *phi, const double *g, const double *u,
                            const double *v, const int32_t *solid, const double *xf,
                            const double *yf, const double *kin, const double *ring, double dt,
                            double nu, int nx, int ny, int slip_body, double *terms,
                            double *vif_density, double *omega);
int ffpm_partition_terms_2d_dev(const double *phi, const double *g, const double *u,
                                const double *v, const int32_t *solid, const double *xf,
                                const double *yf, const double *kin, const double *ring,
                                double dt, double nu, int nx, int ny, int slip_body,
                                double *terms, double *vif_density, double *omega);

/* stairstep_residual_3d, src/fpm_stairstep_mpi_3d.f90:403-419 (value out) */
int ffpm_stairstep_residual_3d(
    const double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double *value);
int ffpm_stairstep_residual_3d_dev(
    const double *phi, const double *rhs, const int32_t *solid, const double *body_grad,
    const double *xlo, const double *xhi, const double *ylo, const double *yhi,
    const double *zlo, const double *zhi, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, int nx, int ny, int nz, int west,
    int east, int south, int north, double *value);

/* compute_stairstep_fpm_terms_3d, src/fpm_stairstep_mpi_3d.f90:426-469 */
int ffpm_compute_stairstep_fpm_terms_3d(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *solid, const double *body_grad, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, const double *bounds, double *gx, double *gy, double *gz, double *vortex,
    double *advection, double *totals, int nx, int ny, int nz, int west, int east, int south,
    int north);
int ffpm_compute_stairstep_fpm_terms_3d_dev(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *solid, const double *body_grad, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, const double *xlo,
    const double *xhi, const double *ylo, const double *yhi, const double *zlo,
    const double *zhi, const double *bounds, double *gx, double *gy, double *gz, double *vortex,
    double *advection, double *totals, int nx, int ny, int nz, int west, int east, int south,
    int north);

/* compute_stairstep_fpm_terms_3d (src/fpm_stairstep_mpi_3d.f90:426-469) for
 * ndir (1-3) directions of one frame in one pass: the nine flow fields are
 * read once for all the directions' potentials (the reference calls the
 * routine once per direction). Per-direction arguments are host arrays of
 * ndir device pointers (phi, the outer face data xlo..zhi, the outputs
 * gx..advection); body_grad holds 3 ndir values (direction d at 3d) and
 * totals receives 2 ndir (direction d at 2d). Each direction's outputs and
 * totals equal those of ffpm_compute_stairstep_fpm_terms_3d_dev for it, bit
 * for bit. */
int ffpm_compute_stairstep_fpm_terms_3d_dirs_dev(
    int ndir, const double *const *phi, const double *u, const double *v, const double *w,
    const double *ox, const double *oy, const double *oz, const double *ax, const double *ay,
    const double *az, const int32_t *solid, const double *body_grad, const double *xc,
    const double *yc, const double *zc, const double *dx, const double *dy, const double *dz,
    const double *const *xlo, const double *const *xhi, const double *const *ylo,
    const double *const *yhi, const double *const *zlo, const double *const *zhi,
    const double *bounds, double *const *gx, double *const *gy, double *const *gz,
    double *const *vortex, double *const *advection, double *totals, int nx, int ny, int nz,
    int west, int east, int south, int north);

/* No counterpart in the reference: the variant of
 * ffpm_compute_stairstep_fpm_terms_3d_dirs_dev that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * same kernels, same results bit for bit. Its scratch memory (the 1-D stencil
 * coefficients, the integrals' per-workgroup partials and totals; 640 KB) is
 * a fixed array the library holds on each GPU it runs on until the process
 * ends; grids with nx + ny + nz > 16384 return FFPM_ERR_ARGUMENT. Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * threads or streams) race and must be serialised by the caller. */
int ffpm_compute_stairstep_fpm_terms_3d_dirs_noalloc_dev(
    int ndir, const double *const *phi, const double *u, const double *v, const double *w,
    const double *ox, const double *oy, const double *oz, const double *ax, const double *ay,
    const double *az, const int32_t *solid, const double *body_grad, const double *xc,
    const double *yc, const double *zc, const double *dx, const double *dy, const double *dz,
    const double *const *xlo, const double *const *xhi, const double *const *ylo,
    const double *const *yhi, const double *const *zlo, const double *const *zhi,
    const double *bounds, double *const *gx, double *const *gy, double *const *gz,
    double *const *vortex, double *const *advection, double *totals, int nx, int ny, int nz,
    int west, int east, int south, int north);

/* ------------------------------------------------------------------------ *
 * No-body pure-Neumann solvers, src/fpm_laplace_3d.f90
 * ------------------------------------------------------------------------ */

/* enforce_neumann_compatibility_3d, src/fpm_laplace_3d.f90:9-53 */
int ffpm_enforce_neumann_compatibility_3d(const double *rhs, double *xlo, double *xhi,
                                          double *ylo, double *yhi, double *zlo, double *zhi,
                                          int nx, int ny, int nz, double dx, double dy,
                                          double dz, double *correction, double *imbalance);
int ffpm_enforce_neumann_compatibility_3d_dev(const double *rhs, double *xlo, double *xhi,
                                              double *ylo, double *yhi, double *zlo,
                                              double *zhi, int nx, int ny, int nz, double dx,
                                              double dy, double dz, double *correction,
                                              double *imbalance);

/* solve_neumann_poisson_3d, src/fpm_laplace_3d.f90:55-107; phi is
 * (0:nx+1, 0:ny+1, 0:nz+1) with a z halo as well. */
int ffpm_solve_neumann_poisson_3d(double *phi, const double *rhs, const double *xlo,
                                  const double *xhi, const double *ylo, const double *yhi,
                                  const double *zlo, const double *zhi, int nx, int ny, int nz,
                                  double dx, double dy, double dz, double tolerance,
                                  int max_iterations, int check_interval, int *iterations,
                                  double *final_residual);
int ffpm_solve_neumann_poisson_3d_dev(double *phi, const double *rhs, const double *xlo,
                                      const double *xhi, const double *ylo, const double *yhi,
                                      const double *zlo, const double *zhi, int nx, int ny,
                                      int nz, double dx, double dy, double dz, double tolerance,
                                      int max_iterations, int check_interval, int *iterations,
                                      double *final_residual);

/* enforce_neumann_compatibility_nonuniform_3d, src/fpm_laplace_3d.f90:150-217
 * (note the reference's argument order: dx, dy, dz before nx, ny, nz). */
int ffpm_enforce_neumann_compatibility_nonuniform_3d(
    const double *rhs, double *xlo, double *xhi, double *ylo, double *yhi, double *zlo,
    double *zhi, const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double *correction, double *imbalance);
int ffpm_enforce_neumann_compatibility_nonuniform_3d_dev(
    const double *rhs, double *xlo, double *xhi, double *ylo, double *yhi, double *zlo,
    double *zhi, const double *dx, const double *dy, const double *dz, int nx, int ny, int nz,
    double *correction, double *imbalance);

/* solve_neumann_poisson_nonuniform_3d, src/fpm_laplace_3d.f90:219-293;
 * phi is (nx, ny, nz); xc (nx), yc (ny), zc (nz). */
int ffpm_solve_neumann_poisson_nonuniform_3d(
    double *phi, const double *rhs, const double *xlo, const double *xhi, const double *ylo,
    const double *yhi, const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, int nx, int ny,
    int nz, double tolerance, int max_iterations, int check_interval, int *iterations,
    double *final_residual);
int ffpm_solve_neumann_poisson_nonuniform_3d_dev(
    double *phi, const double *rhs, const double *xlo, const double *xhi, const double *ylo,
    const double *yhi, const double *zlo, const double *zhi, const double *xc, const double *yc,
    const double *zc, const double *dx, const double *dy, const double *dz, int nx, int ny,
    int nz, double tolerance, int max_iterations, int check_interval, int *iterations,
    double *final_residual);

/* compute_fpm_terms_uniform_3d, src/fpm_terms_3d.f90:8-37; phi is
 * (0:nx+1, 0:ny+1, 0:nz+1), fluid (nx, ny, nz) int32. */
int ffpm_compute_fpm_terms_uniform_3d(const double *phi, const double *u, const double *v,
                                      const double *w, const double *ox, const double *oy,
                                      const double *oz, const double *ax, const double *ay,
                                      const double *az, const int32_t *fluid, double *gx,
                                      double *gy, double *gz, double *vortex, double *advection,
                                      double *totals, int nx, int ny, int nz, double dx,
                                      double dy, double dz);
int ffpm_compute_fpm_terms_uniform_3d_dev(const double *phi, const double *u, const double *v,
                                          const double *w, const double *ox, const double *oy,
                                          const double *oz, const double *ax, const double *ay,
                                          const double *az, const int32_t *fluid, double *gx,
                                          double *gy, double *gz, double *vortex,
                                          double *advection, double *totals, int nx, int ny,
                                          int nz, double dx, double dy, double dz);

/* compute_fpm_terms_nonuniform_3d, src/fpm_terms_3d.f90:39-87; phi (nx, ny, nz). */
int ffpm_compute_fpm_terms_nonuniform_3d(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *fluid, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, const double *xlo, const double *xhi,
    const double *ylo, const double *yhi, const double *zlo, const double *zhi,
    const double *bounds, double *gx, double *gy, double *gz, double *vortex, double *advection,
    double *totals, int nx, int ny, int nz);
int ffpm_compute_fpm_terms_nonuniform_3d_dev(
    const double *phi, const double *u, const double *v, const double *w, const double *ox,
    const double *oy, const double *oz, const double *ax, const double *ay, const double *az,
    const int32_t *fluid, const double *xc, const double *yc, const double *zc,
    const double *dx, const double *dy, const double *dz, const double *xlo, const double *xhi,
    const double *ylo, const double *yhi, const double *zlo, const double *zhi,
    const double *bounds, double *gx, double *gy, double *gz, double *vortex, double *advection,
    double *totals, int nx, int ny, int nz);

/* ------------------------------------------------------------------------ *
 * Kernel timing (HIP events on the library stream around each launch of the
 * named hot kernels; no counterpart in the reference).
 * ------------------------------------------------------------------------ */
int ffpm_timing_enable(int on);
int ffpm_timing_reset(void);
/* Kernel names: "pcg_dir_spmv", "pcg_update", "jacobi", "spmv_dot", "update_dot", "direction",
 * "true_residual", "residual_max", "terms", "lap_sweep". count launches, total ms. */
int ffpm_timing_query(const char *kernel, long *count, double *total_ms);
/* Bandwidth of a plain device copy (read + write) of the given size, averaged
 * over reps launches (16-byte nontemporal loads and stores, one per thread):
 * the achievable streaming figure on this device. */
int ffpm_copy_bandwidth(size_t bytes, int reps, double *gbps, double *avg_us);

#ifdef __cplusplus
}
#endif
#endif /* FFPM_H */
