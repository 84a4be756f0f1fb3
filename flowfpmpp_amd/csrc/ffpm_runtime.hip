// Error state, stream, workspaces and kernel timing of the library.
#include "ffpm_runtime.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ffpm.h"

namespace ffpm {

namespace {
thread_local std::string g_error;
hipStream_t g_stream = nullptr;
}  // namespace

void set_error(int code, const std::string &msg) {
  (void)code;
  g_error = msg;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s at %s:%d", (int)e,
                hipGetErrorString(e), what, file, line);
  set_error(ERR_DEVICE, buf);
  return ERR_DEVICE;
}

hipStream_t stream() { return g_stream; }

namespace {
int env_int(const char *name, int dflt) {
  const char *e = std::getenv(name);
  return e && *e ? std::atoi(e) : dflt;
}
}  // namespace

Tiling make_tiling(int nx, int ny, int nz, int zchunk, int ty, int cells_x) {
  Tiling t;
  // 16 rows (1024 threads) only on request: most kernels are built for
  // workgroups of at most 512 threads.
  if (ty <= 0) {
    ty = env_int("FFPM_TILE_Y", 4);
    if (ty > 8) ty = 4;
  }
  if (ty != 1 && ty != 2 && ty != 4 && ty != 8 && ty != 16) ty = 4;
  t.ty = ty;
  t.tiles_x = (nx + cells_x - 1) / cells_x;
  t.tiles_y = (ny + ty - 1) / ty;
  if (zchunk <= 0) zchunk = env_int("FFPM_ZCHUNK", 0);
  // 16-plane chunks measured best for the paired-cell PCG kernels on a
  // 512x384x256 frame (tile heights 1-8, chunks 8-64 and both tile orders
  // within 10%); smaller grids fall back to the occupancy rule below.
  if (zchunk <= 0 && (long)t.tiles_x * t.tiles_y * ((nz + 15) / 16) >= 2048) zchunk = 16;
  if (zchunk <= 0) {
    // Enough workgroups to fill 256 CUs several times over, and z chunks of
    // at least 4 planes so the z halo re-read stays small.
    const long want = 4096;
    const long txy = (long)t.tiles_x * t.tiles_y;
    long chunks = (want + txy - 1) / txy;
    if (chunks < 1) chunks = 1;
    zchunk = (int)((nz + chunks - 1) / chunks);
    if (zchunk < 4) zchunk = 4;
  }
  if (zchunk > nz) zchunk = nz;
  t.zchunk = zchunk;
  t.chunks_z = (nz + zchunk - 1) / zchunk;
  t.total = t.tiles_x * t.tiles_y * t.chunks_z;
  t.xcd_map = env_int("FFPM_XCD_MAP", 0) != 0;
  t.per_xcd = (t.total + kXcds - 1) / kXcds;
  t.blocks = t.xcd_map ? t.per_xcd * kXcds : t.total;
  return t;
}

void *Workspace::get(int slot, size_t bytes, int *status) {
  if (bytes == 0) bytes = 256;
  Buf &b = bufs_[slot];
  if (b.bytes >= bytes) return b.ptr;
  if (b.ptr) {
    hipError_t e = hipFree(b.ptr);
    if (e != hipSuccess) {
      *status = hip_fail(e, "hipFree", __FILE__, __LINE__);
      b.ptr = nullptr;
      b.bytes = 0;
      return nullptr;
    }
    b.ptr = nullptr;
    b.bytes = 0;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    *status = hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    return nullptr;
  }
  b.ptr = p;
  b.bytes = bytes;
  return p;
}

void Workspace::release_all() {
  for (auto &kv : bufs_)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  bufs_.clear();
}

Workspace::~Workspace() {
  // Process teardown: the runtime may already be gone, so leave the memory
  // to the driver.
}

Workspace &workspace() {
  static Workspace w;
  return w;
}

KernelTimer &timer() {
  static KernelTimer t;
  return t;
}

hipEvent_t KernelTimer::take() {
  if (!free_events.empty()) {
    hipEvent_t e = free_events.back();
    free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void KernelTimer::start(const char *name, hipEvent_t *a) {
  (void)name;
  *a = take();
  if (*a) (void)hipEventRecord(*a, stream());
}

void KernelTimer::stop(const char *name, hipEvent_t a) {
  if (!a) return;
  hipEvent_t b = take();
  if (!b) return;
  (void)hipEventRecord(b, stream());
  pending.push_back({name, a, b});
  if (pending.size() > 4096) (void)collect();
}

int KernelTimer::collect() {
  if (pending.empty()) return 0;
  FFPM_HIP(hipEventSynchronize(pending.back().b));
  for (auto &r : pending) {
    float ms = 0.f;
    FFPM_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto &t = totals[r.name];
    t.first += 1;
    t.second += ms;
    free_events.push_back(r.a);
    free_events.push_back(r.b);
  }
  pending.clear();
  return 0;
}

// Plain copy of n doubles, 16-byte vector loads and stores: the achievable
// streaming bandwidth the stencil kernels are compared with.
__global__ __launch_bounds__(256) void k_copy(size_t n2, const double2 *__restrict__ a,
                                              double2 *__restrict__ b) {
  for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < n2;
       c += (size_t)gridDim.x * blockDim.x)
    b[c] = a[c];
}

}  // namespace ffpm

extern "C" {

int ffpm_copy_bandwidth(size_t bytes, int reps, double *gbps, double *avg_us) {
  void *a = nullptr, *b = nullptr;
  const size_t n2 = bytes / 16;
  FFPM_HIP(hipMalloc(&a, n2 * 16));
  FFPM_HIP(hipMalloc(&b, n2 * 16));
  FFPM_HIP(hipMemsetAsync(a, 0, n2 * 16, ffpm::stream()));
  hipEvent_t e0, e1;
  FFPM_HIP(hipEventCreate(&e0));
  FFPM_HIP(hipEventCreate(&e1));
  size_t blocks = (n2 + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  ffpm::k_copy<<<(int)blocks, 256, 0, ffpm::stream()>>>(n2, (const double2 *)a, (double2 *)b);
  FFPM_HIP(hipEventRecord(e0, ffpm::stream()));
  for (int r = 0; r < reps; ++r)
    ffpm::k_copy<<<(int)blocks, 256, 0, ffpm::stream()>>>(n2, (const double2 *)a, (double2 *)b);
  FFPM_HIP(hipEventRecord(e1, ffpm::stream()));
  FFPM_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  FFPM_HIP(hipEventElapsedTime(&ms, e0, e1));
  *avg_us = 1e3 * ms / reps;
  *gbps = 2.0 * (double)(n2 * 16) / (*avg_us * 1e-6) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  FFPM_HIP(hipFree(a));
  FFPM_HIP(hipFree(b));
  return FFPM_OK;
}


const char *ffpm_last_error(void) { return ffpm::g_error.c_str(); }

int ffpm_set_stream(void *s) {
  ffpm::g_stream = (hipStream_t)s;
  return FFPM_OK;
}

int ffpm_release(void) {
  FFPM_HIP(hipDeviceSynchronize());
  ffpm::workspace().release_all();
  return FFPM_OK;
}

int ffpm_malloc(void **ptr, size_t bytes) {
  if (!ptr) {
    ffpm::set_error(FFPM_ERR_ARGUMENT, "null output pointer");
    return FFPM_ERR_ARGUMENT;
  }
  *ptr = nullptr;
  FFPM_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return FFPM_OK;
}

int ffpm_free(void *ptr) {
  if (ptr) FFPM_HIP(hipFree(ptr));
  return FFPM_OK;
}

int ffpm_memcpy_htod(void *dst, const void *src, size_t bytes) {
  FFPM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ffpm::stream()));
  FFPM_HIP(hipStreamSynchronize(ffpm::stream()));
  return FFPM_OK;
}

int ffpm_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  FFPM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ffpm::stream()));
  FFPM_HIP(hipStreamSynchronize(ffpm::stream()));
  return FFPM_OK;
}

int ffpm_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  FFPM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ffpm::stream()));
  return FFPM_OK;
}

int ffpm_memset_zero(void *ptr, size_t bytes) {
  FFPM_HIP(hipMemsetAsync(ptr, 0, bytes, ffpm::stream()));
  return FFPM_OK;
}

int ffpm_synchronize(void) {
  FFPM_HIP(hipStreamSynchronize(ffpm::stream()));
  return FFPM_OK;
}

int ffpm_device_count(int *count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = e == hipSuccess ? n : 0;
  return FFPM_OK;
}

int ffpm_timing_enable(int on) {
  ffpm::timer().enabled = on != 0;
  return FFPM_OK;
}

int ffpm_timing_reset(void) {
  int st = ffpm::timer().collect();
  ffpm::timer().totals.clear();
  return st;
}

int ffpm_timing_query(const char *kernel, long *count, double *total_ms) {
  int st = ffpm::timer().collect();
  if (st) return st;
  auto it = ffpm::timer().totals.find(kernel ? kernel : "");
  if (it == ffpm::timer().totals.end()) {
    *count = 0;
    *total_ms = 0.0;
  } else {
    *count = it->second.first;
    *total_ms = it->second.second;
  }
  return FFPM_OK;
}

}  // extern "C"
