// Error state, stream, workspaces and kernel timing of the library.
#include "ffpm_runtime.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ffpm.h"

// One bound frame: host arrays with their device copies.
struct ffpm_frame {
  struct Array {
    const char *host;
    size_t bytes;
    void *dev;
    int copy_out;
  };
  std::vector<Array> arrays;
};

namespace ffpm {

namespace {
std::mutex g_frames_mutex;
std::vector<ffpm_frame *> g_frames;
std::atomic<long long> g_h2d_count{0}, g_h2d_bytes{0}, g_d2h_count{0}, g_d2h_bytes{0};

// the bound array holding [p, p + bytes) in any live frame; overlap without
// containment is reported through *partial
const ffpm_frame::Array *find_bound(const char *p, size_t bytes, bool *partial) {
  *partial = false;
  for (const ffpm_frame *f : g_frames)
    for (const auto &a : f->arrays) {
      const char *lo = a.host, *hi = a.host + a.bytes;
      if (p >= lo && p + bytes <= hi) return &a;
      if (p < hi && p + bytes > lo) *partial = true;
    }
  return nullptr;
}
}  // namespace

void *present_device(const void *host, size_t bytes, int *status) {
  if (!host || bytes == 0) return nullptr;
  std::lock_guard<std::mutex> lock(g_frames_mutex);
  if (g_frames.empty()) return nullptr;
  bool partial = false;
  const char *p = static_cast<const char *>(host);
  const ffpm_frame::Array *a = find_bound(p, bytes, &partial);
  if (a) return static_cast<char *>(a->dev) + (p - a->host);
  if (partial) {
    set_error(ERR_ARGUMENT, "an array argument overlaps an array bound to a frame without lying "
                            "inside it");
    *status = ERR_ARGUMENT;
  }
  return nullptr;
}

void count_h2d(size_t bytes) {
  g_h2d_count += 1;
  g_h2d_bytes += (long long)bytes;
}

void count_d2h(size_t bytes) {
  g_d2h_count += 1;
  g_d2h_bytes += (long long)bytes;
}

namespace {
thread_local std::string g_error;
// each thread launches on its own stream (the null stream until it sets one)
thread_local hipStream_t g_stream = nullptr;
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

// One workspace per stream, kept until the process ends: a solve's buffers
// belong to the stream it runs on, so solves on different streams (from
// different threads) never share them.
namespace {
std::mutex g_ws_mutex;
std::map<hipStream_t, Workspace *> g_workspaces;
}  // namespace

Workspace &workspace() {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  Workspace *&w = g_workspaces[stream()];
  if (!w) w = new Workspace();
  return *w;
}

void release_workspaces() {
  std::lock_guard<std::mutex> lock(g_ws_mutex);
  for (auto &kv : g_workspaces) kv.second->release_all();
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
  std::lock_guard<std::mutex> lock(m);
  *a = take();
  if (*a) (void)hipEventRecord(*a, stream());
}

void KernelTimer::stop(const char *name, hipEvent_t a) {
  if (!a) return;
  std::lock_guard<std::mutex> lock(m);
  hipEvent_t b = take();
  if (!b) return;
  (void)hipEventRecord(b, stream());
  pending.push_back({name, a, b});
  if (pending.size() > 4096) (void)collect_locked();
}

int KernelTimer::collect() {
  std::lock_guard<std::mutex> lock(m);
  return collect_locked();
}

int KernelTimer::collect_locked() {
  if (pending.empty()) return 0;
  for (auto &r : pending) {
    // records of several streams: wait for each one's end
    FFPM_HIP(hipEventSynchronize(r.b));
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
// The copy of known size that calibrates the bandwidth figures: one 16-byte
// element per thread, nontemporal loads and stores (the fastest plain copy
// measured on MI355X: 6.7-6.8 TB/s at 402,653,184 bytes each way, against
// 5.5 TB/s for a grid-stride copy with plain accesses).
typedef double copy_v2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy(size_t n2, const copy_v2 *__restrict__ a,
                                              copy_v2 *__restrict__ b) {
  const size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (c < n2) __builtin_nontemporal_store(__builtin_nontemporal_load(a + c), b + c);
}

}  // namespace ffpm

extern "C" {

int ffpm_frame_create(ffpm_frame **frame) {
  if (!frame) {
    ffpm::set_error(ffpm::ERR_ARGUMENT, "null frame pointer");
    return ffpm::ERR_ARGUMENT;
  }
  auto *f = new ffpm_frame();
  std::lock_guard<std::mutex> lock(ffpm::g_frames_mutex);
  ffpm::g_frames.push_back(f);
  *frame = f;
  return FFPM_OK;
}

int ffpm_frame_bind(ffpm_frame *frame, void *host, size_t bytes, int copy_out) {
  using namespace ffpm;
  if (!frame || !host || bytes == 0) {
    set_error(ERR_ARGUMENT, "frame binding needs a frame, a host array and its size");
    return ERR_ARGUMENT;
  }
  {
    std::lock_guard<std::mutex> lock(g_frames_mutex);
    bool partial = false;
    if (find_bound(static_cast<const char *>(host), bytes, &partial) || partial) {
      set_error(ERR_ARGUMENT, "array already bound to a frame");
      return ERR_ARGUMENT;
    }
  }
  void *dev = nullptr;
  FFPM_HIP(hipMalloc(&dev, bytes));
  const hipError_t e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, stream());
  const hipError_t e2 = e == hipSuccess ? hipStreamSynchronize(stream()) : e;
  if (e2 != hipSuccess) {
    (void)hipFree(dev);
    return hip_fail(e2, "frame upload", __FILE__, __LINE__);
  }
  count_h2d(bytes);
  std::lock_guard<std::mutex> lock(g_frames_mutex);
  frame->arrays.push_back({static_cast<const char *>(host), bytes, dev, copy_out ? 1 : 0});
  return FFPM_OK;
}

namespace {
const ffpm_frame::Array *frame_array(ffpm_frame *frame, const void *host) {
  if (!frame) return nullptr;
  for (const auto &a : frame->arrays)
    if (a.host == static_cast<const char *>(host)) return &a;
  return nullptr;
}
}  // namespace

int ffpm_frame_upload(ffpm_frame *frame, const void *host) {
  using namespace ffpm;
  const ffpm_frame::Array *a = frame_array(frame, host);
  if (!a) {
    set_error(ERR_ARGUMENT, "array not bound to this frame");
    return ERR_ARGUMENT;
  }
  FFPM_HIP(hipMemcpyAsync(a->dev, a->host, a->bytes, hipMemcpyHostToDevice, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  count_h2d(a->bytes);
  return FFPM_OK;
}

int ffpm_frame_download(ffpm_frame *frame, void *host) {
  using namespace ffpm;
  const ffpm_frame::Array *a = frame_array(frame, host);
  if (!a) {
    set_error(ERR_ARGUMENT, "array not bound to this frame");
    return ERR_ARGUMENT;
  }
  FFPM_HIP(hipMemcpyAsync(host, a->dev, a->bytes, hipMemcpyDeviceToHost, stream()));
  FFPM_HIP(hipStreamSynchronize(stream()));
  count_d2h(a->bytes);
  return FFPM_OK;
}

int ffpm_frame_destroy(ffpm_frame *frame) {
  using namespace ffpm;
  if (!frame) return FFPM_OK;
  int status = FFPM_OK;
  for (const auto &a : frame->arrays) {
    if (a.copy_out && status == FFPM_OK) {
      hipError_t e = hipMemcpyAsync(const_cast<char *>(a.host), a.dev, a.bytes,
                                    hipMemcpyDeviceToHost, stream());
      if (e == hipSuccess) e = hipStreamSynchronize(stream());
      if (e != hipSuccess) status = hip_fail(e, "frame download", __FILE__, __LINE__);
      else count_d2h(a.bytes);
    }
  }
  {
    std::lock_guard<std::mutex> lock(g_frames_mutex);
    for (size_t i = 0; i < g_frames.size(); ++i)
      if (g_frames[i] == frame) {
        g_frames.erase(g_frames.begin() + (long)i);
        break;
      }
  }
  for (const auto &a : frame->arrays) (void)hipFree(a.dev);
  delete frame;
  return status;
}

int ffpm_transfer_counts(long long *counts) {
  if (!counts) {
    ffpm::set_error(ffpm::ERR_ARGUMENT, "null output pointer");
    return ffpm::ERR_ARGUMENT;
  }
  counts[0] = ffpm::g_h2d_count;
  counts[1] = ffpm::g_h2d_bytes;
  counts[2] = ffpm::g_d2h_count;
  counts[3] = ffpm::g_d2h_bytes;
  return FFPM_OK;
}

int ffpm_trace_push(const char *name) {
  if (!name) {
    ffpm::set_error(ffpm::ERR_ARGUMENT, "null trace range name");
    return ffpm::ERR_ARGUMENT;
  }
  roctxRangePushA(name);
  return FFPM_OK;
}

int ffpm_trace_pop(void) {
  roctxRangePop();
  return FFPM_OK;
}

int ffpm_copy_bandwidth(size_t bytes, int reps, double *gbps, double *avg_us) {
  void *a = nullptr, *b = nullptr;
  const size_t n2 = bytes / 16;
  const size_t blocks = (n2 + 255) / 256;
  if (n2 == 0 || blocks > 0x7fffffff || reps < 1 || !gbps || !avg_us) {
    ffpm::set_error(ffpm::ERR_ARGUMENT, "copy bandwidth: bad size, repetitions or output");
    return ffpm::ERR_ARGUMENT;
  }
  FFPM_HIP(hipMalloc(&a, n2 * 16));
  FFPM_HIP(hipMalloc(&b, n2 * 16));
  FFPM_HIP(hipMemsetAsync(a, 0, n2 * 16, ffpm::stream()));
  hipEvent_t e0, e1;
  FFPM_HIP(hipEventCreate(&e0));
  FFPM_HIP(hipEventCreate(&e1));
  using ffpm::copy_v2;
  ffpm::k_copy<<<(int)blocks, 256, 0, ffpm::stream()>>>(n2, (const copy_v2 *)a, (copy_v2 *)b);
  FFPM_HIP(hipEventRecord(e0, ffpm::stream()));
  for (int r = 0; r < reps; ++r)
    ffpm::k_copy<<<(int)blocks, 256, 0, ffpm::stream()>>>(n2, (const copy_v2 *)a, (copy_v2 *)b);
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
  ffpm::release_workspaces();
  return FFPM_OK;
}

int ffpm_stream_create(void **stream) {
  if (!stream) {
    ffpm::set_error(FFPM_ERR_ARGUMENT, "null output pointer");
    return FFPM_ERR_ARGUMENT;
  }
  hipStream_t s = nullptr;
  // default priority; non-blocking: no implicit ordering with the null stream
  FFPM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return FFPM_OK;
}

int ffpm_stream_destroy(void *stream) {
  if (stream) {
    FFPM_HIP(hipStreamSynchronize((hipStream_t)stream));
    FFPM_HIP(hipStreamDestroy((hipStream_t)stream));
  }
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
  ffpm::count_h2d(bytes);
  return FFPM_OK;
}

int ffpm_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  FFPM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ffpm::stream()));
  FFPM_HIP(hipStreamSynchronize(ffpm::stream()));
  ffpm::count_d2h(bytes);
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
  std::lock_guard<std::mutex> lock(ffpm::timer().m);
  ffpm::timer().totals.clear();
  return st;
}

int ffpm_timing_query(const char *kernel, long *count, double *total_ms) {
  int st = ffpm::timer().collect();
  if (st) return st;
  std::lock_guard<std::mutex> lock(ffpm::timer().m);
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
