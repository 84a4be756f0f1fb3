// Host-side runtime of the library: error state, the stream the library
// launches on, grow-only device workspaces, and optional HIP-event timing of
// the hot kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "ffpm_kernels.hpp"

namespace ffpm {

// A named roctx range over a scope (rocprofv3 --marker-trace shows it on the
// timeline): the ranges the reference emits through its trace module,
// 'stair-step setup', 'stair-step iterations', 'stair-step gauge' and
// 'stair-step FPM terms' (src/fpm_stairstep_mpi_3d.f90:131,162,169,441).
class TraceRange {
 public:
  explicit TraceRange(const char *name) { roctxRangePushA(name); }
  ~TraceRange() {
    if (open_) roctxRangePop();
  }
  // ends this range and opens the next one
  void next(const char *name) {
    if (open_) roctxRangePop();
    roctxRangePushA(name);
    open_ = true;
  }
  TraceRange(const TraceRange &) = delete;
  TraceRange &operator=(const TraceRange &) = delete;

 private:
  bool open_ = true;
};

hipStream_t stream();

// Host arrays bound to device copies by ffpm_frame_bind (the role of the
// caller's OpenACC data region, src/fpm_stairstep_mpi_3d.f90:145-148):
// the device copy of [host, host + bytes) when that range lies inside one
// bound array, else nullptr. A range that overlaps a bound array without
// lying inside it sets *status to ERR_ARGUMENT.
void *present_device(const void *host, size_t bytes, int *status);
// Host<->device copies made by the library (staging and frames), counted
// for ffpm_transfer_counts.
void count_h2d(size_t bytes);
void count_d2h(size_t bytes);

// Grow-only device buffers addressed by slot id; reallocated only when a
// larger size is requested. One workspace per process (the library is not
// re-entrant on one device; separate threads must serialise their calls).
class Workspace {
 public:
  void *get(int slot, size_t bytes, int *status);
  void release_all();
  ~Workspace();

 private:
  struct Buf {
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<int, Buf> bufs_;
};

Workspace &workspace();
void release_workspaces();  // every stream's (no call may be in flight)

enum Slot : int {
  S_COEF = 1, S_PHI, S_NEXT, S_SRC, S_DIAG, S_RDIAG, S_R, S_Q, S_PARTIALS, S_SCALARS,
  S_ISOLATED, S_CODE, S_Z, S_COEF32, S_E32, S_R32, S_Z32, S_P32A, S_P32B,
  // host-pointer entry points stage their arguments here
  S_U_PHI = 40, S_U_RHS, S_U_SOLID, S_U_XLO, S_U_XHI, S_U_YLO, S_U_YHI, S_U_ZLO, S_U_ZHI,
  S_U_XC, S_U_YC, S_U_ZC, S_U_DX, S_U_DY, S_U_DZ, S_U_F0, S_U_F1, S_U_F2, S_U_F3, S_U_F4,
  S_U_F5, S_U_F6, S_U_F7, S_U_F8, S_U_O0, S_U_O1, S_U_O2, S_U_O3, S_U_O4,
  // no-body Neumann solver
  S_L_NEXT = 80, S_L_COEF,
  // batched solves (one arena)
  S_BATCH = 90,
  // flexible CG: the network's source field and the two-term preconditioner
  S_FCG_SRC = 91, S_FCG_U, S_FCG_AU, S_FCG_AN, S_FCG_N, S_FCG_RING,
  // 2-D force partition
  S_P2_AXES = 100, S_P2_WORK, S_P2_PART, S_P2_SC, S_P2_HOST, S_P2_SOLID,
  // multigrid hierarchy of the flexible CG (one arena), the learned level's buffers
  S_MG = 110, S_MG_LRN,
  // densities of several directions in one pass: partials, totals
  S_TD_PART = 120, S_TD_OUT,
};

// ---- kernel timing with HIP events on the library stream --------------------
// One timer for the process: the events of a launch are recorded on the
// launching thread's stream, the records are kept under a mutex.
struct KernelTimer {
  std::atomic<bool> enabled{false};
  std::mutex m;
  struct Rec {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> free_events;
  std::map<std::string, std::pair<long, double>> totals;  // count, ms
  hipEvent_t take();
  void start(const char *name, hipEvent_t *a);
  void stop(const char *name, hipEvent_t a);
  int collect();
  int collect_locked();  // with m held
};

KernelTimer &timer();

// The calling thread's stream (ffpm_set_stream sets it per thread) and the
// grow-only workspace of that stream: calls on different streams may overlap.

// Brackets one launch with events when timing is on.
struct TimedLaunch {
  const char *name;
  hipEvent_t a{};
  explicit TimedLaunch(const char *n) : name(n) {
    if (timer().enabled) timer().start(name, &a);
  }
  ~TimedLaunch() {
    if (timer().enabled) timer().stop(name, a);
  }
};

}  // namespace ffpm
