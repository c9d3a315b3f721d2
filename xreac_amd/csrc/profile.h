// Optional per-kernel timing with hipEvents on the launching stream.
//
// When enabled on a context, launches wrapped in XR_TIMED are bracketed by a
// pair of pooled events; elapsed times are summed per timer after the
// evaluation's final synchronisation.  Disabled, XR_TIMED costs one branch.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

enum XrTimer {
  XT_NBR_COUNT, XT_NBR_FILL, XT_EDGE, XT_CG_ITERATION, XT_CG_INIT, XT_NONBONDED, XT_BOND_FORWARD, XT_MANYBODY,
  XT_BACKWARD, XT_COUNT
};

// Timing events need no system-scope fence: their timestamps are read after a
// stream synchronisation, and a system-scope release at the end of every timed
// dispatch would write back L2 and slow the very kernel being timed.
constexpr unsigned kTimerEventFlags = hipEventDisableSystemFence;

struct XrProfiler {
  bool on = false;
  unsigned mask = 0;  // bit k: timer k is recorded
  int every = 1;      // record one dispatch in `every` of each timer
  long long seen[XT_COUNT] = {};
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<int, size_t>> pending;
  double total_ms[XT_COUNT] = {};
  long long launches[XT_COUNT] = {};

  bool wants(int id) const { return on && ((mask >> id) & 1u); }
  // Whether this dispatch of timer id is recorded (one in `every`).
  bool sample(int id) { return wants(id) && (seen[id]++ % every) == 0; }
  // Start event of a timed group; returns false (and records nothing) when not sampled.
  bool begin(int id, hipStream_t s) {
    if (!sample(id)) return false;
    if (used + 2 > pool.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, kTimerEventFlags) != hipSuccess) {
          on = false;
          return false;
        }
        pool.push_back(e);
      }
    }
    pending.emplace_back(id, used);
    (void)hipEventRecord(pool[used], s);
    used += 2;
    return true;
  }
  // Event pair for a launch that records its own timestamps (hipExtLaunchKernelGGL).
  bool slot(int id, hipEvent_t* a, hipEvent_t* b) {
    if (!sample(id)) return false;
    if (used + 2 > pool.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, kTimerEventFlags) != hipSuccess) {
          on = false;
          return false;
        }
        pool.push_back(e);
      }
    }
    pending.emplace_back(id, used);
    *a = pool[used];
    *b = pool[used + 1];
    used += 2;
    return true;
  }
  void end(hipStream_t s) {
    if (!on || pending.empty()) return;
    (void)hipEventRecord(pool[pending.back().second + 1], s);
  }
  // Call after the stream has been synchronised.
  void resolve() {
    for (auto& p : pending) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, pool[p.second], pool[p.second + 1]) == hipSuccess) {
        total_ms[p.first] += ms;
        launches[p.first] += 1;
      }
    }
    pending.clear();
    used = 0;
  }
  void reset() {
    for (int k = 0; k < XT_COUNT; ++k) {
      total_ms[k] = 0.0;
      launches[k] = 0;
      seen[k] = 0;
    }
  }
  ~XrProfiler() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

// The profiler of the context currently evaluating (set by the engine).
extern thread_local XrProfiler* xr_active_profiler;

#define XR_TIMED(id, stream, launch)                                  \
  do {                                                                \
    XrProfiler* p__ = xr_active_profiler;                             \
    const bool t__ = p__ && p__->begin((id), (stream));               \
    launch;                                                           \
    if (t__) p__->end(stream);                                        \
  } while (0)

// Launch a kernel; when profiling, the dispatch itself records the start and
// stop events (no separate event packets on the stream).
#define XR_LAUNCH(id, kernel, grid, block, shmem, stream, ...)                                  \
  do {                                                                                          \
    XrProfiler* p__ = xr_active_profiler;                                                       \
    hipEvent_t a__, b__;                                                                        \
    if (p__ && p__->slot((id), &a__, &b__)) {                                                   \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, a__, b__, 0, __VA_ARGS__);       \
    } else {                                                                                    \
      hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                      \
    }                                                                                           \
  } while (0)
