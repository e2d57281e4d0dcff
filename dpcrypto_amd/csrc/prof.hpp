// Optional per-kernel timing with HIP events on the launching stream.
//
// The engine installs a Profiler for the calling thread while it launches a
// chunk; every launcher brackets its kernels with DPKEM_PROF_BEGIN/END.  When
// no profiler is installed (the default) the macros cost one thread-local load.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace dpkem {

struct ProfRecord {
  const char* name;
  hipEvent_t start, stop;
};

struct Profiler {
  std::vector<ProfRecord> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
};

inline thread_local Profiler* tls_profiler = nullptr;
inline thread_local hipEvent_t tls_prof_start = nullptr;

inline void prof_begin(hipStream_t st) {
  if (!tls_profiler) return;
  tls_prof_start = tls_profiler->get();
  if (tls_prof_start) (void)hipEventRecord(tls_prof_start, st);
}

inline void prof_end(hipStream_t st, const char* name) {
  if (!tls_profiler || !tls_prof_start) return;
  hipEvent_t stop = tls_profiler->get();
  if (!stop) return;
  (void)hipEventRecord(stop, st);
  tls_profiler->pending.push_back({name, tls_prof_start, stop});
  tls_prof_start = nullptr;
}

}  // namespace dpkem

// Launch a kernel and, when profiling, time it under `name`.
#define DPKEM_LAUNCH(name, kern, grid, block, st, ...)               \
  do {                                                               \
    ::dpkem::prof_begin(st);                                         \
    hipLaunchKernelGGL(kern, grid, block, 0, st, __VA_ARGS__);       \
    ::dpkem::prof_end(st, name);                                     \
  } while (0)
