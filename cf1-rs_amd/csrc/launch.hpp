// Kernel launch wrapper: error check plus optional per-kernel HIP-event timing
// on the launch stream.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace cf1g {

// Per-kernel timing, recorded with hipEvents on the stream of the launch when
// enabled (cf1g_set_profiling). Summaries are read with cf1g_kernel_times.
void prof_begin(const char* name, hipStream_t st);
void prof_end(const char* name, hipStream_t st);

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace cf1g

#define CF1G_LAUNCH(name, kern, grid, block, shmem, st, ...)                       \
  do {                                                                             \
    ::cf1g::prof_begin(name, st);                                                  \
    hipLaunchKernelGGL(kern, grid, block, shmem, st, __VA_ARGS__);                 \
    ::cf1g::hip_check(hipGetLastError(), "launch " name);                          \
    ::cf1g::prof_end(name, st);                                                    \
  } while (0)
