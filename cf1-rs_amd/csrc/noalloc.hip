// Fixed device scratch for the entry point that allocates nothing
// (cf1g_mphf_lookup_noalloc): callers under graph capture or in serving
// loops that must not call hipMalloc. The array lives in the code object, so
// each device the library runs on holds its own copy until the process ends.
// In host code the array's name is only a host-side stand-in: the device
// address comes from hipGetSymbolAddress, taken once per device when a
// context is created on it (noalloc_setup), and only that address is ever
// handed to kernels or copies.
#include <mutex>

#include "launch.hpp"

namespace cf1g {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(32) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

// With `device` current: record the device address of this device's scratch.
void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) throw HipError("noalloc_setup: device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (!p) throw HipError("hipGetSymbolAddress(scratch) returned no address");
  g_scratch_ptr[device] = p;
}

// The device address of `device`'s scratch (nullptr before noalloc_setup)
// and its size.
void* noalloc_scratch(int device, size_t* bytes) {
  if (bytes) *bytes = NOALLOC_SCRATCH_BYTES;
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace cf1g
