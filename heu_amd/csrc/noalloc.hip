// noalloc.hip -- heu_sum_noalloc: heu_sum for callers that cannot allocate
// device memory.  Its scratch (chunk lists, the partial products of every
// reduction level, the segment totals) is carved from a fixed array in device
// memory instead of the context's growable pool, so the call makes no device
// allocation.  The array is a file-scope __device__ variable: each GPU holds its
// own copy, at its own address, from the loading of this code object until the
// process ends.  In host code `scratch` names a host-side stand-in, not device
// memory, so every host use goes through the address hipGetSymbolAddress
// returns for the current device, looked up once per device when the first
// context on it is created.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine_internal.hpp"

namespace {

constexpr size_t kScratchBytes = HEU_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_addr[kMaxDevices];         // this device's address of `scratch`
hipError_t g_status[kMaxDevices];  // the lookup's status (hipSuccess until tried)
bool g_tried[kMaxDevices];

}  // namespace

namespace heu_engine {

hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_tried[device]) return g_status[device];
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_tried[device] = true;
  g_status[device] = e;
  g_addr[device] = e == hipSuccess ? p : nullptr;
  return e;
}

}  // namespace heu_engine

extern "C" int heu_sum_noalloc(heu_key* key, const uint32_t* ct, const int64_t* offsets, size_t segments,
                               uint32_t* out) {
  using heu_engine::fail;
  if (!key || !key->ctx) return fail(HEU_EINVAL, "null key");
  const int dev = key->ctx->device;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev < 0 || dev >= kMaxDevices || !g_tried[dev])
      return fail(HEU_EDEVICE, "no fixed scratch was set up for device %d", dev);
    if (g_status[dev] != hipSuccess)
      return fail(HEU_EDEVICE, "fixed scratch of device %d: hipGetSymbolAddress: %s", dev,
                  hipGetErrorString(g_status[dev]));
    base = g_addr[dev];
  }
  return heu_engine::sum_fixed_scratch(key, ct, offsets, segments, out, static_cast<char*>(base), kScratchBytes);
}
