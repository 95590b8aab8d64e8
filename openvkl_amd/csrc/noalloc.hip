// Fixed device scratch for vklComputeSampleNNoAlloc (openvkl/openvkl_hip.h).
//
// The binned sampling kernels need scratch (point records, segment counters,
// an overflow list). vklComputeSampleN allocates it per sampler on demand;
// callers that must not allocate get this file-scope array instead. Each GPU
// holds its own copy of the array for the life of the process, at its own
// address: the host-side name `scratch` is only a stand-in for that address,
// so host code looks the address up once per device (hipGetSymbolAddress,
// with the device current) and passes that pointer to the kernels.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "noalloc.h"
#include "objects.h"

namespace vklhip {

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

struct PerDevice
{
  std::once_flag once;
  void *ptr      = nullptr;
  hipError_t err = hipSuccess;
  std::mutex lock;
};

PerDevice &perDevice(int device)
{
  static PerDevice table[kMaxDevices];
  if (device < 0 || device >= kMaxDevices)
    throw VklError("HIP device ordinal " + std::to_string(device) + " is out of range", VKL_INVALID_ARGUMENT);
  return table[device];
}

}  // namespace

NoAllocScratch noAllocScratch(int device)
{
  PerDevice &p = perDevice(device);
  std::call_once(p.once, [&] {
    int prev = 0;
    p.err    = hipGetDevice(&prev);
    if (p.err != hipSuccess) return;
    if (prev != device) p.err = hipSetDevice(device);
    if (p.err == hipSuccess) p.err = hipGetSymbolAddress(&p.ptr, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (p.err != hipSuccess) p.ptr = nullptr;
  });
  if (p.err != hipSuccess || !p.ptr)
    throw VklError(std::string("hipGetSymbolAddress of the fixed scratch failed: ") + hipGetErrorString(p.err));
  return NoAllocScratch{static_cast<uint8_t *>(p.ptr), kNoAllocScratchBytes, &p.lock};
}

}  // namespace vklhip
