// Fixed per-device scratch of the allocation-free sampling entry point.
#pragma once

#include <stdint.h>

#include <mutex>

namespace vklhip {

// 64 MiB per GPU, held from the library's first use of the device until the
// process ends.
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20;

struct NoAllocScratch
{
  uint8_t *ptr;       // device address of this GPU's copy
  uint64_t bytes;     // kNoAllocScratchBytes
  std::mutex *lock;   // serialises this process's calls on the GPU
};

// The device address of `device`'s copy, looked up on first use with that
// device current. Throws VklError if the lookup fails.
NoAllocScratch noAllocScratch(int device);

}  // namespace vklhip
