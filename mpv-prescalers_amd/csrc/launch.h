// Host-side launchers of the HIP kernels (one per kernel family), called by abi.cpp.
#pragma once

#include <atomic>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.h"

namespace mpx {

// Compute units of the current device (256 on MI355X), queried once per device (callers
// may be on several threads).
inline int device_cu_count() {
  static std::atomic<int> cache[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (!n) {
    int v = 0;
    n = hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0 ? v : 256;
    cache[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

// Plane copy with element-type conversion (the output of a hook whose WHEN is false).
hipError_t launch_copy(const Io& io, int planes, hipStream_t st);

// RAVU-lite(-ar), r in {2,3,4}; lut = FP16 [288][ceil(n^2/2)][4].
hipError_t launch_ravu_lite(int r, bool ar, const Io& io, const __half* lut, hipStream_t st);

// Separable resampler (filter MPX_FILTER_*), io.W x io.H -> out_w x out_h, with offsets.
hipError_t launch_resample(const Io& io, int out_w, int out_h, int planes, int filter, float off_x, float off_y,
                           hipStream_t st);

}  // namespace mpx
