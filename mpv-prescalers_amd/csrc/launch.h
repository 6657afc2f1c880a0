// Host-side launchers of the HIP kernels (one per kernel family), called by abi.cpp.
#pragma once

#include <atomic>

#include <hip/hip_ext.h>
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

// Launch with an optional completion event carried by the kernel's own dispatch
// (hipExtLaunchKernel's stop event) instead of a separate event record after it: the
// stream's next work then follows this kernel with no marker packet in between.  (The
// event completes when this kernel has, hence after all earlier work of the stream.)
template <typename K, typename... A>
inline hipError_t launch_done(K kern, dim3 grid, dim3 block, hipStream_t st, hipEvent_t done, A... args) {
  if (done)
    hipExtLaunchKernelGGL(kern, grid, block, 0, st, nullptr, done, 0, args...);
  else
    hipLaunchKernelGGL(kern, grid, block, 0, st, args...);
  return hipGetLastError();
}

// Tile shapes each RAVU family's kernels are built for: one block per tile, one thread per
// pixel of it (input pixels for lite / 3x / full RAVU, output pixels for zoom).  A plan
// launches candidate `tile` of its family's list (0 = the default, the best on average in
// A/B runs over frame sizes and channel forms; the others win for some sizes and forms by
// 2-6 %), chosen by mpx_plan_tune or set with mpx_plan_set_tile.
struct TileShape {
  int w, h;
};
constexpr TileShape kLiteTiles[] = {{64, 4}, {32, 8}, {32, 4}};
constexpr TileShape k3xTiles[] = {{32, 8}, {64, 4}};
constexpr TileShape kRavuTiles[] = {{32, 8}, {32, 16}, {64, 8}};
constexpr TileShape kZoomTiles[] = {{32, 8}, {16, 16}};
constexpr int kLiteTileCount = sizeof(kLiteTiles) / sizeof(kLiteTiles[0]);
constexpr int k3xTileCount = sizeof(k3xTiles) / sizeof(k3xTiles[0]);
constexpr int kRavuTileCount = sizeof(kRavuTiles) / sizeof(kRavuTiles[0]);
constexpr int kZoomTileCount = sizeof(kZoomTiles) / sizeof(kZoomTiles[0]);

// Plane copy with element-type conversion (the output of a hook whose WHEN is false).
hipError_t launch_copy(const Io& io, int planes, hipStream_t st);

// RAVU-lite(-ar), r in {2,3,4}; lut = FP16 [288][ceil(n^2/2)][4]; tile: index into kLiteTiles.
hipError_t launch_ravu_lite(int r, bool ar, const Io& io, const __half* lut, int tile, hipStream_t st);

// Separable resampler (filter MPX_FILTER_*), io.W x io.H -> out_w x out_h, with offsets.
hipError_t launch_resample(const Io& io, int out_w, int out_h, int planes, int filter, float off_x, float off_y,
                           hipStream_t st);

}  // namespace mpx
