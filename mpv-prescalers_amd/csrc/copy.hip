// Plane copy with element-type conversion: what a hook leaves in place when its //!WHEN
// is false (mpv then passes the plane through unchanged).
#include "common.h"
#include "launch.h"

namespace mpx {

__global__ __launch_bounds__(256) void copy_kernel(Io io, int planes) {
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int y = blockIdx.y;
  const int fc = blockIdx.z;  // frame * planes + plane
  if (x >= io.W) return;
  const int frame = fc / planes, c = fc - frame * planes;
  const void* in = static_cast<const char*>(frame_in(io, frame)) + c * io.in_cstride * elem_bytes(io.in_dt);
  void* out = static_cast<char*>(frame_out(io, frame)) + c * io.out_cstride * elem_bytes(io.out_dt);
  const float v = load_elem(in, static_cast<int64_t>(y) * io.in_pitch + x, io.in_dt, io.in_max);
  store_elem(out, static_cast<int64_t>(y) * io.out_pitch + x, io.out_dt, v);
}

hipError_t launch_copy(const Io& io, int planes, hipStream_t st) {
  dim3 grid((io.W + 255) / 256, io.H, io.batch * planes);
  hipLaunchKernelGGL(copy_kernel, grid, dim3(256), 0, st, io, planes);
  return hipGetLastError();
}

}  // namespace mpx
