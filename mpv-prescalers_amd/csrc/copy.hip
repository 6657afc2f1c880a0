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
  const int esz_in = io.in_dt == DT_F32 ? 4 : (io.in_dt == DT_U8 ? 1 : 2);
  const int esz_out = io.out_dt == DT_F32 ? 4 : 2;
  const void* in = static_cast<const char*>(io.in) + (frame * io.in_bstride + c * io.in_cstride) * esz_in;
  void* out = static_cast<char*>(io.out) + (frame * io.out_bstride + c * io.out_cstride) * esz_out;
  const float v = load_elem(in, static_cast<int64_t>(y) * io.in_pitch + x, io.in_dt, io.in_max);
  store_elem(out, static_cast<int64_t>(y) * io.out_pitch + x, io.out_dt, v);
}

hipError_t launch_copy(const Io& io, int planes, hipStream_t st) {
  dim3 grid((io.W + 255) / 256, io.H, io.batch * planes);
  hipLaunchKernelGGL(copy_kernel, grid, dim3(256), 0, st, io, planes);
  return hipGetLastError();
}

}  // namespace mpx
