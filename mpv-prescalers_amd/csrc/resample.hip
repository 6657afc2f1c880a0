// The scaler after a prescaler: resamples a prescaled plane to the display size and
// consumes the prescaler's OFFSET (mpv applies a hook's accumulated OFFSET in the next
// scaling pass, e.g. nnedi3-nns256-win8x6.hook:349-352 `//!OFFSET 0.000000 -0.500000`).
//
// Geometry: display pixel j (centre j + 1/2) sits at u = (j + 1/2) * in_w / out_w in the
// prescaled plane's continuous coordinate; the prescaled content is shifted by the
// offset, so the sample point is t = u + offset, and taps run over the texels whose
// centres lie within the filter radius of t (texel i has its centre at i + 1/2).  Weights
// are normalised to sum 1; reads are clamped to the plane (clamp-to-edge).
// Separable filters: bilinear (tent, r = 1), Catmull-Rom and Mitchell-Netravali cubics
// (r = 2), spline36 and Lanczos-3 (r = 3).  No widening when downscaling.
#include <math.h>

#include "common.h"
#include "launch.h"

namespace mpx {

__device__ __forceinline__ float filter_w(int filter, float x) {
  x = fabsf(x);
  switch (filter) {
    case 0: return fmaxf(0.f, 1.f - x);
    case 1:
    case 2: {  // cubic, (B, C) = (0, 1/2) Catmull-Rom, (1/3, 1/3) Mitchell
      const float B = filter == 1 ? 0.f : 1.f / 3.f, C = filter == 1 ? 0.5f : 1.f / 3.f;
      if (x < 1.f) return ((12.f - 9.f * B - 6.f * C) * x * x * x + (-18.f + 12.f * B + 6.f * C) * x * x + (6.f - 2.f * B)) / 6.f;
      if (x < 2.f)
        return ((-B - 6.f * C) * x * x * x + (6.f * B + 30.f * C) * x * x + (-12.f * B - 48.f * C) * x + (8.f * B + 24.f * C)) / 6.f;
      return 0.f;
    }
    case 3: {  // spline36
      if (x < 1.f) return ((13.f / 11.f * x - 453.f / 209.f) * x - 3.f / 209.f) * x + 1.f;
      if (x < 2.f) {
        x -= 1.f;
        return ((-6.f / 11.f * x + 270.f / 209.f) * x - 156.f / 209.f) * x;
      }
      if (x < 3.f) {
        x -= 2.f;
        return ((1.f / 11.f * x - 45.f / 209.f) * x + 26.f / 209.f) * x;
      }
      return 0.f;
    }
    default: {  // Lanczos-3: sinc(x) sinc(x/3)
      if (x >= 3.f) return 0.f;
      if (x < 1e-6f) return 1.f;
      const float px = 3.14159265358979f * x;
      return 3.f * sinf(px) * sinf(px / 3.f) / (px * px);
    }
  }
}

constexpr int kMaxTaps = 6;

struct Taps {
  int base;
  float w[kMaxTaps];
};

// taps for display coordinate j along one axis
__device__ __forceinline__ Taps axis_taps(int filter, int radius, int j, int in_n, int out_n, float off) {
  const float t = (static_cast<float>(j) + 0.5f) * static_cast<float>(in_n) / static_cast<float>(out_n) + off;
  const float p = t - 0.5f;  // in texel-index units
  Taps r;
  const float fl = floorf(p);
  r.base = static_cast<int>(fl) - radius + 1;
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < kMaxTaps; ++i) {
    const float w = i < 2 * radius ? filter_w(filter, p - (fl - radius + 1 + i)) : 0.f;
    r.w[i] = w;
    sum += w;
  }
  const float inv = 1.f / sum;
#pragma unroll
  for (int i = 0; i < kMaxTaps; ++i) r.w[i] *= inv;
  return r;
}

__global__ __launch_bounds__(256) void resample_kernel(Io io, int out_w, int out_h, int planes, int filter,
                                                       int radius, float off_x, float off_y) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int fc = blockIdx.z;  // frame * planes + plane
  if (x >= out_w || y >= out_h) return;
  const int f = fc / planes, c = fc - f * planes;
  const Taps tx = axis_taps(filter, radius, x, io.W, out_w, off_x);
  const Taps ty = axis_taps(filter, radius, y, io.H, out_h, off_y);
  const char* in = static_cast<const char*>(frame_in(io, f)) + c * io.in_cstride * elem_bytes(io.in_dt);
  float acc = 0.f;
  for (int j = 0; j < 2 * radius; ++j) {
    const int yy = min(max(ty.base + j, 0), io.H - 1);
    float row = 0.f;
    for (int i = 0; i < 2 * radius; ++i) {
      const int xx = min(max(tx.base + i, 0), io.W - 1);
      row = fmaf(tx.w[i], load_elem(in, static_cast<int64_t>(yy) * io.in_pitch + xx, io.in_dt, io.in_max), row);
    }
    acc = fmaf(ty.w[j], row, acc);
  }
  char* out = static_cast<char*>(frame_out(io, f)) + c * io.out_cstride * elem_bytes(io.out_dt);
  store_elem(out, static_cast<int64_t>(y) * io.out_pitch + x, io.out_dt, acc);
}

hipError_t launch_resample(const Io& io, int out_w, int out_h, int planes, int filter, float off_x, float off_y,
                           hipStream_t st) {
  static const int kRadius[5] = {1, 2, 2, 3, 3};
  dim3 grid((out_w + 63) / 64, (out_h + 3) / 4, io.batch * planes);
  hipLaunchKernelGGL(resample_kernel, grid, dim3(256), 0, st, io, out_w, out_h, planes, filter, kRadius[filter],
                     off_x, off_y);
  return hipGetLastError();
}

}  // namespace mpx
