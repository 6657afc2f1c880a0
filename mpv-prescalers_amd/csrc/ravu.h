// Launchers of the full RAVU, RAVU-3x and RAVU-zoom kernels.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "common.h"

struct mpx_ctx;
struct mpx_plan;

namespace mpx {

// Per-plan table of ravu-zoom: the output phase repeats along each axis with period
// out / gcd(in, out), so for small periods the bilinear LUT fetches are interpolated once per
// (bucket row, phase pair) at plan creation instead of once per output pixel.
//
// Every plan also holds the phase of each output column and row (int4 {base pixel, FP32
// bits of the sub-pixel phase f, index of the phase within its period, 0}: out_w entries
// for x, then out_h for y), computed once at plan creation in exact integer arithmetic, so
// the kernel reads them instead of dividing per output pixel.
struct ZoomPlan {
  float* wtab = nullptr;       // device: pre-interpolated weights (nullptr: per-pixel)
  size_t wtab_bytes = 0;
  int nphx = 0, nphy = 0;      // phase periods along x and y
  int4* phase = nullptr;       // device: out_w + out_h phase entries
  size_t phase_bytes = 0;
};

int zoom_plan_create(mpx_ctx* ctx, mpx_plan* plan, int in_w, int in_h, int out_w, int out_h);
void zoom_plan_destroy(mpx_ctx* ctx, mpx_plan* plan);

// RAVU-3x, r in {2,3,4}; channels MPX_CH_*; lut = FP16 [216][2*ceil(n^2/2)][4]; tile: index
// into k3xTiles.
hipError_t launch_ravu_3x(int r, int channels, const Io& io, const __half* lut, int tile, hipStream_t st);

// RAVU full: int11 into `int11` scratch (frame stride int11_frame floats), then int10/int01
// and the interleaved 2x output.  lut = FP16 [648][n^2/8][4]; tile: index into kRavuTiles;
// done (optional): an event the last kernel's dispatch completes (launch_done).
hipError_t launch_ravu(int r, int channels, const Io& io, const __half* lut, float* int11, int64_t int11_frame,
                       int tile, hipStream_t st, hipEvent_t done = nullptr);

// RAVU-zoom(-ar) to out_w x out_h; tile: index into kZoomTiles (output pixels per block).
hipError_t launch_ravu_zoom(int r, bool ar, int channels, const Io& io, const __half* lut, const __half* lut_ar,
                            const ZoomPlan& zp, int out_w, int out_h, int tile, hipStream_t st);

}  // namespace mpx
