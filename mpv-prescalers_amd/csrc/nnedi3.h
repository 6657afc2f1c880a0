// NNEDI3 weights as packed for the device and the launcher of the two doubling passes.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

struct mpx_ctx;

namespace mpx {

// Device copies of one NNEDI3 network, packed for a precision.  The softmax half is
// pre-scaled by log2(e) so the kernels use the hardware exp2.
//   FP32:      w  = float [nns][2K+2]  rows (w1*log2e, w2, b1*log2e, b2)
//   FP16/BF16: w  = 16-bit [nns][2][K] high parts of wscale*(w1*log2e, w2),
//              w_lo = the same layout holding the residuals (w - high), for the
//              three-product split; bias = float [nns][2] xscale*wscale*(b1*log2e, b2)
//   FP8:       w  = OCP e4m3 [nns][2][128 slots]: slots [0,K) and [K,2K) both hold
//              w * 2^-e (e per neuron and matrix, so max|w| * 2^-e is in (224, 448]),
//              the rest zero; wexp = int [nns][2] E8M0 scales 127 + e; bias as above
struct NnediWeights {
  int nns = 0, wh = 0, precision = 0;
  float xscale = 1.f, wscale = 1.f;  // power-of-two operand scales of the MFMA forms
  const void* w = nullptr;
  const void* w_lo = nullptr;
  const float* bias = nullptr;
  const int* wexp = nullptr;
  // launch shape (set by the autotuner; 0 = defaults): persistent blocks per CU (capped at
  // the kernel's occupancy), and for nns = 64 / 128 the neuron split over the block's 4
  // waves (0: fewer slices, more neuron tiles per wave; 1: more slices).  16x16 tiles --
  // nns 64: 1 x 4 or 2 x 2 (slices x tiles), nns 128: 2 x 4 or 4 x 2; 32x32 tiles -- nns 64:
  // 1 x 2 or 2 x 1, nns 128: 2 x 2 or 4 x 1
  int grid_per_cu = 0;
  int split = 0;
};

int nnedi3_prepare(mpx_ctx* ctx, int nns, int wh, int precision, NnediWeights* out);

// double_y (if dy) then double_x (if dx); `tmp` holds the W x 2H double_y output when both;
// done (optional): an event the last pass's dispatch completes (launch_done).
hipError_t launch_nnedi3(const NnediWeights& nw, const Io& io, bool dx, bool dy, float* tmp, int64_t tmp_frame,
                         hipStream_t st, hipEvent_t done = nullptr);

}  // namespace mpx
