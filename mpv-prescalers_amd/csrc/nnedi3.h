// NNEDI3 weights as packed for the device and the launcher of the two doubling passes.
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

struct mpx_ctx;

namespace mpx {

struct NnediWeights {
  int nns = 0, wh = 0, precision = 0;
  const void* w = nullptr;      // device table, layout depends on precision
  const float* aux = nullptr;   // device: per-neuron FP32 terms (biases, scales, sums)
};

int nnedi3_prepare(mpx_ctx* ctx, int nns, int wh, int precision, NnediWeights* out);

// double_y (if dy) then double_x (if dx); `tmp` holds the W x 2H double_y output when both.
hipError_t launch_nnedi3(const NnediWeights& nw, const Io& io, bool dx, bool dy, float* tmp, int64_t tmp_frame,
                         hipStream_t st);

}  // namespace mpx
