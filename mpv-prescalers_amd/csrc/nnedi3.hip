// NNEDI3 luma doubling (nnedi3-nns256-win8x6.hook:15-706): the double_y and double_x
// passes with their combine passes folded in (each kernel writes the kept row/column and
// the predicted one).
//
// Per predicted pixel (nnedi3-nns16-win8x4.hook:20-36): window x (K = 8*WH samples),
// m = mean, v = variance, is = v >= eps ? v^-1/2 : 0;
// e_n = exp(dot(w1_n, x)*is + b1_n), t_n = dot(w2_n, x)*is + b2_n,
// out = clamp(m + 5 * sum(e_n t_n / (1 + |t_n|)) / sum(e_n) * v*is, 0, 1).
//
// Both kernels compute on the centred, normalised window x' = (x - m)*is, which equals
// the reference because every neuron's weights sum to zero (|sum w| < 2e-6, so the
// dropped term m*is*sum(w) moves the output by at most 5*m*|sum w| < 1e-5).  The variance
// is taken as mean((x-m)^2) (two passes), the same value as the reference's
// E[x^2]-m^2 without its FP32 cancellation.
//
// Two forms:
//  * nnedi3_valu: plain FP32, one thread per pixel, weights read as wave-uniform scalars.
//  * nnedi3_mfma: the [neurons x K] x [K x pixels] contraction on MFMA
//    (v_mfma_f32_16x16x32_{f16,bf16}; K = 48 as two 32-chunks, zero padded), FP32
//    accumulation, with the biases as the accumulator's initial value; weights and inputs
//    optionally split into high+low 16-bit parts (three products) so the contraction
//    keeps ~22 (FP16) / ~16 (BF16) bits.  The exp2 /
//    reciprocal / blend epilogue stays FP32.  Persistent blocks keep their neuron slice's
//    weights in VGPRs for the whole launch.
#include "common.h"
#include "launch.h"
#include "nnedi3.h"

namespace mpx {

typedef _Float16 half8v __attribute__((ext_vector_type(8)));
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef __bf16 bf4v __attribute__((ext_vector_type(4)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

struct NPass {
  const void* src;
  void* dst;
  const void* tab_src[kMaxTab];  // per-frame bases when ntab_src > 0
  void* tab_dst[kMaxTab];        // per-frame bases when ntab_dst > 0
  int ntab_src, ntab_dst;
  int64_t src_pitch, src_bstride, dst_pitch, dst_bstride;
  int src_dt, dst_dt;
  float src_max;
  int W, H;  // source plane
  int batch;
};

constexpr int NX = 64;  // predicted pixels per block tile: NX x NY
constexpr int NY = 4;
constexpr float kVarEps = 1.192092896e-7f;  // nnedi3-nns16-win8x4.hook:28

// Window / tile geometry.  AXIS 0 = double_y (new pixel between rows y and y+1), window
// dx in [-3,4] (8 long, outer index a), dy in [LO, LO+WH-1] (inner index b);
// AXIS 1 = double_x, the transpose.  Canonical index k = a*WH + b for both.
template <int WH, int AXIS>
struct Geo {
  static constexpr int K = 8 * WH;
  static constexpr int LO = -(WH / 2 - 1);
  static constexpr int TW = AXIS == 0 ? NX + 7 : NX + WH - 1;
  static constexpr int TH = AXIS == 0 ? NY + WH - 1 : NY + 7;
  static constexpr int OX = AXIS == 0 ? -3 : LO;  // tile origin relative to block origin
  static constexpr int OY = AXIS == 0 ? LO : -3;
  __device__ static int at(int tx, int ty, int a, int b) {
    return AXIS == 0 ? (ty + b) * TW + tx + a : (ty + a) * TW + tx + b;
  }
  __device__ static int centre(int tx, int ty) { return (ty - OY) * TW + tx - OX; }
};

template <int WH, int AXIS>
using NTile = HaloTile<1, Geo<WH, AXIS>::TH, Geo<WH, AXIS>::TW, 256>;

// mean, 1/std (0 below the variance threshold) and std*... of a K-sample window held in
// registers; the window is replaced by its centred, normalised form.
template <int K>
__device__ __forceinline__ void normalise(float* x, float& m, float& sd) {
  float sum = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) sum += x[k];
  m = sum / static_cast<float>(K);
  float ss = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    x[k] -= m;
    ss += x[k] * x[k];
  }
  const float v = ss / static_cast<float>(K);
  const float is = v >= kVarEps ? 1.0f / sqrtf(v) : 0.f;
  sd = v * is;
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] *= is;
}

template <int AXIS>
__device__ __forceinline__ void store_result(const NPass& p, void* dst, int x, int y, float orig, float pred) {
  if (AXIS == 0) {  // combine_y: row 2y = input row y, row 2y+1 = prediction
    const int64_t o = static_cast<int64_t>(2 * y) * p.dst_pitch + x;
    store_elem(dst, o, p.dst_dt, orig);
    store_elem(dst, o + p.dst_pitch, p.dst_dt, pred);
  } else {          // combine_x: column 2x = input, 2x+1 = prediction
    store_pair(dst, static_cast<int64_t>(y) * p.dst_pitch + 2 * x, p.dst_dt, orig, pred);
  }
}

__device__ __forceinline__ const void* frame_src(const NPass& p, int f) {
  if (p.ntab_src) return p.tab_src[f];
  return static_cast<const char*>(p.src) + static_cast<int64_t>(f) * p.src_bstride * elem_bytes(p.src_dt);
}
__device__ __forceinline__ void* frame_dst(const NPass& p, int f) {
  if (p.ntab_dst) return p.tab_dst[f];
  return static_cast<char*>(p.dst) + static_cast<int64_t>(f) * p.dst_bstride * elem_bytes(p.dst_dt);
}

// ----------------------------------------------------------------------------- FP32 VALU

template <int WH, int AXIS>
__global__ __launch_bounds__(256) void nnedi3_valu(NPass p, const float* __restrict__ wt, int nns) {
  using G = Geo<WH, AXIS>;
  constexpr int K = G::K;
  __shared__ float tile[G::TH * G::TW];
  const int f = blockIdx.z;
  const int x0 = blockIdx.x * NX, y0 = blockIdx.y * NY;
  {
    NTile<WH, AXIS> ht;
    ht.issue(frame_src(p, f), p.src_dt, p.src_pitch, 0, p.W, p.H, x0 + G::OX, y0 + G::OY);
    ht.commit(tile, p.src_dt, p.src_max);
  }
  __syncthreads();
  const int tx = threadIdx.x % NX, ty = threadIdx.x / NX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= p.W || y >= p.H) return;
  float xv[K];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < WH; ++b) xv[a * WH + b] = tile[G::at(tx, ty, a, b)];
  float m, sd;
  normalise<K>(xv, m, sd);
  float wsum = 0.f, vsum = 0.f;
  for (int n = 0; n < nns; ++n) {
    const float* w = wt + n * (2 * K + 2);
    float s1 = w[2 * K], s2 = w[2 * K + 1];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      s1 = fmaf(w[k], xv[k], s1);
      s2 = fmaf(w[K + k], xv[k], s2);
    }
    const float e = __builtin_amdgcn_exp2f(s1);
    const float q = s2 * __builtin_amdgcn_rcpf(1.0f + fabsf(s2));
    wsum += e;
    vsum = fmaf(e, q, vsum);
  }
  const float pred = fminf(fmaxf(m + 5.0f * vsum / wsum * sd, 0.f), 1.f);
  store_result<AXIS>(p, frame_dst(p, f), x, y, tile[G::centre(tx, ty)], pred);
}

// ----------------------------------------------------------------------------- MFMA

// 16-bit operand types: F16 (IS_BF16 = false) or BF16.
template <bool IS_BF16> struct Op;
template <> struct Op<false> {
  typedef _Float16 T;
  typedef half8v V8;
  __device__ static float4v mma(V8 a, V8 b, float4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return static_cast<_Float16>(v); }
  __device__ static float back(T v) { return static_cast<float>(v); }
};
template <> struct Op<true> {
  typedef __bf16 T;
  typedef bf8v V8;
  __device__ static float4v mma(V8 a, V8 b, float4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return static_cast<__bf16>(v); }
  __device__ static float back(T v) { return static_cast<float>(v); }
};

// K in chunks of 32 (v_mfma_f32_16x16x32): K = 32 -> 1 chunk, K = 48 -> 2 chunks with the
// second half of the last chunk zero (weights and window both zero there).
template <int K> struct KChunks {
  static constexpr int KC = (K + 31) / 32;
  static constexpr int KP = 32 * KC;
  static constexpr int XS = KP + 8;  // LDS row stride of the window matrix (padding vs bank conflicts)
};

// Operand fragments of one neuron tile (16 neurons) for one matrix (w1 or w2): lane l
// holds row (neuron) l%16, k = 32*c + 8*(l/16) + 0..7, high and low parts.
template <int KC, bool IS_BF16>
struct WFrag {
  typename Op<IS_BF16>::V8 hi[KC], lo[KC];
};

template <int KC, bool IS_BF16, int TERMS>
__device__ __forceinline__ float4v contract(const WFrag<KC, IS_BF16>& w, const typename Op<IS_BF16>::V8* bh,
                                            const typename Op<IS_BF16>::V8* bl, float4v acc) {
  using O = Op<IS_BF16>;
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    acc = O::mma(w.hi[c], bh[c], acc);
    if (TERMS >= 3) {
      acc = O::mma(w.hi[c], bl[c], acc);
      acc = O::mma(w.lo[c], bh[c], acc);
    }
  }
  return acc;
}

// Block: 4 waves.  The NT = nns/16 neuron tiles are split into NSL slices (NSL = min(NT,4));
// waves with the same neuron slice share the 16 pixel groups of a tile, 4/NSL pixel
// slices.  TPW = NT/NSL neuron tiles per wave, held in VGPRs for the whole launch.
// TERMS = 1: one product of 16-bit operands; TERMS = 3: high*high + high*low + low*high.
// The window operand is x' * XSCALE and the weights carry WSCALE (host side), so the
// accumulators hold ACC = XSCALE*WSCALE*(s + b): powers of two that keep the FP16 low
// parts out of the subnormal range (MFMA flushes FP16 subnormal inputs).
template <int WH, int AXIS, bool IS_BF16, int TERMS, int NSL, int TPW, bool SCALED>
__global__ __launch_bounds__(256, 2) void nnedi3_mfma(NPass p, const void* __restrict__ w_hi,
                                                      const void* __restrict__ w_lo,
                                                      const float* __restrict__ bias, float xscale, float inv_acc,
                                                      int tiles_x, int tiles_y, int total_tiles) {
  using G = Geo<WH, AXIS>;
  using O = Op<IS_BF16>;
  using T = typename O::T;
  constexpr int K = G::K;
  constexpr int KC = KChunks<K>::KC;
  constexpr int KP = KChunks<K>::KP;
  constexpr int XS = KChunks<K>::XS;
  constexpr int NPX = NX * NY;            // 256 pixels per tile
  constexpr int NGROUP = NPX / 16;        // 16 pixel groups
  constexpr int PSL = 4 / NSL;            // pixel slices
  __shared__ float tile[G::TH * G::TW];
  __shared__ __attribute__((aligned(16))) T xhi[NPX * XS];
  __shared__ __attribute__((aligned(16))) T xlo[TERMS >= 3 ? NPX * XS : 8];
  __shared__ float part[NSL][NPX][2];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nslice = wave % NSL;
  const int pslice = wave / NSL;
  const int row = lane & 15, quad = lane >> 4;
  const T zero = static_cast<T>(0.0f);

  // this wave's weight fragments and the accumulator initial values (scaled biases)
  WFrag<KC, IS_BF16> w1[TPW], w2[TPW];
  float4v b1[TPW], b2[TPW];
  const T* Wh = static_cast<const T*>(w_hi);
  const T* Wl = static_cast<const T*>(w_lo);
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int n = (nslice * TPW + t) * 16 + row;          // neuron of this lane's A row
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 32 * c + 8 * quad + j;
        const bool in = k < K;
        w1[t].hi[c][j] = in ? Wh[(static_cast<int64_t>(n) * 2 + 0) * K + k] : zero;
        w2[t].hi[c][j] = in ? Wh[(static_cast<int64_t>(n) * 2 + 1) * K + k] : zero;
        w1[t].lo[c][j] = (TERMS >= 3 && in) ? Wl[(static_cast<int64_t>(n) * 2 + 0) * K + k] : zero;
        w2[t].lo[c][j] = (TERMS >= 3 && in) ? Wl[(static_cast<int64_t>(n) * 2 + 1) * K + k] : zero;
      }
    // D rows (neurons) of this lane: 4*quad + i
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int nn = (nslice * TPW + t) * 16 + 4 * quad + i;
      b1[t][i] = bias[2 * nn];
      b2[t][i] = bias[2 * nn + 1];
    }
  }
  // accumulator scale: exp2 argument acc*inv_acc, and 1 + |t| -> acc_one + |acc|
  const float acc_one = SCALED ? 1.0f / inv_acc : 1.0f;

  // The halo of the block's next tile is loaded into registers while this one computes.
  const int per_frame = tiles_x * tiles_y;
  NTile<WH, AXIS> ht;
  auto issue = [&](int id) {
    const int ff = id / per_frame, rr = id - ff * per_frame;
    ht.issue(frame_src(p, ff), p.src_dt, p.src_pitch, 0, p.W, p.H, (rr % tiles_x) * NX + G::OX,
             (rr / tiles_x) * NY + G::OY);
  };
  if (static_cast<int>(blockIdx.x) < total_tiles) issue(blockIdx.x);
  for (int tileid = blockIdx.x; tileid < total_tiles; tileid += gridDim.x) {
    const int f = tileid / per_frame;
    const int rem = tileid - f * per_frame;
    const int y0 = (rem / tiles_x) * NY, x0 = (rem % tiles_x) * NX;
    __syncthreads();  // previous iteration done with the LDS
    ht.commit(tile, p.src_dt, p.src_max);
    __syncthreads();
    if (tileid + static_cast<int>(gridDim.x) < total_tiles) issue(tileid + gridDim.x);

    // one thread per pixel: normalised window -> 16-bit rows (hi, lo), zero padded to KP
    const int tx = threadIdx.x % NX, ty = threadIdx.x / NX;
    float m, sd;
    float orig;
    {
      float xv[K];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < WH; ++b) xv[a * WH + b] = tile[G::at(tx, ty, a, b)];
      orig = tile[G::centre(tx, ty)];
      normalise<K>(xv, m, sd);
      T* rh = xhi + threadIdx.x * XS;
      T* rl = xlo + (TERMS >= 3 ? threadIdx.x * XS : 0);
#pragma unroll
      for (int k = 0; k < KP; ++k) {
        const float v = k < K ? (SCALED ? xv[k] * xscale : xv[k]) : 0.f;
        const T h = O::cvt(v);
        rh[k] = h;
        if (TERMS >= 3) rl[k] = O::cvt(v - O::back(h));
      }
    }
    __syncthreads();

    // MFMA over this wave's pixel groups and neuron tiles
    for (int g = pslice; g < NGROUP; g += PSL) {
      const int px = g * 16 + row;  // B column of this lane
      typename O::V8 bh[KC], bl[KC];
#pragma unroll
      for (int c = 0; c < KC; ++c) {
        bh[c] = *reinterpret_cast<const typename O::V8*>(xhi + px * XS + 32 * c + 8 * quad);
        if (TERMS >= 3) bl[c] = *reinterpret_cast<const typename O::V8*>(xlo + px * XS + 32 * c + 8 * quad);
      }
      float ws = 0.f, vs = 0.f;
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        const float4v a1 = contract<KC, IS_BF16, TERMS>(w1[t], bh, bl, b1[t]);
        const float4v a2 = contract<KC, IS_BF16, TERMS>(w2[t], bh, bl, b2[t]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float e = __builtin_amdgcn_exp2f(SCALED ? a1[i] * inv_acc : a1[i]);
          // t/(1+|t|) with t = a2/acc_one  ==  a2/(acc_one + |a2|)
          const float q = a2[i] * __builtin_amdgcn_rcpf(acc_one + fabsf(a2[i]));
          ws += e;
          vs = fmaf(e, q, vs);
        }
      }
      // lanes l, l^16, l^32, l^48 hold the same pixel
      ws += __shfl_xor(ws, 16);
      vs += __shfl_xor(vs, 16);
      ws += __shfl_xor(ws, 32);
      vs += __shfl_xor(vs, 32);
      if (quad == 0) {
        part[nslice][px][0] = ws;
        part[nslice][px][1] = vs;
      }
    }
    __syncthreads();

    const int x = x0 + tx, y = y0 + ty;
    if (x < p.W && y < p.H) {
      float wsum = 0.f, vsum = 0.f;
#pragma unroll
      for (int s = 0; s < NSL; ++s) {
        wsum += part[s][threadIdx.x][0];
        vsum += part[s][threadIdx.x][1];
      }
      const float pred = fminf(fmaxf(m + 5.0f * vsum / wsum * sd, 0.f), 1.f);
      store_result<AXIS>(p, frame_dst(p, f), x, y, orig, pred);
    }
  }
}

// ----------------------------------------------------------------------------- launcher

static int cu_count() {
  static int n = 0;
  if (!n) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
      n = 256;
  }
  return n;
}

template <int WH, int AXIS>
static hipError_t run_pass(const NnediWeights& nw, const NPass& p, hipStream_t st) {
  const int tiles_x = (p.W + NX - 1) / NX, tiles_y = (p.H + NY - 1) / NY;
  if (nw.precision == 0) {
    dim3 grid(tiles_x, tiles_y, p.batch);
    hipLaunchKernelGGL((nnedi3_valu<WH, AXIS>), grid, dim3(256), 0, st, p, static_cast<const float*>(nw.w), nw.nns);
    return hipGetLastError();
  }
  const int total = tiles_x * tiles_y * p.batch;
  const int nt = nw.nns / 16;
  const float inv_acc = 1.0f / (nw.xscale * nw.wscale);
  // Neuron tiles are split over NSL waves only when one wave cannot hold them all
  // (nns >= 128); otherwise every wave owns all neurons for its own pixel groups.
  // Persistent grid: as many blocks as the occupancy of the instance allows per CU.
#define MPX_NN(BF, TERMS, NSL, TPW, SC)                                                               \
  {                                                                                                   \
    auto kern = nnedi3_mfma<WH, AXIS, BF, TERMS, NSL, TPW, SC>;                                       \
    int per_cu = 0;                                                                                   \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess || per_cu < 1) \
      per_cu = 1;                                                                                     \
    const int grid = min(total, cu_count() * per_cu);                                                 \
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, st, p, nw.w, nw.w_lo, nw.bias, nw.xscale,       \
                       inv_acc, tiles_x, tiles_y, total);                                             \
  }
#define MPX_NN_NT(BF, TERMS, SC)                    \
  switch (nt) {                                     \
    case 1: MPX_NN(BF, TERMS, 1, 1, SC) break;      \
    case 2: MPX_NN(BF, TERMS, 1, 2, SC) break;      \
    case 4: MPX_NN(BF, TERMS, 1, 4, SC) break;      \
    case 8: MPX_NN(BF, TERMS, 2, 4, SC) break;      \
    case 16: MPX_NN(BF, TERMS, 4, 4, SC) break;     \
    default: return hipErrorInvalidValue;           \
  }
  if (nw.precision == 1) {
    MPX_NN_NT(false, 1, false)
  } else if (nw.precision == 4) {
    MPX_NN_NT(false, 3, true)
  } else if (nw.precision == 2) {
    MPX_NN_NT(true, 3, false)
  } else {
    return hipErrorInvalidValue;
  }
#undef MPX_NN_NT
#undef MPX_NN
  return hipGetLastError();
}

hipError_t launch_nnedi3(const NnediWeights& nw, const Io& io, bool dx, bool dy, float* tmp, int64_t tmp_frame,
                         hipStream_t st) {
  hipError_t e = hipSuccess;
  if (dy) {
    NPass p{};
    p.src = io.in;
    p.src_pitch = io.in_pitch;
    p.src_bstride = io.in_bstride;
    p.src_dt = io.in_dt;
    p.src_max = io.in_max;
    p.W = io.W;
    p.H = io.H;
    p.batch = io.batch;
    if (io.ntab) {
      p.ntab_src = io.ntab;
      for (int i = 0; i < io.ntab; ++i) p.tab_src[i] = io.tab_in[i];
      if (!dx) {
        p.ntab_dst = io.ntab;
        for (int i = 0; i < io.ntab; ++i) p.tab_dst[i] = io.tab_out[i];
      }
    }
    if (dx) {
      p.dst = tmp;
      p.dst_pitch = io.W;
      p.dst_bstride = tmp_frame;
      p.dst_dt = DT_F32;
    } else {
      p.dst = io.out;
      p.dst_pitch = io.out_pitch;
      p.dst_bstride = io.out_bstride;
      p.dst_dt = io.out_dt;
    }
    e = nw.wh == 4 ? run_pass<4, 0>(nw, p, st) : run_pass<6, 0>(nw, p, st);
    if (e != hipSuccess) return e;
  }
  if (dx) {
    NPass p{};
    if (dy) {
      p.src = tmp;
      p.src_pitch = io.W;
      p.src_bstride = tmp_frame;
      p.src_dt = DT_F32;
      p.src_max = 1.f;
      p.H = 2 * io.H;
    } else {
      p.src = io.in;
      p.src_pitch = io.in_pitch;
      p.src_bstride = io.in_bstride;
      p.src_dt = io.in_dt;
      p.src_max = io.in_max;
      p.H = io.H;
      if (io.ntab) {
        p.ntab_src = io.ntab;
        for (int i = 0; i < io.ntab; ++i) p.tab_src[i] = io.tab_in[i];
      }
    }
    if (io.ntab) {
      p.ntab_dst = io.ntab;
      for (int i = 0; i < io.ntab; ++i) p.tab_dst[i] = io.tab_out[i];
    }
    p.W = io.W;
    p.batch = io.batch;
    p.dst = io.out;
    p.dst_pitch = io.out_pitch;
    p.dst_bstride = io.out_bstride;
    p.dst_dt = io.out_dt;
    e = nw.wh == 4 ? run_pass<4, 1>(nw, p, st) : run_pass<6, 1>(nw, p, st);
  }
  return e;
}

}  // namespace mpx
