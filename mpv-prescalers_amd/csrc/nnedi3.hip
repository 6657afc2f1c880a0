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
//  * nnedi3_mfma / nnedi3_mfma32: the [neurons x K] x [K x pixels] contraction on MFMA
//    (v_mfma_f32_16x16x32_{f16,bf16}, K = 48 as two 32-chunks, zero padded; or
//    v_mfma_f32_32x32x16_{f16,bf16}, K = 48 as three 16-chunks), FP32
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

constexpr int NX = 64;  // predicted pixels per block tile: NX x NY
constexpr int NY = 4;
// Rows of the persistent MFMA kernels' staged tile (template argument TY): 8 = two NY
// sub-tiles per halo load and block barrier; the one-product FP16 forms also have 16-row
// instances (four sub-tiles), launched where they keep the occupancy and give every block
// at least two tiles (run_pass).
constexpr int NYS = 8;

struct NPass {
  const void* src;
  void* dst;
  const void* tab_src[kMaxTab];  // per-frame bases when ntab_src > 0
  void* tab_dst[kMaxTab];        // per-frame bases when ntab_dst > 0
  int ntab_src, ntab_dst;
  int64_t src_pitch, src_bstride, dst_pitch, dst_bstride;
  int src_dt, dst_dt;
  float src_max, src_rmax;  // integer full scale and RN(1 / full scale)
  int W, H;  // source plane
  int batch;
  // frames of different sizes (sized != 0, ntab_src = ntab_dst frames): each frame's
  // source size and row pitches; tiles = prefix of the pass's tiles per frame
  int sized;
  int fW[kMaxTab], fH[kMaxTab];
  int64_t fsp[kMaxTab], fdp[kMaxTab];
  int tiles[kMaxTab + 1];
};

// The NX x TY tile t of a pass: frame, origin and that frame's geometry.
struct PassTile {
  int f, x0, y0, W, H;
  int64_t sp, dp;
};
template <int TY>
__device__ __forceinline__ PassTile pass_tile(const NPass& p, int t) {
  PassTile r;
  if (!p.sized) {
    const int nx = (p.W + NX - 1) / NX, per = nx * ((p.H + TY - 1) / TY);
    r.f = t / per;
    const int q = t - r.f * per;
    r.x0 = (q % nx) * NX;
    r.y0 = (q / nx) * TY;
    r.W = p.W;
    r.H = p.H;
    r.sp = p.src_pitch;
    r.dp = p.dst_pitch;
    return r;
  }
  int f = 0;
  while (f + 1 < p.ntab_src && t >= p.tiles[f + 1]) ++f;
  const int q = t - p.tiles[f], nx = (p.fW[f] + NX - 1) / NX;
  r.f = f;
  r.x0 = (q % nx) * NX;
  r.y0 = (q / nx) * TY;
  r.W = p.fW[f];
  r.H = p.fH[f];
  r.sp = p.fsp[f];
  r.dp = p.fdp[f];
  return r;
}

constexpr float kVarEps = 1.192092896e-7f;  // nnedi3-nns16-win8x4.hook:28

// Window / tile geometry.  AXIS 0 = double_y (new pixel between rows y and y+1), window
// dx in [-3,4] (8 long, outer index a), dy in [LO, LO+WH-1] (inner index b);
// AXIS 1 = double_x, the transpose.  Canonical index k = a*WH + b for both.
template <int WH, int AXIS, int TY = NY>
struct Geo {
  static constexpr int K = 8 * WH;
  static constexpr int LO = -(WH / 2 - 1);
  static constexpr int TW = AXIS == 0 ? NX + 7 : NX + WH - 1;
  static constexpr int TH = AXIS == 0 ? TY + WH - 1 : TY + 7;
  static constexpr int OX = AXIS == 0 ? -3 : LO;  // tile origin relative to block origin
  static constexpr int OY = AXIS == 0 ? LO : -3;
  __device__ static int at(int tx, int ty, int a, int b) {
    return AXIS == 0 ? (ty + b) * TW + tx + a : (ty + a) * TW + tx + b;
  }
  __device__ static int centre(int tx, int ty) { return (ty - OY) * TW + tx - OX; }
};

template <int WH, int AXIS, int TY = NY>
using NTile = HaloTile<1, Geo<WH, AXIS, TY>::TH, Geo<WH, AXIS, TY>::TW, 256>;

// mean, 1/std (0 below the variance threshold) and std of a K-sample window held in
// registers; the window is replaced by its centred form, and by its centred, normalised
// form when SCALE (otherwise the caller applies 1/std, e.g. inside its 16-bit conversion).
template <int K, bool SCALE = true>
__device__ __forceinline__ void normalise(float* x, float& m, float& sd, float& is) {
  float sum = x[0];
#pragma unroll
  for (int k = 1; k < K; ++k) sum += x[k];
  m = sum * (1.0f / static_cast<float>(K));
  float ss = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    x[k] -= m;
    ss += x[k] * x[k];
  }
  const float v = ss * (1.0f / static_cast<float>(K));
  // hardware reciprocal square root (within 1 ulp); the threshold test is exact
  is = v >= kVarEps ? __builtin_amdgcn_rsqf(v) : 0.f;
  sd = v * is;
  if constexpr (SCALE) {
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] *= is;
  }
}

// Window statistics shared across the pixels of a staged tile.  A window is 8 lines of WH
// samples -- AXIS 0: columns tx+a (a < 8) over rows ty..ty+WH-1; AXIS 1: rows ty+a over
// columns tx..tx+WH-1 -- and neighbouring pixels share lines, so each line's mean and
// centred sum of squares is computed once per staged tile (`build`), and a window combines
// its 8 lines by the pairwise formula M2 = sum_a M2_a + WH * sum_a (mean_a - m)^2 (every term
// non-negative: no cancellation in FP32), v = M2 / K, the same quantity as the two-pass
// variance mean((x - m)^2).
template <int WH, int AXIS, int TY>
struct LineStats {
  using G = Geo<WH, AXIS, TY>;
  static constexpr int COLS = AXIS == 0 ? G::TW : NX;
  static constexpr int N = (AXIS == 0 ? TY : G::TH) * COLS;
  __device__ static void build(const float* tile, float2* tab) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
      const int r = i / COLS, c = i - r * COLS;
      float x[WH];
#pragma unroll
      for (int b = 0; b < WH; ++b) x[b] = AXIS == 0 ? tile[(r + b) * G::TW + c] : tile[r * G::TW + c + b];
      float sum = x[0];
#pragma unroll
      for (int b = 1; b < WH; ++b) sum += x[b];
      const float mu = sum * (1.0f / static_cast<float>(WH));
      float m2 = 0.f;
#pragma unroll
      for (int b = 0; b < WH; ++b) {
        const float d = x[b] - mu;
        m2 = fmaf(d, d, m2);
      }
      tab[i] = make_float2(mu, m2);
    }
  }
  // mean m and variance v (mean squared deviation) of pixel (tx, ty)'s window
  __device__ static void window(const float2* tab, int tx, int ty, float& m, float& v) {
    float2 e[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) e[a] = AXIS == 0 ? tab[ty * COLS + tx + a] : tab[(ty + a) * COLS + tx];
    float sm = e[0].x, m2 = e[0].y;
#pragma unroll
    for (int a = 1; a < 8; ++a) {
      sm += e[a].x;
      m2 += e[a].y;
    }
    m = sm * 0.125f;
    float dev = 0.f;
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const float d = e[a].x - m;
      dev = fmaf(d, d, dev);
    }
    v = fmaf(static_cast<float>(WH), dev, m2) * (1.0f / static_cast<float>(8 * WH));
  }
};

template <int AXIS>
__device__ __forceinline__ void store_result(const NPass& p, void* dst, int64_t pitch, int x, int y, float orig,
                                             float pred) {
  if (AXIS == 0) {  // combine_y: row 2y = input row y, row 2y+1 = prediction
    const int64_t o = static_cast<int64_t>(2 * y) * pitch + x;
    store_elem(dst, o, p.dst_dt, orig);
    store_elem(dst, o + pitch, p.dst_dt, pred);
  } else {          // combine_x: column 2x = input, 2x+1 = prediction
    store_pair(dst, static_cast<int64_t>(y) * pitch + 2 * x, p.dst_dt, orig, pred);
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
__global__ __launch_bounds__(256) void nnedi3_valu(NPass p, const float* __restrict__ wt, int nns, int total) {
  using G = Geo<WH, AXIS>;
  constexpr int K = G::K;
  __shared__ float tile[G::TH * G::TW];
  const PassTile pt = pass_tile<NY>(p, xcd_tile(blockIdx.x, total));
  const int f = pt.f;
  const int x0 = pt.x0, y0 = pt.y0;
  {
    NTile<WH, AXIS> ht;
    ht.issue(frame_src(p, f), p.src_dt, pt.sp, 0, pt.W, pt.H, x0 + G::OX, y0 + G::OY);
    ht.commit(tile, p.src_dt, p.src_max, p.src_rmax);
  }
  __syncthreads();
  const int tx = threadIdx.x % NX, ty = threadIdx.x / NX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= pt.W || y >= pt.H) return;
  float xv[K];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < WH; ++b) xv[a * WH + b] = tile[G::at(tx, ty, a, b)];
  float m, sd;
  float is;
  normalise<K>(xv, m, sd, is);
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
  const float pred = fminf(fmaxf(m + 5.0f * vsum * __builtin_amdgcn_rcpf(wsum) * sd, 0.f), 1.f);
  store_result<AXIS>(p, frame_dst(p, f), pt.dp, x, y, tile[G::centre(tx, ty)], pred);
}

// ----------------------------------------------------------------------------- MFMA

// Operand forms of the contraction.
enum Form : int { F_F16 = 0, F_BF16 = 1, F_FP8 = 2 };

template <int FORM> struct Op;
template <> struct Op<F_F16> {
  typedef _Float16 T;
  typedef half8v V8;
  __device__ static float4v mma(V8 a, V8 b, float4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return static_cast<_Float16>(v); }
  __device__ static float back(T v) { return static_cast<float>(v); }
};
template <> struct Op<F_BF16> {
  typedef __bf16 T;
  typedef bf8v V8;
  __device__ static float4v mma(V8 a, V8 b, float4v c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
  __device__ static T cvt(float v) { return static_cast<__bf16>(v); }
  __device__ static float back(T v) { return static_cast<float>(v); }
};

// The contraction of one 16-neuron tile with one 16-pixel group, per operand form.
//
// F16 / BF16 (v_mfma_f32_16x16x32_{f16,bf16}): K in chunks of 32, zero padded (K = 48 ->
// two chunks); lane l holds A row (neuron) l%16 and B column (pixel) l%16 at
// k = 32*c + 8*(l/16) + 0..7.  TERMS = 2 adds high*low (the window split into high and
// low parts, weights single); TERMS = 3 also low*high (both operands split).  The window
// rows sit in LDS as 16-bit values, XS elements apart.
template <int FORM, int K, int TERMS>
struct Mma {
  using O = Op<FORM>;
  using T = typename O::T;
  static constexpr int KC = (K + 31) / 32;
  static constexpr int KP = 32 * KC;
  static constexpr int XS = KP + 8;  // LDS row stride (elements; padding vs bank conflicts)
  static constexpr int LO_ROWS = TERMS >= 2;
  // unscaled FP16 forms: the window arrives centred but unscaled and 1/std is applied
  // inside the conversions (one fused multiply-convert per part, v_fma_mix: high =
  // f16(x * is), low = f16(x * is - high))
  static constexpr bool FUSED_SCALE = FORM == F_F16 && TERMS <= 2;
  // FP16 forms with spare K slots (K = 48 in two 32-chunks): the bias rides in slots K and
  // K+1 as FP16 high and low parts of b * 2^10 against window values 2^-10 (exact in
  // FP16, |b| <= 4.4 over the shipped networks), so the accumulators start at zero and
  // no bias registers are held (win8x6 nns64: 183 -> 151 VGPRs, one more wave per SIMD)
  static constexpr bool BIAS_IN_K = FORM == F_F16 && TERMS <= 2 && K % 32 != 0;
  static constexpr float kBiasX = 1.0f / 1024.0f;
  struct W {
    typename O::V8 hi[KC], lo[KC];
  };
  struct B {
    typename O::V8 h[KC], l[KC];
  };
  __device__ static void load_w(W& w, const void* wh, const void* wl, const int* /*sc*/, const float* bias, int n,
                                int mtx, int quad) {
    const T* H = static_cast<const T*>(wh);
    const T* L = static_cast<const T*>(wl);
    const T zero = static_cast<T>(0.0f);
    T bh = zero, bl = zero;
    if constexpr (BIAS_IN_K) {
      const float b = bias[2 * n + mtx] * 1024.0f;
      bh = O::cvt(b);
      bl = O::cvt(b - O::back(bh));
    }
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = 32 * c + 8 * quad + j;
        const bool in = k < K;
        w.hi[c][j] = in ? H[(static_cast<int64_t>(n) * 2 + mtx) * K + k]
                        : (BIAS_IN_K && k == K ? bh : (BIAS_IN_K && k == K + 1 ? bl : zero));
        w.lo[c][j] = (TERMS >= 3 && in) ? L[(static_cast<int64_t>(n) * 2 + mtx) * K + k] : zero;
      }
  }
  // one pixel's window (raw samples xv, mean m, is = 1/std) -> its LDS row(s): the
  // normalised window x' = (x - m) * is (times xscale when scaled), split into 16-bit high
  // and low parts.  FUSED_SCALE forms centre inside the conversion, x * is - m * is in one
  // fused multiply-add: the rounding of m * is (<= m * is * 2^-24) reaches the output only
  // multiplied by the window's std = 1 / is, so it adds < 1e-6 at most.
  __device__ static void write_row(T* rh, T* rl, const float* xv, float m, float is, float xscale, bool scaled) {
    const float c = -m * is;
    const float sc = scaled ? is * xscale : is;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      if constexpr (FUSED_SCALE) {
        if (BIAS_IN_K && (k == K || k == K + 1)) {  // the bias slots
          rh[k] = O::cvt(kBiasX);
          if (TERMS >= 2) rl[k] = O::cvt(0.0f);
          continue;
        }
        if (k >= K) {
          rh[k] = O::cvt(0.0f);
          if (TERMS >= 2) rl[k] = O::cvt(0.0f);
          continue;
        }
        const float v = fmaf(xv[k], is, c);
        const T h = O::cvt(v);
        rh[k] = h;
        if (TERMS >= 2) rl[k] = O::cvt(v - O::back(h));
      } else {
        const float v = k < K ? (xv[k] - m) * sc : 0.f;
        const T h = O::cvt(v);
        rh[k] = h;
        if (TERMS >= 2) rl[k] = O::cvt(v - O::back(h));
      }
    }
  }
  __device__ static void load_b(B& b, const T* xh, const T* xl, int px, int quad) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      b.h[c] = *reinterpret_cast<const typename O::V8*>(xh + px * XS + 32 * c + 8 * quad);
      if (TERMS >= 2) b.l[c] = *reinterpret_cast<const typename O::V8*>(xl + px * XS + 32 * c + 8 * quad);
    }
  }
  __device__ static float4v contract(const W& w, const B& b, float4v acc) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      acc = O::mma(w.hi[c], b.h[c], acc);
      if (TERMS >= 2) acc = O::mma(w.hi[c], b.l[c], acc);
      if (TERMS >= 3) acc = O::mma(w.lo[c], b.h[c], acc);
    }
    return acc;
  }
};

// FP8 (v_mfma_scale_f32_16x16x128_f8f6f4, OCP e4m3 on both sides, one instruction per
// tile): the 128 K slots hold [window high parts (K) | window low parts (K) | zeros]
// against [weights (K) | the same weights (K) | zeros], so one product carries the split
// window.  Lane l holds A row l%16 / B column l%16 in slots 32*(l/16) + 0..31 (the same
// slots on both sides, so the slot order inside the instruction does not matter).
// Weights are stored as e4m3 of w * 2^-e_n with a power-of-two scale per neuron and
// matrix, applied by the instruction's per-lane E8M0 scale (127 + e_n); the window is
// stored as e4m3 of x' * 2^5 and scaled back by 2^-5 the same way.
constexpr int kFp8XExp = 5;
typedef int int8v __attribute__((ext_vector_type(8)));

template <int K, int TERMS>
struct Mma<F_FP8, K, TERMS> {
  typedef uint8_t T;
  static constexpr int XS = 128 + 16;  // bytes per LDS row
  static constexpr int LO_ROWS = 0;
  struct W {
    int8v a;
    int sc;
  };
  struct B {
    int8v b;
  };
  static constexpr bool BIAS_IN_K = false;
  __device__ static void load_w(W& w, const void* wh, const void* /*wl*/, const int* sc, const float* /*bias*/, int n,
                                int mtx, int quad) {
    const int4* src = reinterpret_cast<const int4*>(static_cast<const uint8_t*>(wh) +
                                                    (static_cast<int64_t>(n) * 2 + mtx) * 128 + 32 * quad);
    const int4 lo = src[0], hi = src[1];
    w.a = int8v{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
    w.sc = sc[n * 2 + mtx];
  }
  static constexpr bool FUSED_SCALE = false;
  __device__ static void write_row(T* rh, T* /*rl*/, const float* xv, float m, float is, float /*xscale*/,
                                   bool /*scaled*/) {
    const float XSC = static_cast<float>(1 << kFp8XExp) * is;
    uint32_t hv[(K + 3) / 4], lv[(K + 3) / 4];
#pragma unroll
    for (int k = 0; k < K; k += 2) {
      const float v0 = (xv[k] - m) * XSC, v1 = (xv[k + 1] - m) * XSC;
      const uint32_t h = __builtin_amdgcn_cvt_pk_fp8_f32(v0, v1, 0, false);
      const float r0 = v0 - __builtin_amdgcn_cvt_f32_fp8(static_cast<int>(h), 0);
      const float r1 = v1 - __builtin_amdgcn_cvt_f32_fp8(static_cast<int>(h), 1);
      const uint32_t l = __builtin_amdgcn_cvt_pk_fp8_f32(r0, r1, 0, false);
      if (k % 4 == 0) {
        hv[k / 4] = h & 0xffffu;
        lv[k / 4] = l & 0xffffu;
      } else {
        hv[k / 4] |= h << 16;
        lv[k / 4] |= l << 16;
      }
    }
    // slots [0, K) high parts, [K, 2K) low parts, [2K, 128) zero; K is a multiple of 16
    uint32_t* r = reinterpret_cast<uint32_t*>(rh);
#pragma unroll
    for (int i = 0; i < 32; ++i) r[i] = i < K / 4 ? hv[i] : (i < K / 2 ? lv[i - K / 4] : 0u);
  }
  __device__ static void load_b(B& b, const T* xh, const T* /*xl*/, int px, int quad) {
    const int4* src = reinterpret_cast<const int4*>(xh + px * XS + 32 * quad);
    const int4 lo = src[0], hi = src[1];
    b.b = int8v{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
  }
  __device__ static float4v contract(const W& w, const B& b, float4v acc) {
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(w.a, b.b, acc, 0, 0, 0, w.sc, 0, 127 - kFp8XExp);
  }
};

// Orders this wave's LDS writes before its later LDS reads (a wave's LDS operations run in
// issue order; this keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Cross-lane sums in VGPRs (v_permlane{16,32}_swap, no LDS round trip), two values at a
// time, each result scattered to the lanes that own it.  pair32_sum: lanes 0-31 get
// a[l] + a[l ^ 32], lanes 32-63 b[l] + b[l ^ 32].  pair16_sum, by rows of 16 lanes: rows 0
// and 2 get a[l] + a[l ^ 16], rows 1 and 3 b[l] + b[l ^ 16].  (The swap exchanges the upper
// half / odd rows of its first operand with the lower half / even rows of its second.)
// Each sum adds the same two operands as v + shfl_xor(v, 32 / 16) does.
__device__ __forceinline__ float pair32_sum(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ float pair16_sum(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
// Four values' sums over the lanes l, l^16, l^32, l^48, scattered: lanes of row q (16q ..
// 16q+15) get the sum of p[q], added as (p(l) + p(l^16)) + (p(l^32) + p(l^48)) like two
// shfl_xor steps (16, then 32).
__device__ __forceinline__ float quad_sum_scatter(const float p[4]) {
  return pair32_sum(pair16_sum(p[0], p[1]), pair16_sum(p[2], p[3]));
}

// Block: 4 waves; a staged tile of NX x TY = 64 x 8 (or 64 x 16) pixels is computed as
// TY / NY sub-tiles of NX x NY = 64 x 4, wave w owning pixel row w of each (pixel groups
// 4w..4w+3 of 16): one halo load and one block barrier serve 512 (1024) pixels (win8x6
// forms: 4-6 % faster than 64 x 4 tiles; 64 x 16: headline -3 %, nns64 -5 to -10 %).  The NT = nns/16 neuron tiles are split into NSL slices; TPW = NT/NSL
// neuron tiles per wave, held in VGPRs for the whole launch.
//  * NSL = 1 (nns <= 64): every wave runs all neurons for its own 64 pixels, so the
//    normalised rows it writes to LDS are read back by the same wave and the softmax sums
//    stay in registers: one block barrier per staged tile (after the halo tile lands).
//  * NSL > 1: waves with the same neuron slice share the 16 pixel groups (4/NSL pixel
//    slices); partial sums meet in LDS.
// TERMS = 1: one product of 16-bit operands; TERMS = 2: high*high + high*low (split
// window); TERMS = 3: high*high + high*low + low*high.
// The window operand is x' * XSCALE and the weights carry WSCALE (host side), so the
// accumulators hold ACC = XSCALE*WSCALE*(s + b): powers of two that keep the FP16 low
// parts out of the subnormal range (MFMA flushes FP16 subnormal inputs).
// The halo tile is double-buffered: the block's next tile is fetched into registers while
// this one computes and lands in the other buffer.
template <int WH, int AXIS, int FORM, int TERMS, int NSL, int TPW, bool SCALED, int TY = NYS>
__global__ __launch_bounds__(256, 2) void nnedi3_mfma(NPass p, const void* __restrict__ w_hi,
                                                      const void* __restrict__ w_lo, const int* __restrict__ wsc,
                                                      const float* __restrict__ bias, float xscale, float inv_acc,
                                                      int total_tiles) {
  using G = Geo<WH, AXIS, TY>;
  constexpr int K = G::K;
  using M = Mma<FORM, K, TERMS>;
  using T = typename M::T;
  constexpr int XS = M::XS;
  constexpr int NPX = NX * NY;            // 256 pixels per tile
  constexpr int NGROUP = NPX / 16;        // 16 pixel groups
  constexpr int PSL = 4 / NSL;            // pixel slices
  using LS = LineStats<WH, AXIS, TY>;
  // the one-product FP16 forms on 8x6 windows take the window statistics from shared line
  // statistics (their variance is good to a few ulp; the split forms, which resolve ~22 bits,
  // keep the per-pixel two-pass variance, good to one).  On 8x4 windows the table's build
  // pass and barrier cost more than the 32-term sums they save where a wave holds one or two
  // neuron tiles (nns 16 / 32: 2-3 % slower) and less from four on (nns 64: -2 %, nns 256:
  // -1 %).
  constexpr bool LINES = M::FUSED_SCALE && TERMS == 1 && (WH == 6 || NSL * TPW >= 4);
  __shared__ float tile[2][G::TH * G::TW];
  __shared__ float2 stab[LINES ? LS::N : 1];   // line statistics of the staged tile
  __shared__ __attribute__((aligned(16))) T xhi[NPX * XS];
  __shared__ __attribute__((aligned(16))) T xlo[M::LO_ROWS ? NPX * XS : 16];
  __shared__ float part[NSL > 1 ? NSL : 1][NSL > 1 ? NPX : 1][2];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nslice = wave % NSL;
  const int pslice = wave / NSL;
  const int row = lane & 15, quad = lane >> 4;

  // this wave's weight fragments and the accumulator initial values (scaled biases)
  typename M::W w1[TPW], w2[TPW];
  float4v b1[TPW], b2[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int n = (nslice * TPW + t) * 16 + row;          // neuron of this lane's A row
    M::load_w(w1[t], w_hi, w_lo, wsc, bias, n, 0, quad);
    M::load_w(w2[t], w_hi, w_lo, wsc, bias, n, 1, quad);
    // D rows (neurons) of this lane: 4*quad + i; the accumulators start at the biases
    // unless the bias rides in the contraction's spare K slots
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int nn = (nslice * TPW + t) * 16 + 4 * quad + i;
      b1[t][i] = M::BIAS_IN_K ? 0.f : bias[2 * nn];
      b2[t][i] = M::BIAS_IN_K ? 0.f : bias[2 * nn + 1];
    }
  }
  // Softmax-weighted elliott sum with one reciprocal per 8 neurons (4 when a wave holds
  // one neuron tile): sum_i e_i t_i / p_i with p_i = 1 + |t_i| is carried as a fraction
  // N / D, combined pairwise ((N1 D2 + N2 D1) / (D1 D2)).  Bounds: with the window
  // normalised (sum x' = 0, sum x'^2 = K), |s| <= sqrt(K) |w| + |b| (Cauchy-Schwarz), at
  // most 63 for the elliott half and 25.2 for the exp2 half over every shipped network,
  // so D <= 64^8 = 2^48 and N <= 8 * 2^26 * 64^8 = 2^77: far inside FP32.
  // For SCALED forms the accumulators carry ACC = s * acc_scale: exp2 takes acc*inv_acc,
  // p = 1 + |acc| * inv_acc, and N (hence vs) carries acc_scale, removed at the end.
  auto quad_frac = [&](const float4v& a1, const float4v& a2, float& ws, float& n, float& d) {
    float e[4], q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      e[i] = __builtin_amdgcn_exp2f(SCALED ? a1[i] * inv_acc : a1[i]);
      ws += e[i];
      q[i] = SCALED ? fmaf(fabsf(a2[i]), inv_acc, 1.0f) : 1.0f + fabsf(a2[i]);
    }
    const float n01 = fmaf(e[0] * a2[0], q[1], e[1] * a2[1] * q[0]), d01 = q[0] * q[1];
    const float n23 = fmaf(e[2] * a2[2], q[3], e[3] * a2[3] * q[2]), d23 = q[2] * q[3];
    n = fmaf(n01, d23, n23 * d01);
    d = d01 * d23;
  };

  // softmax-weighted sums of one pixel group over this wave's neuron tiles: this lane's
  // part (4 neurons per tile) of the sums of pixel g*16 + row; lanes l, l^16, l^32, l^48
  // hold the parts of the same pixel
  auto group_sums = [&](int g, float& ws, float& vs) {
    typename M::B bx;
    M::load_b(bx, xhi, xlo, g * 16 + row, quad);  // B column of this lane: pixel g*16 + row
    ws = 0.f;
    vs = 0.f;
    float n = 0.f, d = 1.f;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const float4v a1 = M::contract(w1[t], bx, b1[t]);
      const float4v a2 = M::contract(w2[t], bx, b2[t]);
      // The epilogue (exp2, fractions, sums) issues at raised wave priority: the SIMD then
      // prefers it over another wave's contraction, whose MFMAs fill the gaps.  Measured
      // 3-5 % faster (headline -3 %); the four-slice instances (nns 256, FP8) gain nothing.
      if constexpr (NSL < 4) __builtin_amdgcn_s_setprio(1);
      float nq, dq;
      quad_frac(a1, a2, ws, nq, dq);
      if (t % 2 == 0) {
        n = nq;
        d = dq;
      } else {
        n = fmaf(n, dq, nq * d);
        d *= dq;
      }
      if (t % 2 == 1 || t == TPW - 1) vs = fmaf(n, __builtin_amdgcn_rcpf(d), vs);
      if constexpr (NSL < 4) __builtin_amdgcn_s_setprio(0);
    }
    if constexpr (SCALED) vs *= inv_acc;
  };

  NTile<WH, AXIS, TY> ht;
  auto issue = [&](int id) {
    const PassTile q = pass_tile<TY>(p, id);
    ht.issue(frame_src(p, q.f), p.src_dt, q.sp, 0, q.W, q.H, q.x0 + G::OX, q.y0 + G::OY);
  };
  // launch-order index L = blockIdx + i*gridDim walks [0, total); tile = xcd_tile(L)
  if (static_cast<int>(blockIdx.x) < total_tiles) issue(xcd_tile(blockIdx.x, total_tiles));
  int buf = 0;
  for (int L = blockIdx.x; L < total_tiles; L += gridDim.x, buf ^= 1) {
    const int tileid = xcd_tile(L, total_tiles);
    const PassTile pt = pass_tile<TY>(p, tileid);
    const int f = pt.f;
    const int y0 = pt.y0, x0 = pt.x0;
    // The other buffer was last read two tiles ago, before the barrier below was last
    // passed by every wave.
    ht.commit(tile[buf], p.src_dt, p.src_max, p.src_rmax);
    __syncthreads();
    if (L + static_cast<int>(gridDim.x) < total_tiles) issue(xcd_tile(L + gridDim.x, total_tiles));
    if constexpr (LINES) {
      // (the previous tile's reads of stab ended before the barrier above)
      LS::build(tile[buf], stab);
      __syncthreads();
    }

    // the staged tile holds TY rows: TY / NY sub-tiles of NX x NY pixels, one pixel per
    // thread each, sharing one halo load and one block barrier
#pragma unroll 1
    for (int sub = 0; sub < TY / NY; ++sub) {
      // one thread per pixel: normalised window -> 16-bit rows (hi, lo), zero padded to KP
      const int tx = threadIdx.x % NX, ty = threadIdx.x / NX + sub * NY;
      float m, sd;
      float orig;
      {
        const float* tl = tile[buf];
        float xv[K];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
          for (int b = 0; b < WH; ++b) xv[a * WH + b] = tl[G::at(tx, ty, a, b)];
        orig = tl[G::centre(tx, ty)];
        float is;
        if constexpr (LINES) {
          float v;
          LS::window(stab, tx, ty, m, v);
          // hardware reciprocal square root (within 1 ulp); the threshold test is exact
          is = v >= kVarEps ? __builtin_amdgcn_rsqf(v) : 0.f;
          sd = v * is;
        } else {
          // two-pass: the window comes back centred (FUSED_SCALE) or normalised
          normalise<K, !M::FUSED_SCALE>(xv, m, sd, is);
        }
        M::write_row(xhi + threadIdx.x * XS, xlo + (M::LO_ROWS ? threadIdx.x * XS : 0), xv, LINES ? m : 0.f,
                     (LINES || M::FUSED_SCALE) ? is : 1.f, xscale, SCALED);
      }

      float wsum = 0.f, vsum = 0.f;
      if constexpr (NSL == 1) {
        // this wave's own rows: groups 4*wave + j; the lanes of quad j own pixel
        // (4*wave + j)*16 + row, their thread's pixel
        wave_lds_sync();
        float ws[4], vs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) group_sums(wave * 4 + j, ws[j], vs[j]);
        wsum = quad_sum_scatter(ws);
        vsum = quad_sum_scatter(vs);
      } else {
        // (also orders the previous sub-tile's reads of `part` before this one's writes)
        __syncthreads();
        static_assert((NGROUP / PSL) % 4 == 0, "groups reduced four at a time");
        for (int g = pslice; g < NGROUP; g += 4 * PSL) {
          float ws[4], vs[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) group_sums(g + j * PSL, ws[j], vs[j]);
          // the lanes of quad j hold the sums of group g + j*PSL
          const int gq = g + quad * PSL;
          part[nslice][gq * 16 + row][0] = quad_sum_scatter(ws);
          part[nslice][gq * 16 + row][1] = quad_sum_scatter(vs);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NSL; ++s) {
          wsum += part[s][threadIdx.x][0];
          vsum += part[s][threadIdx.x][1];
        }
      }

      const int x = x0 + tx, y = y0 + ty;
      if (x < pt.W && y < pt.H) {
        const float pred = fminf(fmaxf(m + 5.0f * vsum * __builtin_amdgcn_rcpf(wsum) * sd, 0.f), 1.f);
        store_result<AXIS>(p, frame_dst(p, f), pt.dp, x, y, orig, pred);
      }
      // this sub-tile's row reads before the next sub-tile's (or tile's) writes
      if constexpr (NSL == 1) wave_lds_sync();
    }
  }
}

// ----------------------------------------------------------------------------- MFMA, 32x32 tiles

typedef float float16v __attribute__((ext_vector_type(16)));

// The contraction on v_mfma_f32_32x32x16_{f16,bf16}: K in chunks of 16, so the win8x6
// window (K = 48) takes exactly three chunks where the 16x16x32 form pads it to 64 (a
// quarter of its matrix work is zeros).  Lane l holds A row (neuron) l%32 and B column
// (pixel) l%32 at k = 16*c + 8*(l/32) + 0..7, the same slots on both sides.  D: lane l
// holds column (pixel) l%32, rows 8*(r/4) + 4*(l/32) + r%4 for r = 0..15: 16 neurons of
// one pixel.  No spare K slots: the biases are the accumulators' initial value (C).
template <int FORM, int K, int TERMS>
struct Mma32 {
  using O = Op<FORM>;
  using T = typename O::T;
  using V8 = typename O::V8;
  static_assert(K % 16 == 0, "K must be a multiple of 16");
  static constexpr int KC = K / 16;
  static constexpr int XS = K + 8;  // LDS row stride (elements): 16-byte aligned rows, spread banks
  static constexpr int LO_ROWS = TERMS >= 2;
  static constexpr bool FUSED_SCALE = FORM == F_F16 && TERMS <= 2;
  struct W {
    V8 hi[KC], lo[KC];
  };
  struct B {
    V8 h[KC], l[KC];
  };
  __device__ static float16v mma(V8 a, V8 b, float16v c) {
    if constexpr (FORM == F_F16) return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static void load_w(W& w, const void* wh, const void* wl, int n, int mtx, int half) {
    const T* H = static_cast<const T*>(wh);
    const T* L = static_cast<const T*>(wl);
#pragma unroll
    for (int c = 0; c < KC; ++c)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t i = (static_cast<int64_t>(n) * 2 + mtx) * K + 16 * c + 8 * half + j;
        w.hi[c][j] = H[i];
        w.lo[c][j] = TERMS >= 3 ? L[i] : static_cast<T>(0.0f);
      }
  }
  // one pixel's window (raw samples, mean m, is = 1/std) -> its LDS row(s), as Mma::write_row
  __device__ static void write_row(T* rh, T* rl, const float* xv, float m, float is, float xscale, bool scaled) {
    const float c = -m * is;
    const float sc = scaled ? is * xscale : is;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float v = FUSED_SCALE ? fmaf(xv[k], is, c) : (xv[k] - m) * sc;
      const T h = O::cvt(v);
      rh[k] = h;
      if (TERMS >= 2) rl[k] = O::cvt(v - O::back(h));
    }
  }
  __device__ static void load_b(B& b, const T* xh, const T* xl, int px, int half) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      b.h[c] = *reinterpret_cast<const V8*>(xh + px * XS + 16 * c + 8 * half);
      if (TERMS >= 2) b.l[c] = *reinterpret_cast<const V8*>(xl + px * XS + 16 * c + 8 * half);
    }
  }
  __device__ static float16v contract(const W& w, const B& b, float16v acc) {
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      acc = mma(w.hi[c], b.h[c], acc);
      if (TERMS >= 2) acc = mma(w.hi[c], b.l[c], acc);
      if (TERMS >= 3) acc = mma(w.lo[c], b.h[c], acc);
    }
    return acc;
  }
};

// nnedi3_mfma on 32x32 tiles (F16 / BF16 forms).  Same block, staged tile, persistent
// loop and double-buffered halo as nnedi3_mfma; the 256 pixels of a sub-tile form 8
// groups of 32, and the nns/32 neuron tiles are split into NSL slices of TPW tiles per
// wave (weights and biases in VGPRs for the launch).  NSL = 1: wave w owns groups 2w and
// 2w+1, its own pixel row, so its LDS rows and sums stay within the wave; NSL > 1: waves
// of one slice share the groups, partial sums meet in LDS.
template <int WH, int AXIS, int FORM, int TERMS, int NSL, int TPW, bool SCALED, int TY = NYS>
__global__ __launch_bounds__(256, 2) void nnedi3_mfma32(NPass p, const void* __restrict__ w_hi,
                                                        const void* __restrict__ w_lo, const float* __restrict__ bias,
                                                        float xscale, float inv_acc, int total_tiles) {
  using G = Geo<WH, AXIS, TY>;
  constexpr int K = G::K;
  using M = Mma32<FORM, K, TERMS>;
  using T = typename M::T;
  constexpr int XS = M::XS;
  constexpr int NPX = NX * NY;      // 256 pixels per sub-tile
  constexpr int NGROUP = NPX / 32;  // 8 pixel groups
  constexpr int PSL = 4 / NSL;      // pixel slices
  static_assert(NSL * PSL == 4, "4 waves per block");
  using LS = LineStats<WH, AXIS, TY>;
  // the one-product FP16 forms on 8x6 windows take the window statistics from shared line
  // statistics (their variance is good to a few ulp; the split forms, which resolve ~22 bits,
  // keep the per-pixel two-pass variance, good to one).  On 8x4 windows the table's build
  // pass and barrier cost more than the 32-term sums they save (measured 1-3% slower).
  constexpr bool LINES = M::FUSED_SCALE && TERMS == 1 && WH == 6;
  __shared__ float tile[2][G::TH * G::TW];
  __shared__ float2 stab[LINES ? LS::N : 1];   // line statistics of the staged tile
  __shared__ __attribute__((aligned(16))) T xhi[NPX * XS];
  __shared__ __attribute__((aligned(16))) T xlo[M::LO_ROWS ? NPX * XS : 8];
  __shared__ float part[NSL > 1 ? NSL : 1][NSL > 1 ? NPX : 1][2];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nslice = wave % NSL;
  const int pslice = wave / NSL;
  const int col = lane & 31, half = lane >> 5;

  typename M::W w1[TPW], w2[TPW];
  float16v b1[TPW], b2[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int n0 = (nslice * TPW + t) * 32;
    M::load_w(w1[t], w_hi, w_lo, n0 + col, 0, half);
    M::load_w(w2[t], w_hi, w_lo, n0 + col, 1, half);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int nn = n0 + 8 * (r / 4) + 4 * half + (r % 4);  // D row of this lane's value r
      b1[t][r] = bias[2 * nn];
      b2[t][r] = bias[2 * nn + 1];
    }
  }
  // Four neurons' softmax weights (added to ws) and elliott terms as one fraction n / d
  // (bounds as in nnedi3_mfma: d <= 64^4, n <= 4 * 2^26 * 64^4 per quad).
  auto quad_frac = [&](const float16v& a1, const float16v& a2, int r0, float& ws, float& n, float& d) {
    float e[4], q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      e[i] = __builtin_amdgcn_exp2f(SCALED ? a1[r0 + i] * inv_acc : a1[r0 + i]);
      ws += e[i];
      q[i] = SCALED ? fmaf(fabsf(a2[r0 + i]), inv_acc, 1.0f) : 1.0f + fabsf(a2[r0 + i]);
    }
    const float n01 = fmaf(e[0] * a2[r0], q[1], e[1] * a2[r0 + 1] * q[0]), d01 = q[0] * q[1];
    const float n23 = fmaf(e[2] * a2[r0 + 2], q[3], e[3] * a2[r0 + 3] * q[2]), d23 = q[2] * q[3];
    n = fmaf(n01, d23, n23 * d01);
    d = d01 * d23;
  };
  // softmax-weighted sums of pixel group g over this wave's neuron tiles: this lane's part
  // (16 neurons per tile) of the sums of pixel g*32 + l%32, the other part in lane l^32
  auto group_sums = [&](int g, float& ws, float& vs) {
    typename M::B bx;
    M::load_b(bx, xhi, xlo, g * 32 + col, half);
    ws = 0.f;
    vs = 0.f;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const float16v a1 = M::contract(w1[t], bx, b1[t]);
      const float16v a2 = M::contract(w2[t], bx, b2[t]);
      __builtin_amdgcn_s_setprio(1);  // epilogue at raised wave priority, as in nnedi3_mfma
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // one reciprocal per 8 neurons
        float n0, d0, n1, d1;
        quad_frac(a1, a2, 8 * h, ws, n0, d0);
        quad_frac(a1, a2, 8 * h + 4, ws, n1, d1);
        vs = fmaf(fmaf(n0, d1, n1 * d0), __builtin_amdgcn_rcpf(d0 * d1), vs);
      }
      __builtin_amdgcn_s_setprio(0);
    }
    if constexpr (SCALED) vs *= inv_acc;
  };

  NTile<WH, AXIS, TY> ht;
  auto issue = [&](int id) {
    const PassTile q = pass_tile<TY>(p, id);
    ht.issue(frame_src(p, q.f), p.src_dt, q.sp, 0, q.W, q.H, q.x0 + G::OX, q.y0 + G::OY);
  };
  if (static_cast<int>(blockIdx.x) < total_tiles) issue(xcd_tile(blockIdx.x, total_tiles));
  int buf = 0;
  for (int L = blockIdx.x; L < total_tiles; L += gridDim.x, buf ^= 1) {
    const PassTile pt = pass_tile<TY>(p, xcd_tile(L, total_tiles));
    const int f = pt.f;
    const int y0 = pt.y0, x0 = pt.x0;
    // the other buffer was last read two tiles ago, before the barrier below was last
    // passed by every wave
    ht.commit(tile[buf], p.src_dt, p.src_max, p.src_rmax);
    __syncthreads();
    if (L + static_cast<int>(gridDim.x) < total_tiles) issue(xcd_tile(L + gridDim.x, total_tiles));
    if constexpr (LINES) {
      // (the previous tile's reads of stab ended before the barrier above)
      LS::build(tile[buf], stab);
      __syncthreads();
    }

#pragma unroll 1
    for (int sub = 0; sub < TY / NY; ++sub) {
      const int tx = threadIdx.x % NX, ty = threadIdx.x / NX + sub * NY;
      float m, sd, orig;
      {
        const float* tl = tile[buf];
        float xv[K];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
          for (int b = 0; b < WH; ++b) xv[a * WH + b] = tl[G::at(tx, ty, a, b)];
        orig = tl[G::centre(tx, ty)];
        float is;
        if constexpr (LINES) {
          float v;
          LS::window(stab, tx, ty, m, v);
          // hardware reciprocal square root (within 1 ulp); the threshold test is exact
          is = v >= kVarEps ? __builtin_amdgcn_rsqf(v) : 0.f;
          sd = v * is;
        } else {
          // two-pass: the window comes back centred (FUSED_SCALE) or normalised
          normalise<K, !M::FUSED_SCALE>(xv, m, sd, is);
        }
        M::write_row(xhi + threadIdx.x * XS, xlo + (M::LO_ROWS ? threadIdx.x * XS : 0), xv, LINES ? m : 0.f,
                     (LINES || M::FUSED_SCALE) ? is : 1.f, xscale, SCALED);
      }

      float wsum = 0.f, vsum = 0.f;
      if constexpr (NSL == 1) {
        // groups 2*wave + j; the lanes of half j own pixel (2*wave + j)*32 + col, their
        // thread's pixel.  The three-product forms sit at the register limit, where the
        // swap's second result register makes them spill more (BF16 nns64 +10 %): they
        // keep the LDS shuffle, one group at a time.
        wave_lds_sync();
        if constexpr (TERMS == 3) {
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            float ws, vs;
            group_sums(wave * 2 + j, ws, vs);
            ws += __shfl_xor(ws, 32);
            vs += __shfl_xor(vs, 32);
            if (half == j) {
              wsum = ws;
              vsum = vs;
            }
          }
        } else {
          float ws[2], vs[2];
#pragma unroll
          for (int j = 0; j < 2; ++j) group_sums(wave * 2 + j, ws[j], vs[j]);
          wsum = pair32_sum(ws[0], ws[1]);
          vsum = pair32_sum(vs[0], vs[1]);
        }
      } else {
        // (also orders the previous sub-tile's reads of `part` before this one's writes)
        __syncthreads();
        static_assert((NGROUP / PSL) % 2 == 0, "groups reduced two at a time");
        for (int g = pslice; g < NGROUP; g += 2 * PSL) {
          float ws[2], vs[2];
#pragma unroll
          for (int j = 0; j < 2; ++j) group_sums(g + j * PSL, ws[j], vs[j]);
          // the lanes of half j hold the sums of group g + j*PSL
          const int gh = g + half * PSL;
          part[nslice][gh * 32 + col][0] = pair32_sum(ws[0], ws[1]);
          part[nslice][gh * 32 + col][1] = pair32_sum(vs[0], vs[1]);
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < NSL; ++s) {
          wsum += part[s][threadIdx.x][0];
          vsum += part[s][threadIdx.x][1];
        }
      }

      const int x = x0 + tx, y = y0 + ty;
      if (x < pt.W && y < pt.H) {
        const float pred = fminf(fmaxf(m + 5.0f * vsum * __builtin_amdgcn_rcpf(wsum) * sd, 0.f), 1.f);
        store_result<AXIS>(p, frame_dst(p, f), pt.dp, x, y, orig, pred);
      }
      if constexpr (NSL == 1) wave_lds_sync();
    }
  }
}

// ----------------------------------------------------------------------------- launcher

template <int WH, int AXIS>
static hipError_t run_pass(const NnediWeights& nw, NPass& p, hipStream_t st, hipEvent_t done) {
  // tiles of NX x ty pixels (ty = NY for the plain kernel, the staged rows TY for the
  // persistent MFMA kernels); fills the per-frame prefix of sized launches
  auto count_tiles = [&](int ty) {
    if (!p.sized) return ((p.W + NX - 1) / NX) * ((p.H + ty - 1) / ty) * p.batch;
    p.tiles[0] = 0;
    for (int f = 0; f < p.ntab_src; ++f) p.tiles[f + 1] = p.tiles[f] + ((p.fW[f] + NX - 1) / NX) * ((p.fH[f] + ty - 1) / ty);
    return p.tiles[p.ntab_src];
  };
  if (nw.precision == 0) {
    const int tiles = count_tiles(NY);
    return launch_done(nnedi3_valu<WH, AXIS>, dim3(tiles), dim3(256), st, done, p, static_cast<const float*>(nw.w),
                       nw.nns, tiles);
  }
  int total = 0;                         // tiles of the launched MFMA instance (launch_inst)
  hipError_t le = hipErrorInvalidValue;  // the MFMA launch's status
  const int nt = nw.nns / 16;
  const float inv_acc = 1.0f / (nw.xscale * nw.wscale);
  // Neuron tiles are split over NSL waves only when one wave cannot hold them all
  // (nns >= 128); otherwise every wave owns all neurons for its own pixel groups.
  // Persistent grid: as many blocks as the occupancy of the instance allows per CU (or the
  // tuned blocks per CU below it).
  auto launch_inst = [&](auto kern, int per_cu, int ty, auto... args) {
    total = count_tiles(ty);
    const int k = nw.grid_per_cu > 0 ? min(nw.grid_per_cu, per_cu) : per_cu;
    const int grid = min(total, device_cu_count() * k);
    return launch_done(kern, dim3(grid), dim3(256), st, done, p, args..., total);
  };
  // The one-product FP16 forms also have 16-row staged tiles: taken where the instance keeps
  // its occupancy (LDS, VGPRs) and the grid's blocks still get 1.5 tiles each or more (a
  // 1080p frame's double_y pass: 2040 tiles for 1024 blocks; small frames keep 8 rows,
  // which spreads them over more blocks).
  auto tall_pays = [&](int occ8, int occ16) {
    const int k = nw.grid_per_cu > 0 ? min(nw.grid_per_cu, occ16) : occ16;
    return occ16 >= occ8 && 2 * count_tiles(16) >= 3 * device_cu_count() * k;
  };
  // Occupancy of an instance, queried once (no runtime queries while capturing); callers
  // may be on several threads: an atomic cache (a race only repeats the same query).
#define MPX_OCC(var, kern)                                                                           \
  static std::atomic<int> var##_cache{0};                                                            \
  int var = var##_cache.load(std::memory_order_relaxed);                                             \
  if (!var) {                                                                                        \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&var, kern, 256, 0) != hipSuccess || var < 1)   \
      var = 1;                                                                                       \
    var##_cache.store(var, std::memory_order_relaxed);                                               \
  }
#define MPX_NN(FM, TERMS, NSL, TPW, SC)                                                               \
  {                                                                                                   \
    auto k8 = nnedi3_mfma<WH, AXIS, FM, TERMS, NSL, TPW, SC, 8>;                                      \
    MPX_OCC(o8, k8)                                                                                   \
    bool tall = false;                                                                                \
    if constexpr (FM == F_F16 && TERMS == 1) {                                                        \
      auto k16 = nnedi3_mfma<WH, AXIS, FM, TERMS, NSL, TPW, SC, 16>;                                  \
      MPX_OCC(o16, k16)                                                                               \
      if ((tall = tall_pays(o8, o16)))                                                                \
        le = launch_inst(k16, o16, 16, nw.w, nw.w_lo, nw.wexp, nw.bias, nw.xscale, inv_acc);          \
    }                                                                                                 \
    if (!tall) le = launch_inst(k8, o8, 8, nw.w, nw.w_lo, nw.wexp, nw.bias, nw.xscale, inv_acc);      \
  }
#define MPX_NN_NT(FM, TERMS, SC)                    \
  switch (nt) {                                     \
    case 1: MPX_NN(FM, TERMS, 1, 1, SC) break;      \
    case 2: MPX_NN(FM, TERMS, 1, 2, SC) break;      \
    case 4:                                         \
      if (nw.split) MPX_NN(FM, TERMS, 2, 2, SC)     \
      else MPX_NN(FM, TERMS, 1, 4, SC)              \
      break;                                        \
    case 8:                                         \
      if (nw.split) MPX_NN(FM, TERMS, 4, 2, SC)     \
      else MPX_NN(FM, TERMS, 2, 4, SC)              \
      break;                                        \
    case 16: MPX_NN(FM, TERMS, 4, 4, SC) break;     \
    default: return hipErrorInvalidValue;           \
  }
  // 32x32x16 tiles for the F16 / BF16 forms with nns >= 32 when they pay (measured, 8 x
  // 1080p): win8x6 (K = 48 in three chunks of 16 instead of two of 32 half padded; nns256
  // FP16 +8 %, BF16 +25 %, nns128 FP16 +61 %) and the three-product win8x4 forms (BF16
  // nns128 +13 %); the one-product win8x4 forms keep 16x16x32 (K = 32 unpadded there, and
  // the 32x32 form measured 3-9 % slower)
#define MPX_NN32(FM, TERMS, NSL, TPW, SC)                                                             \
  {                                                                                                   \
    auto k8 = nnedi3_mfma32<WH, AXIS, FM, TERMS, NSL, TPW, SC, 8>;                                    \
    MPX_OCC(o8, k8)                                                                                   \
    bool tall = false;                                                                                \
    if constexpr (FM == F_F16 && TERMS == 1) {                                                        \
      auto k16 = nnedi3_mfma32<WH, AXIS, FM, TERMS, NSL, TPW, SC, 16>;                                \
      MPX_OCC(o16, k16)                                                                               \
      if ((tall = tall_pays(o8, o16))) le = launch_inst(k16, o16, 16, nw.w, nw.w_lo, nw.bias, nw.xscale, inv_acc); \
    }                                                                                                 \
    if (!tall) le = launch_inst(k8, o8, 8, nw.w, nw.w_lo, nw.bias, nw.xscale, inv_acc);               \
  }
#define MPX_NN32_NT(FM, TERMS, SC)                  \
  switch (nw.nns / 32) {                            \
    case 1: MPX_NN32(FM, TERMS, 1, 1, SC) break;    \
    case 2:                                         \
      if (nw.split) MPX_NN32(FM, TERMS, 2, 1, SC)   \
      else MPX_NN32(FM, TERMS, 1, 2, SC)            \
      break;                                        \
    case 4:                                         \
      if (nw.split) MPX_NN32(FM, TERMS, 4, 1, SC)   \
      else MPX_NN32(FM, TERMS, 2, 2, SC)            \
      break;                                        \
    case 8: MPX_NN32(FM, TERMS, 4, 2, SC) break;    \
    default: return hipErrorInvalidValue;           \
  }
  {
    const bool three = nw.precision == 2 || nw.precision == 4;
    if (nw.nns >= 32 && (three || (WH == 6 && nw.precision == 1))) {
      if (nw.precision == 1) {
        if (nw.nns == 128) MPX_NN32_NT(F_F16, 2, false)
        else MPX_NN32_NT(F_F16, 1, false)
      } else if (nw.precision == 4) {
        MPX_NN32_NT(F_F16, 3, true)
      } else {
        MPX_NN32_NT(F_BF16, 3, false)
      }
      return le;
    }
  }
#undef MPX_NN32_NT
#undef MPX_NN32
  if (nw.precision == 1) {
    // FP16 form: one product, except nns 128, whose weights make the window's FP16
    // rounding alone exceed 2^-10 max-abs on some inputs (1.2e-3 measured): there the
    // window is split into high and low parts (2 products; error then 6.4e-4)
    if (nw.nns == 128) {
      if (nw.split) MPX_NN(F_F16, 2, 4, 2, false)
      else MPX_NN(F_F16, 2, 2, 4, false)
    } else {
      MPX_NN_NT(F_F16, 1, false)
    }
  } else if (nw.precision == 4) {
    MPX_NN_NT(F_F16, 3, true)
  } else if (nw.precision == 2) {
    MPX_NN_NT(F_BF16, 3, false)
  } else if (nw.precision == 3) {
    MPX_NN_NT(F_FP8, 1, false)
  } else {
    return hipErrorInvalidValue;
  }
#undef MPX_NN_NT
#undef MPX_NN
#undef MPX_OCC
  return le;
}

hipError_t launch_nnedi3(const NnediWeights& nw, const Io& io, bool dx, bool dy, float* tmp, int64_t tmp_frame,
                         hipStream_t st, hipEvent_t done) {
  hipError_t e = hipSuccess;
  if (dy) {
    NPass p{};
    p.src = io.in;
    p.src_pitch = io.in_pitch;
    p.src_bstride = io.in_bstride;
    p.src_dt = io.in_dt;
    p.src_max = io.in_max;
    p.src_rmax = io.in_rmax;
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
      p.dst_dt = io.store16 ? DT_F16 : DT_F32;
    } else {
      p.dst = io.out;
      p.dst_pitch = io.out_pitch;
      p.dst_bstride = io.out_bstride;
      p.dst_dt = io.out_dt;
    }
    if (io.sized) {  // frames of different sizes: per-frame geometry, intermediates through tables
      p.sized = 1;
      p.ntab_dst = io.ntab;
      for (int f = 0; f < io.ntab; ++f) {
        p.fW[f] = io.fw[f];
        p.fH[f] = io.fh[f];
        p.fsp[f] = io.fin_pitch[f];
        p.fdp[f] = dx ? io.fw[f] : io.fout_pitch[f];
        p.tab_dst[f] = dx ? static_cast<void*>(tmp + f * tmp_frame) : io.tab_out[f];
      }
    }
    // (the completion event goes with the last pass)
    hipEvent_t dn = dx ? nullptr : done;
    e = nw.wh == 4 ? run_pass<4, 0>(nw, p, st, dn) : run_pass<6, 0>(nw, p, st, dn);
    if (e != hipSuccess) return e;
  }
  if (dx) {
    NPass p{};
    if (dy) {
      p.src = tmp;
      p.src_pitch = io.W;
      p.src_bstride = tmp_frame;
      p.src_dt = io.store16 ? DT_F16 : DT_F32;
      p.src_max = 1.f;
      p.src_rmax = 1.f;
      p.H = 2 * io.H;
    } else {
      p.src = io.in;
      p.src_pitch = io.in_pitch;
      p.src_bstride = io.in_bstride;
      p.src_dt = io.in_dt;
      p.src_max = io.in_max;
      p.src_rmax = io.in_rmax;
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
    if (io.sized) {
      p.sized = 1;
      p.ntab_src = io.ntab;
      for (int f = 0; f < io.ntab; ++f) {
        p.fW[f] = io.fw[f];
        p.fH[f] = dy ? 2 * io.fh[f] : io.fh[f];
        p.fsp[f] = dy ? io.fw[f] : io.fin_pitch[f];
        p.fdp[f] = io.fout_pitch[f];
        p.tab_src[f] = dy ? static_cast<const void*>(tmp + f * tmp_frame) : io.tab_in[f];
      }
    }
    e = nw.wh == 4 ? run_pass<4, 1>(nw, p, st, done) : run_pass<6, 1>(nw, p, st, done);
  }
  return e;
}

}  // namespace mpx
