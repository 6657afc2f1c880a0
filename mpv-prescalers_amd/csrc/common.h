// Shared device code for the RAVU family: element I/O, the structure tensor, the
// eigen-analysis hash with its FP64 re-decision near bucket boundaries.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gauss_tables.h"

namespace mpx {

enum DType : int { DT_U8 = 0, DT_U16 = 1, DT_F16 = 2, DT_F32 = 3 };

constexpr int kMaxTab = 16;  // frames per launch through a pointer table

// Device-side description of one launch's planes (strides in elements).  Frames are
// either `bstride` apart from `in` / `out`, or (ntab > 0) at the addresses of the
// pointer tables, which travel by value in the kernel arguments.
struct Io {
  const void* in;
  void* out;
  const void* tab_in[kMaxTab];
  void* tab_out[kMaxTab];
  int ntab;
  int64_t in_pitch, in_cstride, in_bstride;
  int64_t out_pitch, out_cstride, out_bstride;
  int32_t* buckets;   // optional, [batch][sites] rows
  int64_t bucket_bstride;
  int in_dt, out_dt;
  float in_max;       // integer full-scale value for U8/U16 inputs (255, 1023, ...)
  float in_rmax;      // RN(1 / in_max), computed on the host
  int W, H;           // input plane size
  int batch;
  // Frames of different sizes in one launch (sized != 0; ntab frames through the pointer
  // tables): each frame's input size and row pitches, planes contiguous.  `tiles` is the
  // prefix sum of the launch's tiles per frame, filled by the launcher for its tile shape.
  int sized;
  int fw[kMaxTab], fh[kMaxTab];
  int64_t fin_pitch[kMaxTab], fout_pitch[kMaxTab];
  int tiles[kMaxTab + 1];
  int sx, sy;         // output scale of the launch (for the output plane stride)
  // intermediates (NNEDI3's double_y plane, RAVU's int11) rounded to FP16 as mpv's
  // default rgba16f FBOs store them (MPX_STORE_FP16)
  int store16;
};

// One tile of a launch: its frame, origin and that frame's geometry.
struct FrameTile {
  int f, x0, y0;
  int W, H;
  int64_t in_pitch, out_pitch, in_cstride, out_cstride;
};

__device__ __host__ __forceinline__ int elem_bytes(int dt) { return dt == DT_F32 ? 4 : (dt == DT_U8 ? 1 : 2); }

__device__ __forceinline__ const void* frame_in(const Io& io, int f) {
  if (io.ntab) return io.tab_in[f];
  return static_cast<const char*>(io.in) + static_cast<int64_t>(f) * io.in_bstride * elem_bytes(io.in_dt);
}
__device__ __forceinline__ void* frame_out(const Io& io, int f) {
  if (io.ntab) return io.tab_out[f];
  return static_cast<char*>(io.out) + static_cast<int64_t>(f) * io.out_bstride * elem_bytes(io.out_dt);
}

// Each of the 8 XCDs (own L2 each) takes thread blocks round-robin in launch order.
// xcd_tile maps launch-order block L of T to a tile so that every XCD walks one contiguous
// range of tiles: neighbouring tiles, which share halo rows, then meet in the same L2.
// A bijection on [0, T) for any T.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int L, int T) {
  const int q = T / kXcds, r = T % kXcds;
  const int x = L % kXcds, j = L / kXcds;
  return x * q + min(x, r) + j;
}

// Tile (x, y) and frame z of this block in a 1-D launch of nx * ny * nz blocks.
struct BlockXYZ {
  int x, y, z;
};
__device__ __forceinline__ BlockXYZ block_xyz(int nx, int ny, int nz) {
  const int per = nx * ny;
  const int t = xcd_tile(static_cast<int>(blockIdx.x), per * nz);
  const int z = t / per, r = t - z * per;
  return {r % nx, r / nx, z};
}

// The TX x TY input tile of block L (launch order, XCD-aware) among `total` tiles.
template <int TX, int TY>
__device__ __forceinline__ FrameTile frame_tile(const Io& io, int L, int total) {
  const int t = xcd_tile(L, total);
  FrameTile ft;
  if (!io.sized) {
    const int nx = (io.W + TX - 1) / TX, per = nx * ((io.H + TY - 1) / TY);
    ft.f = t / per;
    const int r = t - ft.f * per;
    ft.x0 = (r % nx) * TX;
    ft.y0 = (r / nx) * TY;
    ft.W = io.W;
    ft.H = io.H;
    ft.in_pitch = io.in_pitch;
    ft.out_pitch = io.out_pitch;
    ft.in_cstride = io.in_cstride;
    ft.out_cstride = io.out_cstride;
    return ft;
  }
  int f = 0;
  while (f + 1 < io.ntab && t >= io.tiles[f + 1]) ++f;
  const int r = t - io.tiles[f];
  const int nx = (io.fw[f] + TX - 1) / TX;
  ft.f = f;
  ft.x0 = (r % nx) * TX;
  ft.y0 = (r / nx) * TY;
  ft.W = io.fw[f];
  ft.H = io.fh[f];
  ft.in_pitch = io.fin_pitch[f];
  ft.out_pitch = io.fout_pitch[f];
  ft.in_cstride = io.fin_pitch[f] * io.fh[f];
  ft.out_cstride = io.fout_pitch[f] * io.fh[f] * io.sy;
  return ft;
}

// Host: tiles of a launch for a TX x TY tile (fills the per-frame prefix of sized launches).
template <int TX, int TY>
inline int launch_tiles(Io& io) {
  if (!io.sized) return ((io.W + TX - 1) / TX) * ((io.H + TY - 1) / TY) * io.batch;
  io.tiles[0] = 0;
  for (int f = 0; f < io.ntab; ++f)
    io.tiles[f + 1] = io.tiles[f] + ((io.fw[f] + TX - 1) / TX) * ((io.fh[f] + TY - 1) / TY);
  return io.tiles[io.ntab];
}

// v / m for an integer sample v (0..65535) and full scale m = 2^bits - 1, rounded exactly
// as the division rounds, given rm = RN(1/m): q0 = v * rm lies within one ulp of v/m, the
// residual v - m*q0 is exact under fma, and one corrected step gives RN(v/m) (Markstein's
// theorem).  Three operations instead of the division's ten; equal to the division for
// every 16-bit v and every bits 1..16 (checked exhaustively).
__device__ __forceinline__ float unorm_quotient(float v, float m, float rm) {
  const float q0 = v * rm;
  return fmaf(fmaf(-q0, m, v), rm, q0);
}

__device__ __forceinline__ float load_elem(const void* p, int64_t idx, int dt, float in_max) {
  switch (dt) {
    case DT_F32: return static_cast<const float*>(p)[idx];
    case DT_F16: return __half2float(static_cast<const __half*>(p)[idx]);
    case DT_U8: return static_cast<float>(static_cast<const uint8_t*>(p)[idx]) / in_max;
    default: return static_cast<float>(static_cast<const uint16_t*>(p)[idx]) / in_max;
  }
}

// Halo tile of P planes, PH x PW samples each, taken at (ox + col, oy + r) clamped to the
// W x H plane (clamp-to-edge sampling) and staged to LDS as tile[c][r][col], by NT
// threads.  `issue` starts every load of the calling thread with no branch on the
// element type: each sample is fetched as the aligned dword that holds it (which never
// leaves the sample's own page), its bit offset kept beside it.  `commit` extracts and
// converts the samples and writes LDS.  A tile therefore costs one memory latency, and a
// persistent block can issue its next tile while it computes this one.
// Sample addresses are 32-bit element offsets from the frame (plane c * cstride + row *
// pitch + column, the row term a 24-bit multiply): the host refuses frames that span 2^32
// elements or more, or whose pitch reaches 2^24 elements (kMaxSpan, kMaxPitch).
// A pointer into global memory (address space 1): loads through it are global loads, never
// flat ones (which would also count against the LDS wait counter).
typedef const __attribute__((address_space(1))) uint32_t* GlobalU32;

template <int P, int PH, int PW, int NT>
struct HaloTile {
  static constexpr int PN = PH * PW;
  static constexpr int N = P * PN;
  static constexpr int E = (N + NT - 1) / NT;
  uint32_t v[E];
  uint32_t sh[E];

  __device__ __forceinline__ void issue(const void* base, int dt, int64_t pitch, int64_t cstride, int W, int H,
                                        int ox, int oy) {
    const int lsh = dt == DT_F32 ? 2 : (dt == DT_U8 ? 0 : 1);  // log2 element bytes
    const uintptr_t b = reinterpret_cast<uintptr_t>(base);
    const uint32_t ps = static_cast<uint32_t>(pitch), cs = static_cast<uint32_t>(cstride);
    if (ox >= 0 && oy >= 0 && ox + PW <= W && oy + PH <= H) {
      // interior tile (uniform branch): no clamping, and the per-thread part of each
      // offset (c*cstride + r*pitch + col) does not depend on the tile
      const uintptr_t t0 = b + (static_cast<uintptr_t>(__umul24(static_cast<uint32_t>(oy), ps) + ox) << lsh);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = min(static_cast<int>(threadIdx.x) + e * NT, N - 1);
        const int c = P == 1 ? 0 : i / PN;
        const int j = i - c * PN;
        const int r = j / PW, col = j - r * PW;
        const uint32_t off = static_cast<uint32_t>(c) * cs + __umul24(static_cast<uint32_t>(r), ps) + col;
        const uintptr_t a = t0 + (static_cast<uintptr_t>(off) << lsh);
        v[e] = *reinterpret_cast<GlobalU32>(a & ~static_cast<uintptr_t>(3));
        sh[e] = static_cast<uint32_t>(a & 3) * 8;
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = min(static_cast<int>(threadIdx.x) + e * NT, N - 1);  // tail lanes repeat the last sample
      const int c = P == 1 ? 0 : i / PN;
      const int j = i - c * PN;
      const int r = j / PW, col = j - r * PW;
      const int gx = min(max(ox + col, 0), W - 1);
      const int gy = min(max(oy + r, 0), H - 1);
      const uint32_t off = static_cast<uint32_t>(c) * cs + __umul24(static_cast<uint32_t>(gy), ps) + gx;
      const uintptr_t a = b + (static_cast<uintptr_t>(off) << lsh);
      v[e] = *reinterpret_cast<GlobalU32>(a & ~static_cast<uintptr_t>(3));
      sh[e] = static_cast<uint32_t>(a & 3) * 8;
    }
  }
  // in_rmax = RN(1 / in_max) (from the host): integer samples cost three operations each
  __device__ __forceinline__ void commit(float* tile, int dt, float in_max, float in_rmax) const {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int i = static_cast<int>(threadIdx.x) + e * NT;
      if (E * NT == N || i < N) {
        const uint32_t w = v[e] >> sh[e];
        float f;
        switch (dt) {
          case DT_F32: f = __uint_as_float(v[e]); break;
          case DT_F16: f = __half2float(__ushort_as_half(static_cast<unsigned short>(w & 0xffffu))); break;
          case DT_U8: f = unorm_quotient(static_cast<float>(w & 0xffu), in_max, in_rmax); break;
          default: f = unorm_quotient(static_cast<float>(w & 0xffffu), in_max, in_rmax); break;
        }
        tile[i] = f;
      }
    }
  }
};

__device__ __forceinline__ void store_elem(void* p, int64_t idx, int dt, float v) {
  if (dt == DT_F32)
    static_cast<float*>(p)[idx] = v;
  else
    static_cast<__half*>(p)[idx] = __float2half_rn(v);
}

// Two horizontally adjacent outputs (idx even-aligned by construction of the callers).
__device__ __forceinline__ void store_pair(void* p, int64_t idx, int dt, float v0, float v1) {
  if (dt == DT_F32) {
    float* q = static_cast<float*>(p) + idx;
    q[0] = v0;
    q[1] = v1;
  } else {
    __half* q = static_cast<__half*>(p) + idx;
    q[0] = __float2half_rn(v0);
    q[1] = __float2half_rn(v1);
  }
}

// two FP16 values packed in one dword (low half first)
__device__ __forceinline__ __half2 bits_h2(uint32_t v) {
  __half2 h;
  __builtin_memcpy(&h, &v, sizeof(h));
  return h;
}

__device__ __forceinline__ float4 load_lut4(const __half* lut, int64_t texel) {
  const __half2* p = reinterpret_cast<const __half2*>(lut + 4 * texel);
  float2 lo = __half22float2(p[0]);
  float2 hi = __half22float2(p[1]);
  return make_float4(lo.x, lo.y, hi.x, hi.y);
}

template <int N>
__device__ __host__ constexpr double gauss_w(int i) {
  return N == 3 ? kGauss3[i] : N == 4 ? kGauss4[i] : N == 5 ? kGauss5[i]
       : N == 6 ? kGauss6[i] : N == 7 ? kGauss7[i] : kGauss8[i];
}

// x / b for a constant b, rounded exactly as the division rounds (to nearest), given
// rb = RN(1/b).  FP64 takes three operations instead of the division's eleven (one of
// them a quarter-rate reciprocal): q0 = x * rb lies within one ulp of x/b, the residual
// x - b*q0 is exact under fma, and one corrected step gives RN(x/b) (Markstein's theorem;
// the operands here are far from underflow and overflow).  A zero quotient may come out
// +0 where the division gives -0, which no caller can tell apart (squares and sums).
template <typename T>
__device__ __forceinline__ T div_const(T x, T b, T rb) {
  if constexpr (sizeof(T) == 8) {
    const T q0 = x * rb;
    return __builtin_fma(__builtin_fma(-q0, b, x), rb, q0);
  }
  return x / b;
}

// Derivative along one window line at position p, in the shaders' formula forms
// (ravu-lite-r3.hook:50 central; ravu-r3.hook:64 five-point; ravu-zoom-r2.hook:50
// forward/backward at the window border).  `s(k)` returns the line's k-th sample.
template <int N, bool FIVE, typename T, typename Line>
__device__ __forceinline__ T deriv(Line s, int p) {
#pragma clang fp contract(off)
  if (p == 0) return s(1) - s(0);
  if (p == N - 1) return s(N - 1) - s(N - 2);
  // FP32 (the fast path, whose decisions are trusted only away from bucket boundaries)
  // multiplies by the reciprocals: exact for 1/2, within one rounding of the quotient for
  // 1/12.  FP64 (the re-decision) gives the correctly rounded quotient, as the oracle's
  // division does.
  if (FIVE && p >= 2 && p + 2 <= N - 1) {
    if constexpr (sizeof(T) == 4)  // same value, fused: (s[p-2] - s[p+2]) + 8 (s[p+1] - s[p-1])
      return fmaf(8.0f, s(p + 1) - s(p - 1), s(p - 2) - s(p + 2)) * (1.0f / 12.0f);
    const T num = -s(p + 2) + T(8) * s(p + 1) - T(8) * s(p - 1) + s(p - 2);
    return div_const(num, T(12), T(1) / T(12));
  }
  if constexpr (sizeof(T) == 4) return (s(p + 1) - s(p - 1)) * T(0.5);
  return (s(p + 1) - s(p - 1)) / T(2);
}

// The same derivatives without their stencil's scale (FP32 fast path), and that scale:
// 1 at the window border (forward/backward), 1/12 for the five-point form, 1/2 central.
template <int N, bool FIVE>
__device__ __host__ constexpr double deriv_scale(int p) {
  return (p == 0 || p == N - 1) ? 1.0 : (FIVE && p >= 2 && p + 2 <= N - 1) ? 1.0 / 12.0 : 0.5;
}
template <int N, bool FIVE, typename Line>
__device__ __forceinline__ float deriv_unscaled(Line s, int p) {
  if (p == 0) return s(1) - s(0);
  if (p == N - 1) return s(N - 1) - s(N - 2);
  if (FIVE && p >= 2 && p + 2 <= N - 1) return fmaf(8.0f, s(p + 1) - s(p - 1), s(p - 2) - s(p + 2));
  return s(p + 1) - s(p - 1);
}

// abd = sum over the gradient grid of (gx*gx, gx*gy, gy*gy) * G, grid X outer, Y inner
// (ravu-lite-r3.hook:48-76).  get(i) returns window sample i = X*N + Y.  FP32 (the fast
// path, trusted only away from bucket boundaries) fuses multiply-adds; FP64 (the
// re-decision) keeps every rounding step of the formula, like the oracle.
// LEAN (the full-RAVU kernels, whose windows are largest) trims both: FP32 takes the
// derivatives unscaled and folds each point's stencil scales into its Gaussian weight
// (compile-time constants), and where gx and gy carry the same scale one product w'*gx
// serves both a and b; FP64 is computed one derivative at a time (scheduling barriers,
// the partial sums pinned), so the rare re-decision does not hold the window in FP64
// registers and set the occupancy of every wave (r4: 122 -> 58 VGPRs).  The sums keep
// the oracle's order either way.
template <int N, bool FIVE, typename T, bool LEAN, typename Get>
__device__ __forceinline__ void structure_tensor(Get get, T& a, T& b, T& d) {
#pragma clang fp contract(off)
  constexpr int P0 = N <= 4 ? 0 : 1;
  constexpr int P1 = N <= 4 ? N : N - 1;
  a = b = d = T(0);
#pragma unroll
  for (int X = P0; X < P1; ++X) {
#pragma unroll
    for (int Y = P0; Y < P1; ++Y) {
      const int gi = (X - P0) * (P1 - P0) + (Y - P0);
      if constexpr (sizeof(T) == 4 && LEAN) {
        const float gx = deriv_unscaled<N, FIVE>([&](int k) { return get(k * N + Y); }, X);
        const float gy = deriv_unscaled<N, FIVE>([&](int k) { return get(X * N + k); }, Y);
        const double cx = deriv_scale<N, FIVE>(X), cy = deriv_scale<N, FIVE>(Y), w = gauss_w<N>(gi);
        const float u = gx * static_cast<float>(w * cx * cx);
        a = fmaf(u, gx, a);
        b = cx == cy ? fmaf(u, gy, b) : fmaf(gx * static_cast<float>(w * cx * cy), gy, b);
        d = fmaf(gy * static_cast<float>(w * cy * cy), gy, d);
        continue;
      }
      T gx = deriv<N, FIVE, T>([&](int k) { return get(k * N + Y); }, X);
      if constexpr (sizeof(T) == 8 && LEAN) {
        asm volatile("" : "+v"(gx));
        __builtin_amdgcn_sched_barrier(0);
      }
      T gy = deriv<N, FIVE, T>([&](int k) { return get(X * N + k); }, Y);
      const T w = T(gauss_w<N>(gi));
      if constexpr (sizeof(T) == 4) {  // fast path: fused accumulation
        a = fmaf(gx * gx, w, a);
        b = fmaf(gx * gy, w, b);
        d = fmaf(gy * gy, w, d);
      } else {                         // re-decision: the reference's rounding steps
        a = a + gx * gx * w;
        b = b + gx * gy * w;
        d = d + gy * gy * w;
        if constexpr (LEAN) {
          asm volatile("" : "+v"(a), "+v"(b), "+v"(d));
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
}

enum Rule : int { RULE_LITE = 0, RULE_3X = 1, RULE_LOG2 = 2 };

template <int RULE> struct RuleInfo;
template <> struct RuleInfo<RULE_LITE> { static constexpr int S = 4; };   // ravu-lite-r3.hook:86,88
template <> struct RuleInfo<RULE_3X> { static constexpr int S = 3; };     // compute/ravu-3x-r3.hook:95,97
template <> struct RuleInfo<RULE_LOG2> { static constexpr int S = 9; };   // ravu-r3.hook:117,119

struct Bucket {
  int angle, strength, coherence;
};

constexpr float kEps = 1.192092896e-7f;
constexpr double kEpsD = 1.192092896e-7;
constexpr double kPiD = 3.141592653589793;
constexpr float kPiF = 3.141592653589793f;

template <int RULE, typename T>
__device__ __forceinline__ int strength_of(T lambda) {
  if (RULE == RULE_LITE)  // ravu-lite-r3.hook:86
    return lambda >= T(0.016) ? (lambda >= T(0.05) ? 3 : 2) : (lambda >= T(0.004) ? 1 : 0);
  if (RULE == RULE_3X)    // compute/ravu-3x-r3.hook:95
    return lambda >= T(0.02) ? 2 : (lambda >= T(0.005) ? 1 : 0);
  // ravu-r3.hook:117: clamp(floor(log2(lambda*2000 + eps)), 0, 8)
  T f = floor(log2(lambda * T(2000) + T(1.192092896e-7)));
  return static_cast<int>(fmin(fmax(f, T(0)), T(8)));
}

// Eigen-analysis and bucket exactly as the reference writes it (ravu-lite-r3.hook:77-87),
// in the reference's operation order, no contraction.  Used in FP64 for re-decisions;
// mirrors the FP64 oracle operation for operation.
template <int RULE, typename T>
__device__ __forceinline__ Bucket hash_reference(T a, T b, T d) {
#pragma clang fp contract(off)
  const T eps = T(1.192092896e-7);
  const T pi = T(3.141592653589793);
  T Tr = a + d;
  T D = a * d - b * b;
  T delta = sqrt(fmax(Tr * Tr / T(4) - D, T(0)));
  T L1 = Tr / T(2) + delta;
  T L2 = Tr / T(2) - delta;
  T sqrtL1 = sqrt(L1);
  T sqrtL2 = sqrt(L2);  // NaN when L2 < 0, as in the reference
  T at = atan2(L1 - a, b) + pi;
  T theta = at - pi * floor(div_const(at, pi, T(1) / pi));  // GLSL mod
  if (fabs(b) < eps) theta = T(0);
  T ssum = sqrtL1 + sqrtL2;
  T mu = (sqrtL1 - sqrtL2) / ssum;
  if (ssum < eps) mu = T(0);
  Bucket bk;
  bk.angle = static_cast<int>(floor(div_const(theta * T(24), pi, T(1) / pi)));
  bk.strength = strength_of<RULE, T>(sqrtL1);
  bk.coherence = mu >= T(0.5) ? 2 : (mu >= T(0.25) ? 1 : 0);  // NaN -> 0
  return bk;
}

// FP32 fast path.  Same quantities through cancellation-free forms (mathematically equal
// to the reference's): delta = sqrt(((a-d)/2)^2 + b^2) instead of sqrt(T^2/4 - D), and
// L1 - a = b^2 / ((a-d)/2 + delta) when d < a.  Sets *near when a decision lies within a
// relative 1e-5 of its bucket boundary, widened for the angle by an FP32 error bound of
// theta (~16 ulp(T) / (L1 - L2)); exact cases (flat window T = 0, b = 0 exactly) are
// never near.
template <int RULE>
__device__ __forceinline__ Bucket hash_fast(float a, float b, float d, bool* near) {
  const float eps = 1.192092896e-7f;
  const float pi = 3.141592653589793f;
  const float T = a + d;
  const float h = (a - d) * 0.5f;
  const float delta = sqrtf(h * h + b * b);
  const float L1 = T * 0.5f + delta;
  const float L2 = T * 0.5f - delta;
  // quotients below go through the hardware reciprocal (within 1 ulp of the division):
  // every decision they feed is re-made in FP64 when it lies within the margins
  const float y = h <= 0.f ? delta - h : b * b * __builtin_amdgcn_rcpf(h + delta);  // L1 - a >= 0
  const float sqrtL1 = sqrtf(L1);
  const float sqrtL2 = sqrtf(L2);
  const float at = atan2f(y, b) + pi;
  float theta = at - pi * floorf(at * (1.0f / 3.141592653589793f));
  const bool bsmall = fabsf(b) < eps;
  if (bsmall) theta = 0.f;
  const float ssum = sqrtL1 + sqrtL2;
  float mu = (sqrtL1 - sqrtL2) * __builtin_amdgcn_rcpf(ssum);
  if (ssum < eps) mu = 0.f;
  const float u = theta * static_cast<float>(24.0 / 3.141592653589793);
  Bucket bk;
  bk.angle = static_cast<int>(floorf(u));
  bk.strength = strength_of<RULE, float>(sqrtL1);
  bk.coherence = mu >= 0.5f ? 2 : (mu >= 0.25f ? 1 : 0);
  const float rel = 1e-5f, ulp16 = 16.0f * 5.9604645e-8f;
  bool n = fabsf(fabsf(b) - eps) <= rel * eps + ulp16 * T;
  n |= fabsf(L2) <= (rel + ulp16) * T && !(b == 0.f && fminf(a, d) == 0.f);
  if (!bsmall) {
    const float cond = (24.0f / pi) * ulp16 * T * __builtin_amdgcn_rcpf(fmaxf(2.0f * delta, 1e-30f));
    n |= fabsf(u - rintf(u)) <= 24.0f * rel + cond;
  }
  if (RULE == RULE_LITE) {
    n |= fabsf(sqrtL1 - 0.004f) <= rel * 0.004f || fabsf(sqrtL1 - 0.016f) <= rel * 0.016f ||
         fabsf(sqrtL1 - 0.05f) <= rel * 0.05f;
  } else if (RULE == RULE_3X) {
    n |= fabsf(sqrtL1 - 0.005f) <= rel * 0.005f || fabsf(sqrtL1 - 0.02f) <= rel * 0.02f;
  } else {
    const float f = log2f(sqrtL1 * 2000.0f + eps);
    n |= f > 0.5f && f < 8.5f && fabsf(f - rintf(f)) <= rel * 8.0f;
  }
  n |= fabsf(mu - 0.25f) <= rel || fabsf(mu - 0.5f) <= rel;
  *near = n && T > 0.f;
  return bk;
}

// LUT row of a bucket as NEAREST + clamp-to-edge sampling hits it:
// texel row floor(((angle*S + strength)*3 + coherence + 0.5)) clamped to [0, rows-1]
// (angle can be 24 or -1 at the wrap of mod(atan(..)+pi, pi)).
template <int RULE>
__device__ __forceinline__ int bucket_row(Bucket bk, int rows) {
  int row = (bk.angle * RuleInfo<RULE>::S + bk.strength) * 3 + bk.coherence;
  return min(max(row, 0), rows - 1);
}

// Hash of one site: FP32, re-decided in FP64 when near a boundary.  getf(i) -> float
// and getd(i) -> double return window sample i (getd widens the same stored values, and
// for RGB forms it recomputes the BT.709 luma in FP64).  LEAN picks the tensor forms.
template <int N, bool FIVE, int RULE, bool LEAN = false, typename GetF, typename GetD>
__device__ __forceinline__ Bucket ravu_hash(GetF getf, GetD getd) {
  float a, b, d;
  structure_tensor<N, FIVE, float, LEAN>(getf, a, b, d);
  bool near = false;
  Bucket bk = hash_fast<RULE>(a, b, d, &near);
  if (near) {
    double ad, bd, dd;
    structure_tensor<N, FIVE, double, LEAN>(getd, ad, bd, dd);
    bk = hash_reference<RULE, double>(ad, bd, dd);
  }
  return bk;
}

}  // namespace mpx
