// RAVU-zoom(-ar) to an arbitrary output size, luma / yuv / rgb
// (ravu-zoom-r2.hook:15-134, ravu-zoom-ar-r2-rgb.hook:15-224, compute tiling
// compute/ravu-zoom-r3.hook:24-45).
//
// Per output pixel o: pos = (o + 1/2) * in/out (OFFSET ALIGN); subpix f = fract(pos - 1/2);
// base = pos - f.  The window is the 2r x 2r input block at offsets -(r-1)..r around base.
// The hash depends on the base pixel only, so a block first hashes every base pixel its
// output tile (32 x 8 by default) touches (into LDS), then each output pixel reads its
// base's bucket; AR variants also tabulate the per-sample powers of the anti-ringing
// weights there.
// Base and phase are computed in exact integer arithmetic:
//   pos - 1/2 = ((2o+1)*in - out) / (2*out)  ->  base = floor, f = remainder / (2*out),
// once per output column and row at plan creation (zoom_phase_kernel); the kernel reads
// them from that table.
// LUT (FILTER LINEAR): group g of 4 taps at texel x = 9g + 8f, row y = 9*bucket + 8f_y
// (LUTPOS, :21-31), bilinear between neighbouring phase texels, rows clamped to the
// texture; taps p < n^2/2 at phase f, tap n^2-1-p at phase 1-f.  AR variants use the
// separate learned ``_ar`` LUT with the same addressing (all taps), per channel, strength
// 0.8, no final clamp.  The LUTs (0.37 / 0.93 MB FP16) stay in L2.
#include "common.h"
#include "launch.h"
#include "ravu.h"
#include "internal.h"

#include <numeric>

namespace mpx {


__device__ __forceinline__ int floor_div(int64_t a, int64_t b) {  // b > 0
  int64_t q = a / b;
  if ((a % b) != 0 && (a < 0)) --q;
  return static_cast<int>(q);
}

struct Phase {
  int base;
  float f;
};
__device__ __forceinline__ Phase phase_of(int o, int in, int out) {
  const int64_t num = static_cast<int64_t>(2 * o + 1) * in - out;
  const int64_t den = 2 * static_cast<int64_t>(out);
  Phase p;
  p.base = floor_div(num, den);
  p.f = static_cast<float>(num - static_cast<int64_t>(p.base) * den) / static_cast<float>(den);
  return p;
}

// FILTER LINEAR fetch of group g at texel coordinates (9g + tx, 9*row + ty), tx, ty in
// [0, 8]; the integer row base is kept apart so the phase fraction keeps full FP32
// precision.  Rows clamp to the texture (row can be -1 or past the last bucket).
// The arithmetic is spelled out (explicit fmaf, no contraction) so that every kernel that
// inlines it rounds the same way: the plan-time weight table and the per-pixel fetch agree
// bit for bit.
__device__ __forceinline__ float4 lut_bilinear(const __half* lut, int width, int rows, int g, float tx, int row,
                                               float ty) {
#pragma clang fp contract(off)
  const float fx0 = floorf(tx), fy0 = floorf(ty);
  const float ax = tx - fx0, ay = ty - fy0;
  const int x0 = 9 * g + static_cast<int>(fx0);
  const int x1 = min(x0 + 1, width - 1);
  const int y0i = 9 * row + static_cast<int>(fy0);
  const int y0 = min(max(y0i, 0), rows - 1), y1 = min(max(y0i + 1, 0), rows - 1);
  const float4 v00 = load_lut4(lut, static_cast<int64_t>(y0) * width + x0);
  const float4 v01 = load_lut4(lut, static_cast<int64_t>(y0) * width + x1);
  const float4 v10 = load_lut4(lut, static_cast<int64_t>(y1) * width + x0);
  const float4 v11 = load_lut4(lut, static_cast<int64_t>(y1) * width + x1);
  const float bx = 1.f - ax, by = 1.f - ay;
  float4 r;
  r.x = fmaf(fmaf(v00.x, bx, v01.x * ax), by, fmaf(v10.x, bx, v11.x * ax) * ay);
  r.y = fmaf(fmaf(v00.y, bx, v01.y * ax), by, fmaf(v10.y, bx, v11.y * ax) * ay);
  r.z = fmaf(fmaf(v00.z, bx, v01.z * ax), by, fmaf(v10.z, bx, v11.z * ax) * ay);
  r.w = fmaf(fmaf(v00.w, bx, v01.w * ax), by, fmaf(v10.w, bx, v11.w * ax) * ay);
  return r;
}

template <int CH>
__device__ __forceinline__ float zluma_f(const float* t, int o, int n) {
  if (CH == 2) return 0.2126f * t[o] + 0.7152f * t[n + o] + 0.0722f * t[2 * n + o];
  return t[o];
}
template <int CH>
__device__ __forceinline__ double zluma_d(const float* t, int o, int n) {
#pragma clang fp contract(off)
  if (CH == 2)
    return 0.2126 * static_cast<double>(t[o]) + 0.7152 * static_cast<double>(t[n + o]) +
           0.0722 * static_cast<double>(t[2 * n + o]);
  return static_cast<double>(t[o]);
}

// FP64 re-decision of one base pixel's hash, out of line (the kernels without anti-ringing):
// it runs for a small fraction of sites, and inlined its registers (the double atan2 alone
// takes 56) would set the occupancy of every wave (zoom-r2: 96 -> 51 VGPRs, r3 rgb: 137 ->
// 61).  Its register budget follows its callers' (6 waves per SIMD, below), so the lean
// tensor form keeps it from spilling.  `t` is the staged tile, `o` the window's first sample.
// (one instance per calling kernel, TAB, so each inherits its own caller's budget)
template <int N, int CH, int PW, int PN, bool TAB>
__device__ __attribute__((noinline)) int zoom_redecide(const float* t, int o) {
  double a, b, d;
  structure_tensor<N, true, double, true>([&](int s) { return zluma_d<CH>(t, o + (s % N) * PW + s / N, PN); },
                                          a, b, d);
  const Bucket bk = hash_reference<RULE_LITE, double>(a, b, d);
  return (bk.angle * 4 + bk.strength) * 3 + bk.coherence;
}

// Stage the part of the PH x PW staging area that a tile's outputs read: the uh x uw
// samples per plane around its base pixels (at 2x 8 x 20 of 12 x 36), clamp-to-edge, at
// the area's row stride PW.  AR forms store each sample's anti-ringing powers beside it
// ((0.1 + l)^32, (1.1 - l)^32 and the ^33 terms, once per sample instead of once per
// output pixel that reads it).  Sample i of the dense P x uh x uw order is split with
// high-half multiply quotients: floor(i / d) = mulhi(i, floor((2^32 - 1) / d) + 1), exact
// while i * d < 2^32.
template <int P, int PH, int PW, int NT, bool AR>
__device__ __forceinline__ void stage_used(float* tile, float4* pw, const void* base, int dt, float in_max,
                                           float in_rmax, int64_t pitch, int64_t cstride, int W, int H, int ox,
                                           int oy, int uw, int uh) {
  constexpr int PN = PH * PW, E = (P * PN + NT - 1) / NT;
  static_assert(static_cast<int64_t>(P) * PN * PN < (int64_t{1} << 32), "high-half quotients exact");
  const int un = uw * uh, total = P * un;
  const uint32_t mn = 0xffffffffu / static_cast<uint32_t>(un) + 1u;
  const uint32_t mw = 0xffffffffu / static_cast<uint32_t>(uw) + 1u;
  const int lsh = dt == DT_F32 ? 2 : (dt == DT_U8 ? 0 : 1);  // log2 element bytes
  const uintptr_t b = reinterpret_cast<uintptr_t>(base);
  uint32_t v[E], sh[E];
  int li[E];
  __builtin_amdgcn_s_setprio(1);  // the loads go out ahead of other waves' arithmetic
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int i = static_cast<int>(threadIdx.x) + NT * e;
    li[e] = -1;
    if (i < total) {
      const int c = P == 1 ? 0 : static_cast<int>(__umulhi(static_cast<uint32_t>(i), mn));
      const int j = i - c * un;
      const int r = static_cast<int>(__umulhi(static_cast<uint32_t>(j), mw));
      const int col = j - r * uw;
      const int gx = min(max(ox + col, 0), W - 1), gy = min(max(oy + r, 0), H - 1);
      // 32-bit element offset, as the halo loaders (the host bounds the frame's span)
      const uint32_t off = static_cast<uint32_t>(c) * static_cast<uint32_t>(cstride) +
                           __umul24(static_cast<uint32_t>(gy), static_cast<uint32_t>(pitch)) + gx;
      const uintptr_t a = b + (static_cast<uintptr_t>(off) << lsh);
      v[e] = *reinterpret_cast<GlobalU32>(a & ~static_cast<uintptr_t>(3));
      sh[e] = static_cast<uint32_t>(a & 3) * 8;
      li[e] = c * PN + r * PW + col;
    }
  }
  __builtin_amdgcn_s_setprio(0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (li[e] >= 0) {
      const uint32_t w = v[e] >> sh[e];
      float l;
      switch (dt) {
        case DT_F32: l = __uint_as_float(v[e]); break;
        case DT_F16: l = __half2float(__ushort_as_half(static_cast<unsigned short>(w & 0xffffu))); break;
        case DT_U8: l = unorm_quotient(static_cast<float>(w & 0xffu), in_max, in_rmax); break;
        default: l = unorm_quotient(static_cast<float>(w & 0xffffu), in_max, in_rmax); break;
      }
      tile[li[e]] = l;
      if constexpr (AR) {
        const float p0 = 0.1f + l, m0 = 1.1f - l;
        float p32 = p0, m32 = m0;
#pragma unroll
        for (int k = 0; k < 5; ++k) { p32 *= p32; m32 *= m32; }
        pw[li[e]] = make_float4(p32, m32, p32 * p0, m32 * m0);
      }
    }
  }
}

// Pre-interpolated weights.  The phase of output column o repeats with period
// out / gcd(in, out) (2 at 2x, 3 at 1.5x), so for such plans every (bucket row, phase pair)
// is interpolated once, at plan creation, by the same lut_bilinear in the same operations:
// table entry (row, py, px, g) holds the group's float4s [wf, wi (, af, ai)].  Bucket rows
// below 0 fetch the same texels as row -1 and rows past the last bucket the same as
// bucket 288 (the texel rows clamp), so rows -1..288 cover every bucket value.
constexpr int kZoomRowLo = -1, kZoomRows = 290;
constexpr int kZoomMaxPhasePairs = 64;

template <int R, bool AR>
__global__ __launch_bounds__(256) void zoom_table_kernel(const __half* __restrict__ lut,
                                                         const __half* __restrict__ lut_ar, float4* wtab, int in_w,
                                                         int in_h, int out_w, int out_h, int nphx, int nphy) {
  constexpr int N = 2 * R;
  constexpr int GROUPS = (N * N / 2 + 3) / 4;
  constexpr int WIDTH = 9 * GROUPS, ROWS = 2592;
  constexpr int NT = AR ? 4 : 2;
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kZoomRows * nphy * nphx * GROUPS) return;
  const int g = i % GROUPS;
  int t = i / GROUPS;
  const int px = t % nphx;
  t /= nphx;
  const int py = t % nphy;
  const int row = t / nphy + kZoomRowLo;
  const Phase phx = phase_of(px, in_w, out_w), phy = phase_of(py, in_h, out_h);
  const float txf = 8.f * phx.f, txi = 8.f * (1.f - phx.f);
  const float tyf = 8.f * phy.f, tyi = 8.f * (1.f - phy.f);
  float4* e = wtab + static_cast<int64_t>(i) * NT;
  e[0] = lut_bilinear(lut, WIDTH, ROWS, g, txf, row, tyf);
  e[1] = lut_bilinear(lut, WIDTH, ROWS, g, txi, row, tyi);
  if (AR) {
    e[2] = lut_bilinear(lut_ar, WIDTH, ROWS, g, txf, row, tyf);
    e[3] = lut_bilinear(lut_ar, WIDTH, ROWS, g, txi, row, tyi);
  }
}

// Phase table of a plan: entry o < out_w is output column o, entry out_w + o row o:
// {base, bits of f, o mod period, 0}.
__global__ __launch_bounds__(256) void zoom_phase_kernel(int4* tab, int in_w, int in_h, int out_w, int out_h,
                                                         int nphx, int nphy) {
  const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= out_w + out_h) return;
  const bool x = i < out_w;
  const int o = x ? i : i - out_w;
  const Phase ph = x ? phase_of(o, in_w, out_w) : phase_of(o, in_h, out_h);
  tab[i] = make_int4(ph.base, __float_as_int(ph.f), o % (x ? nphx : nphy), 0);
}

// One block per TX x TY output tile, one thread per output pixel (kZoomTiles).  (Kernels
// without anti-ringing: at least 6 waves per SIMD, a register budget their out-of-line
// re-decision inherits; r2 anti-ringing kernels and r3 ones on 16 x 16 tiles: 5, 96 VGPRs
// with a few bytes of scratch, measured 1-8 % faster than 4 waves at 118-126.)  TAB: the plan has the pre-interpolated
// weight table; separate instances keep the per-pixel bilinear fetch's registers out of the
// table path.
template <int R, bool AR, int CH, bool TAB, int TX, int TY>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(AR ? ((R == 2 || TX == 16) ? 5 : 1) : 6))) void ravu_zoom_kernel(Io io, const __half* __restrict__ lut,
                                                        const __half* __restrict__ lut_ar,
                                                        const float4* __restrict__ wtab, int nphx, int nphy,
                                                        const int4* __restrict__ phase, int out_w, int out_h) {
  constexpr int P = CH == 0 ? 1 : 3;
  constexpr int N = 2 * R;
  constexpr int HALF = N * N / 2;
  constexpr int GROUPS = (HALF + 3) / 4;
  constexpr int WIDTH = 9 * GROUPS;      // LUT texels per row
  constexpr int ROWS = 2592;             // 288 buckets x 9 phases
  constexpr int BX = TX + 1, BY = TY + 1;  // base pixels a tile can touch (ratio in/out < 1)
  constexpr int PW = BX + N - 1, PH = BY + N - 1, PN = PW * PH;
  __shared__ float tile[P * PN];
  __shared__ int bucket[BY * BX];
  // AR: per staged sample (l + 0.1)^32, (1.1 - l)^32 and the ^33 terms, computed once per
  // sample instead of once per output pixel that reads it (about 25 at 2x)
  __shared__ float4 pw[AR ? P * PN : 1];

  const BlockXYZ bk3 = block_xyz((out_w + TX - 1) / TX, (out_h + TY - 1) / TY, io.batch);
  const int f = bk3.z;
  const int ox0 = bk3.x * TX, oy0 = bk3.y * TY;
  const int4* phy_tab = phase + out_w;
  const int bx0 = phase[ox0].x;
  const int by0 = phy_tab[oy0].x;
  // base pixels this tile's (in-range) outputs touch: nbx x nby <= BX x BY (in < out)
  const int nbx = phase[min(ox0 + TX, out_w) - 1].x - bx0 + 1;
  const int nby = phy_tab[min(oy0 + TY, out_h) - 1].x - by0 + 1;
  const void* in = frame_in(io, f);
  stage_used<P, PH, PW, TX * TY, AR>(tile, pw, in, io.in_dt, io.in_max, io.in_rmax, io.in_pitch, io.in_cstride, io.W, io.H,
                            bx0 - (R - 1), by0 - (R - 1), nbx + N - 1, nby + N - 1);
  __syncthreads();
  // hash every base pixel the tile's outputs use (unclamped LUT row: bucket 9-row blocks
  // are clamped per texel by the bilinear fetch, as the sampler does)
  for (int i = threadIdx.x; i < nbx * nby; i += blockDim.x) {
    const int by = i / nbx, bx = i - by * nbx;
    auto off = [&](int s) { return (by + s % N) * PW + bx + s / N; };
    if constexpr (AR) {
      // the anti-ringing tap loop sets these kernels' register count: the re-decision stays
      // inline, in the lean tensor forms for r3 (r3 rgb: 216 -> 123 VGPRs)
      const Bucket bk = ravu_hash<N, true, RULE_LITE, (R >= 3)>([&](int s) { return zluma_f<CH>(tile, off(s), PN); },
                                                      [&](int s) { return zluma_d<CH>(tile, off(s), PN); });
      bucket[by * BX + bx] = (bk.angle * 4 + bk.strength) * 3 + bk.coherence;
    } else {
      float ta, tb, td;
      structure_tensor<N, true, float, (R >= 3)>([&](int s) { return zluma_f<CH>(tile, off(s), PN); }, ta, tb, td);
      bool near = false;
      const Bucket bk = hash_fast<RULE_LITE>(ta, tb, td, &near);
      bucket[by * BX + bx] = near ? zoom_redecide<N, CH, PW, PN, TAB>(tile, off(0))
                                    : (bk.angle * 4 + bk.strength) * 3 + bk.coherence;
    }
  }
  __syncthreads();

  const int ox = ox0 + threadIdx.x % TX, oy = oy0 + threadIdx.x / TX;
  if (ox >= out_w || oy >= out_h) return;
  const int4 phx = phase[ox], phy = phy_tab[oy];
  const float fx = __int_as_float(phx.y), fy = __int_as_float(phy.y);
  const int lbx = phx.x - bx0, lby = phy.x - by0;  // in [0, nbx) x [0, nby) by construction
  const int row = bucket[lby * BX + lbx];
  if (io.buckets) io.buckets[f * io.bucket_bstride + static_cast<int64_t>(oy) * out_w + ox] = row;
  auto si = [&](int c, int i) { return c * PN + (lby + i % N) * PW + lbx + i / N; };
  const float txf = 8.f * fx, txi = 8.f * (1.f - fx);
  const float tyf = 8.f * fy, tyi = 8.f * (1.f - fy);
  // this pixel's entries of the pre-interpolated table (when the plan has one)
  constexpr int NT = AR ? 4 : 2;
  const float4* wt = nullptr;
  if constexpr (TAB) {
    const int trow = min(max(row, kZoomRowLo), kZoomRowLo + kZoomRows - 1) - kZoomRowLo;
    wt = wtab + static_cast<int64_t>(((trow * nphy + phy.z) * nphx + phx.z) * GROUPS) * NT;
  }
  float res[P], hi[P], lo[P], hi2[P], lo2[P];
#pragma unroll
  for (int c = 0; c < P; ++c) res[c] = hi[c] = lo[c] = hi2[c] = lo2[c] = 0.f;
#pragma unroll
  for (int g = 0; g < GROUPS; ++g) {
    // one group's operands live at a time (its sums complete before the next group)
    __builtin_amdgcn_sched_barrier(0);
    float4 wf4, wi4, a4, b4;
    if constexpr (TAB) {
      wf4 = wt[g * NT];
      wi4 = wt[g * NT + 1];
      if (AR) {
        a4 = wt[g * NT + 2];
        b4 = wt[g * NT + 3];
      }
    } else {
      wf4 = lut_bilinear(lut, WIDTH, ROWS, g, txf, row, tyf);
      wi4 = lut_bilinear(lut, WIDTH, ROWS, g, txi, row, tyi);
      if (AR) {
        a4 = lut_bilinear(lut_ar, WIDTH, ROWS, g, txf, row, tyf);
        b4 = lut_bilinear(lut_ar, WIDTH, ROWS, g, txi, row, tyi);
      }
    }
    const float wf[4] = {wf4.x, wf4.y, wf4.z, wf4.w}, wi[4] = {wi4.x, wi4.y, wi4.z, wi4.w};
    float af[4] = {0.f, 0.f, 0.f, 0.f}, ai[4] = {0.f, 0.f, 0.f, 0.f};
    if (AR) {
      af[0] = a4.x; af[1] = a4.y; af[2] = a4.z; af[3] = a4.w;
      ai[0] = b4.x; ai[1] = b4.y; ai[2] = b4.z; ai[3] = b4.w;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      // two taps' LDS reads (samples and their powers, three planes) in flight at a time
      if (AR && k == 2) __builtin_amdgcn_sched_barrier(0);
      const int p = 4 * g + k;
      if (p < HALF) {
        const int q = N * N - 1 - p;
#pragma unroll
        for (int c = 0; c < P; ++c) {
          // one fused multiply-add per tap and sum (the sums' rounding order is free: the
          // results are held to the FP64 oracle within a tolerance)
          const float lp = tile[si(c, p)], lq = tile[si(c, q)];
          res[c] = fmaf(lq, wi[k], fmaf(lp, wf[k], res[c]));
          if (AR) {
            // ravu-zoom-ar-r2-rgb.hook:152-222: (0.1+l)^32, (1.1-l)^32 and the ^33 terms
            const float4 a = pw[si(c, p)], b = pw[si(c, q)];
            hi[c] = fmaf(b.x, ai[k], fmaf(a.x, af[k], hi[c]));
            lo[c] = fmaf(b.y, ai[k], fmaf(a.y, af[k], lo[c]));
            hi2[c] = fmaf(b.z, ai[k], fmaf(a.z, af[k], hi2[c]));
            lo2[c] = fmaf(b.w, ai[k], fmaf(a.w, af[k], lo2[c]));
          }
        }
      }
    }
#pragma unroll
    for (int c = 0; c < P; ++c) {
      asm volatile("" : "+v"(res[c]));
      if (AR) asm volatile("" : "+v"(hi[c]), "+v"(lo[c]), "+v"(hi2[c]), "+v"(lo2[c]));
    }
  }
  const int esz_out = elem_bytes(io.out_dt);
  char* ob = static_cast<char*>(frame_out(io, f));
#pragma unroll
  for (int c = 0; c < P; ++c) {
    float v = res[c];
    if (AR) {
      const float h = hi2[c] * __builtin_amdgcn_rcpf(hi[c]) - 0.1f, l = 1.1f - lo2[c] * __builtin_amdgcn_rcpf(lo[c]);
      const float cl = fminf(fmaxf(v, l), h);
      v = v * (1.0f - 0.8f) + cl * 0.8f;  // mix(res, clamp(res, lo, hi), 0.8), no final clamp
    } else {
      v = fminf(fmaxf(v, 0.f), 1.f);
    }
    store_elem(ob + c * io.out_cstride * esz_out, static_cast<int64_t>(oy) * io.out_pitch + ox, io.out_dt, v);
  }
}

template <int T>
static hipError_t launch_zoom_tile(int r, bool ar, int channels, const Io& io, const __half* lut,
                                   const __half* lut_ar, const ZoomPlan& zp, int out_w, int out_h, hipStream_t st) {
  constexpr int TX = kZoomTiles[T].w, TY = kZoomTiles[T].h;
  const dim3 grid(((out_w + TX - 1) / TX) * ((out_h + TY - 1) / TY) * io.batch), block(TX * TY);
  const float4* wtab = reinterpret_cast<const float4*>(zp.wtab);
#define MPX_ZOOM(R, A, CH)                                                                                   \
  if (r == R && ar == A && channels == CH) {                                                                 \
    if (zp.wtab)                                                                                             \
      hipLaunchKernelGGL((ravu_zoom_kernel<R, A, CH, true, TX, TY>), grid, block, 0, st, io, lut, lut_ar, wtab, \
                         zp.nphx, zp.nphy, zp.phase, out_w, out_h);                                          \
    else                                                                                                     \
      hipLaunchKernelGGL((ravu_zoom_kernel<R, A, CH, false, TX, TY>), grid, block, 0, st, io, lut, lut_ar,    \
                         nullptr, 0, 0, zp.phase, out_w, out_h);                                             \
    return hipGetLastError();                                                                                \
  }
  MPX_ZOOM(2, false, 0) MPX_ZOOM(2, false, 1) MPX_ZOOM(2, false, 2) MPX_ZOOM(3, false, 0) MPX_ZOOM(3, false, 1)
  MPX_ZOOM(3, false, 2) MPX_ZOOM(2, true, 0) MPX_ZOOM(2, true, 1) MPX_ZOOM(2, true, 2) MPX_ZOOM(3, true, 0)
  MPX_ZOOM(3, true, 1) MPX_ZOOM(3, true, 2)
#undef MPX_ZOOM
  return hipErrorInvalidValue;
}

hipError_t launch_ravu_zoom(int r, bool ar, int channels, const Io& io, const __half* lut, const __half* lut_ar,
                            const ZoomPlan& zp, int out_w, int out_h, int tile, hipStream_t st) {
  if (!zp.phase) return hipErrorInvalidValue;
  static_assert(kZoomTileCount == 2, "one case per tile shape");
  switch (tile) {
    case 0: return launch_zoom_tile<0>(r, ar, channels, io, lut, lut_ar, zp, out_w, out_h, st);
    case 1: return launch_zoom_tile<1>(r, ar, channels, io, lut, lut_ar, zp, out_w, out_h, st);
    default: return hipErrorInvalidValue;
  }
}

int zoom_plan_create(mpx_ctx* ctx, mpx_plan* plan, int in_w, int in_h, int out_w, int out_h) {
  ZoomPlan& zp = plan->zoom;
  zp = ZoomPlan{};
  const int nphx = out_w / std::gcd(in_w, out_w), nphy = out_h / std::gcd(in_h, out_h);
  {
    const size_t pbytes = static_cast<size_t>(out_w + out_h) * sizeof(int4);
    void* pp = nullptr;
    if (int rc = ctx->alloc(pbytes, &pp)) return rc;
    mpx_host::DeviceScope ds;
    hipError_t e = hipErrorInvalidDevice;
    if (!ds.enter(ctx->device)) {
      hipLaunchKernelGGL(zoom_phase_kernel, dim3((out_w + out_h + 255) / 256), dim3(256), 0, nullptr,
                         static_cast<int4*>(pp), in_w, in_h, out_w, out_h, nphx, nphy);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    }
    if (e != hipSuccess) {
      ctx->release(pp, pbytes);
      return mpx_host::fail(MPX_ERR_HIP, std::string("ravu-zoom phase table: ") + hipGetErrorString(e));
    }
    zp.phase = static_cast<int4*>(pp);
    zp.phase_bytes = pbytes;
  }
  if (static_cast<int64_t>(nphx) * nphy > kZoomMaxPhasePairs) return MPX_OK;  // per-pixel interpolation
  const int r = plan->v->radius;
  const bool ar = plan->v->ar != 0;
  const int groups = (2 * r * 2 * r / 2 + 3) / 4;
  const int entries = kZoomRows * nphy * nphx * groups;
  const size_t bytes = static_cast<size_t>(entries) * (ar ? 4 : 2) * sizeof(float4);
  void* p = nullptr;
  if (ctx->alloc(bytes, &p)) return MPX_OK;  // over the budget: per-pixel interpolation
  mpx_host::DeviceScope ds;
  int rc;
  if ((rc = ds.enter(ctx->device))) {
    ctx->release(p, bytes);
    zoom_plan_destroy(ctx, plan);
    return rc;
  }
  const dim3 grid((entries + 255) / 256);
  float4* tab = static_cast<float4*>(p);
  if (r == 2 && !ar)
    hipLaunchKernelGGL((zoom_table_kernel<2, false>), grid, dim3(256), 0, nullptr, plan->lut, plan->lut_ar, tab, in_w,
                       in_h, out_w, out_h, nphx, nphy);
  else if (r == 2)
    hipLaunchKernelGGL((zoom_table_kernel<2, true>), grid, dim3(256), 0, nullptr, plan->lut, plan->lut_ar, tab, in_w,
                       in_h, out_w, out_h, nphx, nphy);
  else if (!ar)
    hipLaunchKernelGGL((zoom_table_kernel<3, false>), grid, dim3(256), 0, nullptr, plan->lut, plan->lut_ar, tab, in_w,
                       in_h, out_w, out_h, nphx, nphy);
  else
    hipLaunchKernelGGL((zoom_table_kernel<3, true>), grid, dim3(256), 0, nullptr, plan->lut, plan->lut_ar, tab, in_w,
                       in_h, out_w, out_h, nphx, nphy);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    ctx->release(p, bytes);
    zoom_plan_destroy(ctx, plan);
    return mpx_host::fail(MPX_ERR_HIP, std::string("ravu-zoom weight table: ") + hipGetErrorString(e));
  }
  zp.wtab = static_cast<float*>(p);
  zp.wtab_bytes = bytes;
  zp.nphx = nphx;
  zp.nphy = nphy;
  return MPX_OK;
}

void zoom_plan_destroy(mpx_ctx* ctx, mpx_plan* plan) {
  ctx->release(plan->zoom.wtab, plan->zoom.wtab_bytes);
  ctx->release(plan->zoom.phase, plan->zoom.phase_bytes);
  plan->zoom = ZoomPlan{};
}

}  // namespace mpx
