// RAVU-lite(-ar) 2x, luma: one thread per input pixel, 2x2 output block.
//
// Follows ravu-lite-r3.hook:22-131 (step1 hash + taps, step2 depth-to-space) and
// ravu-lite-ar-r3.hook:90-182 (anti-ringing); the compute flavour
// compute/ravu-lite-r3.hook:15-105 is the same math with a shared-memory tile.
//
// MI355X layout: a TX x TY pixel tile per block of TX * TY threads (64 x 4 by default: one
// wave per tile row, so LDS reads of the window are consecutive words across lanes) with
// an (r-1)-pixel clamp-to-edge halo staged once in LDS as FP32.  The FP16 LUT (rows of
// ceil(n^2/2) RGBA texels) stays in HBM and is served from L1/L2: 13 texels x 8 B per
// pixel for r3.  Outputs are written as two horizontal pairs per thread (rows 2y, 2y+1),
// coalesced across the wave.
#include "common.h"
#include "launch.h"

namespace mpx {

template <int R, bool AR, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void ravu_lite_kernel(Io io, const __half* __restrict__ lut, int total) {
  constexpr int NT = TX * TY;  // threads per block
  constexpr int N = 2 * R - 1;
  constexpr int HALO = R - 1;
  constexpr int PW = TX + 2 * HALO;
  constexpr int PH = TY + 2 * HALO;
  constexpr int NTEX = (N * N + 1) / 2;
  constexpr int C = (N * N - 1) / 2;
  __shared__ float tile[PH * PW];
  // AR: per staged sample (l + 0.1)^32, (1.1 - l)^32 and the ^33 terms, computed once per
  // sample instead of once per pixel whose anti-ringing window reads it (up to 13)
  __shared__ float4 pw[AR ? PH * PW : 1];

  const FrameTile ft = frame_tile<TX, TY>(io, blockIdx.x, total);
  const int frame = ft.f;
  const int x0 = ft.x0;
  const int y0 = ft.y0;
  const void* in = frame_in(io, frame);

  {
    HaloTile<1, PH, PW, NT> h;
    __builtin_amdgcn_s_setprio(1);  // the loads go out ahead of other waves' arithmetic
    h.issue(in, io.in_dt, ft.in_pitch, 0, ft.W, ft.H, x0 - HALO, y0 - HALO);
    __builtin_amdgcn_s_setprio(0);
    h.commit(tile, io.in_dt, io.in_max, io.in_rmax);
  }
  if constexpr (AR) {
    // each thread takes the samples it staged itself (the commit's index pattern), so the
    // one block barrier below orders both tables
    for (int i = threadIdx.x; i < PH * PW; i += NT) {
      const float l = tile[i];
      const float p0 = 0.1f + l, m0 = 1.1f - l;
      float p32 = p0, m32 = m0;
#pragma unroll
      for (int k = 0; k < 5; ++k) { p32 *= p32; m32 *= m32; }
      pw[i] = make_float4(p32, m32, p32 * p0, m32 * m0);
    }
  }
  __syncthreads();

  const int tx = threadIdx.x % TX;
  const int ty = threadIdx.x / TX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= ft.W || y >= ft.H) return;

  // window sample i = X*N + Y at offset (X - HALO, Y - HALO).  r <= 3 hold the window in
  // registers from the hash to the taps; r4 (49 samples) reads it from LDS where it is
  // used, with the lean tensor forms, so the rare FP64 re-decision does not stack its
  // registers on a live window (82 -> 61 VGPRs, 5 -> 8 waves per SIMD)
  constexpr bool FROM_LDS = R >= 4;
  auto lds = [&](int i) { return tile[(ty + i % N) * PW + tx + i / N]; };
  float s[FROM_LDS ? 1 : N * N];
  if constexpr (!FROM_LDS) {
#pragma unroll
    for (int X = 0; X < N; ++X)
#pragma unroll
      for (int Y = 0; Y < N; ++Y) s[X * N + Y] = tile[(ty + Y) * PW + tx + X];
  }
  auto sample = [&](int i) {
    if constexpr (FROM_LDS) return lds(i);
    else return s[i];
  };
  const Bucket bk = ravu_hash<N, false, RULE_LITE, FROM_LDS>(
      [&](int i) { return sample(i); }, [&](int i) { return static_cast<double>(sample(i)); });
  const int row = bucket_row<RULE_LITE>(bk, 288);
  if (io.buckets)
    io.buckets[frame * io.bucket_bstride + static_cast<int64_t>(y) * ft.W + x] = row;

  const __half* lrow = lut + static_cast<int64_t>(row) * NTEX * 4;
  float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 hi = res, lo = res, hi2 = res, lo2 = res;
  // One step per LUT texel, software-pipelined: the next texel is fetched (raw FP16,
  // converted inside the multiply-adds) while this one computes, and each step's sums
  // complete before the next step starts, so one step's operands are live at a time.
  const uint2* lraw = reinterpret_cast<const uint2*>(lrow);
  uint2 tc = lraw[0];
#pragma unroll
  for (int i = 0; i <= C; ++i) {
    __builtin_amdgcn_sched_barrier(0);
    uint2 tn = tc;
    if (i < C) tn = lraw[i + 1];
    const __half2 h0 = bits_h2(tc.x), h1 = bits_h2(tc.y);
    const float4 w = make_float4(__low2float(h0), __high2float(h0), __low2float(h1), __high2float(h1));
    const float li = sample(i);
    res.x = fmaf(li, w.x, res.x);
    res.y = fmaf(li, w.y, res.y);
    res.z = fmaf(li, w.z, res.z);
    res.w = fmaf(li, w.w, res.w);
    if (i < C) {  // the mirrored sample takes the texel reversed (w.wzyx)
      const float lj = sample(N * N - 1 - i);
      res.x = fmaf(lj, w.w, res.x);
      res.y = fmaf(lj, w.z, res.y);
      res.z = fmaf(lj, w.y, res.z);
      res.w = fmaf(lj, w.x, res.w);
    }
    if (AR) {
      const int X = i / N - HALO, Y = i % N - HALO;
      if (X * X + Y * Y <= 4) {  // anti-ringing taps (ravu-lite-ar-r3.hook:95-181)
        const float4 wg = make_float4(fmaxf(w.x, 0.f), fmaxf(w.y, 0.f), fmaxf(w.z, 0.f), fmaxf(w.w, 0.f));
        const float4 pa = pw[(ty + i % N) * PW + tx + i / N];  // sample i's powers
        const float p32 = pa.x, m32 = pa.y, p33 = pa.z, m33 = pa.w;
        hi.x += p32 * wg.x; hi.y += p32 * wg.y; hi.z += p32 * wg.z; hi.w += p32 * wg.w;
        lo.x += m32 * wg.x; lo.y += m32 * wg.y; lo.z += m32 * wg.z; lo.w += m32 * wg.w;
        hi2.x += p33 * wg.x; hi2.y += p33 * wg.y; hi2.z += p33 * wg.z; hi2.w += p33 * wg.w;
        lo2.x += m33 * wg.x; lo2.y += m33 * wg.y; lo2.z += m33 * wg.z; lo2.w += m33 * wg.w;
        if (i < C) {
          const int j = N * N - 1 - i;
          const float4 pb = pw[(ty + j % N) * PW + tx + j / N];  // the mirrored sample's
          const float q32 = pb.x, n32 = pb.y, q33 = pb.z, n33 = pb.w;
          // mirrored tap uses wg.wzyx
          hi.x += q32 * wg.w; hi.y += q32 * wg.z; hi.z += q32 * wg.y; hi.w += q32 * wg.x;
          lo.x += n32 * wg.w; lo.y += n32 * wg.z; lo.z += n32 * wg.y; lo.w += n32 * wg.x;
          hi2.x += q33 * wg.w; hi2.y += q33 * wg.z; hi2.z += q33 * wg.y; hi2.w += q33 * wg.x;
          lo2.x += n33 * wg.w; lo2.y += n33 * wg.z; lo2.z += n33 * wg.y; lo2.w += n33 * wg.x;
        }
        asm volatile("" : "+v"(hi.x), "+v"(hi.y), "+v"(hi.z), "+v"(hi.w), "+v"(lo.x), "+v"(lo.y), "+v"(lo.z),
                     "+v"(lo.w));
        asm volatile("" : "+v"(hi2.x), "+v"(hi2.y), "+v"(hi2.z), "+v"(hi2.w), "+v"(lo2.x), "+v"(lo2.y),
                     "+v"(lo2.z), "+v"(lo2.w));
      }
    }
    asm volatile("" : "+v"(res.x), "+v"(res.y), "+v"(res.z), "+v"(res.w));
    tc = tn;
  }
  float r4[4] = {res.x, res.y, res.z, res.w};
  if (AR) {
    // quotients through the hardware reciprocal (within 1 ulp of the division)
    const float h[4] = {hi2.x * __builtin_amdgcn_rcpf(hi.x) - 0.1f, hi2.y * __builtin_amdgcn_rcpf(hi.y) - 0.1f,
                        hi2.z * __builtin_amdgcn_rcpf(hi.z) - 0.1f, hi2.w * __builtin_amdgcn_rcpf(hi.w) - 0.1f};
    const float l[4] = {1.1f - lo2.x * __builtin_amdgcn_rcpf(lo.x), 1.1f - lo2.y * __builtin_amdgcn_rcpf(lo.y),
                        1.1f - lo2.z * __builtin_amdgcn_rcpf(lo.z), 1.1f - lo2.w * __builtin_amdgcn_rcpf(lo.w)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float cl = fminf(fmaxf(r4[k], l[k]), h[k]);
      r4[k] = r4[k] * (1.0f - 0.8f) + cl * 0.8f;   // mix(res, clamp(res, lo, hi), 0.8)
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) r4[k] = fminf(fmaxf(r4[k], 0.f), 1.f);
  }
  // res[2*dx + dy] -> output (2x+dx, 2y+dy)
  void* out = frame_out(io, frame);
  const int64_t o0 = static_cast<int64_t>(2 * y) * ft.out_pitch + 2 * x;
  store_pair(out, o0, io.out_dt, r4[0], r4[2]);
  store_pair(out, o0 + ft.out_pitch, io.out_dt, r4[1], r4[3]);
}

template <int T>
static hipError_t launch_lite_tile(int r, bool ar, Io io, const __half* lut, hipStream_t st) {
  constexpr int TX = kLiteTiles[T].w, TY = kLiteTiles[T].h;
  const int total = launch_tiles<TX, TY>(io);
#define MPX_LITE(R, A)                                                                                  \
  if (r == R && ar == A) {                                                                              \
    hipLaunchKernelGGL((ravu_lite_kernel<R, A, TX, TY>), dim3(total), dim3(TX * TY), 0, st, io, lut, total); \
    return hipGetLastError();                                                                           \
  }
  MPX_LITE(2, false) MPX_LITE(2, true) MPX_LITE(3, false) MPX_LITE(3, true) MPX_LITE(4, false) MPX_LITE(4, true)
#undef MPX_LITE
  return hipErrorInvalidValue;
}

hipError_t launch_ravu_lite(int r, bool ar, const Io& io, const __half* lut, int tile, hipStream_t st) {
  static_assert(kLiteTileCount == 3, "one case per tile shape");
  switch (tile) {
    case 0: return launch_lite_tile<0>(r, ar, io, lut, st);
    case 1: return launch_lite_tile<1>(r, ar, io, lut, st);
    case 2: return launch_lite_tile<2>(r, ar, io, lut, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mpx
