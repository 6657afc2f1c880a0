// RAVU (full, 2x with three predictions) and RAVU-3x, luma / yuv / rgb.
//
// RAVU-3x (compute/ravu-3x-r3.hook:22-163): one thread per input pixel, 3x3 output block
// (centre = input pixel), 24x3x3 buckets, two LUT texels per tap.
// RAVU (ravu-r3.hook:15-437): kernel 1 predicts int11 (x+1/2, y+1/2) from the 2r x 2r
// window into an FP32 scratch plane; kernel 2 predicts int10 (x+1/2, y) and then int01
// (x, y+1/2) on the 45-degree lattice mixing the input and int11, from one staging of
// both tiles, and writes the interleaved 2x output (the combine pass, :426-437), each
// prediction with the output row it completes.  All read LDS tiles staged with
// clamp-to-edge halos, each texture clamped on its own size as mpv's sampler does.
//
// Channel forms (ravu-r3-yuv.hook, ravu-r3-rgb.hook): the hash sees plane 0 (yuv) or
// the BT.709 luma (rgb, `dot(rgb, vec3(0.2126, 0.7152, 0.0722))`); taps apply to all
// three planes with the same weights.
#include "common.h"
#include "launch.h"
#include "ravu.h"

namespace mpx {

// Tiles: TX x TY pixels per block of TX * TY threads, one input pixel (3x) or site (full
// RAVU) per thread, from the families' shape lists (k3xTiles, kRavuTiles).  Full RAVU's halo
// is up to 9 rows, so its tiles are taller than 3x's: 32 x 8 by default (r4 lattice: 41 x 17
// staged samples per 256 sites, against 73 x 13 for a 64 x 4 tile), 32 x 16 and 64 x 8 on
// 512 threads as alternatives.


// Stage P planes of a frame into LDS: tile[c][r][col] = plane c at
// (clamp(ox + col), clamp(oy + r)), PH x PW per plane; all loads in flight at once.
template <int P, int PH, int PW, int NT>
__device__ __forceinline__ void stage_planes(float* tile, const void* base, int dt, float in_max, float in_rmax,
                                             int64_t pitch, int64_t cstride, int W, int H, int ox, int oy) {
  HaloTile<P, PH, PW, NT> h;
  __builtin_amdgcn_s_setprio(1);  // the loads go out ahead of other waves' arithmetic
  h.issue(base, dt, pitch, cstride, W, H, ox, oy);
  __builtin_amdgcn_s_setprio(0);
  h.commit(tile, dt, in_max, in_rmax);
}

// Planes staged per tile: the P input planes, plus for rgb a fourth holding the BT.709
// luma the hash reads (computed once per staged sample, not once per window use).
template <int CH>
constexpr int staged_planes() { return CH == 0 ? 1 : (CH == 1 ? 3 : 4); }

template <int CH>
__device__ __forceinline__ float luma_of(const float* t, int o, int n) {
  return 0.2126f * t[o] + 0.7152f * t[n + o] + 0.0722f * t[2 * n + o];
}

// rgb: fill plane 3 of a staged tile (n samples per plane) with the luma of planes 0-2;
// the caller synchronises the block before and after
template <int CH>
__device__ __forceinline__ void stage_luma(float* t, int n) {
  if constexpr (CH == 2) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) t[3 * n + i] = luma_of<CH>(t, i, n);
  }
}

// Hash-plane value from the staged planes at LDS offset o (plane stride n).
template <int CH>
__device__ __forceinline__ float luma_f(const float* t, int o, int n) {
  if (CH == 2) return t[3 * n + o];
  return t[o];
}
template <int CH>
__device__ __forceinline__ double luma_d(const float* t, int o, int n) {
#pragma clang fp contract(off)
  if (CH == 2)
    return 0.2126 * static_cast<double>(t[o]) + 0.7152 * static_cast<double>(t[n + o]) +
           0.0722 * static_cast<double>(t[2 * n + o]);
  return static_cast<double>(t[o]);
}

// Hash of the site whose N x N window starts at staged-tile offset o (sample i at
// o + (i % N) * PW + i / N), with the FP64 re-decision out of line: it is rare, and inlined
// its registers (the double atan2 alone takes 56) would set the occupancy of every wave
// (r4 rgb step1: 96 -> 51 VGPRs).  Its register budget follows its callers' (the kernel asks
// for 6 waves per SIMD), so it takes the lean tensor form to stay out of scratch.  Used
// where it pays (full-RAVU step1); ravu-3x measured 2-3 % slower with it.
template <int N, bool FIVE, int RULE, int CH, int PW, int PN>
__device__ __attribute__((noinline)) Bucket redecide_site(const float* t, int o) {
  double a, b, d;
  structure_tensor<N, FIVE, double, true>([&](int s) { return luma_d<CH>(t, o + (s % N) * PW + s / N, PN); }, a, b,
                                          d);
  return hash_reference<RULE, double>(a, b, d);
}
template <int N, bool FIVE, int RULE, int CH, int PW, int PN, bool LEAN>
__device__ __forceinline__ Bucket hash_site(const float* t, int o) {
  float a, b, d;
  structure_tensor<N, FIVE, float, LEAN>([&](int s) { return luma_f<CH>(t, o + (s % N) * PW + s / N, PN); }, a, b, d);
  bool near = false;
  const Bucket bk = hash_fast<RULE>(a, b, d, &near);
  return near ? redecide_site<N, FIVE, RULE, CH, PW, PN>(t, o) : bk;
}

// ----------------------------------------------------------------------------- RAVU-3x

template <int R, int CH, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void ravu_3x_kernel(Io io, const __half* __restrict__ lut, int total) {
  constexpr int P = CH == 0 ? 1 : 3;
  constexpr int N = 2 * R - 1;
  constexpr int HALO = R - 1;
  constexpr int PW = TX + 2 * HALO, PH = TY + 2 * HALO, PN = PW * PH;
  constexpr int C = (N * N - 1) / 2;
  constexpr int NTEX = 2 * (C + 1);
  __shared__ float tile[staged_planes<CH>() * PN];
  const FrameTile ft = frame_tile<TX, TY>(io, blockIdx.x, total);
  const int f = ft.f;
  const int x0 = ft.x0, y0 = ft.y0;
  const void* in = frame_in(io, f);
  stage_planes<P, PH, PW, TX * TY>(tile, in, io.in_dt, io.in_max, io.in_rmax, ft.in_pitch, ft.in_cstride, ft.W, ft.H, x0 - HALO, y0 - HALO);
  __syncthreads();
  if constexpr (CH == 2) {
    stage_luma<CH>(tile, PN);
    __syncthreads();
  }
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= ft.W || y >= ft.H) return;
  auto off = [&](int i) { return (ty + i % N) * PW + tx + i / N; };  // sample i = X*N + Y
  const Bucket bk = ravu_hash<N, false, RULE_3X>([&](int i) { return luma_f<CH>(tile, off(i), PN); },
                                                 [&](int i) { return luma_d<CH>(tile, off(i), PN); });
  const int row = bucket_row<RULE_3X>(bk, 216);
  if (io.buckets) io.buckets[f * io.bucket_bstride + static_cast<int64_t>(y) * ft.W + x] = row;
  const __half* lrow = lut + static_cast<int64_t>(row) * NTEX * 4;
  float r0[P][4], r1[P][4];
#pragma unroll
  for (int c = 0; c < P; ++c)
#pragma unroll
    for (int k = 0; k < 4; ++k) r0[c][k] = r1[c][k] = 0.f;
  // One step per tap pair (two LUT texels), software-pipelined: the next step's texels
  // are fetched (as raw FP16, converted inside the multiply-adds) while this step
  // computes, and each step's sums complete before the next step's samples are read, so
  // only one step's operands are live (r3 luma: 46 VGPRs instead of 134).
  const uint2* lraw = reinterpret_cast<const uint2*>(lrow);
  uint2 t0 = lraw[0], t1 = lraw[1];
#pragma unroll
  for (int i = 0; i <= C; ++i) {
    __builtin_amdgcn_sched_barrier(0);
    uint2 n0 = t0, n1 = t1;
    if (i < C) {
      n0 = lraw[2 * i + 2];
      n1 = lraw[2 * i + 3];
    }
    const __half2 h00 = bits_h2(t0.x), h01 = bits_h2(t0.y), h10 = bits_h2(t1.x), h11 = bits_h2(t1.y);
    const float a0[4] = {__low2float(h00), __high2float(h00), __low2float(h01), __high2float(h01)};
    const float a1[4] = {__low2float(h10), __high2float(h10), __low2float(h11), __high2float(h11)};
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const float li = tile[c * PN + off(i)];
      if (i < C) {
        const float lj = tile[c * PN + off(N * N - 1 - i)];
        // res0 += l_i*w0 + l_j*w1.wzyx, res1 += l_i*w1 + l_j*w0.wzyx, as four fused
        // multiply-adds (the LUT operand converts inside them, v_fma_mix)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          r0[c][k] = fmaf(lj, a1[3 - k], fmaf(li, a0[k], r0[c][k]));
          r1[c][k] = fmaf(lj, a0[3 - k], fmaf(li, a1[k], r1[c][k]));
        }
      } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          r0[c][k] = fmaf(li, a0[k], r0[c][k]);
          r1[c][k] = fmaf(li, a1[k], r1[c][k]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < P; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k) asm volatile("" : "+v"(r0[c][k]), "+v"(r1[c][k]));
    t0 = n0;
    t1 = n1;
  }
  const int esz_out = elem_bytes(io.out_dt);
  char* ob = static_cast<char*>(frame_out(io, f));
#pragma unroll
  for (int c = 0; c < P; ++c) {
    void* o = ob + c * ft.out_cstride * esz_out;
    const float centre = tile[c * PN + off(C)];
#pragma unroll
    for (int dy = 0; dy < 3; ++dy)
#pragma unroll
      for (int dx = 0; dx < 3; ++dx) {
        const int k = 3 * dx + dy;  // compute/ravu-3x-r3.hook:154-162
        float v = k < 4 ? r0[c][k] : (k == 4 ? centre : r1[c][k - 5]);
        if (k != 4) v = fminf(fmaxf(v, 0.f), 1.f);
        store_elem(o, static_cast<int64_t>(3 * y + dy) * ft.out_pitch + 3 * x + dx, io.out_dt, v);
      }
  }
}

template <int T>
static hipError_t launch_3x_tile(int r, int channels, Io io, const __half* lut, hipStream_t st) {
  constexpr int TX = k3xTiles[T].w, TY = k3xTiles[T].h;
  const int total = launch_tiles<TX, TY>(io);
#define MPX_3X(R, CH)                                                                                  \
  if (r == R && channels == CH) {                                                                      \
    hipLaunchKernelGGL((ravu_3x_kernel<R, CH, TX, TY>), dim3(total), dim3(TX * TY), 0, st, io, lut, total); \
    return hipGetLastError();                                                                          \
  }
  MPX_3X(2, 0) MPX_3X(2, 1) MPX_3X(2, 2) MPX_3X(3, 0) MPX_3X(3, 1) MPX_3X(3, 2) MPX_3X(4, 0) MPX_3X(4, 1) MPX_3X(4, 2)
#undef MPX_3X
  return hipErrorInvalidValue;
}

hipError_t launch_ravu_3x(int r, int channels, const Io& io, const __half* lut, int tile, hipStream_t st) {
  static_assert(k3xTileCount == 2, "one case per tile shape");
  switch (tile) {
    case 0: return launch_3x_tile<0>(r, channels, io, lut, st);
    case 1: return launch_3x_tile<1>(r, channels, io, lut, st);
    default: return hipErrorInvalidValue;
  }
}

// ----------------------------------------------------------------------------- RAVU full

// res = sum_{p < N^2/2} (s_p + s_{N^2-1-p}) * w_p, w_p = texel p/4, component p%4,
// clamped to [0,1] (ravu-r3.hook:120-146).  s(c, i) = plane c sample i.
template <int N> struct RavuLut {
  static constexpr int NTEX = (N * N / 2 + 3) / 4;  // texels per LUT row: 2 / 5 / 8 for r2 / r3 / r4
};

template <int N, int P, typename S>
__device__ __forceinline__ void ravu_taps(const __half* lrow, S s, float* res) {
#pragma unroll
  for (int c = 0; c < P; ++c) res[c] = 0.f;
#pragma unroll
  for (int t = 0; t < RavuLut<N>::NTEX; ++t) {
    // large windows (r4: 64 samples per plane): one LUT texel's taps at a time, so the
    // loads of all samples are not hoisted together (register pressure, hence occupancy;
    // small windows lose by it)
    if constexpr (N * N * P >= 64) __builtin_amdgcn_sched_barrier(0);
    const float4 w = load_lut4(lrow, t);
    const float a[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = 4 * t + j;
      if (p < N * N / 2) {
#pragma unroll
        for (int c = 0; c < P; ++c) res[c] += (s(c, p) + s(c, N * N - 1 - p)) * a[j];
      }
    }
    // ... and this texel's sums complete before the next texel's loads: without this the
    // multiply-adds sink below every load and all N*N*P samples are live at once (r4 yuv
    // lattice: 216 VGPRs, 2 waves per SIMD; with it 128, 4 waves)
    if constexpr (N * N * P >= 64) {
#pragma unroll
      for (int c = 0; c < P; ++c) asm volatile("" : "+v"(res[c]));
    }
  }
#pragma unroll
  for (int c = 0; c < P; ++c) res[c] = fminf(fmaxf(res[c], 0.f), 1.f);
}

// step1: int11 from the N = 2R window at offsets -(R-1)..R (ravu-r3.hook:22-146).
template <int R, int CH, int TX, int TY>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(6))) void ravu_step1_kernel(Io io, const __half* __restrict__ lut, float* int11,
                                                         int64_t int11_frame, int total) {
  constexpr int P = CH == 0 ? 1 : 3;
  constexpr int N = 2 * R;
  constexpr int PW = TX + N - 1, PH = TY + N - 1, PN = PW * PH;
  __shared__ float tile[staged_planes<CH>() * PN];
  const FrameTile ft = frame_tile<TX, TY>(io, blockIdx.x, total);
  const int f = ft.f;
  const int x0 = ft.x0, y0 = ft.y0;
  const void* in = frame_in(io, f);
  stage_planes<P, PH, PW, TX * TY>(tile, in, io.in_dt, io.in_max, io.in_rmax, ft.in_pitch, ft.in_cstride, ft.W, ft.H, x0 - (R - 1),
                  y0 - (R - 1));
  __syncthreads();
  if constexpr (CH == 2) {
    stage_luma<CH>(tile, PN);
    __syncthreads();
  }
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= ft.W || y >= ft.H) return;
  auto off = [&](int i) { return (ty + i % N) * PW + tx + i / N; };
  const Bucket bk = hash_site<N, true, RULE_LOG2, CH, PW, PN, true>(tile, off(0));
  const int row = bucket_row<RULE_LOG2>(bk, 648);
  if (io.buckets) io.buckets[f * io.bucket_bstride + static_cast<int64_t>(y) * ft.W + x] = row;
  float res[P];
  ravu_taps<N, P>(lut + static_cast<int64_t>(row) * RavuLut<N>::NTEX * 4,
                  [&](int c, int i) { return tile[c * PN + off(i)]; }, res);
  const int64_t plane = static_cast<int64_t>(ft.W) * ft.H;
  float* o = int11 + f * int11_frame + static_cast<int64_t>(y) * ft.W + x;
#pragma unroll
  for (int c = 0; c < P; ++c) o[c * plane] = io.store16 ? static_cast<float>(static_cast<_Float16>(res[c])) : res[c];
}

// Lattice position of sample (X, Y) of step 2 (int10, centre (x+1/2, y)) / step 3
// (int01, centre (x, y+1/2)): sample at centre + (X-(N-1)/2)*(1/2,-1/2) + (Y-(N-1)/2)*(1/2,1/2)
// (ravu-r3.hook:155-191, 289-325).  Returns doubled coordinates.
template <int N, int STEP>
__device__ __host__ constexpr int lat2x(int X, int Y) { return STEP == 2 ? X + Y - N + 2 : X + Y - N + 1; }
template <int N, int STEP>
__device__ __host__ constexpr int lat2y(int X, int Y) { return STEP == 2 ? Y - X : Y - X + 1; }

// One site's step 2 (int10, STEP = 2) or step 3 (int01, STEP = 3) on the lattice, and
// the combine for one output row: STEP 2 writes row 2y = (input, int10) pairs, STEP 3
// row 2y+1 = (int01, int11) pairs (ravu-r3.hook:426-437), so each step stores contiguous
// pairs.  hk / i11: the staged input and int11 tiles, base: the site's LDS offset.
template <int R, int CH, int STEP, int PW, int PN>
__device__ __forceinline__ void lattice_site(const Io& io, const __half* __restrict__ lut, const float* hk,
                                             const float* i11, int base, const FrameTile& ft, int x, int y,
                                             int64_t plane) {
  constexpr int P = CH == 0 ? 1 : 3;
  constexpr int N = 2 * R;
  auto src = [&](int i) -> const float* {
    const int X = i / N, Y = i % N;
    const int px = lat2x<N, STEP>(X, Y), py = lat2y<N, STEP>(X, Y);
    return (px & 1) ? i11 + base + ((py - 1) >> 1) * PW + ((px - 1) >> 1) : hk + base + (py >> 1) * PW + (px >> 1);
  };
  const Bucket bk = ravu_hash<N, true, RULE_LOG2, true>([&](int i) { return luma_f<CH>(src(i), 0, PN); },
                                                  [&](int i) { return luma_d<CH>(src(i), 0, PN); });
  const int row = bucket_row<RULE_LOG2>(bk, 648);
  if (io.buckets)
    io.buckets[ft.f * io.bucket_bstride + (STEP - 1) * plane + static_cast<int64_t>(y) * ft.W + x] = row;
  float res[P];
  ravu_taps<N, P>(lut + static_cast<int64_t>(row) * RavuLut<N>::NTEX * 4, [&](int c, int i) { return src(i)[c * PN]; },
                  res);
  const int esz_out = elem_bytes(io.out_dt);
  char* ob = static_cast<char*>(frame_out(io, ft.f));
#pragma unroll
  for (int c = 0; c < P; ++c) {
    void* o = ob + c * ft.out_cstride * esz_out;
    if (STEP == 2)
      store_pair(o, static_cast<int64_t>(2 * y) * ft.out_pitch + 2 * x, io.out_dt, hk[c * PN + base], res[c]);
    else
      store_pair(o, static_cast<int64_t>(2 * y + 1) * ft.out_pitch + 2 * x, io.out_dt, res[c], i11[c * PN + base]);
  }
}

// Steps 2 and 3 of a tile: both read the same staged input and int11 tiles (lattice
// offsets -R..R), so one kernel stages them once and runs the two steps one after the
// other per site; a scheduling barrier between them keeps one step's registers live at a
// time.
template <int R, int CH, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void ravu_lattice_kernel(Io io, const __half* __restrict__ lut,
                                                           const float* __restrict__ int11, int64_t int11_frame,
                                                           int total) {
  constexpr int P = CH == 0 ? 1 : 3;
  constexpr int PW = TX + 2 * R + 1, PH = TY + 2 * R + 1, PN = PW * PH;  // offsets -R..R
  __shared__ float hk[staged_planes<CH>() * PN];   // input (HOOKED)
  __shared__ float i11[staged_planes<CH>() * PN];  // ravu_int11
  const FrameTile ft = frame_tile<TX, TY>(io, blockIdx.x, total);
  const int f = ft.f;
  const int x0 = ft.x0, y0 = ft.y0;
  const void* in = frame_in(io, f);
  const int64_t plane = static_cast<int64_t>(ft.W) * ft.H;
  {
    HaloTile<P, PH, PW, TX * TY> a, b;  // both tiles' loads in flight together
    __builtin_amdgcn_s_setprio(1);  // the loads go out ahead of other waves' arithmetic
    a.issue(in, io.in_dt, ft.in_pitch, ft.in_cstride, ft.W, ft.H, x0 - R, y0 - R);
    b.issue(int11 + f * int11_frame, DT_F32, ft.W, plane, ft.W, ft.H, x0 - R, y0 - R);
    __builtin_amdgcn_s_setprio(0);
    a.commit(hk, io.in_dt, io.in_max, io.in_rmax);
    b.commit(i11, DT_F32, 1.f, 1.f);
  }
  __syncthreads();
  if constexpr (CH == 2) {
    stage_luma<CH>(hk, PN);
    stage_luma<CH>(i11, PN);
    __syncthreads();
  }
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int x = x0 + tx, y = y0 + ty;
  if (x >= ft.W || y >= ft.H) return;
  const int base = (ty + R) * PW + tx + R;  // LDS offset of (x, y)
  lattice_site<R, CH, 2, PW, PN>(io, lut, hk, i11, base, ft, x, y, plane);
  __builtin_amdgcn_sched_barrier(0);
  lattice_site<R, CH, 3, PW, PN>(io, lut, hk, i11, base, ft, x, y, plane);
}

template <int T>
static hipError_t launch_ravu_tile(int r, int channels, Io io, const __half* lut, float* int11, int64_t int11_frame,
                                   hipStream_t st, hipEvent_t done) {
  constexpr int TX = kRavuTiles[T].w, TY = kRavuTiles[T].h;
  const int total = launch_tiles<TX, TY>(io);
  const dim3 grid(total), block(TX * TY);
#define MPX_RAVU(R, CH)                                                                                         \
  if (r == R && channels == CH) {                                                                               \
    hipLaunchKernelGGL((ravu_step1_kernel<R, CH, TX, TY>), grid, block, 0, st, io, lut, int11, int11_frame, total); \
    hipError_t e = hipGetLastError();                                                                           \
    if (e != hipSuccess) return e;                                                                              \
    return launch_done(ravu_lattice_kernel<R, CH, TX, TY>, grid, block, st, done, io, lut, int11, int11_frame,      \
                       total);                                                                                  \
  }
  MPX_RAVU(2, 0) MPX_RAVU(2, 1) MPX_RAVU(2, 2) MPX_RAVU(3, 0) MPX_RAVU(3, 1) MPX_RAVU(3, 2) MPX_RAVU(4, 0)
  MPX_RAVU(4, 1) MPX_RAVU(4, 2)
#undef MPX_RAVU
  return hipErrorInvalidValue;
}

hipError_t launch_ravu(int r, int channels, const Io& io, const __half* lut, float* int11, int64_t int11_frame,
                       int tile, hipStream_t st, hipEvent_t done) {
  static_assert(kRavuTileCount == 3, "one case per tile shape");
  switch (tile) {
    case 0: return launch_ravu_tile<0>(r, channels, io, lut, int11, int11_frame, st, done);
    case 1: return launch_ravu_tile<1>(r, channels, io, lut, int11, int11_frame, st, done);
    case 2: return launch_ravu_tile<2>(r, channels, io, lut, int11, int11_frame, st, done);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mpx
