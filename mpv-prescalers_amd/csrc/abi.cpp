// The C ABI of include/mpx.h: variant registry, context (device + resident weights +
// HBM budget), plans (sizes, scratch, kernel choice) and dispatch.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpx.h"
#include "launch.h"
#include "nnedi3.h"
#include "ravu.h"
#include "internal.h"

namespace {

// ----------------------------------------------------------------------------- registry

using mpx_host::Variant;

const std::vector<Variant>& registry() {
  static const std::vector<Variant> v = [] {
    std::vector<Variant> out;
    const char* chs[3] = {"", "-yuv", "-rgb"};
    for (int nns : {16, 32, 64, 128, 256})
      for (int wh : {4, 6})
        out.push_back({"nnedi3-nns" + std::to_string(nns) + "-win8x" + std::to_string(wh),
                       MPX_FAMILY_NNEDI3, MPX_CH_LUMA, 0, nns, wh, 0, 1});
    for (int r : {2, 3, 4})
      for (int c = 0; c < 3; ++c)
        out.push_back({"ravu-r" + std::to_string(r) + chs[c], MPX_FAMILY_RAVU, c, r, 0, 0, 0, 1});
    for (int ar : {0, 1})
      for (int r : {2, 3, 4})
        out.push_back({std::string(ar ? "ravu-lite-ar-r" : "ravu-lite-r") + std::to_string(r),
                       MPX_FAMILY_RAVU_LITE, MPX_CH_LUMA, r, 0, 0, ar, 1});
    for (int r : {2, 3, 4})
      for (int c = 0; c < 3; ++c)
        out.push_back({"ravu-3x-r" + std::to_string(r) + chs[c], MPX_FAMILY_RAVU_3X, c, r, 0, 0, 0, 1});
    for (int ar : {0, 1})
      for (int r : {2, 3})
        for (int c = 0; c < 3; ++c)
          // ravu-zoom-ar-r3*: listed in .MISSING_LARGE_BLOBS:1-7, weights not shipped
          out.push_back({std::string(ar ? "ravu-zoom-ar-r" : "ravu-zoom-r") + std::to_string(r) + chs[c],
                         MPX_FAMILY_RAVU_ZOOM, c, r, 0, 0, ar, (ar && r == 3) ? 0 : 1});
    return out;
  }();
  return v;
}

const Variant* find_variant(const char* name) {
  if (!name) return nullptr;
  for (const Variant& v : registry())
    if (v.name == name) return &v;
  return nullptr;
}

// mpv evaluates //!WHEN in float: `HOOKED.w OUTPUT.w / 0.707106 <`.
bool ratio_below(int in, int target, float thr) {
  return static_cast<float>(in) / static_cast<float>(target) < thr;
}

void fill_info(const Variant& v, int in_w, int in_h, int tw, int th, mpx_variant_info* info) {
  std::memset(info, 0, sizeof(*info));
  info->family = v.family;
  info->channels = v.channels;
  info->planes = v.channels == MPX_CH_LUMA ? 1 : 3;
  info->radius = v.radius;
  info->nns = v.nns;
  info->win_w = v.nns ? 8 : 0;
  info->win_h = v.win_h;
  info->ar = v.ar;
  info->available = v.available;
  std::strcpy(info->hook, v.channels == MPX_CH_LUMA ? "LUMA" : (v.channels == MPX_CH_YUV ? "NATIVE" : "MAIN"));
  info->out_w = in_w;
  info->out_h = in_h;
  switch (v.family) {
    case MPX_FAMILY_NNEDI3: {
      // double_y: nnedi3-nns256-win8x6.hook:19; double_x: :364
      const bool ay = ratio_below(in_h, th, 0.707106f), ax = ratio_below(in_w, tw, 0.707106f);
      info->applies = ax || ay;
      info->scale_x = ax ? 2 : 1;
      info->scale_y = ay ? 2 : 1;
      info->out_w = in_w * info->scale_x;
      info->out_h = in_h * info->scale_y;
      info->offset_x = ax ? -0.5f : 0.f;  // combine_x OFFSET -0.5 0 (:697)
      info->offset_y = ay ? -0.5f : 0.f;  // combine_y OFFSET 0 -0.5 (:351)
      break;
    }
    case MPX_FAMILY_RAVU:
    case MPX_FAMILY_RAVU_LITE: {
      // ravu-r3.hook:20, ravu-lite-r3.hook:20 (both axes, product of the two tests)
      info->applies = ratio_below(in_w, tw, 0.707106f) && ratio_below(in_h, th, 0.707106f);
      info->scale_x = info->scale_y = info->applies ? 2 : 1;
      info->out_w = in_w * info->scale_x;
      info->out_h = in_h * info->scale_y;
      if (v.family == MPX_FAMILY_RAVU && info->applies) info->offset_x = info->offset_y = -0.5f;  // ravu-r3.hook:424
      break;
    }
    case MPX_FAMILY_RAVU_3X:
      // compute/ravu-3x-r3.hook:21
      info->applies = ratio_below(in_w, tw, 0.471404f) && ratio_below(in_h, th, 0.471404f);
      info->scale_x = info->scale_y = info->applies ? 3 : 1;
      info->out_w = in_w * info->scale_x;
      info->out_h = in_h * info->scale_y;
      break;
    case MPX_FAMILY_RAVU_ZOOM:
      // ravu-zoom-r2.hook:18-21: WIDTH OUTPUT.w, HEIGHT OUTPUT.h, OFFSET ALIGN, in < out
      info->applies = in_w < tw && in_h < th;
      info->scale_x = info->scale_y = info->applies ? 0 : 1;
      if (info->applies) {
        info->out_w = tw;
        info->out_h = th;
        info->align = 1;
      }
      break;
  }
}

std::string default_weights_dir() {
  if (const char* e = std::getenv("MPX_WEIGHTS_DIR")) return e;
  Dl_info dl;
  if (dladdr(reinterpret_cast<void*>(&default_weights_dir), &dl) && dl.dli_fname) {
    std::string p(dl.dli_fname);
    auto pos = p.rfind('/');
    std::string dir = pos == std::string::npos ? "." : p.substr(0, pos);
    return dir + "/../weights";
  }
  return "weights";
}

int elem_size(int t) { return t == MPX_F32 ? 4 : (t == MPX_U8 ? 1 : 2); }

// Kernels index tiles and blocks of one launch with 32-bit integers.
constexpr int64_t kMaxLaunchPixels = (int64_t{1} << 31) - 1;

// The staging loaders address an input frame's samples with 32-bit element offsets (plane
// * cstride + row * pitch + column, the row term a 24-bit multiply): a frame may span at
// most 2^32 elements (16 GiB of FP32) with a pitch below 2^24 elements.
constexpr int64_t kMaxSpan = int64_t{1} << 32;
constexpr int64_t kMaxPitch = int64_t{1} << 24;
int check_span(int64_t pitch, int64_t cstride, int planes, int h) {
  const int64_t span = (planes > 1 ? (planes - 1) * cstride : 0) + static_cast<int64_t>(h) * pitch;
  if (pitch >= kMaxPitch || cstride < 0 || span > kMaxSpan)
    return fail(MPX_ERR_SHAPE, "input frame spans 2^32 elements or more, or its pitch 2^24 or more");
  return MPX_OK;
}

}  // namespace

// ----------------------------------------------------------------------------- static scratch

// The intermediates of mpx_run_static.  The array lives in this file's code object, so each
// device that loads the library holds its own copy until the process ends, and no call
// allocates it.  On the host, `scratch` names a host-side stand-in: kernels are given the
// device address that hipGetSymbolAddress returns for the current device, never the name.
static __device__ __align__(256) unsigned char scratch[MPX_STATIC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;  // device -> device address of its `scratch`

// The device address of `device`'s copy of `scratch`, resolved once per device with that
// device current (the caller has made it current).
int scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end()) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(MPX_ERR_HIP, "hipGetSymbolAddress returned a NULL address for the static scratch");
    it = g_scratch_ptr.emplace(device, p).first;
  }
  *out = it->second;
  return MPX_OK;
}

}  // namespace

// ----------------------------------------------------------------------------- context

extern "C" {

int mpx_variant_count(void) { return static_cast<int>(registry().size()); }

const char* mpx_variant_name(int i) {
  if (i < 0 || i >= mpx_variant_count()) return nullptr;
  return registry()[i].name.c_str();
}

int mpx_query(const char* name, int in_w, int in_h, int target_w, int target_h, mpx_variant_info* info) {
  const Variant* v = find_variant(name);
  if (!v) return fail(MPX_ERR_UNKNOWN_VARIANT, std::string("unknown variant '") + (name ? name : "(null)") + "'");
  if (!info) return fail(MPX_ERR_ARG, "info is NULL");
  if (in_w <= 0 || in_h <= 0 || target_w <= 0 || target_h <= 0) return fail(MPX_ERR_SHAPE, "sizes must be positive");
  fill_info(*v, in_w, in_h, target_w, target_h, info);
  return MPX_OK;
}

int mpx_weights_verify(const char* weights_dir, const char* name) {
  const Variant* v = find_variant(name);
  if (!v) return fail(MPX_ERR_UNKNOWN_VARIANT, std::string("unknown variant '") + (name ? name : "(null)") + "'");
  const std::string dir = weights_dir ? weights_dir : default_weights_dir();
  for (const std::string& t : mpx_host::variant_tables(*v)) {
    if (!v->available && t.find("_ar") != std::string::npos) {
      FILE* f = std::fopen((dir + "/" + t + ".bin").c_str(), "rb");
      if (!f)
        return fail(MPX_ERR_UNAVAILABLE, std::string(name) + ": " + t + ".bin is not shipped by the reference");
      std::fclose(f);
    }
    std::vector<float> data;
    std::vector<uint32_t> dims;
    if (int rc = mpx_host::load_table(dir, t, data, dims)) return rc;
  }
  return MPX_OK;
}

int mpx_create(int device, uint64_t hbm_budget, const char* weights_dir, mpx_ctx** out) {
  if (!out) return fail(MPX_ERR_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(MPX_ERR_ARG, "device " + std::to_string(device) + " out of range");
  void* static_scratch = nullptr;
  {
    mpx_host::DeviceScope dscope;
    if (int drc = dscope.enter(device)) return drc;
    if (int src = scratch_address(device, &static_scratch)) return src;
  }
  auto* c = new mpx_ctx();
  c->device = device;
  c->static_scratch = static_scratch;
  c->budget = hbm_budget;
  c->wdir = weights_dir ? weights_dir : default_weights_dir();
  *out = c;
  return MPX_OK;
}

void mpx_destroy(mpx_ctx* ctx) {
  if (!ctx) return;
  for (auto& kv : ctx->weights) ctx->release(kv.second.ptr, kv.second.bytes);
  delete ctx;
}

uint64_t mpx_ctx_bytes(const mpx_ctx* ctx) { return ctx ? ctx->used : 0; }

int mpx_plan_create(mpx_ctx* ctx, const char* name, int in_w, int in_h, int target_w, int target_h, int in_type,
                    int bits, int out_type, int precision, int max_batch, mpx_plan** out) {
  if (!ctx || !out) return fail(MPX_ERR_ARG, "ctx/out is NULL");
  *out = nullptr;
  const Variant* v = find_variant(name);
  if (!v) return fail(MPX_ERR_UNKNOWN_VARIANT, std::string("unknown variant '") + (name ? name : "(null)") + "'");
  if (!v->available) {
    // The reference does not ship these weights; a user-supplied table in the weights
    // directory (same MPXW format, (2592, 45, 4)) makes the variant usable.
    FILE* fchk = std::fopen((ctx->wdir + "/ravu_zoom_lut" + std::to_string(v->radius) + "_ar.bin").c_str(), "rb");
    if (!fchk)
      return fail(MPX_ERR_UNAVAILABLE, std::string(name) +
                                           ": weights are not shipped by the reference (listed in "
                                           ".MISSING_LARGE_BLOBS); supply ravu_zoom_lut" +
                                           std::to_string(v->radius) + "_ar.bin in the weights directory");
    std::fclose(fchk);
  }
  // every table the variant reads is checked (shape, size, SHA-256 of shipped data) before
  // any of them is uploaded; resident tables were checked when they were loaded
  for (const std::string& t : mpx_host::variant_tables(*v)) {
    if (ctx->has(t)) continue;
    std::vector<float> data;
    std::vector<uint32_t> dims;
    if (int rc = mpx_host::load_table(ctx->wdir, t, data, dims)) return rc;
  }
  if (in_w <= 0 || in_h <= 0 || target_w <= 0 || target_h <= 0 || in_w > 32768 || in_h > 32768 ||
      target_w > 65536 || target_h > 65536)
    return fail(MPX_ERR_SHAPE, "sizes out of range");
  if (in_type < MPX_U8 || in_type > MPX_F32) return fail(MPX_ERR_ARG, "bad in_type");
  if (out_type != MPX_F16 && out_type != MPX_F32) return fail(MPX_ERR_ARG, "out_type must be MPX_F16 or MPX_F32");
  // the pass-through copy launches one grid layer per frame and plane (at most 65535)
  if (max_batch < 1 || max_batch * (v->channels == MPX_CH_LUMA ? 1 : 3) > 65535)
    return fail(MPX_ERR_SHAPE, "max_batch out of range");
  if (in_type == MPX_U16 && (bits < 1 || bits > 16)) return fail(MPX_ERR_ARG, "bits must be 1..16 for MPX_U16");
  const bool store16 = (precision & MPX_STORE_FP16) != 0;
  precision &= ~MPX_STORE_FP16;
  if (precision < MPX_PREC_FP32 || precision > MPX_PREC_FP16X3) return fail(MPX_ERR_ARG, "bad precision");
  if (v->family != MPX_FAMILY_NNEDI3 && precision != MPX_PREC_FP32)
    return fail(MPX_ERR_UNSUPPORTED, "RAVU variants compute in FP32 only");

  std::unique_ptr<mpx_plan> p(new mpx_plan());
  p->ctx = ctx;
  p->v = v;
  fill_info(*v, in_w, in_h, target_w, target_h, &p->info);
  p->in_w = in_w;
  p->in_h = in_h;
  p->in_type = in_type;
  p->out_type = out_type;
  p->precision = precision;
  p->store16 = store16 ? 1 : 0;
  p->max_batch = max_batch;
  p->in_max = in_type == MPX_U8 ? 255.f : (in_type == MPX_U16 ? static_cast<float>((1u << bits) - 1u) : 1.f);
  const bool applies = p->info.applies != 0;
  int rc = MPX_OK;
  if (applies) {
    const std::string r = std::to_string(v->radius);
    switch (v->family) {
      case MPX_FAMILY_RAVU_LITE:
        rc = ctx->lut16("ravu_lite_lut" + r, &p->lut, nullptr);
        break;
      case MPX_FAMILY_RAVU:
        rc = ctx->lut16("ravu_lut" + r, &p->lut, nullptr);
        p->scratch_frame = static_cast<int64_t>(in_w) * in_h * p->info.planes;  // ravu_int11, FP32
        break;
      case MPX_FAMILY_RAVU_3X:
        rc = ctx->lut16("ravu_3x_lut" + r, &p->lut, nullptr);
        break;
      case MPX_FAMILY_RAVU_ZOOM:
        rc = ctx->lut16("ravu_zoom_lut" + r, &p->lut, nullptr);
        if (!rc && v->ar) rc = ctx->lut16("ravu_zoom_lut" + r + "_ar", &p->lut_ar, nullptr);
        if (!rc) rc = mpx::zoom_plan_create(ctx, p.get(), in_w, in_h, p->info.out_w, p->info.out_h);
        break;
      case MPX_FAMILY_NNEDI3:
        rc = mpx::nnedi3_prepare(ctx, v->nns, v->win_h, precision, &p->nw);
        if (p->info.scale_x == 2 && p->info.scale_y == 2)
          p->scratch_frame = static_cast<int64_t>(in_w) * in_h * 2;  // double_y output, FP32
        break;
    }
  }
  if (rc) return rc;
  if (p->scratch_frame) {
    p->scratch_bytes = static_cast<size_t>(p->scratch_frame) * 4 * max_batch;
    rc = ctx->alloc(p->scratch_bytes, &p->scratch);
    if (rc) return rc;
    mpx_host::DeviceScope dscope;
    if (int drc = dscope.enter(ctx->device)) {
      ctx->release(p->scratch, p->scratch_bytes);
      return drc;
    }
    // The completion event only tells the host whether the plan's intermediates are still
    // in use on the device; no host access depends on it, so it carries no system-scope
    // fence (with one, each run ended in an L2 writeback: 5.7 us idle before the next run).
    if (hipEventCreateWithFlags(&p->done, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) {
      ctx->release(p->scratch, p->scratch_bytes);
      return fail(MPX_ERR_HIP, "hipEventCreateWithFlags failed");
    }
  }
  *out = p.release();
  return applies ? MPX_OK : MPX_SKIPPED;
}

int mpx_plan_info(const mpx_plan* plan, mpx_variant_info* info) {
  if (!plan || !info) return fail(MPX_ERR_ARG, "plan/info is NULL");
  *info = plan->info;
  return MPX_OK;
}

int64_t mpx_plan_bucket_count(const mpx_plan* plan) {
  if (!plan || !plan->info.applies) return 0;
  const int64_t px = static_cast<int64_t>(plan->in_w) * plan->in_h;
  switch (plan->v->family) {
    case MPX_FAMILY_RAVU_LITE:
    case MPX_FAMILY_RAVU_3X: return px;
    case MPX_FAMILY_RAVU: return 3 * px;  // int11, int10, int01 sites
    case MPX_FAMILY_RAVU_ZOOM: return static_cast<int64_t>(plan->info.out_w) * plan->info.out_h;
    default: return 0;
  }
}

void mpx_plan_destroy(mpx_plan* plan) {
  if (!plan) return;
  if (plan->scratch) plan->ctx->release(plan->scratch, plan->scratch_bytes);
  if (plan->done) {
    mpx_host::DeviceScope dscope;
    (void)dscope.enter(plan->ctx->device);
    (void)hipEventDestroy(plan->done);
  }
  mpx::zoom_plan_destroy(plan->ctx, plan);
  delete plan;
}

static int dispatch(mpx_plan* plan, const mpx::Io& d, hipStream_t st, void* interm);

int mpx_run_ex(mpx_plan* plan, const mpx_io* io, void* hip_stream) {
  return mpx_host::run_io(plan, io, hip_stream, plan ? plan->scratch : nullptr);
}

int mpx_run_static(mpx_plan* plan, const mpx_io* io, void* hip_stream) {
  if (!plan || !io) return fail(MPX_ERR_ARG, "plan/io is NULL");
  if (io->batch >= 1 && plan->scratch_frame) {
    const int64_t need = plan->scratch_frame * 4 * static_cast<int64_t>(io->batch);
    if (need > static_cast<int64_t>(MPX_STATIC_SCRATCH_BYTES))
      return fail(MPX_ERR_SHAPE, "intermediates of " + std::to_string(io->batch) + " frame(s) need " +
                                     std::to_string(need) + " B, more than the static scratch's " +
                                     std::to_string(MPX_STATIC_SCRATCH_BYTES) + " B");
  }
  if (!plan->ctx->static_scratch) return fail(MPX_ERR_HIP, "the context has no static scratch address");
  return mpx_host::run_io(plan, io, hip_stream, plan->ctx->static_scratch);
}

}  // extern "C"

int mpx_host::run_io(mpx_plan* plan, const mpx_io* io, void* hip_stream, void* interm) {
  if (!plan || !io) return fail(MPX_ERR_ARG, "plan/io is NULL");
  if (!io->in || !io->out) return fail(MPX_ERR_ARG, "in/out is NULL");
  if (io->batch < 1 || io->batch > plan->max_batch)
    return fail(MPX_ERR_SHAPE, "batch " + std::to_string(io->batch) + " outside 1.." + std::to_string(plan->max_batch));
  const mpx_variant_info& inf = plan->info;
  if (io->in_pitch < plan->in_w || io->out_pitch < inf.out_w)
    return fail(MPX_ERR_ARG, "pitch smaller than the row width");
  if (inf.planes > 1 && (io->in_cstride < io->in_pitch * plan->in_h || io->out_cstride < io->out_pitch * inf.out_h))
    return fail(MPX_ERR_ARG, "plane stride smaller than a plane");
  if (io->batch > 1 && (io->in_bstride < io->in_pitch * plan->in_h * inf.planes ||
                        io->out_bstride < io->out_pitch * inf.out_h * inf.planes))
    return fail(MPX_ERR_ARG, "frame stride smaller than a frame");
  if (io->buckets && !inf.applies) return fail(MPX_ERR_ARG, "buckets requested from a skipped plan");
  if (static_cast<int64_t>(inf.out_w) * inf.out_h * io->batch > kMaxLaunchPixels)
    return fail(MPX_ERR_SHAPE, "more than 2^31 output pixels per plane in one launch");
  if (int rc = check_span(io->in_pitch, io->in_cstride, inf.planes, plan->in_h)) return rc;
  mpx_host::DeviceScope dscope;
  if (int drc = dscope.enter(plan->ctx->device)) return drc;
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  mpx::Io d{};
  d.ntab = 0;
  d.in = io->in;
  d.out = io->out;
  d.in_pitch = io->in_pitch;
  d.in_cstride = io->in_cstride;
  d.in_bstride = io->in_bstride;
  d.out_pitch = io->out_pitch;
  d.out_cstride = io->out_cstride;
  d.out_bstride = io->out_bstride;
  d.buckets = io->buckets;
  d.bucket_bstride = mpx_plan_bucket_count(plan);
  d.in_dt = plan->in_type;
  d.out_dt = plan->out_type;
  d.in_max = plan->in_max;
  d.in_rmax = 1.0f / d.in_max;
  d.store16 = plan->store16;
  d.W = plan->in_w;
  d.H = plan->in_h;
  d.batch = io->batch;
  return dispatch(plan, d, st, interm);
}

extern "C" {

int mpx_run_frames(mpx_plan* plan, int n, const void* const* ins, void* const* outs, int64_t in_pitch,
                   int64_t out_pitch, void* hip_stream) {
  if (!plan || !ins || !outs) return fail(MPX_ERR_ARG, "plan/ins/outs is NULL");
  if (n < 1 || n > MPX_MAX_FRAMES || n > plan->max_batch)
    return fail(MPX_ERR_SHAPE, "n = " + std::to_string(n) + " outside 1.." +
                                   std::to_string(std::min(MPX_MAX_FRAMES, plan->max_batch)));
  const mpx_variant_info& inf = plan->info;
  if (in_pitch < plan->in_w || out_pitch < inf.out_w) return fail(MPX_ERR_ARG, "pitch smaller than the row width");
  for (int i = 0; i < n; ++i)
    if (!ins[i] || !outs[i]) return fail(MPX_ERR_ARG, "frame " + std::to_string(i) + " has a NULL plane");
  if (static_cast<int64_t>(inf.out_w) * inf.out_h * n > kMaxLaunchPixels)
    return fail(MPX_ERR_SHAPE, "more than 2^31 output pixels per plane in one launch");
  if (int rc = check_span(in_pitch, in_pitch * plan->in_h, inf.planes, plan->in_h)) return rc;
  mpx_host::DeviceScope dscope;
  if (int drc = dscope.enter(plan->ctx->device)) return drc;
  mpx::Io d{};
  d.ntab = n;
  for (int i = 0; i < n; ++i) {
    d.tab_in[i] = ins[i];
    d.tab_out[i] = outs[i];
  }
  d.in = ins[0];
  d.out = outs[0];
  d.in_pitch = in_pitch;
  d.in_cstride = in_pitch * plan->in_h;
  d.out_pitch = out_pitch;
  d.out_cstride = out_pitch * inf.out_h;
  d.in_dt = plan->in_type;
  d.out_dt = plan->out_type;
  d.in_max = plan->in_max;
  d.in_rmax = 1.0f / d.in_max;
  d.store16 = plan->store16;
  d.W = plan->in_w;
  d.H = plan->in_h;
  d.batch = n;
  return dispatch(plan, d, static_cast<hipStream_t>(hip_stream), plan->scratch);
}

int mpx_run_frames_sized(mpx_plan* plan, int n, const void* const* ins, void* const* outs, const int* widths,
                         const int* heights, const int64_t* in_pitches, const int64_t* out_pitches,
                         void* hip_stream) {
  if (!plan || !ins || !outs || !widths || !heights || !in_pitches || !out_pitches)
    return fail(MPX_ERR_ARG, "NULL argument");
  if (n < 1 || n > MPX_MAX_FRAMES || n > plan->max_batch)
    return fail(MPX_ERR_SHAPE, "n = " + std::to_string(n) + " outside 1.." +
                                   std::to_string(std::min(MPX_MAX_FRAMES, plan->max_batch)));
  const mpx_variant_info& inf = plan->info;
  if (!inf.applies) return fail(MPX_ERR_UNSUPPORTED, "frames of different sizes need a plan whose hook applies");
  if (plan->v->family == MPX_FAMILY_RAVU_ZOOM)
    return fail(MPX_ERR_UNSUPPORTED, "ravu-zoom scales to the plan's target size: one size per launch");
  mpx::Io d{};
  int64_t pixels = 0;
  for (int i = 0; i < n; ++i) {
    if (!ins[i] || !outs[i]) return fail(MPX_ERR_ARG, "frame " + std::to_string(i) + " has a NULL plane");
    if (widths[i] < 1 || heights[i] < 1 || widths[i] > plan->in_w || heights[i] > plan->in_h)
      return fail(MPX_ERR_SHAPE, "frame " + std::to_string(i) + " is not within the plan's input size");
    if (in_pitches[i] < widths[i] || out_pitches[i] < static_cast<int64_t>(widths[i]) * inf.scale_x)
      return fail(MPX_ERR_ARG, "frame " + std::to_string(i) + ": pitch smaller than the row width");
    if (int rc = check_span(in_pitches[i], in_pitches[i] * heights[i], inf.planes, heights[i])) return rc;
    d.tab_in[i] = ins[i];
    d.tab_out[i] = outs[i];
    d.fw[i] = widths[i];
    d.fh[i] = heights[i];
    d.fin_pitch[i] = in_pitches[i];
    d.fout_pitch[i] = out_pitches[i];
    pixels += static_cast<int64_t>(widths[i]) * inf.scale_x * heights[i] * inf.scale_y;
  }
  if (pixels > kMaxLaunchPixels) return fail(MPX_ERR_SHAPE, "more than 2^31 output pixels per plane in one launch");
  mpx_host::DeviceScope dscope;
  if (int drc = dscope.enter(plan->ctx->device)) return drc;
  d.ntab = n;
  d.sized = 1;
  d.sx = inf.scale_x;
  d.sy = inf.scale_y;
  d.in_dt = plan->in_type;
  d.out_dt = plan->out_type;
  d.in_max = plan->in_max;
  d.in_rmax = 1.0f / d.in_max;
  d.store16 = plan->store16;
  d.W = plan->in_w;
  d.H = plan->in_h;
  d.batch = n;
  return dispatch(plan, d, static_cast<hipStream_t>(hip_stream), plan->scratch);
}

// done (optional): an event the run's last kernel completes.
static int launch_all(mpx_plan* plan, const mpx::Io& d, hipStream_t st, void* interm, hipEvent_t done = nullptr);

// The plan's own intermediates serve one stream at a time.  A run on another stream while
// the last run's work may still be in flight would race on them: it is refused before
// anything is enqueued.  Runs captured into a graph are not tracked (the graph's launches
// are ordered by the caller); mpx_run_static's array is the caller's to serialise.
static int dispatch(mpx_plan* plan, const mpx::Io& d, hipStream_t st, void* interm) {
  if (!plan->done || interm != plan->scratch) return launch_all(plan, d, st, interm);
  std::lock_guard<std::mutex> lk(plan->run_mu);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) return fail(MPX_ERR_HIP, "hipStreamIsCapturing failed");
  if (cs != hipStreamCaptureStatusNone) return launch_all(plan, d, st, interm);
  if (plan->ran && st != plan->last) {
    const hipError_t q = hipEventQuery(plan->done);
    if (q == hipErrorNotReady)
      return fail(MPX_ERR_ARG, "the plan's intermediates are in use by its previous run on another stream: use one "
                               "plan per stream, or wait for that stream first");
    if (q != hipSuccess) return fail(MPX_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
  }
  // the run's last kernel completes plan->done (no separate event record after it)
  if (int rc = launch_all(plan, d, st, interm, plan->done)) return rc;
  plan->last = st;
  plan->ran = true;
  return MPX_OK;
}

static int launch_all(mpx_plan* plan, const mpx::Io& d, hipStream_t st, void* interm, hipEvent_t done) {
  const mpx_variant_info& inf = plan->info;
  hipError_t e = hipSuccess;
  bool done_bound = false;  // whether the last launch carried `done`
  if (!inf.applies) {
    e = mpx::launch_copy(d, inf.planes, st);
  } else {
    const Variant& v = *plan->v;
    switch (v.family) {
      case MPX_FAMILY_RAVU_LITE:
        e = mpx::launch_ravu_lite(v.radius, v.ar != 0, d, plan->lut, plan->tile, st);
        break;
      case MPX_FAMILY_RAVU_3X:
        e = mpx::launch_ravu_3x(v.radius, v.channels, d, plan->lut, plan->tile, st);
        break;
      case MPX_FAMILY_RAVU:
        e = mpx::launch_ravu(v.radius, v.channels, d, plan->lut, static_cast<float*>(interm), plan->scratch_frame,
                             plan->tile, st, done);
        done_bound = true;
        break;
      case MPX_FAMILY_RAVU_ZOOM:
        e = mpx::launch_ravu_zoom(v.radius, v.ar != 0, v.channels, d, plan->lut, plan->lut_ar, plan->zoom,
                                  inf.out_w, inf.out_h, plan->tile, st);
        break;
      case MPX_FAMILY_NNEDI3:
        e = mpx::launch_nnedi3(plan->nw, d, inf.scale_x == 2, inf.scale_y == 2, static_cast<float*>(interm),
                               plan->scratch_frame, st, done);
        done_bound = true;
        break;
    }
  }
  if (e != hipSuccess) return fail(MPX_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  if (done && !done_bound) HIP_TRY(hipEventRecord(done, st));
  return MPX_OK;
}

int mpx_run(mpx_plan* plan, const void* in, int64_t in_pitch, int64_t in_bstride, void* out, int64_t out_pitch,
            int64_t out_bstride, int batch, void* hip_stream) {
  if (!plan) return fail(MPX_ERR_ARG, "plan is NULL");
  mpx_io io{};
  io.in = in;
  io.in_pitch = in_pitch;
  io.in_cstride = in_pitch * plan->in_h;
  io.in_bstride = in_bstride;
  io.out = out;
  io.out_pitch = out_pitch;
  io.out_cstride = out_pitch * plan->info.out_h;
  io.out_bstride = out_bstride;
  io.batch = batch;
  io.buckets = nullptr;
  return mpx_run_ex(plan, &io, hip_stream);
}

const char* mpx_last_error(void) { return mpx_host::last_error().c_str(); }

}  // extern "C"

int mpx_resample(const mpx_io* io, int in_w, int in_h, int in_type, int bits, int out_w, int out_h, int out_type,
                 int planes, int filter, float offset_x, float offset_y, void* hip_stream) {
  if (!io || !io->in || !io->out) return fail(MPX_ERR_ARG, "io/in/out is NULL");
  if (filter < MPX_FILTER_BILINEAR || filter > MPX_FILTER_LANCZOS3) return fail(MPX_ERR_ARG, "bad filter");
  if (in_type < MPX_U8 || in_type > MPX_F32 || (out_type != MPX_F16 && out_type != MPX_F32))
    return fail(MPX_ERR_ARG, "bad element type");
  if (in_type == MPX_U16 && (bits < 9 || bits > 16)) return fail(MPX_ERR_ARG, "bits outside 9..16");
  if (planes != 1 && planes != 3) return fail(MPX_ERR_ARG, "planes must be 1 or 3");
  if (in_w < 1 || in_h < 1 || out_w < 1 || out_h < 1 || in_w > 32768 || in_h > 32768 || out_w > 32768 ||
      out_h > 32768)
    return fail(MPX_ERR_SHAPE, "sizes outside 1..32768");
  if (!std::isfinite(offset_x) || !std::isfinite(offset_y) || std::fabs(offset_x) > 64.f || std::fabs(offset_y) > 64.f)
    return fail(MPX_ERR_ARG, "offset outside [-64, 64]");
  if (io->batch < 1 || io->batch * planes > 65535) return fail(MPX_ERR_SHAPE, "batch * planes outside 1..65535");
  if (io->in_pitch < in_w || io->out_pitch < out_w) return fail(MPX_ERR_ARG, "pitch smaller than the row width");
  if (planes > 1 && (io->in_cstride < io->in_pitch * in_h || io->out_cstride < io->out_pitch * out_h))
    return fail(MPX_ERR_ARG, "plane stride smaller than a plane");
  if (io->batch > 1 &&
      (io->in_bstride < io->in_pitch * in_h * planes || io->out_bstride < io->out_pitch * out_h * planes))
    return fail(MPX_ERR_ARG, "frame stride smaller than a frame");
  mpx::Io d{};
  d.in = io->in;
  d.out = io->out;
  d.in_pitch = io->in_pitch;
  d.in_cstride = io->in_cstride;
  d.in_bstride = io->in_bstride;
  d.out_pitch = io->out_pitch;
  d.out_cstride = io->out_cstride;
  d.out_bstride = io->out_bstride;
  d.in_dt = in_type;
  d.out_dt = out_type;
  d.in_max = in_type == MPX_U8 ? 255.f : (in_type == MPX_U16 ? static_cast<float>((1 << bits) - 1) : 1.f);
  d.in_rmax = 1.0f / d.in_max;
  d.W = in_w;
  d.H = in_h;
  d.batch = io->batch;
  HIP_TRY(mpx::launch_resample(d, out_w, out_h, planes, filter, offset_x, offset_y,
                               static_cast<hipStream_t>(hip_stream)));
  return MPX_OK;
}

int mpx_plan_set_shape(mpx_plan* plan, int grid_per_cu, int split) {
  if (!plan) return fail(MPX_ERR_ARG, "plan is NULL");
  const bool nn = plan->v->family == MPX_FAMILY_NNEDI3 && plan->precision != MPX_PREC_FP32;
  if (!nn) {
    if (grid_per_cu != 0 || split != 0) return fail(MPX_ERR_ARG, "this plan has no tunable launch shape");
    return MPX_OK;
  }
  if (grid_per_cu < 0 || grid_per_cu > 64) return fail(MPX_ERR_ARG, "grid_per_cu outside 0..64");
  const bool can_split = plan->v->nns == 64 || plan->v->nns == 128;
  if (split < 0 || split > 1 || (split && !can_split))
    return fail(MPX_ERR_ARG, "split must be 0, or 1 for nns 64 / 128");
  plan->nw.grid_per_cu = grid_per_cu;
  plan->nw.split = split;
  return MPX_OK;
}

// Tile shapes of the plan's family (1 for families without a choice) and the shape of
// candidate t in pixels.
static int tile_count(const mpx_plan* plan) {
  switch (plan->v->family) {
    case MPX_FAMILY_RAVU_LITE: return mpx::kLiteTileCount;
    case MPX_FAMILY_RAVU_3X: return mpx::k3xTileCount;
    case MPX_FAMILY_RAVU: return mpx::kRavuTileCount;
    case MPX_FAMILY_RAVU_ZOOM: return mpx::kZoomTileCount;
    default: return 1;
  }
}
static mpx::TileShape tile_shape(const mpx_plan* plan, int t) {
  switch (plan->v->family) {
    case MPX_FAMILY_RAVU_LITE: return mpx::kLiteTiles[t];
    case MPX_FAMILY_RAVU_3X: return mpx::k3xTiles[t];
    case MPX_FAMILY_RAVU: return mpx::kRavuTiles[t];
    case MPX_FAMILY_RAVU_ZOOM: return mpx::kZoomTiles[t];
    default: return mpx::TileShape{0, 0};
  }
}

int mpx_plan_set_tile(mpx_plan* plan, int tile) {
  if (!plan) return fail(MPX_ERR_ARG, "plan is NULL");
  if (tile < 0 || tile >= tile_count(plan))
    return fail(MPX_ERR_ARG, "tile outside 0.." + std::to_string(tile_count(plan) - 1) + " for this plan");
  plan->tile = tile;
  return MPX_OK;
}

int mpx_plan_tile(const mpx_plan* plan, int* tile, int* count, int* tile_w, int* tile_h) {
  if (!plan) return fail(MPX_ERR_ARG, "plan is NULL");
  const mpx::TileShape sh = tile_shape(plan, plan->tile);
  if (tile) *tile = plan->tile;
  if (count) *count = tile_count(plan);
  if (tile_w) *tile_w = sh.w;
  if (tile_h) *tile_h = sh.h;
  return MPX_OK;
}

int mpx_plan_tune(mpx_plan* plan, const mpx_io* io, void* hip_stream, int iters, int* grid_per_cu, int* split,
                  float* best_us) {
  if (!plan || !io) return fail(MPX_ERR_ARG, "plan/io is NULL");
  if (iters < 1 || iters > 1000) return fail(MPX_ERR_ARG, "iters outside 1..1000");
  if (grid_per_cu) *grid_per_cu = 0;
  if (split) *split = 0;
  if (best_us) *best_us = 0.f;
  const bool nn = plan->v->family == MPX_FAMILY_NNEDI3 && plan->precision != MPX_PREC_FP32;
  const int ntile = tile_count(plan);
  if ((!nn && ntile == 1) || !plan->info.applies) return MPX_OK;
  // candidates: NNEDI3 MFMA plans sweep the persistent grid and the neuron split, RAVU
  // plans their family's tile shapes
  struct Shape {
    int grid, split, tile;
  };
  std::vector<Shape> cand;
  if (nn) {
    const int nsplit = plan->v->nns == 64 || plan->v->nns == 128 ? 2 : 1;
    for (int sp = 0; sp < nsplit; ++sp)
      for (int g : {0, 1, 2, 3, 4, 6, 8}) cand.push_back({g, sp, 0});
  } else {
    for (int t = 0; t < ntile; ++t) cand.push_back({0, 0, t});
  }
  mpx_host::DeviceScope dscope;
  if (int drc = dscope.enter(plan->ctx->device)) return drc;
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(MPX_ERR_HIP, "hipEventCreate failed");
  }
  const mpx::NnediWeights saved = plan->nw;
  const int saved_tile = plan->tile;
  auto set_shape = [&](const Shape& c) {
    if (nn) {
      plan->nw.grid_per_cu = c.grid;
      plan->nw.split = c.split;
    }
    plan->tile = c.tile;
  };
  float best = 0.f;
  Shape best_c{0, 0, 0};
  int rc = MPX_OK;
  for (const Shape& c : cand) {
    set_shape(c);
    rc = mpx_run_ex(plan, io, hip_stream);  // warm-up; also validates io
    if (rc != MPX_OK) break;
    if (hipEventRecord(e0, st) != hipSuccess) {
      rc = fail(MPX_ERR_HIP, "hipEventRecord failed");
      break;
    }
    for (int i = 0; i < iters && rc == MPX_OK; ++i) rc = mpx_run_ex(plan, io, hip_stream);
    if (rc != MPX_OK) break;
    float ms = 0.f;
    if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = fail(MPX_ERR_HIP, "event timing failed");
      break;
    }
    const float us = ms * 1000.f / static_cast<float>(iters);
    if (best == 0.f || us < best) {
      best = us;
      best_c = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != MPX_OK) {
    plan->nw = saved;
    plan->tile = saved_tile;
    return rc;
  }
  set_shape(best_c);
  if (grid_per_cu) *grid_per_cu = best_c.grid;
  if (split) *split = best_c.split;
  if (best_us) *best_us = best;
  return MPX_OK;
}
