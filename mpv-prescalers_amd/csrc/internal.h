// Host-side internals shared by the C ABI and the per-family host code: error state,
// the context (device, HBM budget, resident weights) and the plan.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mpx.h"
#include "nnedi3.h"
#include "ravu.h"

namespace mpx_host {

inline std::string& last_error() {
  thread_local std::string e;
  return e;
}

inline int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      return ::mpx_host::fail(MPX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Makes a device current for a scope and gives the calling thread its previous device
// back at the end (entry points must not change the caller's current device).
class DeviceScope {
 public:
  DeviceScope() = default;
  DeviceScope(const DeviceScope&) = delete;
  DeviceScope& operator=(const DeviceScope&) = delete;
  int enter(int dev) {
    if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
    if (prev_ == dev) return MPX_OK;
    HIP_TRY(hipSetDevice(dev));
    set_ = true;
    return MPX_OK;
  }
  ~DeviceScope() {
    if (set_ && prev_ >= 0) (void)hipSetDevice(prev_);
  }

 private:
  int prev_ = -1;
  bool set_ = false;
};

struct Variant {
  std::string name;
  int family, channels, radius, nns, win_h, ar, available;
};

// Weight files (weights.cpp).  load_table reads `dir`/`name`.bin after checking its header,
// its shape against the shape the kernels index with, its size and, for tables the
// reference ships, the SHA-256 of its data; MPX_ERR_IO (message set) on any mismatch,
// before anything is allocated on the device.
int load_table(const std::string& dir, const std::string& name, std::vector<float>& data,
               std::vector<uint32_t>& dims);
bool expected_dims(const std::string& name, std::vector<uint32_t>& dims);
std::vector<std::string> variant_tables(const Variant& v);
std::string sha256_of(const void* data, size_t len);

}  // namespace mpx_host

struct mpx_plan;
namespace mpx_host {
// mpx_run_ex with the intermediates in `interm` (plan->scratch_frame FP32 elements per
// frame; the caller has checked its size).
int run_io(mpx_plan* plan, const mpx_io* io, void* hip_stream, void* interm);
}  // namespace mpx_host

using mpx_host::fail;

struct DeviceBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
};

struct mpx_ctx {
  int device = 0;
  uint64_t budget = 0;
  uint64_t used = 0;
  std::string wdir;
  void* static_scratch = nullptr;  // device address of this device's array for mpx_run_static
  std::mutex mu;    // weight tables
  std::mutex amu;   // budget accounting (taken inside mu, never the other way round)
  std::map<std::string, DeviceBuf> weights;            // resident weight tables
  std::map<std::string, std::vector<uint32_t>> dims;   // their host-side shapes

  // Device memory under the budget; thread-safe (plans may be made and destroyed from
  // several threads).
  int alloc(size_t bytes, void** out) {
    std::lock_guard<std::mutex> lk(amu);
    if (budget && used + bytes > budget)
      return fail(MPX_ERR_BUDGET, "allocation of " + std::to_string(bytes) + " B exceeds the HBM budget (" +
                                      std::to_string(used) + " of " + std::to_string(budget) + " B used)");
    mpx_host::DeviceScope ds;
    int rc = ds.enter(device);
    if (rc) return rc;
    HIP_TRY(hipMalloc(out, bytes));
    used += bytes;
    return MPX_OK;
  }
  void release(void* p, size_t bytes) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(amu);
    mpx_host::DeviceScope ds;
    (void)ds.enter(device);
    (void)hipFree(p);
    used -= bytes;
  }

  // FP16 copy (RNE, as the rgba16f upload) of a LUT weight file, resident on the device.
  int lut16(const std::string& name, const __half** out, std::vector<uint32_t>* shape) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = weights.find(name);
    if (it == weights.end()) {
      std::vector<float> data;
      std::vector<uint32_t> d;
      if (int rc = mpx_host::load_table(wdir, name, data, d)) return rc;
      std::vector<_Float16> h(data.size());
      for (size_t i = 0; i < data.size(); ++i) h[i] = static_cast<_Float16>(data[i]);
      void* p = nullptr;
      int rc = alloc(h.size() * 2, &p);
      if (rc) return rc;
      HIP_TRY(hipMemcpy(p, h.data(), h.size() * 2, hipMemcpyHostToDevice));
      it = weights.emplace(name, DeviceBuf{p, h.size() * 2}).first;
      dims[name] = d;
    }
    *out = static_cast<const __half*>(it->second.ptr);
    if (shape) *shape = dims[name];
    return MPX_OK;
  }

  // Host copy of a float32 weight file (NNEDI3 tables are repacked per precision).
  int host_table(const std::string& name, std::vector<float>& data, std::vector<uint32_t>& d) {
    return mpx_host::load_table(wdir, name, data, d);
  }

  // Upload an already-packed device table once under `key`.
  int upload(const std::string& key, const void* src, size_t bytes, const void** out) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = weights.find(key);
    if (it == weights.end()) {
      void* p = nullptr;
      int rc = alloc(bytes, &p);
      if (rc) return rc;
      HIP_TRY(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice));
      it = weights.emplace(key, DeviceBuf{p, bytes}).first;
    }
    *out = it->second.ptr;
    return MPX_OK;
  }
  bool has(const std::string& key) {
    std::lock_guard<std::mutex> lk(mu);
    return weights.count(key) != 0;
  }
};

struct mpx_plan {
  mpx_ctx* ctx = nullptr;
  const mpx_host::Variant* v = nullptr;
  mpx_variant_info info{};
  int in_w = 0, in_h = 0, in_type = 0, out_type = 0, precision = 0, max_batch = 0;
  int store16 = 0;                    // MPX_STORE_FP16: FP16 intermediates
  float in_max = 1.f;
  const __half* lut = nullptr;        // main LUT (RAVU family)
  const __half* lut_ar = nullptr;     // zoom-ar LUT
  mpx::NnediWeights nw{};             // NNEDI3 packed weights (and their launch shape)
  mpx::ZoomPlan zoom{};               // zoom phase tables
  int tile = 0;                       // RAVU families: index into the family's tile shapes
  void* scratch = nullptr;            // intermediates, max_batch frames
  size_t scratch_bytes = 0;
  int64_t scratch_frame = 0;          // elements per frame
  // the intermediates serve one stream at a time: the last stream that ran the plan and
  // an event recorded after its work (plans with intermediates only)
  std::mutex run_mu;
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;
  bool ran = false;
};

