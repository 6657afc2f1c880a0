// NNEDI3 weight packing per precision (host side), uploaded once per context.
#include <algorithm>
#include <cmath>

#include "internal.h"
#include "nnedi3.h"

namespace mpx {

namespace {

constexpr double kLog2e = 1.4426950408889634;

// Round-to-nearest-even float -> bfloat16 bits.
uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
float bf16_float(uint16_t b) {
  uint32_t u = static_cast<uint32_t>(b) << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Round-to-nearest-even double -> OCP FP8 e4m3 (bias 7, largest finite 448, saturating).
uint8_t e4m3_bits(double v) {
  const uint8_t sign = v < 0 ? 0x80 : 0;
  const double a = std::fabs(v);
  if (!(a < 448.0)) return sign | 0x7e;
  int e;
  std::frexp(a, &e);                       // a = f * 2^e, f in [0.5, 1)
  int E = e - 1 + 7;                       // biased exponent of a normal encoding
  if (E < 1) {                             // subnormal: multiples of 2^-9
    const double mm = std::nearbyint(a * 512.0);
    return sign | static_cast<uint8_t>(mm >= 8 ? 0x08 : static_cast<int>(mm));
  }
  double mm = std::nearbyint((a / std::ldexp(1.0, e - 1) - 1.0) * 8.0);
  if (mm >= 8) {
    mm = 0;
    ++E;
  }
  if (E > 15 || (E == 15 && mm > 6)) return sign | 0x7e;
  return sign | static_cast<uint8_t>((E << 3) | static_cast<int>(mm));
}

}  // namespace

int nnedi3_prepare(mpx_ctx* ctx, int nns, int wh, int precision, NnediWeights* out) {
  const std::string name = "nnedi3_nns" + std::to_string(nns) + "_win8x" + std::to_string(wh);
  std::vector<float> t;
  std::vector<uint32_t> dims;
  int rc = ctx->host_table(name, t, dims);
  if (rc) return rc;
  const int K = 8 * wh;
  if (dims.size() != 2 || dims[0] != static_cast<uint32_t>(nns) || dims[1] != static_cast<uint32_t>(2 * K + 2))
    return fail(MPX_ERR_IO, "unexpected shape of " + name);
  out->nns = nns;
  out->wh = wh;
  out->precision = precision;
  // The FP16 split form carries power-of-two operand scales (window x256, weights x256)
  // so that the low parts stay normal FP16 numbers (MFMA flushes FP16 subnormal inputs);
  // BF16 has FP32's exponent range, the single-product FP16 form has no low parts.
  const bool f16 = precision == MPX_PREC_FP16 || precision == MPX_PREC_FP16X3;
  out->xscale = precision == MPX_PREC_FP16X3 ? 256.f : 1.f;
  out->wscale = precision == MPX_PREC_FP16X3 ? 256.f : 1.f;
  const double ws = out->wscale, accs = static_cast<double>(out->xscale) * out->wscale;
  auto W1 = [&](int n, int k) { return static_cast<double>(t[n * (2 * K + 2) + k]) * kLog2e; };
  auto W2 = [&](int n, int k) { return static_cast<double>(t[n * (2 * K + 2) + K + k]); };
  auto B1 = [&](int n) { return static_cast<double>(t[n * (2 * K + 2) + 2 * K]) * kLog2e; };
  auto B2 = [&](int n) { return static_cast<double>(t[n * (2 * K + 2) + 2 * K + 1]); };

  std::vector<float> bias(2 * nns);
  const double bscale = precision == MPX_PREC_FP32 ? 1.0 : accs;
  for (int n = 0; n < nns; ++n) {
    bias[2 * n] = static_cast<float>(B1(n) * bscale);
    bias[2 * n + 1] = static_cast<float>(B2(n) * bscale);
  }
  const void* p = nullptr;
  if (precision == MPX_PREC_FP32) {
    std::vector<float> w(static_cast<size_t>(nns) * (2 * K + 2));
    for (int n = 0; n < nns; ++n) {
      float* r = &w[static_cast<size_t>(n) * (2 * K + 2)];
      for (int k = 0; k < K; ++k) {
        r[k] = static_cast<float>(W1(n, k));
        r[K + k] = static_cast<float>(W2(n, k));
      }
      r[2 * K] = bias[2 * n];
      r[2 * K + 1] = bias[2 * n + 1];
    }
    rc = ctx->upload(name + "/fp32", w.data(), w.size() * 4, &p);
    if (rc) return rc;
    out->w = p;
    return MPX_OK;
  }
  rc = ctx->upload(name + "/bias_x" + std::to_string(static_cast<long long>(accs)), bias.data(), bias.size() * 4, &p);
  if (rc) return rc;
  out->bias = static_cast<const float*>(p);
  const size_t nel = static_cast<size_t>(nns) * 2 * K;
  if (f16 || precision == MPX_PREC_BF16) {
    // [n][matrix][k] high part and residual
    std::vector<uint16_t> hi(nel), lo(nel);
    for (int n = 0; n < nns; ++n)
      for (int mtx = 0; mtx < 2; ++mtx)
        for (int k = 0; k < K; ++k) {
          const double v = (mtx == 0 ? W1(n, k) : W2(n, k)) * ws;
          const size_t i = (static_cast<size_t>(n) * 2 + mtx) * K + k;
          if (f16) {
            const _Float16 h = static_cast<_Float16>(v);
            const _Float16 l = static_cast<_Float16>(v - static_cast<double>(h));
            std::memcpy(&hi[i], &h, 2);
            std::memcpy(&lo[i], &l, 2);
          } else {
            const uint16_t h = bf16_bits(static_cast<float>(v));
            hi[i] = h;
            lo[i] = bf16_bits(static_cast<float>(v - static_cast<double>(bf16_float(h))));
          }
        }
    // the key carries the weight scale: FP16 and FP16X3 pack the same table differently
    const std::string tag = (f16 ? "/fp16_x" : "/bf16_x") + std::to_string(static_cast<long long>(ws));
    rc = ctx->upload(name + tag + "_hi", hi.data(), nel * 2, &p);
    if (rc) return rc;
    out->w = p;
    rc = ctx->upload(name + tag + "_lo", lo.data(), nel * 2, &p);
    if (rc) return rc;
    out->w_lo = p;
    return MPX_OK;
  }
  if (precision == MPX_PREC_FP8) {
    std::vector<uint8_t> q(static_cast<size_t>(nns) * 2 * 128, 0);
    std::vector<int> ex(static_cast<size_t>(nns) * 2);
    for (int n = 0; n < nns; ++n)
      for (int mtx = 0; mtx < 2; ++mtx) {
        double mx = 0;
        for (int k = 0; k < K; ++k) mx = std::max(mx, std::fabs(mtx == 0 ? W1(n, k) : W2(n, k)));
        int e = 0;
        if (mx > 0) {
          std::frexp(mx / 448.0, &e);  // mx / 448 = f * 2^e, f in [0.5, 1): mx * 2^-e in [224, 448)
          if (std::ldexp(mx, -e) > 448.0) ++e;
        }
        ex[n * 2 + mtx] = 127 + e;
        uint8_t* r = &q[(static_cast<size_t>(n) * 2 + mtx) * 128];
        for (int k = 0; k < K; ++k) r[k] = r[K + k] = e4m3_bits(std::ldexp(mtx == 0 ? W1(n, k) : W2(n, k), -e));
      }
    rc = ctx->upload(name + "/fp8", q.data(), q.size(), &p);
    if (rc) return rc;
    out->w = p;
    rc = ctx->upload(name + "/fp8_exp", ex.data(), ex.size() * 4, &p);
    if (rc) return rc;
    out->wexp = static_cast<const int*>(p);
    return MPX_OK;
  }
  return fail(MPX_ERR_UNSUPPORTED, "NNEDI3 precision " + std::to_string(precision) + " not built");
}

}  // namespace mpx
