// Weight files: the MPXW reader with shape and integrity checks (host only).
//
// Every table a plan uploads is checked before any device memory is touched:
//   * the header (magic, version, float32, rank) and the dims against the shape the
//     kernels index with, derived from the table's name (e.g. ravu-zoom r3: 2592 rows of
//     9 phase texels x 5 tap groups x RGBA, the contract of the reference's
//     `//!TEXTURE ... SIZE 45 2592 ... FORMAT rgba16f`, ravu-zoom-r3-rgb.hook:214-218);
//   * the file size against header + data, so a truncated or padded file fails;
//   * for the tables the reference ships, the SHA-256 of the float32 data against the
//     value recorded at extraction (weights_manifest.h).  A table the reference does not
//     ship (the supplied ravu-zoom-ar-r3 table, .MISSING_LARGE_BLOBS:1-7) has no recorded
//     hash and is checked by shape and size only.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "weights_manifest.h"

namespace mpx_host {

namespace {

// ---------------------------------------------------------------------------- SHA-256
// FIPS 180-4.  The round constants are the first 32 fraction bits of the cube roots of the
// first 64 primes and the initial state those of the square roots of the first 8: both
// computed here in exact integer arithmetic (integer roots of p * 2^96 and p * 2^64).

using u128 = unsigned __int128;

uint64_t iroot(u128 v, int k) {  // floor(v^(1/k)) for k in {2, 3}, v < 2^110
  uint64_t lo = 0, hi = uint64_t{1} << 56;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    u128 p = mid;
    bool over = false;
    for (int i = 1; i < k && !over; ++i) {
      if (p > (~u128{0}) / mid) over = true;
      else p *= mid;
    }
    if (!over && p <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct ShaConst {
  uint32_t k[64];
  uint32_t h0[8];
  ShaConst() {
    int n = 0;
    for (uint32_t p = 2; n < 64; ++p) {
      bool prime = true;
      for (uint32_t d = 2; d * d <= p; ++d)
        if (p % d == 0) { prime = false; break; }
      if (!prime) continue;
      k[n] = static_cast<uint32_t>(iroot(static_cast<u128>(p) << 96, 3));
      if (n < 8) h0[n] = static_cast<uint32_t>(iroot(static_cast<u128>(p) << 64, 2));
      ++n;
    }
  }
};

inline uint32_t rotr(uint32_t x, int s) { return (x >> s) | (x << (32 - s)); }

std::string sha256_hex(const unsigned char* data, size_t len) {
  static const ShaConst C;
  uint32_t h[8];
  std::memcpy(h, C.h0, sizeof(h));
  const uint64_t bits = static_cast<uint64_t>(len) * 8;
  const size_t total = ((len + 8) / 64 + 1) * 64;  // message + 0x80 + zeros + 64-bit length
  unsigned char blk[64];
  for (size_t off = 0; off < total; off += 64) {
    for (int i = 0; i < 64; ++i) {
      const size_t j = off + i;
      unsigned char c;
      if (j < len) c = data[j];
      else if (j == len) c = 0x80;
      else if (j >= total - 8) c = static_cast<unsigned char>(bits >> (8 * (total - 1 - j)));
      else c = 0;
      blk[i] = c;
    }
    uint32_t w[64];
    for (int t = 0; t < 16; ++t)
      w[t] = (uint32_t{blk[4 * t]} << 24) | (uint32_t{blk[4 * t + 1]} << 16) | (uint32_t{blk[4 * t + 2]} << 8) |
             uint32_t{blk[4 * t + 3]};
    for (int t = 16; t < 64; ++t) {
      const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
      const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
      w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int t = 0; t < 64; ++t) {
      const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + C.k[t] + w[t];
      const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
      const uint32_t maj = (a & b) ^ (a & c) ^ (b & c);
      const uint32_t t2 = S0 + maj;
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }
  static const char* hex = "0123456789abcdef";
  std::string out(64, '0');
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j) out[8 * i + j] = hex[(h[i] >> (28 - 4 * j)) & 15];
  return out;
}

bool starts_with(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }

// Radius r from a "..._lut<r>" / "..._lut<r>_ar" name.
int lut_radius(const std::string& name) {
  const size_t p = name.find("_lut");
  return p == std::string::npos ? 0 : std::atoi(name.c_str() + p + 4);
}

}  // namespace

std::string sha256_of(const void* data, size_t len) {
  return sha256_hex(static_cast<const unsigned char*>(data), len);
}

// The shape the kernels index a table with, from its name; false for an unknown name.
bool expected_dims(const std::string& name, std::vector<uint32_t>& dims) {
  dims.clear();
  if (starts_with(name, "nnedi3_nns")) {
    const int nns = std::atoi(name.c_str() + 10);
    const size_t p = name.find("_win8x");
    if (p == std::string::npos) return false;
    const int wh = std::atoi(name.c_str() + p + 6);
    if ((nns != 16 && nns != 32 && nns != 64 && nns != 128 && nns != 256) || (wh != 4 && wh != 6)) return false;
    dims = {static_cast<uint32_t>(nns), static_cast<uint32_t>(2 * 8 * wh + 2)};
    return true;
  }
  const int r = lut_radius(name);
  if (r < 2 || r > 4) return false;
  if (starts_with(name, "ravu_lite_lut")) {  // 2r-1 window, ceil(n^2/2) texels, 24*4*3 rows
    const int n = 2 * r - 1;
    dims = {288, static_cast<uint32_t>((n * n + 1) / 2), 4};
  } else if (starts_with(name, "ravu_3x_lut")) {  // two texel sets, 24*3*3 rows
    const int n = 2 * r - 1;
    dims = {216, static_cast<uint32_t>(2 * ((n * n + 1) / 2)), 4};
  } else if (starts_with(name, "ravu_zoom_lut")) {  // 9 phases x tap groups, 288 buckets x 9 phases
    if (r > 3) return false;
    const int n = 2 * r;
    dims = {2592, static_cast<uint32_t>(9 * ((n * n / 2 + 3) / 4)), 4};
  } else if (starts_with(name, "ravu_lut")) {  // one weight per symmetric pair, 4 per texel, 24*9*3 rows
    const int n = 2 * r;
    dims = {648, static_cast<uint32_t>((n * n / 2 + 3) / 4), 4};
  } else {
    return false;
  }
  return true;
}

int load_table(const std::string& dir, const std::string& name, std::vector<float>& data,
               std::vector<uint32_t>& dims) {
  const std::string path = dir + "/" + name + ".bin";
  std::vector<uint32_t> want;
  if (!expected_dims(name, want)) return fail(MPX_ERR_IO, "no weight table is named " + name);
  const ShippedTable* shipped = nullptr;
  for (const ShippedTable& t : kShipped)
    if (name == t.name) shipped = &t;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(MPX_ERR_IO, "cannot open weight file " + path);
  auto bad = [&](const std::string& why) {
    std::fclose(f);
    return fail(MPX_ERR_IO, "weight file " + path + ": " + why);
  };
  char magic[4];
  uint32_t hdr[3];
  if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "MPXW", 4) != 0) return bad("not an MPXW file");
  if (std::fread(hdr, 4, 3, f) != 3) return bad("truncated header");
  if (hdr[0] != 1 || hdr[1] != 0) return bad("unsupported version or element type");
  if (hdr[2] != want.size()) return bad("rank " + std::to_string(hdr[2]) + ", expected " + std::to_string(want.size()));
  dims.resize(hdr[2]);
  if (std::fread(dims.data(), 4, hdr[2], f) != hdr[2]) return bad("truncated header");
  std::string got, exp;
  for (size_t i = 0; i < want.size(); ++i) {
    got += (i ? "x" : "") + std::to_string(dims[i]);
    exp += (i ? "x" : "") + std::to_string(want[i]);
  }
  if (dims != want) return bad("shape " + got + ", expected " + exp);
  size_t n = 1;
  for (uint32_t d : dims) n *= d;  // bounded by the expected shape (at most 2592 x 45 x 4)
  const long data_off = 16 + 4 * static_cast<long>(dims.size());
  if (std::fseek(f, 0, SEEK_END) != 0) return bad("cannot seek");
  const long size = std::ftell(f);
  if (size != data_off + static_cast<long>(n * 4))
    return bad("size " + std::to_string(size) + " B, expected " + std::to_string(data_off + n * 4) + " B");
  if (std::fseek(f, data_off, SEEK_SET) != 0) return bad("cannot seek");
  data.resize(n);
  if (std::fread(data.data(), 4, n, f) != n) return bad("truncated data");
  std::fclose(f);
  if (shipped) {
    const std::string h = sha256_of(data.data(), n * 4);
    if (h != shipped->sha256)
      return fail(MPX_ERR_IO, "weight file " + path + ": SHA-256 of the data is " + h + ", the shipped table's is " +
                                  shipped->sha256 + " (damaged or replaced file)");
  }
  return MPX_OK;
}

std::vector<std::string> variant_tables(const Variant& v) {
  const std::string r = std::to_string(v.radius);
  switch (v.family) {
    case MPX_FAMILY_NNEDI3: return {"nnedi3_nns" + std::to_string(v.nns) + "_win8x" + std::to_string(v.win_h)};
    case MPX_FAMILY_RAVU: return {"ravu_lut" + r};
    case MPX_FAMILY_RAVU_LITE: return {"ravu_lite_lut" + r};
    case MPX_FAMILY_RAVU_3X: return {"ravu_3x_lut" + r};
    case MPX_FAMILY_RAVU_ZOOM:
      if (v.ar) return {"ravu_zoom_lut" + r, "ravu_zoom_lut" + r + "_ar"};
      return {"ravu_zoom_lut" + r};
  }
  return {};
}

}  // namespace mpx_host
