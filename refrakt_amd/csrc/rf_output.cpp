// Output formats: 8-bit RGBA conversion, a PNG writer and an OpenEXR writer.
//
// The reference saves a screenshot by reading its tonemapped RGBA32F render
// target back as GL_UNSIGNED_BYTE (src/buffer_objects.hpp:117, a unorm
// conversion: clamp to [0, 1], scale by 255, round to nearest) and handing the
// rows, first texture row first, to stb_image_write's PNG encoder
// (src/main.cpp:590-593).  rf_write_png writes the same pixels as an RGBA8
// PNG (colour type 6, no interlace) whose IDAT is one zlib stream of rows with
// filter byte 0.
//
// rf_write_exr keeps a float RGBA image (the density-estimated accumulator or
// the tonemapped image) without the 8-bit quantisation: a single-part
// scanline OpenEXR file, uncompressed, channels A, B, G, R (the format keeps
// them in name order) as HALF or FLOAT, one scanline per block.
#include <stdint.h>
#include <zlib.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/refrakt_amd.h"

namespace {

void put_u32(std::vector<uint8_t>& v, uint32_t x) {
  v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16));
  v.push_back((uint8_t)(x >> 8)); v.push_back((uint8_t)x);
}

// One chunk: length, type, data, CRC-32 of type + data (PNG spec, section 5.3).
void chunk(std::vector<uint8_t>& out, const char type[4], const uint8_t* data, size_t n) {
  put_u32(out, (uint32_t)n);
  const size_t at = out.size();
  out.insert(out.end(), type, type + 4);
  if (n) out.insert(out.end(), data, data + n);
  const uLong crc = crc32(0L, out.data() + at, (uInt)(4 + n));
  put_u32(out, (uint32_t)crc);
}

// Little-endian fields of the EXR header and scanline blocks.
void le32(std::vector<uint8_t>& v, uint32_t x) {
  for (int k = 0; k < 4; ++k) v.push_back((uint8_t)(x >> (8 * k)));
}
void le64(std::vector<uint8_t>& v, uint64_t x) {
  for (int k = 0; k < 8; ++k) v.push_back((uint8_t)(x >> (8 * k)));
}
void lef(std::vector<uint8_t>& v, float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  le32(v, x);
}
void cstr(std::vector<uint8_t>& v, const char* s) { v.insert(v.end(), s, s + strlen(s) + 1); }
// One header attribute: name, type name, value size, value.
void attr(std::vector<uint8_t>& out, const char* name, const char* type, const std::vector<uint8_t>& val) {
  cstr(out, name);
  cstr(out, type);
  le32(out, (uint32_t)val.size());
  out.insert(out.end(), val.begin(), val.end());
}

}  // namespace

// IEEE 754 binary32 -> binary16, round to nearest even: overflow to
// infinity, subnormal halves below 2^-14, NaN kept quiet with its sign.
extern "C" uint16_t rf_half_from_float(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // >= 65520 rounds past 65504
  if (ax < 0x38800000u) {                                     // below 2^-14: a subnormal half or zero
    if (ax < 0x33000000u) return (uint16_t)sign;              // below 2^-25
    const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;                          // 14 .. 24
    uint32_t r = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
    if (rem > halfway || (rem == halfway && (r & 1u))) ++r;   // may carry into the smallest normal
    return (uint16_t)(sign | r);
  }
  uint32_t r = (ax - 0x38000000u) >> 13;                      // rebias 127 -> 15, 23 -> 10 mantissa bits
  const uint32_t rem = ax & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return (uint16_t)(sign | r);
}

extern "C" int rf_write_exr(const char* path, const float* rgba, uint32_t width, uint32_t height, int32_t half) {
  if (!path || !rgba || width == 0 || height == 0 || width > 0x7fffffffu || height > 0x7fffffffu) return RF_ERR_ARG;
  const uint32_t pt = half ? 1u : 2u;  // HALF / FLOAT
  const uint32_t bpc = half ? 2u : 4u;
  if ((uint64_t)width * 4u * bpc > 0x7fffffffu) return RF_ERR_ARG;
  std::vector<uint8_t> out = {0x76, 0x2f, 0x31, 0x01, 2, 0, 0, 0};  // magic, version 2, single-part scanline
  std::vector<uint8_t> v;
  for (const char* ch : {"A", "B", "G", "R"}) {
    cstr(v, ch);
    le32(v, pt);
    v.push_back(0);  // pLinear
    v.push_back(0); v.push_back(0); v.push_back(0);
    le32(v, 1u);     // x sampling
    le32(v, 1u);     // y sampling
  }
  v.push_back(0);
  attr(out, "channels", "chlist", v);
  attr(out, "compression", "compression", {0});  // NO_COMPRESSION
  v.clear();
  le32(v, 0u); le32(v, 0u); le32(v, width - 1u); le32(v, height - 1u);
  attr(out, "dataWindow", "box2i", v);
  attr(out, "displayWindow", "box2i", v);
  attr(out, "lineOrder", "lineOrder", {0});      // INCREASING_Y
  v.clear(); lef(v, 1.0f);
  attr(out, "pixelAspectRatio", "float", v);
  v.clear(); lef(v, 0.0f); lef(v, 0.0f);
  attr(out, "screenWindowCenter", "v2f", v);
  v.clear(); lef(v, 1.0f);
  attr(out, "screenWindowWidth", "float", v);
  out.push_back(0);  // end of header
  const uint64_t line = (uint64_t)width * 4u * bpc;
  const uint64_t table_at = out.size(), first = table_at + 8ull * height;
  for (uint32_t y = 0; y < height; ++y) le64(out, first + (uint64_t)y * (8u + line));
  FILE* f = fopen(path, "wb");
  if (!f) return RF_ERR_IO;
  bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
  std::vector<uint8_t> blk;
  blk.reserve(8 + line);
  static const int order[4] = {3, 2, 1, 0};  // A, B, G, R from RGBA
  for (uint32_t y = 0; y < height && ok; ++y) {
    blk.clear();
    le32(blk, y);
    le32(blk, (uint32_t)line);
    const float* row = rgba + (size_t)y * width * 4;
    for (int c : order) {
      for (uint32_t x = 0; x < width; ++x) {
        const float val = row[(size_t)x * 4 + c];
        if (half) {
          const uint16_t h = rf_half_from_float(val);
          blk.push_back((uint8_t)h);
          blk.push_back((uint8_t)(h >> 8));
        } else {
          lef(blk, val);
        }
      }
    }
    ok = fwrite(blk.data(), 1, blk.size(), f) == blk.size();
  }
  const int c = fclose(f);
  return (ok && c == 0) ? RF_OK : RF_ERR_IO;
}

extern "C" int rf_rgba8_from_f32(const float* in, uint32_t width, uint32_t height, uint8_t* out) {
  if (!in || !out || width == 0 || height == 0) return RF_ERR_ARG;
  const size_t n = (size_t)width * height * 4;
  for (size_t i = 0; i < n; ++i) {
    float v = in[i];
    v = v > 0.0f ? (v < 1.0f ? v : 1.0f) : 0.0f;   // NaN -> 0
    out[i] = (uint8_t)(v * 255.0f + 0.5f);
  }
  return RF_OK;
}

extern "C" int rf_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height) {
  if (!path || !rgba8 || width == 0 || height == 0 || width > 0x7fffffffu || height > 0x7fffffffu)
    return RF_ERR_ARG;
  const size_t row = (size_t)width * 4;
  std::vector<uint8_t> raw((row + 1) * height);
  for (uint32_t y = 0; y < height; ++y) {
    raw[y * (row + 1)] = 0;  // filter type None
    memcpy(&raw[y * (row + 1) + 1], rgba8 + y * row, row);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<uint8_t> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return RF_ERR_IO;
  std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
  std::vector<uint8_t> ihdr;
  put_u32(ihdr, width);
  put_u32(ihdr, height);
  const uint8_t rest[5] = {8, 6, 0, 0, 0};  // bit depth 8, RGBA, deflate, adaptive filtering, no interlace
  ihdr.insert(ihdr.end(), rest, rest + 5);
  chunk(png, "IHDR", ihdr.data(), ihdr.size());
  chunk(png, "IDAT", z.data(), zlen);
  chunk(png, "IEND", nullptr, 0);
  FILE* f = fopen(path, "wb");
  if (!f) return RF_ERR_IO;
  const size_t w = fwrite(png.data(), 1, png.size(), f);
  const int c = fclose(f);
  return (w == png.size() && c == 0) ? RF_OK : RF_ERR_IO;
}
