// Output formats: 8-bit RGBA conversion and a PNG writer.
//
// The reference saves a screenshot by reading its tonemapped RGBA32F render
// target back as GL_UNSIGNED_BYTE (src/buffer_objects.hpp:117, a unorm
// conversion: clamp to [0, 1], scale by 255, round to nearest) and handing the
// rows, first texture row first, to stb_image_write's PNG encoder
// (src/main.cpp:590-593).  rf_write_png writes the same pixels as an RGBA8
// PNG (colour type 6, no interlace) whose IDAT is one zlib stream of rows with
// filter byte 0.
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

}  // namespace

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
