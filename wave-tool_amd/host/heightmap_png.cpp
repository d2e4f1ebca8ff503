// PNG decoding for the heightmap asset (see heightmap_png.hpp). Written from the PNG
// specification (ISO/IEC 15948): chunk stream, zlib-compressed IDAT, per-row filters
// None / Sub / Up / Average / Paeth; zlib's inflate does the decompression.
#include "heightmap_png.hpp"

#include <zlib.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>

namespace wavetool {
namespace {

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
  const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return a;
  return pb <= pc ? b : c;
}

bool fail(std::string* e, const std::string& m) {
  if (e) *e = m;
  return false;
}

}  // namespace

bool load_heightmap_png(const std::string& path, std::vector<uint8_t>* texels, uint32_t* width, uint32_t* height,
                        std::string* error) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(error, "cannot open " + path);
  const std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (data.size() < 8 || std::memcmp(data.data(), sig, 8) != 0) return fail(error, "not a PNG file");
  uint32_t w = 0, h = 0;
  int depth = 0, ctype = -1;
  std::vector<uint8_t> idat, palette;
  bool ended = false;
  for (size_t p = 8; p + 12 <= data.size();) {
    const uint32_t len = be32(&data[p]);
    if (len > data.size() - p - 12) return fail(error, "truncated chunk");
    const uint8_t* type = &data[p + 4];
    const uint8_t* body = &data[p + 8];
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len != 13) return fail(error, "bad IHDR");
      w = be32(body);
      h = be32(body + 4);
      depth = body[8];
      ctype = body[9];
      if (body[10] != 0 || body[11] != 0) return fail(error, "unknown compression or filter method");
      if (body[12] != 0) return fail(error, "interlaced PNGs are not supported");
    } else if (!std::memcmp(type, "PLTE", 4)) {
      palette.assign(body, body + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), body, body + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      ended = true;
      break;
    }
    p += 12 + (size_t)len;
  }
  if (!ended || w == 0 || h == 0 || w > 16384 || h > 16384) return fail(error, "missing IHDR / IEND or bad size");
  int channels = 0;
  switch (ctype) {
    case 0: channels = 1; break;  // grey
    case 2: channels = 3; break;  // RGB
    case 3: channels = 1; break;  // palette index
    case 4: channels = 2; break;  // grey + alpha
    case 6: channels = 4; break;  // RGBA
    default: return fail(error, "unknown colour type");
  }
  if (!(depth == 8 || (depth == 16 && ctype != 3))) return fail(error, "only 8-bit (or 16-bit non-palette) PNGs");
  if (ctype == 3 && palette.empty()) return fail(error, "palette image without PLTE");
  const size_t bpp = (size_t)channels * (depth / 8);  // bytes per pixel
  const size_t stride = (size_t)w * bpp;
  std::vector<uint8_t> raw((stride + 1) * h);
  uLongf out_len = (uLongf)raw.size();
  if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
    return fail(error, "corrupt image data");
  // undo the per-row filters in place; prev points at the previous reconstructed row
  std::vector<uint8_t> img(stride * h), zero(stride, 0);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t ft = raw[y * (stride + 1)];
    const uint8_t* src = &raw[y * (stride + 1) + 1];
    uint8_t* dst = &img[y * stride];
    const uint8_t* prev = y ? &img[(y - 1) * stride] : zero.data();
    for (size_t x = 0; x < stride; ++x) {
      const int a = x >= bpp ? dst[x - bpp] : 0, b = prev[x], c = x >= bpp ? prev[x - bpp] : 0;
      int v = src[x];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += paeth(a, b, c); break;
        default: return fail(error, "unknown row filter");
      }
      dst[x] = (uint8_t)v;
    }
  }
  // red channel (grey for grey images, the palette entry's red for palette images), the
  // high byte of 16-bit samples, rows flipped so row 0 is the bottom
  texels->resize((size_t)w * h);
  for (uint32_t y = 0; y < h; ++y) {
    const uint8_t* row = &img[(size_t)(h - 1 - y) * stride];
    uint8_t* o = &(*texels)[(size_t)y * w];
    for (uint32_t x = 0; x < w; ++x) {
      const uint8_t s = row[x * bpp];  // first sample, big-endian high byte at 16 bits
      if (ctype == 3) {
        if ((size_t)s * 3 + 2 >= palette.size()) return fail(error, "palette index out of range");
        o[x] = palette[(size_t)s * 3];
      } else {
        o[x] = s;
      }
    }
  }
  *width = w;
  *height = h;
  return true;
}

}  // namespace wavetool
