#include "image.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

namespace rad {
namespace host {

namespace {

std::vector<uint8_t> ReadFile(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open image: " + path);
  return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

[[noreturn]] void Bad(const std::string& path, const std::string& why) {
  throw std::runtime_error("image " + path + ": " + why);
}

uint32_t BE32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }

uint8_t LumaStb(uint8_t r, uint8_t g, uint8_t b) { return uint8_t((r * 77 + g * 150 + 29 * b) >> 8); }

// src has `comp` 8-bit samples per pixel (1 grey, 2 grey + alpha, 3 RGB, 4 RGBA).
Image8 ToChannels(uint32_t w, uint32_t h, int comp, const std::vector<uint8_t>& src, int req) {
  Image8 out;
  out.width = w;
  out.height = h;
  out.channels = uint32_t(req);
  out.px.resize(size_t(w) * h * req);
  const size_t n = size_t(w) * h;
  for (size_t i = 0; i < n; i++) {
    const uint8_t* s = &src[i * comp];
    uint8_t* d = &out.px[i * req];
    if (req == 3) {
      if (comp <= 2) {
        d[0] = d[1] = d[2] = s[0];
      } else {
        d[0] = s[0], d[1] = s[1], d[2] = s[2];
      }
    } else {
      d[0] = comp <= 2 ? s[0] : LumaStb(s[0], s[1], s[2]);
    }
  }
  return out;
}

uint8_t Paeth(int a, int b, int c) {
  int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
  if (pa <= pb && pa <= pc) return uint8_t(a);
  return uint8_t(pb <= pc ? b : c);
}

Image8 ReadPng(const std::string& path, const std::vector<uint8_t>& f, int req) {
  static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) Bad(path, "not a PNG file");
  size_t pos = 8;
  uint32_t W = 0, H = 0;
  int depth = 0, ctype = -1, interlace = 0;
  std::vector<uint8_t> idat, plte;
  bool end = false;
  while (!end) {
    if (pos + 12 > f.size()) Bad(path, "truncated chunk");
    const uint32_t len = BE32(&f[pos]);
    const uint8_t* type = &f[pos + 4];
    if (pos + 12 + size_t(len) > f.size()) Bad(path, "truncated chunk");
    const uint8_t* data = &f[pos + 8];
    const uint32_t crc = BE32(&f[pos + 8 + len]);
    if (uint32_t(crc32(crc32(0L, Z_NULL, 0), type, 4 + len)) != crc) Bad(path, "chunk CRC mismatch");
    if (!std::memcmp(type, "IHDR", 4)) {
      if (len != 13) Bad(path, "bad IHDR");
      W = BE32(data);
      H = BE32(data + 4);
      depth = data[8];
      ctype = data[9];
      if (data[10] != 0 || data[11] != 0) Bad(path, "unknown compression or filter method");
      interlace = data[12];
    } else if (!std::memcmp(type, "PLTE", 4)) {
      plte.assign(data, data + len);
    } else if (!std::memcmp(type, "IDAT", 4)) {
      idat.insert(idat.end(), data, data + len);
    } else if (!std::memcmp(type, "IEND", 4)) {
      end = true;
    } else if (!(type[0] & 0x20)) {
      Bad(path, "unknown critical chunk");
    }
    pos += 12 + size_t(len);
  }
  if (W == 0 || H == 0 || W > (1u << 15) || H > (1u << 15)) Bad(path, "unsupported size");
  if (interlace) Bad(path, "interlaced PNG is not supported");
  int samples;
  switch (ctype) {
    case 0: samples = 1; break;
    case 2: samples = 3; break;
    case 3: samples = 1; break;
    case 4: samples = 2; break;
    case 6: samples = 4; break;
    default: Bad(path, "unknown colour type");
  }
  const bool okDepth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                       (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                       ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
  if (!okDepth) Bad(path, "unsupported bit depth");
  if (ctype == 3 && (plte.empty() || plte.size() % 3)) Bad(path, "missing palette");
  const size_t bitsPerPixel = size_t(samples) * depth;
  const size_t rowBytes = (size_t(W) * bitsPerPixel + 7) / 8;
  const size_t bpp = std::max<size_t>(1, bitsPerPixel / 8);
  std::vector<uint8_t> raw(size_t(H) * (rowBytes + 1));
  uLongf rawLen = uLongf(raw.size());
  if (uncompress(raw.data(), &rawLen, idat.data(), uLong(idat.size())) != Z_OK || rawLen != raw.size())
    Bad(path, "corrupt image data");
  std::vector<uint8_t> img(size_t(H) * rowBytes);
  for (uint32_t y = 0; y < H; y++) {  // undo the per-row filters
    const uint8_t ft = raw[size_t(y) * (rowBytes + 1)];
    const uint8_t* in = &raw[size_t(y) * (rowBytes + 1) + 1];
    uint8_t* cur = &img[size_t(y) * rowBytes];
    const uint8_t* prev = y ? &img[size_t(y - 1) * rowBytes] : nullptr;
    for (size_t i = 0; i < rowBytes; i++) {
      const int a = i >= bpp ? cur[i - bpp] : 0, b = prev ? prev[i] : 0, c = (prev && i >= bpp) ? prev[i - bpp] : 0;
      int v = in[i];
      switch (ft) {
        case 0: break;
        case 1: v += a; break;
        case 2: v += b; break;
        case 3: v += (a + b) >> 1; break;
        case 4: v += Paeth(a, b, c); break;
        default: Bad(path, "unknown row filter");
      }
      cur[i] = uint8_t(v);
    }
  }
  // expand to 8-bit samples (palette -> RGB); 16-bit keeps the high byte and
  // sub-byte grey is rescaled to 0..255, as stb_image does
  const int comp = ctype == 3 ? 3 : samples;
  std::vector<uint8_t> px(size_t(W) * H * comp);
  static const uint8_t scale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
  for (uint32_t y = 0; y < H; y++) {
    const uint8_t* row = &img[size_t(y) * rowBytes];
    for (uint32_t x = 0; x < W; x++) {
      uint8_t* d = &px[(size_t(y) * W + x) * comp];
      if (depth < 8) {
        const size_t bit = size_t(x) * depth;
        const int v = (row[bit / 8] >> (8 - depth - bit % 8)) & ((1 << depth) - 1);
        if (ctype == 3) {
          if (size_t(v) * 3 + 2 >= plte.size()) Bad(path, "palette index out of range");
          d[0] = plte[3 * v], d[1] = plte[3 * v + 1], d[2] = plte[3 * v + 2];
        } else {
          d[0] = uint8_t(v * scale[depth]);
        }
      } else if (ctype == 3) {
        const int v = row[x];
        if (size_t(v) * 3 + 2 >= plte.size()) Bad(path, "palette index out of range");
        d[0] = plte[3 * v], d[1] = plte[3 * v + 1], d[2] = plte[3 * v + 2];
      } else {
        for (int c = 0; c < samples; c++)
          d[c] = depth == 16 ? row[(size_t(x) * samples + c) * 2] : row[size_t(x) * samples + c];
      }
    }
  }
  return ToChannels(W, H, comp, px, req);
}

Image8 ReadPnm(const std::string& path, const std::vector<uint8_t>& f, int req) {
  size_t pos = 2;
  auto token = [&]() {
    for (;;) {
      while (pos < f.size() && std::isspace(f[pos])) pos++;
      if (pos < f.size() && f[pos] == '#') {
        while (pos < f.size() && f[pos] != '\n') pos++;
        continue;
      }
      break;
    }
    long v = 0;
    bool any = false;
    while (pos < f.size() && std::isdigit(f[pos])) {
      v = v * 10 + (f[pos++] - '0');
      any = true;
      if (v > (1 << 20)) Bad(path, "bad header");
    }
    if (!any) Bad(path, "bad header");
    return uint32_t(v);
  };
  const int comp = f[1] == '6' ? 3 : 1;
  const uint32_t W = token(), H = token(), maxv = token();
  if (maxv != 255) Bad(path, "only 8-bit (maxval 255) PGM / PPM files are supported");
  pos++;  // one whitespace byte
  if (W == 0 || H == 0 || pos + size_t(W) * H * comp > f.size()) Bad(path, "truncated data");
  std::vector<uint8_t> px(f.begin() + pos, f.begin() + pos + size_t(W) * H * comp);
  return ToChannels(W, H, comp, px, req);
}

void FlipRows(uint8_t* data, uint32_t rowBytes, uint32_t rows) {
  std::vector<uint8_t> tmp(rowBytes);
  for (uint32_t y = 0; y < rows / 2; y++) {
    uint8_t* a = data + size_t(y) * rowBytes;
    uint8_t* b = data + size_t(rows - 1 - y) * rowBytes;
    std::memcpy(tmp.data(), a, rowBytes);
    std::memcpy(a, b, rowBytes);
    std::memcpy(b, tmp.data(), rowBytes);
  }
}

}  // namespace

Image8 ReadImage8(const std::string& path, int channels, bool flipY) {
  if (channels != 1 && channels != 3) throw std::runtime_error("image " + path + ": channel must be 1 or 3");
  std::vector<uint8_t> f = ReadFile(path);
  Image8 img;
  if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') {
    img = ReadPng(path, f, channels);
  } else if (f.size() >= 3 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) {
    img = ReadPnm(path, f, channels);
  } else {
    Bad(path, "unsupported format (PNG, binary PGM / PPM are read)");
  }
  if (flipY) FlipRows(img.px.data(), img.width * img.channels, img.height);
  return img;
}

// Radiance RGBE as stb_image's float loader decodes it: a texel with exponent
// e != 0 is (r, g, b) * 2^(e - 136), else 0.
ImageF ReadHdr(const std::string& path, bool flipY) {
  std::vector<uint8_t> f = ReadFile(path);
  size_t pos = 0;
  auto line = [&]() {
    std::string s;
    while (pos < f.size() && f[pos] != '\n') s.push_back(char(f[pos++]));
    if (pos >= f.size()) Bad(path, "truncated header");
    pos++;
    return s;
  };
  std::string first = line();
  if (first.rfind("#?RADIANCE", 0) != 0 && first.rfind("#?RGBE", 0) != 0) Bad(path, "not a Radiance HDR file");
  for (;;) {
    std::string s = line();
    if (s.empty()) break;
    if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe") Bad(path, "unsupported HDR format " + s);
  }
  int H = 0, W = 0;
  std::string res = line();
  if (std::sscanf(res.c_str(), "-Y %d +X %d", &H, &W) != 2 || W <= 0 || H <= 0 || W > (1 << 15) || H > (1 << 15))
    Bad(path, "unsupported resolution line (only -Y H +X W)");
  std::vector<uint8_t> rgbe(size_t(W) * H * 4);
  auto need = [&](size_t n) {
    if (pos + n > f.size()) Bad(path, "truncated data");
  };
  for (int y = 0; y < H; y++) {
    uint8_t* row = &rgbe[size_t(y) * W * 4];
    const bool rle = W >= 8 && W < 0x8000 && pos + 4 <= f.size() && f[pos] == 2 && f[pos + 1] == 2 &&
                     !(f[pos + 2] & 0x80);
    if (!rle) {
      need(size_t(W) * 4);
      std::memcpy(row, &f[pos], size_t(W) * 4);
      pos += size_t(W) * 4;
      continue;
    }
    if (((f[pos + 2] << 8) | f[pos + 3]) != W) Bad(path, "bad RLE scanline width");
    pos += 4;
    for (int c = 0; c < 4; c++) {
      int x = 0;
      while (x < W) {
        need(1);
        int count = f[pos++];
        if (count > 128) {
          count -= 128;
          need(1);
          const uint8_t v = f[pos++];
          if (x + count > W) Bad(path, "bad RLE run");
          for (int k = 0; k < count; k++) row[size_t(x++) * 4 + c] = v;
        } else {
          if (count == 0 || x + count > W) Bad(path, "bad RLE run");
          need(size_t(count));
          for (int k = 0; k < count; k++) row[size_t(x++) * 4 + c] = f[pos++];
        }
      }
    }
  }
  if (flipY) FlipRows(rgbe.data(), uint32_t(W) * 4, uint32_t(H));
  ImageF out;
  out.width = uint32_t(W);
  out.height = uint32_t(H);
  out.channels = 3;
  out.px.resize(size_t(W) * H * 3);
  for (size_t i = 0; i < size_t(W) * H; i++) {
    const uint8_t* p = &rgbe[4 * i];
    if (p[3]) {
      const float s = std::ldexp(1.0f, int(p[3]) - (128 + 8));
      out.px[3 * i] = float(p[0]) * s;
      out.px[3 * i + 1] = float(p[1]) * s;
      out.px[3 * i + 2] = float(p[2]) * s;
    } else {
      out.px[3 * i] = out.px[3 * i + 1] = out.px[3 * i + 2] = 0.f;
    }
  }
  return out;
}

}  // namespace host
}  // namespace rad
