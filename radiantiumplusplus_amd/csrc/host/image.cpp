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

namespace {

uint32_t LE32(const uint8_t* p) { return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24); }
uint64_t LE64(const uint8_t* p) { return uint64_t(LE32(p)) | (uint64_t(LE32(p + 4)) << 32); }

// IEEE 754 binary16 -> binary32 (exact).
float HalfToFloat(uint16_t h) {
  const uint32_t s = uint32_t(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
  uint32_t bits;
  if (e == 0) {
    const float v = std::ldexp(float(m), -24);  // zero or subnormal
    return s ? -v : v;
  }
  bits = e == 31 ? (s | 0x7f800000u | (m << 13)) : (s | ((e + 112u) << 23) | (m << 13));
  float f;
  std::memcpy(&f, &bits, 4);
  return f;
}

// binary32 -> binary16, round to nearest, ties to even.
uint16_t FloatToHalf(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint16_t s = uint16_t((x >> 16) & 0x8000u);
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return uint16_t(s | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));  // inf / nan
  if (ax >= 0x477ff000u) return uint16_t(s | 0x7c00u);  // >= 65520 rounds to inf
  if (ax < 0x38800000u) {  // below 2^-14: a multiple of 2^-24, exact after scaling
    float a;
    std::memcpy(&a, &ax, 4);
    return uint16_t(s | uint16_t(std::nearbyint(a * 16777216.0f)));
  }
  uint32_t h = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13);
  const uint32_t rem = ax & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return uint16_t(s | h);
}

// OpenEXR's ZIP / RLE post-processing: a byte-wise delta predictor over data
// whose even and odd bytes were split into two halves.
void ExrUnpredict(std::vector<uint8_t>& t, std::vector<uint8_t>& out) {
  for (size_t i = 1; i < t.size(); i++) t[i] = uint8_t(t[i - 1] + t[i] - 128);
  out.resize(t.size());
  const size_t half = (t.size() + 1) / 2;
  for (size_t k = 0; k < t.size(); k++) out[k] = (k & 1) ? t[half + k / 2] : t[k / 2];
}

}  // namespace

ImageF ReadExr(const std::string& path) {
  const std::vector<uint8_t> f = ReadFile(path);
  if (f.size() < 8 || LE32(f.data()) != 20000630u) Bad(path, "not an OpenEXR file");
  const uint32_t version = LE32(f.data() + 4);
  if ((version & 0xffu) != 2) Bad(path, "unknown OpenEXR version");
  if (version & 0x200u) throw ImageUnsupported("image " + path + ": tiled OpenEXR files are not supported");
  if (version & 0x1800u) throw ImageUnsupported("image " + path + ": deep or multi-part OpenEXR files are not supported");
  size_t pos = 8;
  auto cstr = [&]() {
    std::string r;
    while (pos < f.size() && f[pos]) r.push_back(char(f[pos++]));
    if (pos >= f.size()) Bad(path, "truncated header");
    pos++;
    return r;
  };
  struct Chan {
    std::string name;
    int32_t type;
  };
  std::vector<Chan> chans;
  int32_t comp = -1, x0 = 0, y0 = 0, x1 = -1, y1 = -1;
  bool haveWindow = false;
  for (;;) {
    const std::string name = cstr();
    if (name.empty()) break;
    const std::string type = cstr();
    if (pos + 4 > f.size()) Bad(path, "truncated header");
    const uint32_t size = LE32(f.data() + pos);
    pos += 4;
    if (size > f.size() - pos) Bad(path, "truncated header");
    const uint8_t* v = f.data() + pos;
    if (name == "channels" && type == "chlist") {
      size_t q = 0;
      while (q < size && v[q]) {
        std::string cn;
        while (q < size && v[q]) cn.push_back(char(v[q++]));
        q++;
        if (q + 16 > size) Bad(path, "bad channel list");
        const int32_t pt = int32_t(LE32(v + q)), xs = int32_t(LE32(v + q + 8)), ys = int32_t(LE32(v + q + 12));
        if (pt < 0 || pt > 2) Bad(path, "bad pixel type");
        if (xs != 1 || ys != 1) throw ImageUnsupported("image " + path + ": subsampled OpenEXR channels are not supported");
        chans.push_back({cn, pt});
        q += 16;
      }
    } else if (name == "compression" && size >= 1) {
      comp = v[0];
    } else if (name == "dataWindow" && size == 16) {
      x0 = int32_t(LE32(v)), y0 = int32_t(LE32(v + 4)), x1 = int32_t(LE32(v + 8)), y1 = int32_t(LE32(v + 12));
      haveWindow = true;
    } else if (name == "tiles") {
      throw ImageUnsupported("image " + path + ": tiled OpenEXR files are not supported");
    }
    pos += size;
  }
  if (chans.empty() || comp < 0 || !haveWindow || x1 < x0 || y1 < y0) Bad(path, "missing channels, compression or data window");
  const int64_t W = int64_t(x1) - x0 + 1, H = int64_t(y1) - y0 + 1;
  if (W > (1 << 20) || H > (1 << 20) || W * H > (int64_t(1) << 28)) Bad(path, "image too large");
  int32_t linesPerBlock;
  switch (comp) {
    case 0: case 1: case 2: linesPerBlock = 1; break;  // NONE, RLE, ZIPS
    case 3: linesPerBlock = 16; break;                  // ZIP
    default: throw ImageUnsupported("image " + path + ": OpenEXR compression " + std::to_string(comp) + " is not supported");
  }
  // R / G / B by the lower-cased suffix after the last '.', first occurrence wins
  int rgbChan[3] = {-1, -1, -1};
  for (size_t c = 0; c < chans.size(); c++) {
    std::string n = chans[c].name;
    const size_t dot = n.rfind('.');
    if (dot != std::string::npos) n = n.substr(dot + 1);
    for (char& ch : n) ch = char(std::tolower(static_cast<unsigned char>(ch)));
    const int k = n == "r" ? 0 : n == "g" ? 1 : n == "b" ? 2 : -1;
    if (k >= 0 && rgbChan[k] < 0) rgbChan[k] = int(c);
  }
  const int32_t pixelType = chans[0].type;  // the frame buffer's type for every slice
  if (pixelType == 0) throw ImageUnsupported("image " + path + ": UINT OpenEXR pixels are not supported");
  for (int k = 0; k < 3; k++)
    if (rgbChan[k] >= 0 && chans[size_t(rgbChan[k])].type == 0)
      throw ImageUnsupported("image " + path + ": UINT OpenEXR channels are not supported");
  std::vector<size_t> chanOff(chans.size());
  size_t lineBytes = 0;
  for (size_t c = 0; c < chans.size(); c++) {
    chanOff[c] = lineBytes;
    lineBytes += size_t(W) * (chans[c].type == 1 ? 2 : 4);
  }
  if (lineBytes > (size_t(1) << 26)) Bad(path, "scanline too large");
  const int64_t blocks = (H + linesPerBlock - 1) / linesPerBlock;
  if (pos + size_t(blocks) * 8 > f.size()) Bad(path, "truncated offset table");
  ImageF img;
  img.width = uint32_t(W);
  img.height = uint32_t(H);
  img.channels = 3;
  img.px.assign(size_t(W) * size_t(H) * 3, 0.f);
  std::vector<uint8_t> tmp, data;
  for (int64_t b = 0; b < blocks; b++) {
    const uint64_t off = LE64(f.data() + pos + size_t(b) * 8);
    if (off > f.size() || f.size() - off < 8) Bad(path, "bad chunk offset");
    const int32_t y = int32_t(LE32(f.data() + off));
    const uint32_t packed = LE32(f.data() + off + 4);
    if (packed > f.size() - off - 8) Bad(path, "truncated chunk");
    if (y < y0 || y > y1 || (int64_t(y) - y0) % linesPerBlock != 0) Bad(path, "bad chunk line");
    const int64_t lines = std::min<int64_t>(linesPerBlock, int64_t(y1) - y + 1);
    const size_t want = size_t(lines) * lineBytes;
    const uint8_t* src = f.data() + off + 8;
    if (comp == 0 || packed >= want) {  // stored as is when compression does not pay
      if (packed != want) Bad(path, "bad chunk size");
      data.assign(src, src + want);
    } else if (comp == 1) {  // RLE
      tmp.clear();
      for (size_t i = 0; i < packed;) {
        const int8_t c = int8_t(src[i++]);
        if (c < 0) {
          if (i + size_t(-c) > packed) Bad(path, "bad RLE run");
          tmp.insert(tmp.end(), src + i, src + i + size_t(-c));
          i += size_t(-c);
        } else {
          if (i >= packed) Bad(path, "bad RLE run");
          tmp.insert(tmp.end(), size_t(c) + 1, src[i++]);
        }
        if (tmp.size() > want) Bad(path, "RLE overrun");
      }
      if (tmp.size() != want) Bad(path, "bad RLE size");
      ExrUnpredict(tmp, data);
    } else {  // ZIPS / ZIP
      tmp.resize(want);
      uLongf n = uLongf(want);
      if (uncompress(tmp.data(), &n, src, uLong(packed)) != Z_OK || n != want) Bad(path, "bad zlib chunk");
      ExrUnpredict(tmp, data);
    }
    for (int64_t l = 0; l < lines; l++) {
      const uint8_t* line = data.data() + size_t(l) * lineBytes;
      float* row = img.px.data() + size_t(int64_t(y) - y0 + l) * size_t(W) * 3;
      for (int k = 0; k < 3; k++) {
        if (rgbChan[k] < 0) continue;
        const size_t c = size_t(rgbChan[k]);
        const uint8_t* p = line + chanOff[c];
        for (int64_t x = 0; x < W; x++) {
          float v;
          if (chans[c].type == 1) {
            v = HalfToFloat(uint16_t(p[2 * x] | (p[2 * x + 1] << 8)));
          } else {
            const uint32_t bits = LE32(p + 4 * x);
            std::memcpy(&v, &bits, 4);
            if (pixelType == 1) v = HalfToFloat(FloatToHalf(v));  // a FLOAT channel read into a HALF slice
          }
          row[3 * size_t(x) + size_t(k)] = v;
        }
      }
    }
  }
  return img;
}

}  // namespace host
}  // namespace rad
