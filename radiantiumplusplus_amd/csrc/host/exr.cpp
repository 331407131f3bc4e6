// OpenEXR output of the rendered framebuffer (Renderer::SaveResult,
// sample_renderer.cpp:34-46; ImageReader::WriteExr, core/src/image_reader.cpp:230-255):
// RGB FLOAT channels, one scanline per block, written without compression.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../../include/rad_hip.h"

namespace {

void Put(std::vector<char>& b, const void* p, size_t n) {
  const char* c = static_cast<const char*>(p);
  b.insert(b.end(), c, c + n);
}
void PutStr(std::vector<char>& b, const char* s) { Put(b, s, std::strlen(s) + 1); }
template <class T>
void PutV(std::vector<char>& b, T v) {
  Put(b, &v, sizeof(T));
}
void Attr(std::vector<char>& b, const char* name, const char* type, const std::vector<char>& value) {
  PutStr(b, name);
  PutStr(b, type);
  PutV<int32_t>(b, int32_t(value.size()));
  Put(b, value.data(), value.size());
}

}  // namespace

extern "C" rad_status rad_write_exr(const char* path, const float* rgb, int32_t width, int32_t height) {
  if (!path || !rgb || width <= 0 || height <= 0) return RAD_ERR_INVALID_ARG;
  std::vector<char> h;
  PutV<int32_t>(h, 20000630);  // magic
  PutV<int32_t>(h, 2);         // version 2, scanline image
  {
    std::vector<char> ch;
    for (const char* c : {"B", "G", "R"}) {  // channel list is sorted by name
      PutStr(ch, c);
      PutV<int32_t>(ch, 2);  // FLOAT
      PutV<uint8_t>(ch, 0);  // pLinear
      PutV<uint8_t>(ch, 0);
      PutV<uint8_t>(ch, 0);
      PutV<uint8_t>(ch, 0);
      PutV<int32_t>(ch, 1);  // xSampling
      PutV<int32_t>(ch, 1);  // ySampling
    }
    ch.push_back(0);
    Attr(h, "channels", "chlist", ch);
  }
  {
    std::vector<char> v;
    v.push_back(0);  // NO_COMPRESSION
    Attr(h, "compression", "compression", v);
  }
  {
    std::vector<char> v;
    PutV<int32_t>(v, 0);
    PutV<int32_t>(v, 0);
    PutV<int32_t>(v, width - 1);
    PutV<int32_t>(v, height - 1);
    Attr(h, "dataWindow", "box2i", v);
    Attr(h, "displayWindow", "box2i", v);
  }
  {
    std::vector<char> v;
    v.push_back(0);  // INCREASING_Y
    Attr(h, "lineOrder", "lineOrder", v);
  }
  {
    std::vector<char> v;
    PutV<float>(v, 1.f);
    Attr(h, "pixelAspectRatio", "float", v);
  }
  {
    std::vector<char> v;
    PutV<float>(v, 0.f);
    PutV<float>(v, 0.f);
    Attr(h, "screenWindowCenter", "v2f", v);
  }
  {
    std::vector<char> v;
    PutV<float>(v, 1.f);
    Attr(h, "screenWindowWidth", "float", v);
  }
  {
    std::vector<char> v;
    const char* c = "rad.offline";
    Put(v, c, std::strlen(c));
    Attr(h, "comments", "string", v);
  }
  h.push_back(0);  // end of header
  const int64_t lineBytes = int64_t(width) * 3 * 4;
  const int64_t blockBytes = 8 + lineBytes;
  int64_t offset = int64_t(h.size()) + 8 * int64_t(height);
  for (int32_t y = 0; y < height; y++) PutV<int64_t>(h, offset + int64_t(y) * blockBytes);
  std::ofstream f(path, std::ios::binary);
  if (!f) return RAD_ERR_IO;
  f.write(h.data(), std::streamsize(h.size()));
  std::vector<float> line(size_t(width) * 3);
  for (int32_t y = 0; y < height; y++) {
    int32_t yy = y, n = int32_t(lineBytes);
    f.write(reinterpret_cast<const char*>(&yy), 4);
    f.write(reinterpret_cast<const char*>(&n), 4);
    const float* row = rgb + size_t(y) * width * 3;
    for (int c = 0; c < 3; c++) {  // B, G, R planes
      int src = 2 - c;
      for (int32_t x = 0; x < width; x++) line[size_t(c) * width + x] = row[3 * size_t(x) + src];
    }
    f.write(reinterpret_cast<const char*>(line.data()), std::streamsize(lineBytes));
  }
  return f ? RAD_OK : RAD_ERR_IO;
}
