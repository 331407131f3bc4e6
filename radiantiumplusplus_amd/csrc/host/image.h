// Image readers for the scene loader's image assets (module/rad.core/src/
// asset/image_stb.cpp, image_hdr.cpp): 8- and 16-bit PNG (zlib inflate), binary
// PGM / PPM with maxval 255, and Radiance RGBE (.hdr, flat or RLE scanlines).
// Pixels come back as the reference's stb-based loaders deliver them: 8-bit
// values converted to the requested channel count the way stb_image does
// (grey -> RGB by replication, RGB -> grey by (77 r + 150 g + 29 b) >> 8,
// 16-bit samples >> 8, alpha dropped), optionally flipped so that row 0 is
// the bottom row.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace rad {
namespace host {

struct Image8 {
  uint32_t width = 0, height = 0, channels = 0;  // channels as requested (1 or 3)
  std::vector<uint8_t> px;                       // row-major, row 0 first
};
struct ImageF {
  uint32_t width = 0, height = 0, channels = 0;
  std::vector<float> px;
};

// LDR file (PNG, PGM, PPM) with `channels` 1 or 3; throws std::runtime_error
// (message names the file) on unsupported or malformed input.
Image8 ReadImage8(const std::string& path, int channels, bool flip_y);
// Radiance RGBE file as linear RGB floats.
ImageF ReadHdr(const std::string& path, bool flip_y);

}  // namespace host
}  // namespace rad
