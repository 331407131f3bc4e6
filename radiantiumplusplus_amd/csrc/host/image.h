// Image readers for the scene loader's image assets (module/rad.core/src/
// asset/image_stb.cpp, image_hdr.cpp, image_exr.cpp): 8- and 16-bit PNG (zlib
// inflate), binary PGM / PPM with maxval 255, Radiance RGBE (.hdr, flat or RLE
// scanlines) and OpenEXR scanline files.
// Pixels come back as the reference's stb-based loaders deliver them: 8-bit
// values converted to the requested channel count the way stb_image does
// (grey -> RGB by replication, RGB -> grey by (77 r + 150 g + 29 b) >> 8,
// 16-bit samples >> 8, alpha dropped), optionally flipped so that row 0 is
// the bottom row.
#pragma once

#include <cstdint>
#include <stdexcept>
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

// A well-formed file that uses a feature these readers do not carry.
struct ImageUnsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// OpenEXR scanline file as ImageReader::ReadExr delivers it (core/src/
// image_reader.cpp:118-214): the data window's R, G and B channels (names
// matched case-insensitively after the last '.', the first of each in the
// sorted channel list; a missing one reads 0), converted to the pixel type of
// the file's first channel, row 0 being the top row.  NONE, RLE, ZIPS and ZIP
// compression; other compressions, tiled / deep / multi-part files, UINT
// pixels and subsampled channels throw ImageUnsupported.
ImageF ReadExr(const std::string& path);

}  // namespace host
}  // namespace rad
