// Float RGB images: Radiance .hdr (RGBE, flat or run-length encoded) and PFM
// readers for environment maps; PFM and uncompressed OpenEXR writers for renders.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace nori {

struct ImageRGB {
    uint32_t width = 0, height = 0;
    std::vector<float> rgb; // row-major, row 0 at the top, 3 floats per pixel
};

/// Reads .hdr / .pic (RGBE) or .pfm by extension; throws NoriException on failure.
ImageRGB readImage(const std::string &filename);
/// Writes .pfm or .exr (by extension); throws NoriException on failure.
void writeImage(const std::string &filename, const ImageRGB &img);

/// Decodes a texture file (PNG, baseline JPEG, binary PNM) to RGBA8, row 0 at the top; on
/// failure returns false with the reason in `err`.
bool loadTextureRGBA8(const std::string &path, uint32_t &width, uint32_t &height, std::vector<uint8_t> &rgba,
                      std::string &err);
/// Baseline JPEG decoder behind loadTextureRGBA8.
bool decodeJpegRGBA8(const std::vector<uint8_t> &file, uint32_t &width, uint32_t &height, std::vector<uint8_t> &rgba,
                     std::string &err);

} // namespace nori
