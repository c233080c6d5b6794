// Host side of the textures: mip chains and the device tables of tex_dev.h.
#pragma once
#include "../../include/nori_gpu.h"
#include "tex_dev.h"
#include <vector>

namespace ng {

struct TexUpload {
    std::vector<TexDev> tex;
    std::vector<uint32_t> texel;
    std::vector<float> lut; // 512 entries
};

// Mip chains of n textures (the reference's 2x2 box filter on the 8-bit values,
// src/gpu/DXRApp_Textures.cpp:57-93) packed as tex_dev.h describes; throws
// std::invalid_argument on an empty or oversized texture.
void build_textures(const nori_texture *t, uint32_t n, TexUpload &out);

} // namespace ng
