// Host envmap tables (see envmap.cpp).
#pragma once
#include <cstdint>
#include <vector>

namespace ng {

struct EnvTables {
    std::vector<float> texels;  // W*H*4 RGBA
    std::vector<float> ppix;    // W*H
    std::vector<uint32_t> alias; // W*H*2: float bits of prob, alias index
};

// footprint: texel weights from the 3 x 3 neighbourhood maximum (envmap.cpp);
// false: the reference's per-texel luminance.
void build_env_tables(uint32_t W, uint32_t H, const float *rgb, EnvTables &t, bool footprint = false);

} // namespace ng
