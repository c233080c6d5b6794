// Mip chains and lookup tables of the scene textures.
#include "texture.h"
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace ng {

void build_textures(const nori_texture *t, uint32_t n, TexUpload &out) {
    out.tex.clear();
    out.texel.clear();
    out.lut.assign(512, 0.0f);
    for (int c = 0; c < 256; ++c) {
        // sRGB decode (IEC 61966-2-1) and UNORM8 -> c / 255, each rounded once to float
        const double x = c / 255.0;
        out.lut[c] = (float)(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
        out.lut[256 + c] = (float)x;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const nori_texture &s = t[i];
        if (s.width == 0 || s.height == 0 || !s.rgba)
            throw std::invalid_argument("texture " + std::to_string(i) + ": empty");
        if (s.width > (1u << 15) || s.height > (1u << 15))
            throw std::invalid_argument("texture " + std::to_string(i) + ": larger than 32768 texels a side");
        TexDev d{};
        d.w = s.width;
        d.h = s.height;
        d.srgb = s.srgb ? 1u : 0u;
        uint32_t levels = 1;
        for (uint32_t dim = std::max(s.width, s.height); dim > 1; dim >>= 1) ++levels;
        d.levels = levels;
        // level 0
        std::vector<uint8_t> prev(s.rgba, s.rgba + (size_t)s.width * s.height * 4), cur;
        uint32_t pw = s.width, ph = s.height;
        for (uint32_t l = 0; l < levels; ++l) {
            if (l > 0) {
                const uint32_t mw = std::max(pw / 2, 1u), mh = std::max(ph / 2, 1u);
                cur.assign((size_t)mw * mh * 4, 0);
                for (uint32_t y = 0; y < mh; ++y)
                    for (uint32_t x = 0; x < mw; ++x) {
                        const uint32_t sx = std::min(x * 2, pw - 1), sy = std::min(y * 2, ph - 1);
                        const uint32_t sx1 = std::min(sx + 1, pw - 1), sy1 = std::min(sy + 1, ph - 1);
                        for (int c = 0; c < 4; ++c) {
                            const int v = prev[((size_t)sy * pw + sx) * 4 + c] + prev[((size_t)sy * pw + sx1) * 4 + c] +
                                          prev[((size_t)sy1 * pw + sx) * 4 + c] + prev[((size_t)sy1 * pw + sx1) * 4 + c];
                            cur[((size_t)y * mw + x) * 4 + c] = (uint8_t)(v / 4);
                        }
                    }
                prev.swap(cur);
                pw = mw;
                ph = mh;
            }
            if (out.texel.size() + (size_t)pw * ph > 0xFFFFFFFFull)
                throw std::invalid_argument("textures: more than 2^32 texels in all");
            d.off[l] = (uint32_t)out.texel.size();
            for (size_t k = 0; k < (size_t)pw * ph; ++k)
                out.texel.push_back((uint32_t)prev[4 * k] | ((uint32_t)prev[4 * k + 1] << 8) |
                                    ((uint32_t)prev[4 * k + 2] << 16) | ((uint32_t)prev[4 * k + 3] << 24));
        }
        out.tex.push_back(d);
    }
}

} // namespace ng
