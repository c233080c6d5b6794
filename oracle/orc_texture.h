// CPU restatement of the reference's texture path (test infrastructure only).
//   mips      src/gpu/DXRApp_Textures.cpp:57-93: full chain, level l+1 from the
//             2x2 box average (integer, truncated) of level l's 8-bit values,
//             source coordinates clamped
//   formats   :33-42: sRGB textures decode on sampling (IEC 61966-2-1), the
//             others are UNORM (c / 255); alpha is always UNORM
//   sampler   src/gpu/DXRApp_Pipeline.cpp:48-51: MIN_MAG_MIP_LINEAR, WRAP;
//             texel centres at (i + 1/2) / size; values decoded before filtering
//   LOD       shaders/GeometryUtils.hlsli:120-154
// The filter weights are exact floats (hardware quantises them to a few bits).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace orc {

struct Texture {
    uint32_t w = 0, h = 0;
    bool srgb = false;
    std::vector<std::vector<uint8_t>> level; // RGBA8 per level
    std::vector<uint32_t> lw, lh;

    void build(uint32_t width, uint32_t height, bool is_srgb, const uint8_t *rgba) {
        w = width;
        h = height;
        srgb = is_srgb;
        level.assign(1, std::vector<uint8_t>(rgba, rgba + (size_t)w * h * 4));
        lw.assign(1, w);
        lh.assign(1, h);
        while (lw.back() > 1 || lh.back() > 1) {
            const uint32_t pw = lw.back(), ph = lh.back();
            const uint32_t mw = std::max(pw / 2, 1u), mh = std::max(ph / 2, 1u);
            const std::vector<uint8_t> &src = level.back();
            std::vector<uint8_t> dst((size_t)mw * mh * 4);
            for (uint32_t y = 0; y < mh; ++y)
                for (uint32_t x = 0; x < mw; ++x) {
                    const uint32_t x0 = std::min(2 * x, pw - 1), y0 = std::min(2 * y, ph - 1);
                    const uint32_t x1 = std::min(x0 + 1, pw - 1), y1 = std::min(y0 + 1, ph - 1);
                    for (int c = 0; c < 4; ++c) {
                        const int sum = src[((size_t)y0 * pw + x0) * 4 + c] + src[((size_t)y0 * pw + x1) * 4 + c] +
                                        src[((size_t)y1 * pw + x0) * 4 + c] + src[((size_t)y1 * pw + x1) * 4 + c];
                        dst[((size_t)y * mw + x) * 4 + c] = (uint8_t)(sum / 4);
                    }
                }
            level.push_back(std::move(dst));
            lw.push_back(mw);
            lh.push_back(mh);
        }
    }
};

// Decode tables: [0, 256) sRGB, [256, 512) UNORM.
inline const std::vector<float> &texel_tables() {
    static const std::vector<float> t = [] {
        std::vector<float> v(512);
        for (int c = 0; c < 256; ++c) {
            const double x = c / 255.0;
            v[c] = (float)(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
            v[256 + c] = (float)x;
        }
        return v;
    }();
    return t;
}

struct Rgba {
    float r, g, b, a;
};

inline Rgba texel(const Texture &t, uint32_t l, int x, int y) {
    const int w = (int)t.lw[l], h = (int)t.lh[l];
    x = ((x % w) + w) % w;
    y = ((y % h) + h) % h;
    const uint8_t *p = &t.level[l][((size_t)y * w + x) * 4];
    const std::vector<float> &lut = texel_tables();
    const int o = t.srgb ? 0 : 256;
    return {lut[o + p[0]], lut[o + p[1]], lut[o + p[2]], lut[256 + p[3]]};
}

inline Rgba bilinear(const Texture &t, uint32_t l, float u, float v) {
    const float x = u * (float)t.lw[l] - 0.5f, y = v * (float)t.lh[l] - 0.5f;
    const float fx0 = std::floor(x), fy0 = std::floor(y);
    const float fx = x - fx0, fy = y - fy0, gx = 1.0f - fx, gy = 1.0f - fy;
    const int ix = (int)fx0, iy = (int)fy0;
    const Rgba a = texel(t, l, ix, iy), b = texel(t, l, ix + 1, iy), c = texel(t, l, ix, iy + 1),
               d = texel(t, l, ix + 1, iy + 1);
    auto mix = [&](float p, float q, float r, float s) { return (p * gx + q * fx) * gy + (r * gx + s * fx) * fy; };
    return {mix(a.r, b.r, c.r, d.r), mix(a.g, b.g, c.g, d.g), mix(a.b, b.b, c.b, d.b), mix(a.a, b.a, c.a, d.a)};
}

inline Rgba sample(const Texture &t, float u, float v, float lod) {
    const float top = (float)(t.level.size() - 1);
    const float l = std::min(std::max(lod, 0.0f), top);
    const float l0 = std::floor(l), f = l - l0;
    const uint32_t i0 = (uint32_t)l0;
    Rgba c = bilinear(t, i0, u, v);
    if (f > 0.0f && i0 + 1 < t.level.size()) {
        const Rgba d = bilinear(t, i0 + 1, u, v);
        const float g = 1.0f - f;
        c = {c.r * g + d.r * f, c.g * g + d.g * f, c.b * g + d.b * f, c.a * g + d.a * f};
    }
    return c;
}

inline float lod_for(const Texture &t, float footprint) {
    const float texels = footprint * (float)std::max(t.w, t.h);
    return std::max(std::log2(std::max(texels, 1.0f)), 0.0f);
}

} // namespace orc
