// Texture lookups of the shading and alpha-test code (host and device).
//
// A texture is RGBA8 with a full mip chain built by the reference's 2x2 box
// filter on 8-bit values (src/gpu/DXRApp_Textures.cpp:57-93): level l is
// max(w >> l, 1) x max(h >> l, 1) texels, packed one uint32 per texel (R in the
// low byte) at texel offset off[l] of the scene's texel buffer.  Sampling is
// trilinear with wrapped coordinates and texel centres at (i + 0.5) / size,
// the reference's MIN_MAG_MIP_LINEAR / WRAP sampler (src/gpu/DXRApp_Pipeline.
// cpp:48-51); 8-bit values become floats through 256-entry tables (sRGB decode
// for colour textures, c / 255 otherwise) before filtering, as the sRGB
// formats do.  Filter weights are full FP32 (hardware samplers quantise them;
// that difference is the only one from the reference's lookups).  Evaluated
// without contraction so host and device round the same way.
#pragma once
#include "nori_math.h"

namespace ng {

constexpr int TEX_MAX_LEVELS = 16;

struct TexDev {
    uint32_t w, h, levels, srgb;
    uint32_t off[TEX_MAX_LEVELS]; // texel offset of each level
};

struct TexTables {
    const TexDev *tex;     // per texture
    const uint32_t *texel; // all levels of all textures
    const float *lut;      // [0, 256): sRGB decode, [256, 512): c / 255
    uint32_t count;
};

NG_HD f3 tex_fetch_rgb(const TexTables &t, const TexDev &d, uint32_t level, int x, int y, float &alpha) {
    const int w = (int)(d.w >> level) > 0 ? (int)(d.w >> level) : 1;
    const int h = (int)(d.h >> level) > 0 ? (int)(d.h >> level) : 1;
    x %= w;
    if (x < 0) x += w;
    y %= h;
    if (y < 0) y += h;
    const uint32_t v = t.texel[d.off[level] + (uint32_t)y * (uint32_t)w + (uint32_t)x];
    const float *cl = t.lut + (d.srgb ? 0 : 256);
    alpha = t.lut[256 + (v >> 24)];
    return f3{cl[v & 0xFFu], cl[(v >> 8) & 0xFFu], cl[(v >> 16) & 0xFFu]};
}

// Bilinear lookup of one level; rgb in the return value, alpha in `alpha`.
// Scalar arithmetic only (the f3 operators live outside this pragma's reach).
NG_HD f3 tex_bilinear(const TexTables &t, const TexDev &d, uint32_t level, float u, float v, float &alpha) {
#pragma clang fp contract(off)
    const uint32_t w = d.w >> level > 0 ? d.w >> level : 1u, h = d.h >> level > 0 ? d.h >> level : 1u;
    const float x = u * (float)w - 0.5f, y = v * (float)h - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0;
    const int ix = (int)x0, iy = (int)y0;
    float a00, a10, a01, a11;
    const f3 c00 = tex_fetch_rgb(t, d, level, ix, iy, a00), c10 = tex_fetch_rgb(t, d, level, ix + 1, iy, a10);
    const f3 c01 = tex_fetch_rgb(t, d, level, ix, iy + 1, a01), c11 = tex_fetch_rgb(t, d, level, ix + 1, iy + 1, a11);
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    const float r = (c00.x * gx + c10.x * fx) * gy + (c01.x * gx + c11.x * fx) * fy;
    const float g = (c00.y * gx + c10.y * fx) * gy + (c01.y * gx + c11.y * fx) * fy;
    const float b = (c00.z * gx + c10.z * fx) * gy + (c01.z * gx + c11.z * fx) * fy;
    alpha = (a00 * gx + a10 * fx) * gy + (a01 * gx + a11 * fx) * fy;
    return f3{r, g, b};
}

// Trilinear lookup at level-of-detail `lod` (clamped to the chain).
NG_HD f3 tex_sample(const TexTables &t, uint32_t ti, float u, float v, float lod, float &alpha) {
#pragma clang fp contract(off)
    const TexDev &d = t.tex[ti];
    const float top = (float)(d.levels - 1);
    const float l = fminf(fmaxf(lod, 0.0f), top);
    const float l0 = floorf(l), f = l - l0;
    const uint32_t i0 = (uint32_t)l0;
    f3 c = tex_bilinear(t, d, i0, u, v, alpha);
    if (f > 0.0f && i0 + 1 < d.levels) {
        float a1;
        const f3 c1 = tex_bilinear(t, d, i0 + 1, u, v, a1);
        const float g = 1.0f - f;
        c = f3{c.x * g + c1.x * f, c.y * g + c1.y * f, c.z * g + c1.z * f};
        alpha = alpha * g + a1 * f;
    }
    return c;
}

// The value the alpha test reads at level 0 (shaders/Shaders.hlsl:674-735):
// the bilinear red channel for radiance rays, and for shadow rays the alpha
// channel where it is below 1, else red; the same texels, weights and sums as
// tex_sample at lod 0 on those channels, read from a decoded plane of level 0
// ((red, alpha) per texel, the values the texture tables give them).
NG_HD float alpha_plane_value(const float2 *plane, uint32_t tw, uint32_t th, float u, float v, bool shadow) {
#pragma clang fp contract(off)
    const int w = (int)tw, h = (int)th;
    const float x = u * (float)tw - 0.5f, y = v * (float)th - 0.5f;
    const float x0 = floorf(x), y0 = floorf(y);
    const float fx = x - x0, fy = y - y0, gx = 1.0f - fx, gy = 1.0f - fy;
    int xa = (int)x0 % w, ya = (int)y0 % h;
    if (xa < 0) xa += w;
    if (ya < 0) ya += h;
    const int xb = xa + 1 == w ? 0 : xa + 1, yb = ya + 1 == h ? 0 : ya + 1;
    const float2 *row_a = plane + (uint32_t)ya * (uint32_t)w, *row_b = plane + (uint32_t)yb * (uint32_t)w;
    const float2 v00 = row_a[xa], v10 = row_a[xb], v01 = row_b[xa], v11 = row_b[xb];
    const float red = (v00.x * gx + v10.x * fx) * gy + (v01.x * gx + v11.x * fx) * fy;
    if (!shadow) return red;
    const float a = (v00.y * gx + v10.y * fx) * gy + (v01.y * gx + v11.y * fx) * fy;
    return a < 1.0f ? a : red;
}

// Mip level of a texture for a UV footprint (shaders/GeometryUtils.hlsli:148-154).
NG_HD float tex_lod(const TexTables &t, uint32_t ti, float uv_footprint) {
    const TexDev &d = t.tex[ti];
    const float texels = uv_footprint * (float)(d.w > d.h ? d.w : d.h);
    return fmaxf(log2f(fmaxf(texels, 1.0f)), 0.0f);
}

} // namespace ng
