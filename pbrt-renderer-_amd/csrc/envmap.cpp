// Host construction of the envmap sampling tables.
//
// The per-texel probability is the one the reference's pdf query uses,
// pMar(y) * pCon(x|y) (shaders/Envmap.hlsli:127-131), from the 2D CDF of
// src/gpu/DXRApp_Textures.cpp:281-390: weight = luminance * sin(theta at the row
// centre), rows accumulated in float and normalised by the float row sum, the
// marginal accumulated in double.  The alias table is Vose's method over those
// probabilities; its worklists are stacks filled in texel order, so the table
// (and every draw from it) is a deterministic function of the envmap.
//
// Footprint mode (not in the reference): the luminance of a texel is replaced
// by the largest luminance of its 3 x 3 neighbourhood (U wrapped, V clamped).
// Radiance is interpolated bilinearly between texel centres, so inside a texel
// it never exceeds that maximum; a dim texel next to a bright one (the rim of
// a small sun) is then drawn in proportion to the radiance it can return, and
// next-event samples there no longer carry radiance / pdf ratios hundreds of
// times the mean.
#include "envmap.h"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace ng {

void build_env_tables(uint32_t W, uint32_t H, const float *rgb, EnvTables &t, bool footprint) {
    const float PI_F = 3.14159265358979323846f;
    std::vector<float> lum((size_t)W * H);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const float *p = rgb + i * 3;
        lum[i] = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
    }
    if (footprint) {
        std::vector<float> m(lum.size());
        for (uint32_t y = 0; y < H; ++y)
            for (uint32_t x = 0; x < W; ++x) {
                float v = lum[(size_t)y * W + x];
                for (int dy = -1; dy <= 1; ++dy) {
                    const uint32_t yy = (uint32_t)std::min(std::max((int)y + dy, 0), (int)H - 1);
                    for (int dx = -1; dx <= 1; ++dx) {
                        const uint32_t xx = (x + W + (uint32_t)(dx + 1) - 1u) % W;
                        v = std::max(v, lum[(size_t)yy * W + xx]);
                    }
                }
                m[(size_t)y * W + x] = v;
            }
        lum.swap(m);
    }
    std::vector<float> cond((size_t)H * (W + 1)), rows(H), marg(H + 1);
    for (uint32_t y = 0; y < H; ++y) {
        float st = std::sin(PI_F * (y + 0.5f) / (float)H);
        float *c = &cond[(size_t)y * (W + 1)];
        c[0] = 0.0f;
        float s = 0.0f;
        for (uint32_t x = 0; x < W; ++x) {
            float w = lum[(size_t)y * W + x] * st;
            s += w < 0.0f ? 0.0f : w;
            c[x + 1] = s;
        }
        rows[y] = s;
        if (s > 0.0f) {
            float inv = 1.0f / s;
            for (uint32_t x = 1; x <= W; ++x) c[x] *= inv;
        } else {
            for (uint32_t x = 0; x <= W; ++x) c[x] = (float)x / (float)W;
        }
    }
    double total = 0.0;
    for (uint32_t y = 0; y < H; ++y) total += rows[y];
    marg[0] = 0.0f;
    if (total > 0.0) {
        double acc = 0.0, inv = 1.0 / total;
        for (uint32_t y = 0; y < H; ++y) {
            acc += rows[y];
            marg[y + 1] = (float)(acc * inv);
        }
        marg[H] = 1.0f;
    } else {
        for (uint32_t y = 0; y <= H; ++y) marg[y] = (float)y / (float)H;
    }
    const size_t n = (size_t)W * H;
    t.ppix.resize(n);
    for (uint32_t y = 0; y < H; ++y) {
        float pm = marg[y + 1] - marg[y];
        const float *c = &cond[(size_t)y * (W + 1)];
        for (uint32_t x = 0; x < W; ++x) t.ppix[(size_t)y * W + x] = pm * (c[x + 1] - c[x]);
    }
    // Vose alias table.
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += (double)t.ppix[i];
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        q[i] = sum > 0.0 ? (double)t.ppix[i] * (double)n / sum : 1.0;
        (q[i] < 1.0 ? small : large).push_back((uint32_t)i);
    }
    t.alias.assign(2 * n, 0u);
    auto put = [&](uint32_t i, float prob, uint32_t a) {
        uint32_t bits;
        std::memcpy(&bits, &prob, 4);
        t.alias[2 * i] = bits;
        t.alias[2 * i + 1] = a;
    };
    while (!small.empty() && !large.empty()) {
        uint32_t l = small.back();
        small.pop_back();
        uint32_t g = large.back();
        large.pop_back();
        put(l, (float)q[l], g);
        q[g] = (q[g] + q[l]) - 1.0;
        (q[g] < 1.0 ? small : large).push_back(g);
    }
    for (uint32_t i : large) put(i, 1.0f, i);
    for (uint32_t i : small) put(i, 1.0f, i);
    t.texels.resize(4 * n);
    for (size_t i = 0; i < n; ++i) {
        t.texels[4 * i] = rgb[3 * i];
        t.texels[4 * i + 1] = rgb[3 * i + 1];
        t.texels[4 * i + 2] = rgb[3 * i + 2];
        t.texels[4 * i + 3] = 1.0f;
    }
}

} // namespace ng
