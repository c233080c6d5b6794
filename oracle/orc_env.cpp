// CPU oracle — test infrastructure only.  See orc_env.h.
#include "orc_env.h"
#include <algorithm>

namespace orc {

void build_alias(const float *w, uint32_t n, float *prob, uint32_t *alias) {
    double sum = 0.0;
    for (uint32_t i = 0; i < n; ++i) sum += (double)w[i];
    std::vector<double> q(n);
    std::vector<uint32_t> small, large;
    small.reserve(n); large.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        q[i] = sum > 0.0 ? (double)w[i] * (double)n / sum : 1.0;
        if (q[i] < 1.0) small.push_back(i);
        else large.push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        uint32_t l = small.back(); small.pop_back();
        uint32_t g = large.back(); large.pop_back();
        prob[l] = (float)q[l];
        alias[l] = g;
        q[g] = (q[g] + q[l]) - 1.0;
        if (q[g] < 1.0) small.push_back(g);
        else large.push_back(g);
    }
    for (uint32_t i : large) { prob[i] = 1.0f; alias[i] = i; }
    for (uint32_t i : small) { prob[i] = 1.0f; alias[i] = i; }
}

void Envmap::build(uint32_t w, uint32_t h, const float *rgb3, float sc, bool footprint) {
    W = w; H = h; scale = sc;
    rgb.assign(rgb3, rgb3 + (size_t)w * h * 3);
    const float PI = 3.14159265358979323846f;
    // Texel luminance; in footprint mode (no counterpart in the reference) the
    // largest luminance of the 3 x 3 neighbourhood, U wrapped and V clamped,
    // which bounds the bilinear radiance anywhere inside the texel.
    std::vector<float> lum((size_t)W * H);
    for (size_t i = 0; i < lum.size(); ++i) {
        const float *p = &rgb[i * 3];
        lum[i] = 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2];
    }
    if (footprint) {
        std::vector<float> mx(lum.size());
        for (int y = 0; y < (int)H; ++y)
            for (int x = 0; x < (int)W; ++x) {
                float v = 0.0f;
                for (int yy = std::max(y - 1, 0); yy <= std::min(y + 1, (int)H - 1); ++yy)
                    for (int k = -1; k <= 1; ++k) v = std::max(v, lum[(size_t)yy * W + (size_t)((x + k + (int)W) % (int)W)]);
                mx[(size_t)y * W + x] = v;
            }
        lum.swap(mx);
    }
    conditional.assign((size_t)H * (W + 1), 0.0f);
    std::vector<float> row_sums(H);
    for (uint32_t y = 0; y < H; ++y) {
        float theta = PI * (y + 0.5f) / (float)H;
        float st = std::sin(theta);
        size_t base = (size_t)y * (W + 1);
        conditional[base] = 0.0f;
        float rs = 0.0f;
        for (uint32_t x = 0; x < W; ++x) {
            float wgt = lum[(size_t)y * W + x] * st;
            if (wgt < 0.0f) wgt = 0.0f;
            rs += wgt;
            conditional[base + x + 1] = rs;
        }
        row_sums[y] = rs;
        if (rs > 0.0f) {
            float inv = 1.0f / rs;
            for (uint32_t x = 1; x <= W; ++x) conditional[base + x] *= inv;
        } else {
            for (uint32_t x = 0; x <= W; ++x) conditional[base + x] = (float)x / (float)W;
        }
    }
    marginal.assign(H + 1, 0.0f);
    double total = 0.0;
    for (uint32_t y = 0; y < H; ++y) total += row_sums[y];
    if (total > 0.0) {
        double acc = 0.0, inv = 1.0 / total;
        for (uint32_t y = 0; y < H; ++y) {
            acc += row_sums[y];
            marginal[y + 1] = (float)(acc * inv);
        }
        marginal[H] = 1.0f;
    } else {
        for (uint32_t y = 0; y <= H; ++y) marginal[y] = (float)y / (float)H;
    }
    ppix.resize((size_t)W * H);
    for (uint32_t y = 0; y < H; ++y) {
        float pm = marginal[y + 1] - marginal[y];
        size_t base = (size_t)y * (W + 1);
        for (uint32_t x = 0; x < W; ++x)
            ppix[(size_t)y * W + x] = pm * (conditional[base + x + 1] - conditional[base + x]);
    }
    alias_prob.resize((size_t)W * H);
    alias_idx.resize((size_t)W * H);
    build_alias(ppix.data(), W * H, alias_prob.data(), alias_idx.data());
}

V3f Envmap::bilinear(float u, float v) const {
    float x = u * (float)W - 0.5f, y = v * (float)H - 0.5f;
    float fx0 = std::floor(x), fy0 = std::floor(y);
    float fx = x - fx0, fy = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0;
    int xa = ((x0 % (int)W) + (int)W) % (int)W, xb = (xa + 1) % (int)W;
    int ya = std::min(std::max(y0, 0), (int)H - 1), yb = std::min(std::max(y0 + 1, 0), (int)H - 1);
    V3f a = texel(xa, ya), b = texel(xb, ya), c = texel(xa, yb), d = texel(xb, yb);
    V3f top = a * (1.0f - fx) + b * fx, bot = c * (1.0f - fx) + d * fx;
    return top * (1.0f - fy) + bot * fy;
}

V3f Envmap::eval(V3f dir) const {
    V3f d = normalize(dir);
    const float PI = 3.14159265358979323846f;
    float phi = std::atan2(d.z, d.x);
    if (phi < 0.0f) phi += 2.0f * PI;
    float theta = std::acos(std::min(std::max(d.y, -1.0f), 1.0f));
    return bilinear(phi / (2.0f * PI), theta / PI) * scale;
}

float Envmap::pdf(V3f dir) const {
    V3f d = normalize(dir);
    const float PI = 3.14159265358979323846f;
    float phi = std::atan2(d.z, d.x);
    if (phi < 0.0f) phi += 2.0f * PI;
    float theta = std::acos(std::min(std::max(d.y, -1.0f), 1.0f));
    float st = std::sin(theta);
    float u = phi / (2.0f * PI), v = theta / PI;
    uint32_t x = std::min((uint32_t)(u * (float)W), W - 1);
    uint32_t y = std::min((uint32_t)(v * (float)H), H - 1);
    float pp = ppix[(size_t)y * W + x];
    return pp * (float)(W * H) / (2.0f * PI * PI * std::max(st, 1e-8f));
}

static uint32_t cdf_search(const float *c, uint32_t n, float u) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    return (uint32_t)std::max(0, (int)lo - 1);
}

void Envmap::sample_cdf(float u1, float u2, V3f &dir, V3f &L, float &pdf_out) const {
    const float PI = 3.14159265358979323846f;
    uint32_t y = cdf_search(marginal.data(), H + 1, u1);
    float m0 = marginal[y], m1 = marginal[y + 1], pm = m1 - m0;
    float jv = pm > 1e-20f ? saturate((u1 - m0) / pm) : 0.5f;
    const float *row = &conditional[(size_t)y * (W + 1)];
    uint32_t x = cdf_search(row, W + 1, u2);
    float c0 = row[x], c1 = row[x + 1], pc = c1 - c0;
    float ju = pc > 1e-20f ? saturate((u2 - c0) / pc) : 0.5f;
    float u = ((float)x + ju) / (float)W, v = ((float)y + jv) / (float)H;
    float phi = 2.0f * PI * u, theta = PI * v;
    float st = std::sin(theta), ct = std::cos(theta);
    dir = V3f(st * std::cos(phi), ct, st * std::sin(phi));
    L = bilinear(u, v) * scale;
    pdf_out = pm * pc * (float)(W * H) / (2.0f * PI * PI * std::max(st, 1e-8f));
}

uint32_t Envmap::alias_pick(uint32_t r, float coin) const {
    uint32_t n = W * H;
    uint32_t slot = (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
    return coin < alias_prob[slot] ? slot : alias_idx[slot];
}

void Envmap::sample_alias(uint32_t r, float coin, float jx, float jy, V3f &dir, V3f &L, float &pdf_out) const {
    const float PI = 3.14159265358979323846f;
    uint32_t idx = alias_pick(r, coin);
    uint32_t x = idx % W, y = idx / W;
    float u = ((float)x + jx) / (float)W, v = ((float)y + jy) / (float)H;
    float phi = 2.0f * PI * u, theta = PI * v;
    float st = std::sin(theta), ct = std::cos(theta);
    dir = V3f(st * std::cos(phi), ct, st * std::sin(phi));
    L = bilinear(u, v) * scale;
    pdf_out = ppix[idx] * (float)(W * H) / (2.0f * PI * PI * std::max(st, 1e-8f));
}

} // namespace orc
