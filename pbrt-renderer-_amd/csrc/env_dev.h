// Equirectangular environment map on the device (y up, row 0 = +y).
//   radiance   shaders/Envmap.hlsli:15-24, bilinear with U wrap / V clamp
//              (the sampler of src/gpu/DXRApp_Pipeline.cpp:56-61)
//   pdf        shaders/Envmap.hlsli:109-135 from the per-texel probability
//              pMar*pCon of the 2D CDF (src/gpu/DXRApp_Textures.cpp:281-390)
//   sampling   alias table over the same per-texel probabilities (Vose 1991):
//              one 8-byte {prob, alias} entry per texel, O(1) per draw instead
//              of two dependent binary searches.
#pragma once
#include "nori_math.h"

namespace ng {

struct EnvDev {
    const float4 *texels; // W*H RGBA32F
    const float *ppix;    // W*H
    const uint2 *alias;   // W*H: {float bits of prob, alias index}
    uint32_t W, H;
    float scale;
};

NG_HD f3 env_bilinear(const EnvDev &e, float u, float v) {
    float x = u * (float)e.W - 0.5f, y = v * (float)e.H - 0.5f;
    float fx0 = floorf(x), fy0 = floorf(y);
    float fx = x - fx0, fy = y - fy0;
    int x0 = (int)fx0, y0 = (int)fy0, W = (int)e.W, H = (int)e.H;
    int xa = ((x0 % W) + W) % W, xb = xa + 1 == W ? 0 : xa + 1;
    int ya = min(max(y0, 0), H - 1), yb = min(max(y0 + 1, 0), H - 1);
    float4 a = e.texels[(size_t)ya * W + xa], b = e.texels[(size_t)ya * W + xb];
    float4 c = e.texels[(size_t)yb * W + xa], d = e.texels[(size_t)yb * W + xb];
    f3 A = mk(a.x, a.y, a.z), B = mk(b.x, b.y, b.z), C = mk(c.x, c.y, c.z), D = mk(d.x, d.y, d.z);
    f3 top = A * (1.0f - fx) + B * fx, bot = C * (1.0f - fx) + D * fx;
    return top * (1.0f - fy) + bot * fy;
}

NG_HD void env_dir_uv(f3 dir, float &u, float &v, float &sin_theta) {
    f3 d = normalize(dir);
    float phi = atan2f(d.z, d.x);
    if (phi < 0.0f) phi += 2.0f * PI;
    float theta = acosf(fminf(fmaxf(d.y, -1.0f), 1.0f));
    u = phi / (2.0f * PI);
    v = theta / PI;
    sin_theta = sinf(theta);
}

NG_HD f3 env_eval(const EnvDev &e, f3 dir) {
    float u, v, st;
    env_dir_uv(dir, u, v, st);
    return env_bilinear(e, u, v) * e.scale;
}

NG_HD float env_pdf(const EnvDev &e, f3 dir) {
    float u, v, st;
    env_dir_uv(dir, u, v, st);
    uint32_t x = min((uint32_t)(u * (float)e.W), e.W - 1), y = min((uint32_t)(v * (float)e.H), e.H - 1);
    return e.ppix[(size_t)y * e.W + x] * (float)(e.W * e.H) / (2.0f * PI * PI * fmaxf(st, 1e-8f));
}

NG_HD uint32_t env_alias_pick(const EnvDev &e, uint32_t r, float coin) {
    uint32_t n = e.W * e.H;
    uint32_t slot = (uint32_t)(((uint64_t)r * (uint64_t)n) >> 32);
    uint2 a = e.alias[slot];
#ifdef __HIP_DEVICE_COMPILE__
    float prob = __uint_as_float(a.x);
#else
    float prob;
    __builtin_memcpy(&prob, &a.x, 4);
#endif
    return coin < prob ? slot : a.y;
}

// Alias draw: r picks the slot, coin picks slot or alias, (jx, jy) jitter inside the texel.
NG_HD uint32_t env_sample(const EnvDev &e, uint32_t r, float coin, float jx, float jy, f3 &dir, f3 &L, float &pdf) {
    uint32_t idx = env_alias_pick(e, r, coin);
    uint32_t x = idx % e.W, y = idx / e.W;
    float u = ((float)x + jx) / (float)e.W, v = ((float)y + jy) / (float)e.H;
    float sp, cp, st, ct;
    sincosf(2.0f * PI * u, &sp, &cp);
    sincosf(PI * v, &st, &ct);
    dir = mk(st * cp, ct, st * sp);
    L = env_bilinear(e, u, v) * e.scale;
    pdf = e.ppix[idx] * (float)(e.W * e.H) / (2.0f * PI * PI * fmaxf(st, 1e-8f));
    return idx;
}

} // namespace ng
