// Host/device math for the MI355X path tracer: a small float3 type, the pcg32
// generator and the shading-frame helpers used by every kernel.
//
// Device code is compiled with -ffp-contract=fast except where a function opens
// `#pragma clang fp contract(off)`: the ray/triangle and box tests, whose results
// must match the CPU Möller–Trumbore evaluation bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define NG_HD __host__ __device__ __forceinline__

namespace ng {

struct f3 {
    float x, y, z;
};
NG_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
NG_HD f3 mk(float a) { return f3{a, a, a}; }
NG_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
NG_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
NG_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
NG_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
NG_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
NG_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
NG_HD f3 operator/(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
NG_HD f3 &operator+=(f3 &a, f3 b) { a = a + b; return a; }
NG_HD f3 &operator*=(f3 &a, f3 b) { a = a * b; return a; }
NG_HD f3 &operator*=(f3 &a, float s) { a = a * s; return a; }
NG_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
NG_HD f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
NG_HD float maxc(f3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
NG_HD f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
NG_HD float length(f3 a) { return sqrtf(dot(a, a)); }
NG_HD f3 lerp3(f3 a, f3 b, float t) { return a + (b - a) * t; }
NG_HD float saturatef_(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
NG_HD f3 reflect(f3 i, f3 n) { return i - 2.0f * dot(n, i) * n; }
NG_HD f3 refract(f3 i, f3 n, float eta) {
    float c = dot(n, i);
    float k = 1.0f - eta * eta * (1.0f - c * c);
    if (k < 0.0f) return mk(0.0f);
    return eta * i - (eta * c + sqrtf(k)) * n;
}
NG_HD bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

constexpr float PI = 3.14159265358979323846f;
constexpr float INV_PI = 0.31830988618379067154f;

// Rec. 709 luminance with the weights of shaders/GeometryUtils.hlsli:225-228.
NG_HD float luminance(f3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// Shading frame (shaders/GeometryUtils.hlsli:160-165).
NG_HD void build_onb(f3 n, f3 &t, f3 &b) {
    f3 up = fabsf(n.y) < 0.999f ? mk(0, 1, 0) : mk(1, 0, 0);
    t = normalize(cross(up, n));
    b = cross(n, t);
}
NG_HD f3 to_local(f3 v, f3 t, f3 b, f3 n) { return {dot(v, t), dot(v, b), dot(v, n)}; }
NG_HD f3 to_world(f3 v, f3 t, f3 b, f3 n) { return t * v.x + b * v.y + n * v.z; }
// shaders/GeometryUtils.hlsli:178-183
NG_HD f3 offset_ray_origin(f3 p, f3 ng, f3 n) {
    f3 q = p + n * 0.002f;
    return q + ng * fmaxf(0.0f, -dot(q - p, ng));
}
// shaders/GeometryUtils.hlsli:18-24
NG_HD float balance_heuristic(float a, float b) {
    float s = a + b;
    return s < 1e-20f ? 0.0f : a / s;
}
// shaders/GeometryUtils.hlsli:189-194
NG_HD f3 cosine_sample_hemisphere(float ux, float uy) {
    float r = sqrtf(ux), phi = 2.0f * PI * uy;
    float s, c;
    sincosf(phi, &s, &c);
    return {r * c, r * s, sqrtf(fmaxf(0.0f, 1.0f - ux))};
}
// shaders/GeometryUtils.hlsli:200-219
NG_HD float fresnel_dielectric(float cosi, float etaI, float etaT) {
    if (etaI == etaT) return 0.0f;
    if (cosi < 0.0f) {
        float t = etaI; etaI = etaT; etaT = t;
        cosi = -cosi;
    }
    float eta = etaI / etaT;
    float s2 = eta * eta * (1.0f - cosi * cosi);
    if (s2 > 1.0f) return 1.0f;
    float cost = sqrtf(1.0f - s2);
    float rs = (etaI * cosi - etaT * cost) / (etaI * cosi + etaT * cost);
    float rp = (etaT * cosi - etaI * cost) / (etaT * cosi + etaI * cost);
    return 0.5f * (rs * rs + rp * rp);
}

// ------------------------------------------------------------------ pcg32
// PCG-XSH-RR 64/32 (O'Neill 2014) as used by Nori's Independent sampler
// (src/cpu/independent.cpp:37-54).  Every stream of the renderer uses the
// increment of PCG32_DEFAULT_STREAM, so a generator is one 64-bit state word.
constexpr uint64_t PCG_MULT = 0x5851f42d4c957f2dULL;
constexpr uint64_t PCG_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;
constexpr uint64_t PCG_INC = (PCG_DEFAULT_STREAM << 1u) | 1u;
constexpr int SAMPLE_WINDOW_LOG2 = 20;

NG_HD uint32_t pcg_output(uint64_t old) {
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((~rot + 1u) & 31));
}
NG_HD uint32_t pcg_next(uint64_t &state, uint64_t inc = PCG_INC) {
    uint64_t old = state;
    state = old * PCG_MULT + inc;
    return pcg_output(old);
}
NG_HD float pcg_float(uint64_t &state) {
    uint32_t u = (pcg_next(state) >> 9) | 0x3f800000u;
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(u) - 1.0f;
#else
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f - 1.0f;
#endif
}
NG_HD uint64_t pcg_seed(uint64_t initstate, uint64_t initseq) {
    uint64_t inc = (initseq << 1u) | 1u, st = 0;
    pcg_next(st, inc);
    st += initstate;
    pcg_next(st, inc);
    return st;
}
// LCG jump by `delta` steps (Brown 1994).
NG_HD uint64_t pcg_advance(uint64_t state, uint64_t delta, uint64_t inc = PCG_INC) {
    uint64_t cm = PCG_MULT, cp = inc, am = 1u, ap = 0u;
    while (delta > 0) {
        if (delta & 1) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
        delta >>= 1;
    }
    return am * state + ap;
}
// Pixel stream seed(render_seed ^ pixel, PCG32_DEFAULT_STREAM); sample s owns the
// window starting at advance(s << SAMPLE_WINDOW_LOG2).
NG_HD uint64_t sample_stream(uint64_t render_seed, uint32_t pixel, uint32_t sample) {
    uint64_t st = pcg_seed(render_seed ^ (uint64_t)pixel, PCG_DEFAULT_STREAM);
    return pcg_advance(st, (uint64_t)sample << SAMPLE_WINDOW_LOG2);
}

// ------------------------------------------------- Owen-scrambled Sobol'
// The progressive sampler of shaders/RNG.hlsli:9-99 (Burley 2020, "Practical
// Hash-based Owen Scrambling"): per pixel a scramble seed, per sample its
// index; dimension d draws axis d % 2 of the 2-D Sobol' point of pad d / 2, the
// sample index shuffled per pad and the value Owen-scrambled per pixel, pad
// and axis.  Integer arithmetic throughout, so host and device agree bit for
// bit.
NG_HD uint32_t pcg_hash(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
NG_HD uint32_t reverse_bits32(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_bitreverse32(x);
#else
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
#endif
}
// Sobol' base 2, MSB-aligned: axis 0 is the van der Corput sequence, axis 1
// uses the direction numbers v_{k+1} = v_k ^ (v_k >> 1), v_1 = 2^31.
NG_HD uint32_t sobol2d(uint32_t index, uint32_t axis) {
    if (axis == 0u) return reverse_bits32(index);
    uint32_t r = 0u, dirn = 1u << 31;
    for (uint32_t i = index; i != 0u; i >>= 1u) {
        if (i & 1u) r ^= dirn;
        dirn ^= dirn >> 1u;
    }
    return r;
}
NG_HD uint32_t laine_karras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
NG_HD uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    return reverse_bits32(laine_karras(reverse_bits32(x), seed));
}
struct SobolRng {
    uint32_t pixel_seed, index, dim;
};
// pixel: y * width + x of the view; sample: the progressive sample number.
NG_HD SobolRng sobol_init(uint32_t pixel, uint32_t sample, uint32_t dim = 0) {
    return SobolRng{pcg_hash(pixel ^ 0x9e3779b9u), sample, dim};
}
NG_HD uint32_t sobol_next_u32(SobolRng &r) {
    const uint32_t d = r.dim++;
    const uint32_t axis = d & 1u, pad = d >> 1u;
    const uint32_t sidx = owen_scramble(r.index, pcg_hash(pad * 0x9e3779b9u + 1u));
    return owen_scramble(sobol2d(sidx, axis), pcg_hash(r.pixel_seed + pad * 0x6c8e9cf5u + axis * 0x68bc21ebu));
}
NG_HD float sobol_float(SobolRng &r) {
    return fminf((float)sobol_next_u32(r) * 2.3283064365386963e-10f, 0.99999994f);
}

// The path stream as either generator: rng_float / rng_u32 take one dimension.
NG_HD float rng_float(uint64_t &s) { return pcg_float(s); }
NG_HD uint32_t rng_u32(uint64_t &s) { return pcg_next(s); }
NG_HD float rng_float(SobolRng &s) { return sobol_float(s); }
NG_HD uint32_t rng_u32(SobolRng &s) { return sobol_next_u32(s); }

} // namespace ng
