// CPU restatement of the reference GPU's progressive sampler (test
// infrastructure only): Owen-scrambled Sobol' after Burley 2020,
// shaders/RNG.hlsli:9-99.
//   pcg_hash       shaders/RNG.hlsli:9-14 (orc_pcg32.h)
//   sobol_2d       shaders/RNG.hlsli:17-30 (axis 0 van der Corput, axis 1 with
//                  direction numbers v_{k+1} = v_k xor (v_k >> 1), v_1 = 2^31)
//   laine_karras   shaders/RNG.hlsli:33-41
//   owen_scramble  shaders/RNG.hlsli:43-49
//   SobolRng       InitRNG (:60-70, pixel seed = pcg_hash(pixel ^ 0x9e3779b9),
//                  dimension 0) and NextFloat (:81-97, USE_SOBOL = 1)
#pragma once
#include "orc_pcg32.h"
#include <cstdint>

namespace orc {

inline uint32_t reverse_bits(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) {
        r = (r << 1) | (x & 1u);
        x >>= 1;
    }
    return r;
}

inline uint32_t sobol_2d(uint32_t index, uint32_t axis) {
    if (axis == 0) return reverse_bits(index);
    uint32_t r = 0, v = 0x80000000u;
    for (; index; index >>= 1, v ^= v >> 1)
        if (index & 1u) r ^= v;
    return r;
}

inline uint32_t laine_karras(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}

inline uint32_t owen_scramble(uint32_t x, uint32_t seed) { return reverse_bits(laine_karras(reverse_bits(x), seed)); }

// The stream of one (pixel, sample): dimension d takes axis d % 2 of pad d / 2.
struct SobolRng {
    uint32_t pixel_seed = 0, index = 0, dim = 0;
    SobolRng() = default;
    SobolRng(uint32_t pixel, uint32_t sample) : pixel_seed(pcg_hash(pixel ^ 0x9e3779b9u)), index(sample) {}
    uint32_t next_uint() {
        const uint32_t d = dim++, axis = d & 1u, pad = d >> 1u;
        const uint32_t shuffled = owen_scramble(index, pcg_hash(pad * 0x9e3779b9u + 1u));
        return owen_scramble(sobol_2d(shuffled, axis), pcg_hash(pixel_seed + pad * 0x6c8e9cf5u + axis * 0x68bc21ebu));
    }
    float next_float() {
        const float f = (float)next_uint() * 2.3283064365386963e-10f;
        return f < 0.99999994f ? f : 0.99999994f;
    }
};

} // namespace orc
