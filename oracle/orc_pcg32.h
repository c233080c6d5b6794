// CPU oracle — test infrastructure only.
//
// PCG-XSH-RR 64/32 (M. E. O'Neill, "PCG: A Family of Simple Fast Space-Efficient
// Statistically Good Algorithms for Random Number Generation", 2014), restated
// from the published algorithm.  The reference uses it through its `pcg32`
// submodule (ext/pcg32, not vendored; .gitmodules:16-18) in
// src/cpu/independent.cpp:37-54 (`seed`, `nextFloat`).
//
// Stream convention of the renderer: pixel p of a render with seed S uses
// seed(S ^ p, PCG32_DEFAULT_STREAM); sample s of that pixel starts at
// advance(s << SAMPLE_WINDOW_LOG2), so every sample owns a fixed window of the
// pixel's single stream.
#pragma once
#include <cstdint>

namespace orc {

static const uint64_t PCG32_DEFAULT_STATE = 0x853c49e6748fea9bULL;
static const uint64_t PCG32_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;
static const uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;
static const int SAMPLE_WINDOW_LOG2 = 20;

struct Pcg32 {
    uint64_t state = PCG32_DEFAULT_STATE;
    uint64_t inc = PCG32_DEFAULT_STREAM;

    void seed(uint64_t initstate, uint64_t initseq = 1) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next_uint();
        state += initstate;
        next_uint();
    }
    uint32_t next_uint() {
        uint64_t old = state;
        state = old * PCG32_MULT + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31));
    }
    // 23 random mantissa bits in [0,1): (u >> 9) | 0x3f800000 reinterpreted, minus 1.
    float next_float() {
        union { uint32_t u; float f; } x;
        x.u = (next_uint() >> 9) | 0x3f800000u;
        return x.f - 1.0f;
    }
    // Jump the LCG by `delta` steps in O(log delta) (Brown, "Random number
    // generation with arbitrary strides", 1994).
    void advance(uint64_t delta) {
        uint64_t cur_mult = PCG32_MULT, cur_plus = inc, acc_mult = 1u, acc_plus = 0u;
        while (delta > 0) {
            if (delta & 1) {
                acc_mult *= cur_mult;
                acc_plus = acc_plus * cur_mult + cur_plus;
            }
            cur_plus = (cur_mult + 1) * cur_plus;
            cur_mult *= cur_mult;
            delta >>= 1;
        }
        state = acc_mult * state + acc_plus;
    }
};

inline Pcg32 sample_stream(uint64_t render_seed, uint32_t pixel, uint32_t sample) {
    Pcg32 r;
    r.seed(render_seed ^ (uint64_t)pixel, PCG32_DEFAULT_STREAM);
    r.advance((uint64_t)sample << SAMPLE_WINDOW_LOG2);
    return r;
}

// shaders/RNG.hlsli:9-14 — the integer hash used for the stochastic alpha test.
inline uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

} // namespace orc
