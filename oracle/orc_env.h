// CPU oracle — test infrastructure only.
//
// Equirectangular environment map (y up).
//   evaluation               shaders/Envmap.hlsli:15-24 (bilinear, U wrap / V clamp as
//                            the sampler at src/gpu/DXRApp_Pipeline.cpp:56-61)
//   2D CDF build             src/gpu/DXRApp_Textures.cpp:281-390
//   CDF sampling             shaders/Envmap.hlsli:31-101
//   solid-angle pdf          shaders/Envmap.hlsli:109-135
//   alias table (new)        Vose, "A linear algorithm for generating random numbers with
//                            a given distribution", IEEE TSE 1991, over the per-texel
//                            probabilities pMar*pCon that the CDF pdf uses.
#pragma once
#include "orc_math.h"
#include <vector>

namespace orc {

struct Envmap {
    uint32_t W = 0, H = 0;
    float scale = 1.0f;
    std::vector<float> rgb;          // H*W*3, row 0 = theta 0 (+y)
    std::vector<float> marginal;     // H+1
    std::vector<float> conditional;  // H*(W+1)
    std::vector<float> ppix;         // H*W: pMar(y)*pCon(x|y) in float
    std::vector<float> alias_prob;   // H*W
    std::vector<uint32_t> alias_idx; // H*W

    bool valid() const { return W > 0 && H > 0; }
    // footprint: 3 x 3 neighbourhood maximum weights (orc_env.cpp); false: the reference's
    void build(uint32_t w, uint32_t h, const float *rgb3, float scale, bool footprint = false);
    V3f texel(int x, int y) const {
        const float *p = &rgb[((size_t)y * W + x) * 3];
        return V3f(p[0], p[1], p[2]);
    }
    V3f bilinear(float u, float v) const;
    V3f eval(V3f dir) const;
    float pdf(V3f dir) const;
    // Reference CDF sampler (two uniforms).
    void sample_cdf(float u1, float u2, V3f &dir, V3f &L, float &pdf) const;
    // Alias sampler: r = 32-bit uniform integer picks the texel slot, coin picks
    // slot or alias, (jx, jy) place the direction inside the texel.
    uint32_t alias_pick(uint32_t r, float coin) const;
    void sample_alias(uint32_t r, float coin, float jx, float jy, V3f &dir, V3f &L, float &pdf) const;
};

// Vose alias build over `w` (n entries, any non-negative scale).  Worklists are
// LIFO stacks filled in index order; prob is rounded to float at the end.
void build_alias(const float *w, uint32_t n, float *prob, uint32_t *alias);

} // namespace orc
