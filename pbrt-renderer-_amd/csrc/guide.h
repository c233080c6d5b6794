// Online neural path guiding (no counterpart in the reference, whose integrator
// samples only the BSDF, shaders/Shaders.hlsl:473-511).
//
// Model: 64 input features -> 3 hidden layers of 64 (ReLU) -> 64 outputs, of
// which 41 parameterise an 8-lobe von Mises-Fisher mixture q and the selection
// probability c of one-sample MIS between q and the BSDF:
//   P(w) = c q(w) + (1 - c) b(w).
//   features  per axis of the scene-normalised position, 8 octaves of
//             (sin, cos)(2^k pi p) (48, octaves 1-3 and 5-7 by angle
//             doubling); shading normal (3); direction to the
//             viewer (3); roughness, metallic; 8 zeros
//   outputs   mean_j = normalize(o[3j..3j+2]), kappa_j = exp(clamp(o[24+j], -4, 9)),
//             weights = softmax(o[32..39]), c = 0.1 + 0.5 sigmoid(o[40])
//             (c at most 0.6: a BSDF-sampled direction's weight is at most
//             1 / (1 - c) = 2.5 times the plain estimator's, however wrong q is)
//   training  minimises -W log P(w_s) over samples the renderer records (the
//             sample gradient of KL(target || P); Mueller et al., TOG 2019),
//             W = luminance of the radiance that arrived along w_s divided by
//             the throughput before the sampling vertex, less what an emitter
//             or the envmap hit right at the next vertex gave (light NEE also
//             samples); Adam (Kingma & Ba).
// Matrix products use BF16 MFMA with FP32 accumulation
// (v_mfma_f32_32x32x16_bf16 on gfx950); parameters are FP32 masters.
#pragma once
#include "nori_math.h"

namespace ng {

constexpr int G_DIM = 64;
constexpr int G_LAYERS = 4;
constexpr int G_LAYER_PARAMS = G_DIM * G_DIM + G_DIM;
constexpr int G_PARAMS = G_LAYERS * G_LAYER_PARAMS;
constexpr int G_USED_OUT = 41;

typedef __bf16 bf16_t;

struct GuideDev {
    float *params;  // FP32 masters: per layer W[64 out][64 in] then b[64]
    float *adam_m, *adam_v;
    float *grad;    // FP32 gradient accumulator, same layout
    bf16_t *w;      // BF16 W[l][out][in]
    bf16_t *wt;     // BF16 W[l][in][out]
    float *loss;    // [0] summed loss of the last step
};

// Vertex record written by the shading kernel for training (6 float4).
struct GuideRecord {
    float4 pos_rough;  // world position, roughness
    float4 n_metal;    // shading normal, metallic
    float4 wi_b;       // direction to the viewer, BSDF pdf of the sampled direction
    float4 w_valid;    // sampled direction (world)
    float4 thr;        // throughput before the vertex
    float4 lsnap;      // path radiance collected up to and including this vertex's NEE; bounce index
};

// Training sample after the path has finished (4 float4).
struct GuideSample {
    float4 pos01_rough;
    float4 n_metal;
    float4 wi_b;
    float4 w_weight;
};

// Features 16s + 8h .. 16s + 8h + 7 of one query (s = 0..3, h = 0..1) in the
// layout above: octaves 4h..4h+3 of axis s for s < 3, then the normal,
// direction, roughness and metallic (h = 0) or the zero padding (h = 1).
NG_HD void guide_feature_block(f3 p01, f3 n, f3 wi, float rough, float metal, int s, int h, float *f) {
    if (s < 3) {
        // octave 4h from sincos, the next three by doubling the angle
        // (sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x): each doubling
        // about doubles the absolute error, to a few 1e-7 at the top octave,
        // far below the features' BF16 rounding
        const float p = s == 0 ? p01.x : (s == 1 ? p01.y : p01.z);
        float sn, cs;
        sincosf((float)(1 << (4 * h)) * PI * p, &sn, &cs);
        for (int k = 0; k < 4; ++k) {
            f[2 * k] = sn;
            f[2 * k + 1] = cs;
            const float s2 = 2.0f * sn * cs;
            cs = 1.0f - 2.0f * sn * sn;
            sn = s2;
        }
    } else if (h == 0) {
        f[0] = n.x; f[1] = n.y; f[2] = n.z;
        f[3] = wi.x; f[4] = wi.y; f[5] = wi.z;
        f[6] = rough;
        f[7] = metal;
    } else {
        for (int k = 0; k < 8; ++k) f[k] = 0.0f;
    }
}

NG_HD void guide_features(f3 p01, f3 n, f3 wi, float rough, float metal, float *f) {
    for (int s = 0; s < 4; ++s)
        for (int h = 0; h < 2; ++h) guide_feature_block(p01, n, wi, rough, metal, s, h, f + 16 * s + 8 * h);
}

// Decoded mixture of one query.
struct GuideMix {
    f3 mu[8];
    float kappa[8], norm[8], w[8], c;
};

NG_HD GuideMix guide_decode(const float *o) {
    GuideMix g;
    for (int j = 0; j < 8; ++j) {
        f3 r = mk(o[3 * j], o[3 * j + 1], o[3 * j + 2]);
        float l = fmaxf(sqrtf(dot(r, r)), 1e-12f);
        g.mu[j] = r * (1.0f / l);
        float k = expf(fminf(fmaxf(o[24 + j], -4.0f), 9.0f));
        g.kappa[j] = k;
        g.norm[j] = k / (2.0f * PI * (1.0f - expf(-2.0f * k)));
    }
    float mx = o[32];
    for (int j = 1; j < 8; ++j) mx = fmaxf(mx, o[32 + j]);
    float s = 0.0f;
    for (int j = 0; j < 8; ++j) {
        g.w[j] = expf(o[32 + j] - mx);
        s += g.w[j];
    }
    float inv = 1.0f / s;
    for (int j = 0; j < 8; ++j) g.w[j] *= inv;
    g.c = 0.1f + 0.5f / (1.0f + expf(-o[40]));
    return g;
}

// Mixture of one query as the render loop keeps it between inference and
// shading (96 bytes instead of 64 FP32 outputs): FP16 unit lobe axes,
// concentrations, lobe weights and c.  guide_unpack derives everything the
// sampler and the density use from these stored values alone (axes normalised
// again, weights divided by their sum), so sampling and density evaluation
// describe the same distribution.
struct alignas(16) GuidePacked {
    _Float16 v[48];
};

NG_HD void guide_pack(const float *o, GuidePacked &p) {
    const GuideMix g = guide_decode(o);
    for (int j = 0; j < 8; ++j) {
        p.v[3 * j] = (_Float16)g.mu[j].x;
        p.v[3 * j + 1] = (_Float16)g.mu[j].y;
        p.v[3 * j + 2] = (_Float16)g.mu[j].z;
        p.v[24 + j] = (_Float16)g.kappa[j];
        p.v[32 + j] = (_Float16)g.w[j];
    }
    p.v[40] = (_Float16)g.c;
    for (int k = 41; k < 48; ++k) p.v[k] = (_Float16)0.0f;
}

NG_HD GuideMix guide_unpack(const GuidePacked &p) {
    GuideMix g;
    float s = 0.0f;
    for (int j = 0; j < 8; ++j) {
        const f3 r = mk((float)p.v[3 * j], (float)p.v[3 * j + 1], (float)p.v[3 * j + 2]);
        g.mu[j] = r * (1.0f / fmaxf(sqrtf(dot(r, r)), 1e-12f));
        const float k = (float)p.v[24 + j];
        g.kappa[j] = k;
        g.norm[j] = k / (2.0f * PI * (1.0f - expf(-2.0f * k)));
        g.w[j] = (float)p.v[32 + j];
        s += g.w[j];
    }
    const float inv = 1.0f / s;
    for (int j = 0; j < 8; ++j) g.w[j] *= inv;
    g.c = (float)p.v[40];
    return g;
}

// The same decode one lobe at a time, straight from the packed record, for
// callers that keep the 24-register packed form live instead of the unpacked
// mixture: guide_lobe_* and guide_q_packed / guide_sample_packed compute the
// same values as guide_unpack followed by guide_q / guide_sample.
NG_HD float guide_packed_inv_wsum(const GuidePacked &p) {
    float s = 0.0f;
    for (int j = 0; j < 8; ++j) s += (float)p.v[32 + j];
    return 1.0f / s;
}
NG_HD f3 guide_lobe_mu(const GuidePacked &p, int j) {
    const f3 r = mk((float)p.v[3 * j], (float)p.v[3 * j + 1], (float)p.v[3 * j + 2]);
    return r * (1.0f / fmaxf(sqrtf(dot(r, r)), 1e-12f));
}
NG_HD float guide_q_packed(const GuidePacked &p, float inv_wsum, f3 w) {
    float q = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float k = (float)p.v[24 + j];
        const float norm = k / (2.0f * PI * (1.0f - expf(-2.0f * k)));
        q += ((float)p.v[32 + j] * inv_wsum) * norm * expf(k * (dot(guide_lobe_mu(p, j), w) - 1.0f));
    }
    return q;
}

NG_HD float guide_q(const GuideMix &g, f3 w) {
    float q = 0.0f;
    for (int j = 0; j < 8; ++j) q += g.w[j] * g.norm[j] * expf(g.kappa[j] * (dot(g.mu[j], w) - 1.0f));
    return q;
}

// vMF lobe sample (Jakob 2012, numerically stable form), frame of Duff et al. 2017.
NG_HD f3 vmf_sample(f3 mu, float k, float u1, float u2) {
    float cw = 1.0f + logf(u1 + (1.0f - u1) * expf(-2.0f * k)) / k;
    float sw = sqrtf(fmaxf(0.0f, 1.0f - cw * cw));
    float sp, cp;
    sincosf(2.0f * PI * u2, &sp, &cp);
    float sgn = mu.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sgn + mu.z), b = mu.x * mu.y * a;
    f3 t = mk(1.0f + sgn * mu.x * mu.x * a, sgn * b, -sgn * mu.x);
    f3 s = mk(b, sgn + mu.y * mu.y * a, -mu.y);
    return normalize(mu * cw + (t * cp + s * sp) * sw);
}

// Pick a lobe by weight with u in [0,1): the first j with ul < w_0 + .. + w_j
// (summed in order; the last lobe when rounding leaves ul above the total),
// then sample it.  Unrolled with static indices only, so the mixture stays in
// registers.
NG_HD f3 guide_sample(const GuideMix &g, float ul, float u1, float u2) {
    f3 mu = g.mu[0];
    float k = g.kappa[0], acc = g.w[0];
    bool found = ul < acc;
#pragma unroll
    for (int j = 1; j < 8; ++j) {
        if (!found) {
            mu = g.mu[j];
            k = g.kappa[j];
            acc += g.w[j];
            found = ul < acc;
        }
    }
    return vmf_sample(mu, k, u1, u2);
}
NG_HD f3 guide_sample_packed(const GuidePacked &p, float inv_wsum, float ul, float u1, float u2) {
    int sel = 0;
    float acc = (float)p.v[32] * inv_wsum;
    bool found = ul < acc;
#pragma unroll
    for (int j = 1; j < 8; ++j) {
        if (!found) {
            sel = j;
            acc += (float)p.v[32 + j] * inv_wsum;
            found = ul < acc;
        }
    }
    f3 mu = mk(0.0f);
    float k = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
        if (j == sel) {
            mu = guide_lobe_mu(p, j);
            k = (float)p.v[24 + j];
        }
    return vmf_sample(mu, k, u1, u2);
}

// Loss -W log P and its gradient w.r.t. the outputs (41 used, rest zero).
NG_HD float guide_loss_dout(const float *o, f3 w, float b, float weight, float *g) {
    for (int k = 0; k < 64; ++k) g[k] = 0.0f;
    GuideMix m = guide_decode(o);
    float v[8], cosm[8], q = 0.0f;
    for (int j = 0; j < 8; ++j) {
        cosm[j] = dot(m.mu[j], w);
        v[j] = m.norm[j] * expf(m.kappa[j] * (cosm[j] - 1.0f));
        q += m.w[j] * v[j];
    }
    float P = m.c * q + (1.0f - m.c) * b;
    if (!(P > 0.0f) || weight == 0.0f) return 0.0f;
    float dP = -weight / P;
    float s = (m.c - 0.1f) / 0.5f;
    g[40] = dP * (q - b) * 0.5f * s * (1.0f - s);
    float dq = dP * m.c;
    for (int j = 0; j < 8; ++j) {
        g[32 + j] = dq * m.w[j] * (v[j] - q);
        float k = m.kappa[j], e = expf(-2.0f * k);
        float dlogv = 1.0f / k - 2.0f * e / (1.0f - e) + (cosm[j] - 1.0f);
        float dv = dq * m.w[j];
        float raw = o[24 + j];
        if (raw > -4.0f && raw < 9.0f) g[24 + j] = dv * v[j] * dlogv * k;
        f3 r = mk(o[3 * j], o[3 * j + 1], o[3 * j + 2]);
        float len = fmaxf(sqrtf(dot(r, r)), 1e-12f);
        float dm = dv * v[j] * k / len;
        g[3 * j] = dm * (w.x - cosm[j] * m.mu[j].x);
        g[3 * j + 1] = dm * (w.y - cosm[j] * m.mu[j].y);
        g[3 * j + 2] = dm * (w.z - cosm[j] * m.mu[j].z);
    }
    return -weight * logf(P);
}

// Launchers (guide.hip).  Row counts may live on the device (count != null:
// rows = *count, grid sized for cap rows).
void guide_refresh(const GuideDev &g, hipStream_t st);
// Network outputs (FP32, 64 per row) of the BF16 features X through all four
// layers in one launch (inference: no activations kept).
void guide_infer(const bf16_t *X, const GuideDev &g, float *out, uint32_t cap, const uint32_t *count, hipStream_t st);
// Y[rows x 64] = act(X[rows x 64] M^T (+ bias)), M[64 n][64 k] BF16.
// mode 0: +bias, ReLU -> BF16 Y; mode 1: +bias -> FP32 Y; mode 2: * (Hmask > 0) -> BF16 Y.
void guide_layer(const bf16_t *X, const bf16_t *M, const float *bias, void *Y, const bf16_t *Hmask, int mode,
                 uint32_t cap, const uint32_t *count, hipStream_t st);
// grad_W[64 o][64 i] += dZ^T H ; grad_b[o] += sum dZ (rows from count or cap).
void guide_wgrad(const bf16_t *dZ, const bf16_t *H, float *gradW, float *gradb, uint32_t cap, const uint32_t *count,
                 hipStream_t st);
// dZ4 (BF16) and the summed loss from the FP32 outputs and the samples' direction, BSDF pdf, weight.
void guide_dout(const float *out, const float4 *w_weight, const float4 *wi_b, bf16_t *dZ, float *loss, float scale,
                uint32_t cap, const uint32_t *count, hipStream_t st);
void guide_adam(const GuideDev &g, int step, float lr, float b1, float b2, float eps, hipStream_t st);
void guide_zero(float *p, size_t n, hipStream_t st);
void guide_feats_from_f32(const float *feat, bf16_t *X, uint32_t rows, hipStream_t st);
// Decode + sample (u3: lobe choice, two vMF uniforms) + mixture density at w3 (checks).
void guide_mixture(const float *out64, uint32_t n, const float *u3, const float *w3, const float *bpdf, float *dir,
                   float *P, float *q, float *c, hipStream_t st);

} // namespace ng
