// CPU oracle — test infrastructure only.
//
// Neural path guiding has no counterpart in the reference (its path integrator,
// shaders/Shaders.hlsl:473-511, samples only the BSDF).  This is the FP64 oracle
// of the guiding model the GPU trains:
//
//   features  64: per axis of the scene-normalised position 8 octaves of
//             (sin, cos)(2^k pi p) (48), shading normal (3), direction to the
//             viewer (3), roughness, metallic, 8 zeros
//   MLP       64 -> 64 -> 64 -> 64 (ReLU) -> 64 outputs, of which 41 are used
//   outputs   8 von Mises-Fisher lobes: mean = normalize(o[3j..3j+2]),
//             kappa = exp(clamp(o[24+j], -4, 9)), weight = softmax(o[32..39]);
//             selection probability c = 0.1 + 0.5 sigmoid(o[40])
//   density   P(w) = c q(w) + (1 - c) b(w), q the vMF mixture, b the BSDF pdf
//             (one-sample MIS between the learned mixture and the BSDF)
//   loss      -W log P(w) per training sample (W = throughput-normalised
//             radiance of the sampled direction): the sample estimate of the
//             gradient of KL(target || P) (Mueller et al., "Neural Importance
//             Sampling", ACM TOG 2019)
//   Adam      Kingma & Ba 2015, bias-corrected.
//
// Parameters: four layers, each W[64 out][64 in] then b[64], FP32 masters.
// The GPU computes with BF16 weights and activations; `bf16` emulates that
// rounding (round to nearest even) for forward-pass comparisons.
#pragma once
#include <cstdint>

namespace orc {

constexpr int G_IN = 64, G_HID = 64, G_OUT = 64, G_LAYERS = 4, G_LOBES = 8;
constexpr int G_PARAMS = G_LAYERS * (64 * 64 + 64);

float bf16_round(float x);

void guide_features(const float pos01[3], const float n[3], const float wi[3], float rough, float metal, float *f64);
void guide_forward(const float *params, const float *feat, int bf16, double *out64, double *acts /*4x64 or null*/);
// Density of the decoded outputs at direction w (world), with BSDF pdf b.
double guide_pdf(const double *out, const double w[3], double b, double *q_out, double *c_out);
// Loss and gradient w.r.t. the 64 outputs for one sample.
double guide_loss_dout(const double *out, const double w[3], double b, double weight, double *dout64);
// Full loss and parameter gradient summed over a batch (FP64, no BF16).
double guide_loss_grad(const float *params, const float *feat, const float *w3, const float *b, const float *weight,
                       int n, double *grad);
void guide_adam(float *params, float *m, float *v, const double *grad, int step, float lr, float beta1, float beta2,
                float eps);
// vMF lobe j sample from uniforms (u1, u2): Wenzel Jakob, "Numerically stable
// sampling of the von Mises Fisher distribution on S^2", 2012.
void vmf_sample(const double mu[3], double kappa, double u1, double u2, double w[3]);

} // namespace orc
