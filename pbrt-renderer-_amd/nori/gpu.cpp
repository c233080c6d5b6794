#include "gpu.h"

#include <cstring>
#include <mutex>

NORI_NAMESPACE_BEGIN

namespace {
struct CtxHolder {
    nori_ctx *ctx = nullptr;
    ~CtxHolder() {
        if (ctx) nori_gpu_destroy(ctx);
    }
};
CtxHolder g_ctx;
std::mutex g_ctx_mutex;
} // namespace

nori_ctx *gpuContext() {
    std::lock_guard<std::mutex> lock(g_ctx_mutex);
    if (!g_ctx.ctx) {
        int rc = nori_gpu_create(&g_ctx.ctx, 0);
        if (rc != NORI_OK) {
            g_ctx.ctx = nullptr;
            throw NoriException("No GPU available for the Nori GPU path (nori_gpu_create returned %i)", rc);
        }
    }
    return g_ctx.ctx;
}

void gpuCheck(int rc, const char *what, nori_ctx *ctx) {
    if (rc == NORI_OK) return;
    const char *msg = ctx ? nori_gpu_last_error(ctx) : "";
    throw NoriException("%s failed (%i): %s", what, rc, msg ? msg : "");
}

nori_material toGPUMaterial(const BSDFGPUData &d) {
    nori_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = (uint32_t)d.type;
    for (int k = 0; k < 3; ++k) m.albedo[k] = d.albedo[k];
    m.intIOR = d.intIOR;
    m.extIOR = d.extIOR;
    m.alpha = d.alpha;
    m.albedoTexIndex = m.normalTexIndex = m.roughnessTexIndex = m.metallicTexIndex = NORI_NO_TEXTURE;
    m.specularTexIndex = m.subsurfaceTexIndex = m.alphaTexIndex = NORI_NO_TEXTURE;
    m.roughness = d.roughness;
    m.metallic = d.metallic;
    m.specular = d.specular;
    m.specularTint = d.specularTint;
    m.sheen = d.sheen;
    m.sheenTint = d.sheenTint;
    m.subsurface = d.subsurface;
    m.clearcoat = d.clearcoat;
    m.clearcoatGloss = d.clearcoatGloss;
    m.anisotropic = d.anisotropic;
    m.betaN = d.betaN;
    return m;
}

Color3f gpuBsdfQuery(int op, const BSDFGPUData &d, const Vector3f &wi, const Vector3f &wo, const float *u4, float h,
                     float &pdf, Vector3f *wo_out) {
    nori_ctx *ctx = gpuContext();
    nori_material m = toGPUMaterial(d);
    float a[3] = {wi[0], wi[1], wi[2]}, b[3] = {wo[0], wo[1], wo[2]}, rgb[3] = {0, 0, 0}, w[3] = {0, 0, 0};
    float hh = h;
    gpuCheck(nori_gpu_bsdf_batch(ctx, op, 1, &m, a, op == NORI_BSDF_SAMPLE ? nullptr : b, u4, &hh, rgb, &pdf, w),
             "nori_gpu_bsdf_batch", ctx);
    if (wo_out) *wo_out = Vector3f(w[0], w[1], w[2]);
    return Color3f(rgb[0], rgb[1], rgb[2]);
}

void expandSample(const Point2f &s, float u[4]) {
    u[0] = s.x();
    u[1] = s.y();
    uint32_t bx, by;
    std::memcpy(&bx, &s.v[0], 4);
    std::memcpy(&by, &s.v[1], 4);
    uint64_t state = ((uint64_t)bx << 32) ^ by ^ 0x853c49e6748fea9bull;
    for (int k = 2; k < 4; ++k) {
        // PCG-XSH-RR step
        uint64_t old = state;
        state = old * 6364136223846793005ull + 1442695040888963407ull;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
        uint32_t r = (xs >> rot) | (xs << ((32u - rot) & 31u));
        u[k] = (float)(r >> 8) * (1.0f / 16777216.0f);
    }
}

NORI_NAMESPACE_END
