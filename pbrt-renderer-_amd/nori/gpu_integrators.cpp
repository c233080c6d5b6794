// Integrators registered for the XML scene format, all running on the GPU:
//   "path_mis"     NEE + BSDF sampling with balance-heuristic MIS
//                  (shaders/Shaders.hlsl:25-651); maxDepth (default 32, the
//                  reference's MAX_BOUNCES, shaders/Common.hlsli:151-153),
//                  rrDepth (default 3), seed (default 0);
//   "path_guided"  the same with online neural path guiding (one-sample MIS
//                  between the learned vMF mixture and the BSDF);
//   "direct"       the type the reference's scene files name for its direct-
//                  lighting tests; the reference's renderer path traces every
//                  scene with the same RayGen whatever the integrator type, so
//                  "direct" is path_mis with the same defaults (maxDepth 1
//                  keeps the first vertex's emission and next-event estimate).
// preprocess() uploads the scene; Li() traces one ray through nori_gpu_li;
// render() renders the camera view at the sampler's sample count through
// nori_gpu_render.
#include "camera.h"
#include "gpu.h"
#include "integrator.h"
#include "sampler.h"
#include "scene.h"

#include <cstring>

NORI_NAMESPACE_BEGIN

void Integrator::LiBatch(const Scene *scene, Sampler *sampler, const Ray3f *rays, size_t n, Color3f *out) const {
    for (size_t i = 0; i < n; ++i) out[i] = Li(scene, sampler, rays[i]);
}

void Integrator::render(const Scene *, std::vector<float> &) const {
    throw NoriException("Integrator \"%s\" has no batch render", toString());
}

namespace {

class GpuPathIntegrator : public Integrator {
public:
    GpuPathIntegrator(const PropertyList &p, const char *name, bool guided, int defaultDepth)
        : m_name(name), m_guided(guided) {
        m_maxDepth = p.getInteger("maxDepth", defaultDepth);
        m_rrDepth = p.getInteger("rrDepth", 3);
        m_seed = p.getInteger("seed", 0);
        if (m_maxDepth <= 0) throw NoriException("%s: maxDepth must be positive", name);
        if (m_rrDepth < 0) throw NoriException("%s: rrDepth must not be negative", name);
    }

    void preprocess(const Scene *scene) override { scene->getAccel()->context(); /* upload */ }

    Color3f Li(const Scene *scene, Sampler *sampler, const Ray3f &ray) const override {
        Color3f out;
        LiBatch(scene, sampler, &ray, 1, &out);
        return out;
    }

    // One GPU path per ray; ray i's pcg32 stream is seeded from two sampler draws.
    void LiBatch(const Scene *scene, Sampler *sampler, const Ray3f *rays, size_t n, Color3f *out) const override {
        if (n == 0) return;
        nori_ctx *ctx = scene->getAccel()->context();
        std::vector<nori_ray> rr(n);
        std::vector<uint64_t> seeds(n);
        for (size_t i = 0; i < n; ++i) {
            const Ray3f &ray = rays[i];
            float a = sampler->next1D(), b = sampler->next1D();
            seeds[i] = ((uint64_t)(a * 4294967296.0) << 32) ^ (uint64_t)(b * 4294967296.0);
            rr[i] = nori_ray{{ray.o[0], ray.o[1], ray.o[2]}, ray.mint, {ray.d[0], ray.d[1], ray.d[2]}, ray.maxt};
        }
        std::vector<float> rgb(3 * n);
        gpuCheck(nori_gpu_li(ctx, rr.data(), seeds.data(), (int)n, (uint32_t)m_maxDepth, (uint32_t)m_rrDepth,
                             rgb.data()),
                 "nori_gpu_li", ctx);
        for (size_t i = 0; i < n; ++i) out[i] = Color3f(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    }

    void render(const Scene *scene, std::vector<float> &rgb) const override {
        nori_ctx *ctx = scene->getAccel()->context();
        const Camera *cam = scene->getCamera();
        nori_view v;
        std::memset(&v, 0, sizeof(v));
        v.camera = cam->getGPUCamera();
        v.spp = (uint32_t)scene->getSampler()->getSampleCount();
        v.sample_offset = 0;
        v.max_depth = (uint32_t)m_maxDepth;
        v.rr_depth = (uint32_t)m_rrDepth;
        gpuCheck(nori_gpu_set_option(ctx, "guiding", m_guided ? 1 : 0), "nori_gpu_set_option", ctx);
        gpuCheck(nori_gpu_reset_film(ctx), "nori_gpu_reset_film", ctx);
        const size_t npix = (size_t)v.camera.width * v.camera.height;
        std::vector<float> acc(npix * 4);
        gpuCheck(nori_gpu_render(ctx, &v, 1, (uint64_t)m_seed, acc.data()), "nori_gpu_render", ctx);
        rgb.resize(npix * 3);
        for (size_t i = 0; i < npix; ++i) {
            float w = acc[4 * i + 3] > 0.0f ? 1.0f / acc[4 * i + 3] : 0.0f;
            for (int c = 0; c < 3; ++c) rgb[3 * i + c] = acc[4 * i + c] * w;
        }
    }

    std::string toString() const override {
        return format("%s[maxDepth=%i, rrDepth=%i, seed=%i]", m_name, m_maxDepth, m_rrDepth, m_seed);
    }

private:
    const char *m_name;
    bool m_guided;
    int m_maxDepth, m_rrDepth, m_seed;
};

} // namespace

class PathMisIntegrator : public GpuPathIntegrator {
public:
    explicit PathMisIntegrator(const PropertyList &p) : GpuPathIntegrator(p, "path_mis", false, 32) {}
};
class PathGuidedIntegrator : public GpuPathIntegrator {
public:
    explicit PathGuidedIntegrator(const PropertyList &p) : GpuPathIntegrator(p, "path_guided", true, 32) {}
};
class DirectIntegrator : public GpuPathIntegrator {
public:
    explicit DirectIntegrator(const PropertyList &p) : GpuPathIntegrator(p, "direct", false, 32) {}
};

NORI_REGISTER_CLASS(PathMisIntegrator, "path_mis");
NORI_REGISTER_CLASS(PathGuidedIntegrator, "path_guided");
NORI_REGISTER_CLASS(DirectIntegrator, "direct");

NORI_NAMESPACE_END
