// Integrator plugin interface (include/integrator.h:14-47 of the reference)
// plus the batch render hook the GPU path adds (Nori itself has no render loop
// to hook: include/block.h is unimplemented in the reference).
#pragma once
#include "object.h"

NORI_NAMESPACE_BEGIN

class Scene;
class Sampler;

class Integrator : public NoriObject {
public:
    virtual ~Integrator() {}
    /// Called once the scene is loaded, before rendering.
    virtual void preprocess(const Scene *scene) { (void)scene; }
    /// Radiance arriving along `ray`.
    virtual Color3f Li(const Scene *scene, Sampler *sampler, const Ray3f &ray) const = 0;
    /// Li for n rays at once (the default calls Li per ray).
    virtual void LiBatch(const Scene *scene, Sampler *sampler, const Ray3f *rays, size_t n, Color3f *out) const;
    /// Renders the scene's camera view at the sampler's sample count into
    /// `rgb` (width x height x 3, row 0 at the top).  The default throws.
    virtual void render(const Scene *scene, std::vector<float> &rgb) const;
    EClassType getClassType() const override { return EIntegrator; }
};

NORI_NAMESPACE_END
