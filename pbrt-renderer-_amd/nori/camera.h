// Camera and reconstruction-filter plugin interfaces (include/camera.h,
// include/rfilter.h of the reference).
#pragma once
#include "../../include/nori_gpu.h"
#include "object.h"

NORI_NAMESPACE_BEGIN

class ReconstructionFilter : public NoriObject {
public:
    const Vector2f &getRadius() const { return m_radius; }
    virtual float eval(float x) const = 0;
    EClassType getClassType() const override { return EReconstructionFilter; }

protected:
    Vector2f m_radius;
};

class Camera : public NoriObject {
public:
    /// Ray through an image-plane position (in pixels) and an aperture sample;
    /// returns the importance weight (1 for the perspective camera).
    virtual Color3f sampleRay(Ray3f &ray, const Point2f &samplePosition, const Point2f &apertureSample) const = 0;
    const Vector2i &getOutputSize() const { return m_outputSize; }
    const ReconstructionFilter *getReconstructionFilter() const { return m_rfilter; }
    virtual float getLensRadius() const { return 0.0f; }
    virtual float getFocalDistance() const { return 1.0f; }
    /// The camera as the GPU path's raygen consumes it.
    virtual nori_camera getGPUCamera() const = 0;
    EClassType getClassType() const override { return ECamera; }

protected:
    Vector2i m_outputSize;
    ReconstructionFilter *m_rfilter = nullptr;
};

NORI_NAMESPACE_END
