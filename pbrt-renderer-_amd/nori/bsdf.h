// BSDF plugin interface (include/bsdf.h:20-165 of the reference): query record,
// the GPU parameter record every BSDF packs itself into, and the abstract class.
#pragma once
#include "object.h"

NORI_NAMESPACE_BEGIN

enum EMeasure { EUnknownMeasure = 0, ESolidAngle, EDiscrete };

class AlphaTexture; // the CPU integrators' alpha masks; the GPU path alpha-tests from the material's texture maps

struct BSDFQueryRecord {
    Vector3f wi, wo;
    float eta = 1.0f;
    EMeasure measure = EUnknownMeasure;
    Point2f uv;
    /// Hair: offset across the fiber in [-1, 1] (the GPU path takes it from the hit's UV).
    float h = 0.0f;
    explicit BSDFQueryRecord(const Vector3f &wi) : wi(wi) {}
    BSDFQueryRecord(const Vector3f &wi, const Vector3f &wo, EMeasure measure) : wi(wi), wo(wo), measure(measure) {}
};

/// Parameters of one material as the GPU path consumes them (include/bsdf.h:46-89).
struct BSDFGPUData {
    enum Type { DIFFUSE = 0, MIRROR = 1, DIELECTRIC = 2, MICROFACET = 3, DISNEY = 4, HAIR = 5, SUBSURFACE = 6 };
    int type = DIFFUSE;
    float albedo[3] = {0.5f, 0.5f, 0.5f};
    float intIOR = 1.5046f;
    float extIOR = 1.000277f;
    float alpha = 0.1f;
    std::string albedoTexture, normalTexture, roughnessTexture, metallicTexture, specularTexture, subsurfaceTexture,
        alphaTexture;
    float roughness = 0.5f;
    float metallic = 0.0f;
    float specular = 0.5f;
    float specularTint = 0.0f;
    float sheen = 0.0f;
    float sheenTint = 0.5f;
    float subsurface = 0.0f;
    float clearcoat = 0.0f;
    float clearcoatGloss = 1.0f;
    float anisotropic = 0.0f;
    float betaN = 0.3f;
};

class BSDF : public NoriObject {
public:
    /// Samples wo for bRec.wi; returns f * |cos| / pdf (0 on failure).
    virtual Color3f sample(BSDFQueryRecord &bRec, const Point2f &sample) const = 0;
    virtual Color3f eval(const BSDFQueryRecord &bRec) const = 0;
    virtual float pdf(const BSDFQueryRecord &bRec) const = 0;
    EClassType getClassType() const override { return EBSDF; }
    virtual bool isDiffuse() const { return false; }
    virtual BSDFGPUData getGPUData() const { return BSDFGPUData{}; }
    virtual const AlphaTexture *getAlphaTexture() const { return nullptr; }
};

NORI_NAMESPACE_END
