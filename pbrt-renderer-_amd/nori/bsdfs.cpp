// BSDF plugins with the XML properties and GPU parameter packing of the
// reference (src/cpu/diffuse.cpp, mirror.cpp, dielectric.cpp, microfacet.cpp,
// disney.cpp, hair.cpp:326-436, subsurface.cpp).
//
// sample/eval/pdf of the non-delta BSDFs run on the GPU through
// nori_gpu_bsdf_batch, so a plugin query and the renderer's shading are the
// same code; mirror and dielectric are closed forms (src/cpu/mirror.cpp:30-46,
// src/cpu/dielectric.cpp:43-103).
#include "bsdf.h"
#include "gpu.h"

NORI_NAMESPACE_BEGIN

namespace {

float fresnel(float cosThetaI, float extIOR, float intIOR) {
    // src/cpu/common.cpp:259-288
    float etaI = extIOR, etaT = intIOR;
    if (extIOR == intIOR) return 0.0f;
    if (cosThetaI < 0.0f) {
        std::swap(etaI, etaT);
        cosThetaI = -cosThetaI;
    }
    float eta = etaI / etaT, sinThetaTSqr = eta * eta * (1 - cosThetaI * cosThetaI);
    if (sinThetaTSqr > 1.0f) return 1.0f;
    float cosThetaT = std::sqrt(1.0f - sinThetaTSqr);
    float Rs = (etaI * cosThetaI - etaT * cosThetaT) / (etaI * cosThetaI + etaT * cosThetaT);
    float Rp = (etaT * cosThetaI - etaI * cosThetaT) / (etaT * cosThetaI + etaI * cosThetaT);
    return (Rs * Rs + Rp * Rp) / 2.0f;
}

/// A non-delta BSDF whose queries run on the GPU.
class GpuBSDF : public BSDF {
public:
    Color3f sample(BSDFQueryRecord &bRec, const Point2f &s) const override {
        float u[4], pdf = 0.0f;
        expandSample(s, u);
        Vector3f wo;
        Color3f w = gpuBsdfQuery(NORI_BSDF_SAMPLE, getGPUData(), bRec.wi, Vector3f(), u, bRec.h, pdf, &wo);
        bRec.wo = wo;
        bRec.eta = 1.0f;
        bRec.measure = ESolidAngle;
        return pdf > 0.0f ? w : Color3f(0.0f);
    }
    Color3f eval(const BSDFQueryRecord &bRec) const override {
        if (bRec.measure != ESolidAngle) return Color3f(0.0f);
        float pdf;
        return gpuBsdfQuery(NORI_BSDF_EVAL, getGPUData(), bRec.wi, bRec.wo, nullptr, bRec.h, pdf, nullptr);
    }
    float pdf(const BSDFQueryRecord &bRec) const override {
        if (bRec.measure != ESolidAngle) return 0.0f;
        float pdf = 0.0f;
        gpuBsdfQuery(NORI_BSDF_PDF, getGPUData(), bRec.wi, bRec.wo, nullptr, bRec.h, pdf, nullptr);
        return pdf;
    }
};

void readTextures(const PropertyList &p, BSDFGPUData &d) {
    d.albedoTexture = p.getString("albedoTexture", "");
    d.normalTexture = p.getString("normalTexture", "");
    d.roughnessTexture = p.getString("roughnessTexture", "");
    d.metallicTexture = p.getString("metallicTexture", "");
    d.alphaTexture = p.getString("alphaTexture", "");
}

} // namespace

class Diffuse : public GpuBSDF {
public:
    explicit Diffuse(const PropertyList &p) {
        Color3f a = p.getColor("albedo", Color3f(0.5f));
        for (int k = 0; k < 3; ++k) m_d.albedo[k] = a[k];
        m_d.type = BSDFGPUData::DIFFUSE;
        readTextures(p, m_d);
    }
    bool isDiffuse() const override { return true; }
    BSDFGPUData getGPUData() const override { return m_d; }
    std::string toString() const override {
        return format("Diffuse[\n  albedo = %s\n]", Color3f(m_d.albedo[0], m_d.albedo[1], m_d.albedo[2]).toString());
    }

private:
    BSDFGPUData m_d;
};

class Mirror : public BSDF {
public:
    explicit Mirror(const PropertyList &) {}
    Color3f eval(const BSDFQueryRecord &) const override { return Color3f(0.0f); }
    float pdf(const BSDFQueryRecord &) const override { return 0.0f; }
    Color3f sample(BSDFQueryRecord &bRec, const Point2f &) const override {
        if (bRec.wi.z() <= 0) return Color3f(0.0f);
        bRec.wo = Vector3f(-bRec.wi.x(), -bRec.wi.y(), bRec.wi.z());
        bRec.measure = EDiscrete;
        bRec.eta = 1.0f;
        return Color3f(1.0f);
    }
    BSDFGPUData getGPUData() const override {
        BSDFGPUData d;
        d.type = BSDFGPUData::MIRROR;
        return d;
    }
    std::string toString() const override { return "Mirror[]"; }
};

class Dielectric : public BSDF {
public:
    explicit Dielectric(const PropertyList &p) {
        m_intIOR = p.getFloat("intIOR", 1.5046f);
        m_extIOR = p.getFloat("extIOR", 1.000277f);
        m_tint = p.getColor("tintColor", Color3f(1.0f));
    }
    Color3f eval(const BSDFQueryRecord &) const override { return Color3f(0.0f); }
    float pdf(const BSDFQueryRecord &) const override { return 0.0f; }
    Color3f sample(BSDFQueryRecord &bRec, const Point2f &s) const override {
        float cosI = bRec.wi.z(), Fr = fresnel(cosI, m_extIOR, m_intIOR);
        bRec.measure = EDiscrete;
        if (s.x() < Fr) {
            bRec.wo = Vector3f(-bRec.wi.x(), -bRec.wi.y(), bRec.wi.z());
            bRec.eta = 1.0f;
            return Color3f(1.0f);
        }
        float etaI = m_extIOR, etaT = m_intIOR;
        Vector3f n(0, 0, 1);
        if (cosI < 0) {
            std::swap(etaI, etaT);
            n = Vector3f(0, 0, -1);
            cosI = -cosI;
        }
        float eta = etaI / etaT, sin2T = eta * eta * (1 - cosI * cosI);
        if (sin2T > 1.0f) { // total internal reflection by rounding
            bRec.wo = Vector3f(-bRec.wi.x(), -bRec.wi.y(), bRec.wi.z());
            bRec.eta = 1.0f;
            return Color3f(1.0f);
        }
        float cosT = std::sqrt(1.0f - sin2T);
        bRec.wo = (-bRec.wi) * eta + n * (eta * cosI - cosT);
        bRec.eta = eta;
        return Color3f(1.0f);
    }
    BSDFGPUData getGPUData() const override {
        BSDFGPUData d;
        d.type = BSDFGPUData::DIELECTRIC;
        d.albedo[0] = m_tint.r(); d.albedo[1] = m_tint.g(); d.albedo[2] = m_tint.b();
        d.intIOR = m_intIOR;
        d.extIOR = m_extIOR;
        return d;
    }
    std::string toString() const override {
        return format("Dielectric[\n  intIOR = %f,\n  extIOR = %f\n]", m_intIOR, m_extIOR);
    }

private:
    float m_intIOR, m_extIOR;
    Color3f m_tint;
};

class Microfacet : public GpuBSDF {
public:
    explicit Microfacet(const PropertyList &p) {
        m_d.type = BSDFGPUData::MICROFACET;
        m_d.alpha = p.getFloat("alpha", 0.1f);
        m_d.intIOR = p.getFloat("intIOR", 1.5046f);
        m_d.extIOR = p.getFloat("extIOR", 1.000277f);
        Color3f kd = p.getColor("kd", Color3f(0.5f));
        for (int k = 0; k < 3; ++k) m_d.albedo[k] = kd[k];
        readTextures(p, m_d);
    }
    BSDFGPUData getGPUData() const override { return m_d; }
    std::string toString() const override {
        return format("Microfacet[\n  alpha = %f,\n  intIOR = %f,\n  extIOR = %f,\n  kd = %s\n]", m_d.alpha, m_d.intIOR,
                      m_d.extIOR, Color3f(m_d.albedo[0], m_d.albedo[1], m_d.albedo[2]).toString());
    }

private:
    BSDFGPUData m_d;
};

class Disney : public GpuBSDF {
public:
    explicit Disney(const PropertyList &p) {
        m_d.type = BSDFGPUData::DISNEY;
        // baseColor, else albedo (src/cpu/disney.cpp:29-32)
        Color3f c = p.getColor("baseColor", Color3f(-1.0f)).r() >= 0.0f ? p.getColor("baseColor")
                                                                           : p.getColor("albedo", Color3f(0.8f));
        for (int k = 0; k < 3; ++k) m_d.albedo[k] = c[k];
        m_d.roughness = p.getFloat("roughness", 0.5f);
        m_d.metallic = p.getFloat("metallic", 0.0f);
        m_d.specular = p.getFloat("specular", 0.5f);
        m_d.specularTint = p.getFloat("specularTint", 0.0f);
        m_d.sheen = p.getFloat("sheen", 0.0f);
        m_d.sheenTint = p.getFloat("sheenTint", 0.5f);
        m_d.subsurface = p.getFloat("subsurface", 0.0f);
        m_d.clearcoat = p.getFloat("clearcoat", 0.0f);
        m_d.clearcoatGloss = p.getFloat("clearcoatGloss", 1.0f);
        m_d.anisotropic = p.getFloat("anisotropic", 0.0f);
        readTextures(p, m_d);
        m_d.specularTexture = p.getString("specularTexture", "");
        m_d.subsurfaceTexture = p.getString("subsurfaceTexture", "");
    }
    BSDFGPUData getGPUData() const override { return m_d; }
    std::string toString() const override {
        return format("Disney[\n  baseColor = %s,\n  roughness = %f,\n  metallic = %f,\n  specular = %f\n]",
                      Color3f(m_d.albedo[0], m_d.albedo[1], m_d.albedo[2]).toString(), m_d.roughness, m_d.metallic,
                      m_d.specular);
    }

private:
    BSDFGPUData m_d;
};

class HairBSDF : public GpuBSDF {
public:
    explicit HairBSDF(const PropertyList &p) {
        m_d.type = BSDFGPUData::HAIR;
        Color3f c = p.getColor("color", Color3f(0.28f, 0.15f, 0.06f));
        for (int k = 0; k < 3; ++k) m_d.albedo[k] = c[k];
        m_d.roughness = p.getFloat("beta_m", 0.25f);
        m_d.betaN = p.getFloat("beta_n", 0.3f);
        m_d.alpha = p.getFloat("alpha", 2.0f);
        m_d.intIOR = p.getFloat("eta", 1.55f);
        m_d.alphaTexture = p.getString("alphaTexture", "");
    }
    BSDFGPUData getGPUData() const override { return m_d; }
    std::string toString() const override {
        return format("HairBSDF[\n  color = %s,\n  beta_m = %f,\n  beta_n = %f,\n  alpha = %f,\n  eta = %f\n]",
                      Color3f(m_d.albedo[0], m_d.albedo[1], m_d.albedo[2]).toString(), m_d.roughness, m_d.betaN,
                      m_d.alpha, m_d.intIOR);
    }

private:
    BSDFGPUData m_d;
};

/// "subsurface": the random-walk subsurface material (src/cpu/subsurface.cpp:
/// 1-92; the GPU record reuses fields as the reference does: albedo -> scatter
/// colour A, subsurface -> mean free path, anisotropic -> HG g, sheen /
/// sheenTint / clearcoat -> tint RGB).  eval/pdf give the boundary's Beckmann
/// reflection lobe (the GPU BSDF query); the interior walk is part of the
/// path tracer, so sample() returns no direction.
class Subsurface : public GpuBSDF {
public:
    explicit Subsurface(const PropertyList &p) {
        Color3f a = p.getColor("albedo", Color3f(0.8f, 0.5f, 0.4f));
        for (int k = 0; k < 3; ++k) m_d.albedo[k] = a[k];
        m_d.subsurface = p.getFloat("radius", 0.1f);
        m_d.intIOR = p.getFloat("intIOR", 1.33f);
        m_d.extIOR = p.getFloat("extIOR", 1.000277f);
        m_d.anisotropic = p.getFloat("g", 0.0f);
        m_d.roughness = p.getFloat("roughness", 0.3f);
        Color3f tint = p.getColor("tint", Color3f(1.0f, 1.0f, 1.0f));
        m_d.sheen = tint[0];
        m_d.sheenTint = tint[1];
        m_d.clearcoat = tint[2];
        m_d.specular = p.getFloat("specular", 0.0f);
        readTextures(p, m_d);
        m_d.type = BSDFGPUData::SUBSURFACE;
    }
    bool isDiffuse() const override { return true; }
    BSDFGPUData getGPUData() const override { return m_d; }
    std::string toString() const override {
        return format("Subsurface[\n  albedo = %s,\n  radius = %f,\n  g = %f\n]",
                      Color3f(m_d.albedo[0], m_d.albedo[1], m_d.albedo[2]).toString(), m_d.subsurface,
                      m_d.anisotropic);
    }

private:
    BSDFGPUData m_d;
};

NORI_REGISTER_CLASS(Diffuse, "diffuse");
NORI_REGISTER_CLASS(Mirror, "mirror");
NORI_REGISTER_CLASS(Dielectric, "dielectric");
NORI_REGISTER_CLASS(Microfacet, "microfacet");
NORI_REGISTER_CLASS(Disney, "disney");
NORI_REGISTER_CLASS(HairBSDF, "hair_bsdf");
NORI_REGISTER_CLASS(Subsurface, "subsurface");

NORI_NAMESPACE_END
