// "area" emitter (src/cpu/area.cpp:5-42 semantics; the point sampling uses the
// GPU path's barycentric mapping, see Mesh::samplePosition).
#include "emitter.h"

NORI_NAMESPACE_BEGIN

class AreaLight : public Emitter {
public:
    explicit AreaLight(const PropertyList &p) { m_radiance = p.getColor("radiance"); }
    Color3f getRadiance() const override { return m_radiance; }
    Color3f sample(const Point2f &s, Point3f &p, Normal3f &n, float &pdf) const override {
        if (!m_mesh) throw NoriException("AreaLight: not attached to a mesh");
        m_mesh->samplePosition(s, p, n, pdf);
        return m_radiance;
    }
    float pdf() const override {
        if (!m_mesh) throw NoriException("AreaLight: not attached to a mesh");
        return 1.0f / m_mesh->getDiscretePDF().getSum();
    }
    std::string toString() const override { return format("AreaLight[\n  radiance = %s\n]", m_radiance.toString()); }

private:
    Color3f m_radiance;
};

NORI_REGISTER_CLASS(AreaLight, "area");

NORI_NAMESPACE_END
