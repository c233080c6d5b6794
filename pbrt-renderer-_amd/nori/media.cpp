// "homogeneous" and "heterogeneous" media with the reference's properties
// (src/cpu/homogeneous.cpp, src/cpu/heterogeneous.cpp).
#include "medium.h"

NORI_NAMESPACE_BEGIN

class HomogeneousMedium : public Medium {
public:
    explicit HomogeneousMedium(const PropertyList &p) {
        m_sigmaA = readScalarOrColor(p, "sigmaA", 0.0f);
        m_sigmaS = readScalarOrColor(p, "sigmaS", 0.3f);
        m_g = p.getFloat("g", 0.0f);
        Point3f bmin = p.getPoint("boundsMin", Point3f(1e18f)), bmax = p.getPoint("boundsMax", Point3f(-1e18f));
        m_explicit = bmin.x() < 1e17f && bmax.x() > -1e17f;
        if (m_explicit) m_bounds = BoundingBox3f(bmin, bmax);
    }
    float getSigmaA() const override { return m_sigmaA.r(); }
    float getSigmaS() const override { return m_sigmaS.r(); }
    Color3f getSigmaARGB() const override { return m_sigmaA; }
    Color3f getSigmaSRGB() const override { return m_sigmaS; }
    float getPhaseG() const override { return m_g; }
    bool hasExplicitBounds() const override { return m_explicit; }
    BoundingBox3f getExplicitBounds() const override { return m_bounds; }
    std::string toString() const override {
        return format("HomogeneousMedium[\n  sigmaA = %s,\n  sigmaS = %s,\n  g = %f,\n  explicitBounds = %s\n]",
                      m_sigmaA.toString(), m_sigmaS.toString(), m_g, m_explicit ? "yes" : "no (scene bbox)");
    }

private:
    Color3f m_sigmaA, m_sigmaS;
    float m_g;
    bool m_explicit = false;
    BoundingBox3f m_bounds;
};

class HeterogeneousMedium : public Medium {
public:
    explicit HeterogeneousMedium(const PropertyList &p) {
        m_sigmaA = readScalarOrColor(p, "sigmaA", 0.0f);
        m_sigmaS = readScalarOrColor(p, "sigmaS", 5.0f);
        m_g = p.getFloat("g", 0.85f);
        m_path = p.getString("volume", "");
        m_bounds = BoundingBox3f(p.getPoint("boundsMin"), p.getPoint("boundsMax"));
    }
    float getSigmaA() const override { return m_sigmaA.r(); }
    float getSigmaS() const override { return m_sigmaS.r(); }
    Color3f getSigmaARGB() const override { return m_sigmaA; }
    Color3f getSigmaSRGB() const override { return m_sigmaS; }
    float getPhaseG() const override { return m_g; }
    bool isHeterogeneous() const override { return true; }
    bool hasExplicitBounds() const override { return true; }
    BoundingBox3f getExplicitBounds() const override { return m_bounds; }
    std::string getVolumePath() const override { return m_path; }
    std::string toString() const override {
        return format("HeterogeneousMedium[\n  sigmaA = %s,\n  sigmaS = %s,\n  g = %f,\n  volume = \"%s\"\n]",
                      m_sigmaA.toString(), m_sigmaS.toString(), m_g, m_path.empty() ? "(procedural)" : m_path);
    }

private:
    Color3f m_sigmaA, m_sigmaS;
    float m_g;
    std::string m_path;
    BoundingBox3f m_bounds;
};

NORI_REGISTER_CLASS(HomogeneousMedium, "homogeneous");
NORI_REGISTER_CLASS(HeterogeneousMedium, "heterogeneous");

NORI_NAMESPACE_END
