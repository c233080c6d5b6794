// Participating-medium plugin interface (include/medium.h:14-66 of the reference).
#pragma once
#include "object.h"

NORI_NAMESPACE_BEGIN

/// A colour property, or a float property broadcast to a colour, or the default.
inline Color3f readScalarOrColor(const PropertyList &propList, const std::string &name, float scalarDefault) {
    try {
        return propList.getColor(name);
    } catch (const NoriException &) {
    }
    try {
        return Color3f(propList.getFloat(name));
    } catch (const NoriException &) {
    }
    return Color3f(scalarDefault);
}

class Medium : public NoriObject {
public:
    virtual ~Medium() {}
    virtual float getSigmaA() const = 0;
    virtual float getSigmaS() const = 0;
    virtual Color3f getSigmaARGB() const { return Color3f(getSigmaA()); }
    virtual Color3f getSigmaSRGB() const { return Color3f(getSigmaS()); }
    virtual float getPhaseG() const = 0;
    virtual bool isHeterogeneous() const { return false; }
    virtual std::string getVolumePath() const { return ""; }
    virtual bool hasExplicitBounds() const = 0;
    virtual BoundingBox3f getExplicitBounds() const = 0;
    EClassType getClassType() const override { return EMedium; }
};

NORI_NAMESPACE_END
