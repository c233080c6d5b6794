// Typed property list handed to plugin constructors (include/proplist.h of the
// reference): set/get per type, a getter with a default, and NoriException on
// a missing property or a type mismatch.
#pragma once
#include "common.h"

NORI_NAMESPACE_BEGIN

class PropertyList {
public:
    PropertyList() {}

#define NORI_PROPERTY_DECL(Type, Name)                                                                                 \
    void set##Name(const std::string &name, const Type &value);                                                       \
    Type get##Name(const std::string &name) const;                                                                     \
    Type get##Name(const std::string &name, const Type &defaultValue) const;
    NORI_PROPERTY_DECL(bool, Boolean)
    NORI_PROPERTY_DECL(int, Integer)
    NORI_PROPERTY_DECL(float, Float)
    NORI_PROPERTY_DECL(std::string, String)
    NORI_PROPERTY_DECL(Color3f, Color)
    NORI_PROPERTY_DECL(Point3f, Point)
    NORI_PROPERTY_DECL(Vector3f, Vector)
    NORI_PROPERTY_DECL(Transform, Transform)
#undef NORI_PROPERTY_DECL

    bool has(const std::string &name) const { return m_properties.count(name) != 0; }

private:
    enum Kind { KBoolean, KInteger, KFloat, KString, KColor, KPoint, KVector, KTransform };
    struct Property {
        Kind kind = KBoolean;
        bool b = false;
        int i = 0;
        float f = 0.0f;
        std::string s;
        Color3f c;
        Vector3f v; // point or vector
        Transform t;
    };
    static const char *kindName(Kind k);
    Property &put(const std::string &name, Kind k);
    const Property *find(const std::string &name, Kind k, bool required) const;
    std::map<std::string, Property> m_properties;
};

NORI_NAMESPACE_END
