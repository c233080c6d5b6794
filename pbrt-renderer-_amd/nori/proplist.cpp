// PropertyList accessors: semantics of src/cpu/proplist.cpp (a repeated name
// warns and overwrites; a missing or mistyped property throws).
#include "proplist.h"

NORI_NAMESPACE_BEGIN

const char *PropertyList::kindName(Kind k) {
    switch (k) {
    case KBoolean: return "boolean";
    case KInteger: return "integer";
    case KFloat: return "float";
    case KString: return "string";
    case KColor: return "color";
    case KPoint: return "point";
    case KVector: return "vector";
    default: return "transform";
    }
}

PropertyList::Property &PropertyList::put(const std::string &name, Kind k) {
    if (m_properties.find(name) != m_properties.end())
        cerr << "Property \"" << name << "\" was specified multiple times!" << endl;
    Property &p = m_properties[name];
    p = Property();
    p.kind = k;
    return p;
}

const PropertyList::Property *PropertyList::find(const std::string &name, Kind k, bool required) const {
    auto it = m_properties.find(name);
    if (it == m_properties.end()) {
        if (required) throw NoriException("Property '%s' is missing!", name);
        return nullptr;
    }
    if (it->second.kind != k)
        throw NoriException("Property '%s' has the wrong type! (expected <%s>)!", name, kindName(k));
    return &it->second;
}

#define NORI_PROPERTY_DEF(Type, Name, K, field)                                                                        \
    void PropertyList::set##Name(const std::string &name, const Type &value) { put(name, K).field = value; }          \
    Type PropertyList::get##Name(const std::string &name) const { return find(name, K, true)->field; }                 \
    Type PropertyList::get##Name(const std::string &name, const Type &defaultValue) const {                            \
        const Property *p = find(name, K, false);                                                                      \
        return p ? p->field : defaultValue;                                                                            \
    }
NORI_PROPERTY_DEF(bool, Boolean, KBoolean, b)
NORI_PROPERTY_DEF(int, Integer, KInteger, i)
NORI_PROPERTY_DEF(float, Float, KFloat, f)
NORI_PROPERTY_DEF(std::string, String, KString, s)
NORI_PROPERTY_DEF(Color3f, Color, KColor, c)
NORI_PROPERTY_DEF(Point3f, Point, KPoint, v)
NORI_PROPERTY_DEF(Vector3f, Vector, KVector, v)
NORI_PROPERTY_DEF(Transform, Transform, KTransform, t)
#undef NORI_PROPERTY_DEF

NORI_NAMESPACE_END
