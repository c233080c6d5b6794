// NoriObject base class, the object factory and NORI_REGISTER_CLASS: the plugin
// surface of include/object.h:14-149 (same names, class types and semantics).
#pragma once
#include "proplist.h"

NORI_NAMESPACE_BEGIN

class NoriObject {
public:
    enum EClassType {
        EScene = 0,
        EMesh,
        EBSDF,
        EPhaseFunction,
        EEmitter,
        EMedium,
        ECamera,
        EIntegrator,
        ESampler,
        ETest,
        EReconstructionFilter,
        EClassTypeCount
    };

    virtual ~NoriObject() {}
    virtual EClassType getClassType() const = 0;
    /// Add a child; the default accepts none and throws.
    virtual void addChild(NoriObject *child);
    /// Called once the object has been added to its parent; default: nothing.
    virtual void setParent(NoriObject *parent);
    /// Called by the parser after construction and addChild of all children.
    virtual void activate();
    virtual std::string toString() const = 0;

    static std::string classTypeName(EClassType type);
};

class NoriObjectFactory {
public:
    typedef std::function<NoriObject *(const PropertyList &)> Constructor;
    static void registerClass(const std::string &name, const Constructor &constr);
    /// Throws NoriException when no class of that name was registered.
    static NoriObject *createInstance(const std::string &name, const PropertyList &propList);
    static std::vector<std::string> registeredNames();

private:
    static std::map<std::string, Constructor> &constructors();
};

/// Registers `cls` under the XML type name `name` at static-initialisation time.
#define NORI_REGISTER_CLASS(cls, name)                                                                                 \
    cls *cls##_create(const PropertyList &list) { return new cls(list); }                                             \
    static struct cls##_ {                                                                                             \
        cls##_() { NoriObjectFactory::registerClass(name, cls##_create); }                                             \
    } cls##__NORI_;

NORI_NAMESPACE_END
