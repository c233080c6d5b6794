// Default NoriObject behaviour and the class registry (src/cpu/object.cpp).
#include "object.h"

NORI_NAMESPACE_BEGIN

void NoriObject::addChild(NoriObject *child) {
    throw NoriException("NoriObject::addChild() is not implemented for objects of type '%s'!",
                        classTypeName(getClassType()));
    (void)child;
}

void NoriObject::setParent(NoriObject *) {}

void NoriObject::activate() {}

std::string NoriObject::classTypeName(EClassType type) {
    switch (type) {
    case EScene: return "scene";
    case EMesh: return "mesh";
    case EBSDF: return "bsdf";
    case EPhaseFunction: return "phase";
    case EEmitter: return "emitter";
    case EMedium: return "medium";
    case ECamera: return "camera";
    case EIntegrator: return "integrator";
    case ESampler: return "sampler";
    case ETest: return "test";
    case EReconstructionFilter: return "rfilter";
    default: return "<unknown>";
    }
}

std::map<std::string, NoriObjectFactory::Constructor> &NoriObjectFactory::constructors() {
    static std::map<std::string, Constructor> table;
    return table;
}

void NoriObjectFactory::registerClass(const std::string &name, const Constructor &constr) {
    constructors()[name] = constr;
}

NoriObject *NoriObjectFactory::createInstance(const std::string &name, const PropertyList &propList) {
    auto &t = constructors();
    auto it = t.find(name);
    if (it == t.end()) throw NoriException("A constructor for class \"%s\" could not be found!", name);
    return it->second(propList);
}

std::vector<std::string> NoriObjectFactory::registeredNames() {
    std::vector<std::string> out;
    for (auto &kv : constructors()) out.push_back(kv.first);
    return out;
}

NORI_NAMESPACE_END
