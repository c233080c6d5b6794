// XML scene loader with the tag set, nesting checks, attribute checks, property
// conversions and transform composition of src/cpu/parser.cpp:17-337:
//   * object tags (scene, mesh, bsdf, emitter, medium, phase, camera,
//     integrator, sampler, rfilter, test) instantiate the plugin registered
//     under their `type`, check its class type against the tag, add every
//     child with addChild + setParent, then activate();
//   * property tags (boolean, integer, float, string, point, vector, color,
//     transform) fill the parent's PropertyList;
//   * inside <transform>, translate / scale / rotate / matrix / lookat compose
//     as op * current, starting from the identity.
#include "parser.h"

#include "xml.h"

#include <fstream>
#include <set>

NORI_NAMESPACE_BEGIN

namespace {

enum ETag {
    TScene = NoriObject::EScene,
    TMesh = NoriObject::EMesh,
    TBSDF = NoriObject::EBSDF,
    TPhase = NoriObject::EPhaseFunction,
    TEmitter = NoriObject::EEmitter,
    TMedium = NoriObject::EMedium,
    TCamera = NoriObject::ECamera,
    TIntegrator = NoriObject::EIntegrator,
    TSampler = NoriObject::ESampler,
    TTest = NoriObject::ETest,
    TRFilter = NoriObject::EReconstructionFilter,
    TBoolean = NoriObject::EClassTypeCount,
    TInteger,
    TFloat,
    TString,
    TPoint,
    TVector,
    TColor,
    TTransform,
    TTranslate,
    TMatrix,
    TRotate,
    TScale,
    TLookAt,
    TInvalid
};

const std::map<std::string, int> &tagTable() {
    static const std::map<std::string, int> t = {
        {"scene", TScene},     {"mesh", TMesh},         {"bsdf", TBSDF},           {"emitter", TEmitter},
        {"camera", TCamera},   {"medium", TMedium},     {"phase", TPhase},         {"integrator", TIntegrator},
        {"sampler", TSampler}, {"rfilter", TRFilter},   {"test", TTest},           {"boolean", TBoolean},
        {"integer", TInteger}, {"float", TFloat},       {"string", TString},       {"point", TPoint},
        {"vector", TVector},   {"color", TColor},       {"transform", TTransform}, {"translate", TTranslate},
        {"matrix", TMatrix},   {"rotate", TRotate},     {"scale", TScale},         {"lookat", TLookAt}};
    return t;
}

Transform fromRows(const float r[16]) { return Transform(r); }

// Rotation by `angle` radians about the unit `axis` (Rodrigues).
Transform rotation(float angle, const Vector3f &a) {
    float c = std::cos(angle), s = std::sin(angle), t = 1.0f - c;
    float m[16] = {t * a[0] * a[0] + c,        t * a[0] * a[1] - s * a[2], t * a[0] * a[2] + s * a[1], 0,
                   t * a[0] * a[1] + s * a[2], t * a[1] * a[1] + c,        t * a[1] * a[2] - s * a[0], 0,
                   t * a[0] * a[2] - s * a[1], t * a[1] * a[2] + s * a[0], t * a[2] * a[2] + c,        0,
                   0,                          0,                          0,                          1};
    return fromRows(m);
}

struct Loader {
    const std::string &text, &filename;
    Transform transform;

    std::string where(size_t off) const { return xmlPosition(text, off); }

    void checkAttributes(const XmlNode &node, std::set<std::string> attrs) {
        for (auto &a : node.attributes) {
            auto it = attrs.find(a.first);
            if (it == attrs.end())
                throw NoriException("Error while parsing \"%s\": unexpected attribute \"%s\" in \"%s\" at %s", filename,
                                    a.first, node.name, where(node.offset));
            attrs.erase(it);
        }
        if (!attrs.empty())
            throw NoriException("Error while parsing \"%s\": missing attribute \"%s\" in \"%s\" at %s", filename,
                                *attrs.begin(), node.name, where(node.offset));
    }
    static std::string attr(const XmlNode &n, const char *k) {
        const std::string *v = n.attribute(k);
        return v ? *v : std::string();
    }

    NoriObject *parseTag(const XmlNode &node, PropertyList &list, int parentTag) {
        if (node.type == XmlNode::Comment || node.type == XmlNode::Declaration) return nullptr;
        if (node.type != XmlNode::Element)
            throw NoriException("Error while parsing \"%s\": unexpected content at %s", filename, where(node.offset));
        auto it = tagTable().find(node.name);
        if (it == tagTable().end())
            throw NoriException("Error while parsing \"%s\": unexpected tag \"%s\" at %s", filename, node.name,
                                where(node.offset));
        const int tag = it->second;
        const bool hasParent = parentTag != TInvalid;
        const bool parentIsObject = hasParent && parentTag < NoriObject::EClassTypeCount;
        const bool currentIsObject = tag < NoriObject::EClassTypeCount;
        const bool parentIsTransform = parentTag == TTransform;
        const bool currentIsTransformOp =
            tag == TTranslate || tag == TRotate || tag == TScale || tag == TLookAt || tag == TMatrix;
        if (!hasParent && !currentIsObject)
            throw NoriException("Error while parsing \"%s\": root element \"%s\" must be a Nori object (at %s)",
                                filename, node.name, where(node.offset));
        if (parentIsTransform != currentIsTransformOp)
            throw NoriException("Error while parsing \"%s\": transform nodes can only contain transform operations "
                                "(at %s)",
                                filename, where(node.offset));
        if (hasParent && !parentIsObject && !(parentIsTransform && currentIsTransformOp))
            throw NoriException("Error while parsing \"%s\": node \"%s\" requires a Nori object as parent (at %s)",
                                filename, node.name, where(node.offset));
        if (tag == TTransform) transform.setIdentity();

        PropertyList props;
        std::vector<NoriObject *> children;
        try {
            for (auto &ch : node.children) {
                NoriObject *c = parseTag(*ch, props, tag);
                if (c) children.push_back(c);
            }
        } catch (...) {
            for (auto *c : children) delete c;
            throw;
        }

        NoriObject *result = nullptr;
        try {
            if (currentIsObject) {
                std::string type;
                if (tag == TScene) {
                    checkAttributes(node, {});
                    type = "scene";
                } else {
                    checkAttributes(node, {"type"});
                    type = attr(node, "type");
                }
                result = NoriObjectFactory::createInstance(type, props);
                if ((int)result->getClassType() != tag) {
                    std::string got = NoriObject::classTypeName(result->getClassType()), desc = result->toString();
                    delete result;
                    result = nullptr;
                    throw NoriException("Unexpectedly constructed an object of type <%s> (expected type <%s>): %s", got,
                                        NoriObject::classTypeName((NoriObject::EClassType)tag), desc);
                }
                for (size_t k = 0; k < children.size(); ++k) {
                    NoriObject *c = children[k];
                    children[k] = nullptr;
                    result->addChild(c);
                    c->setParent(result);
                }
                result->activate();
            } else {
                switch (tag) {
                case TString:
                    checkAttributes(node, {"name", "value"});
                    list.setString(attr(node, "name"), attr(node, "value"));
                    break;
                case TFloat:
                    checkAttributes(node, {"name", "value"});
                    list.setFloat(attr(node, "name"), toFloat(attr(node, "value")));
                    break;
                case TInteger:
                    checkAttributes(node, {"name", "value"});
                    list.setInteger(attr(node, "name"), toInt(attr(node, "value")));
                    break;
                case TBoolean:
                    checkAttributes(node, {"name", "value"});
                    list.setBoolean(attr(node, "name"), toBool(attr(node, "value")));
                    break;
                case TPoint:
                    checkAttributes(node, {"name", "value"});
                    list.setPoint(attr(node, "name"), toVector3f(attr(node, "value")));
                    break;
                case TVector:
                    checkAttributes(node, {"name", "value"});
                    list.setVector(attr(node, "name"), toVector3f(attr(node, "value")));
                    break;
                case TColor:
                    checkAttributes(node, {"name", "value"});
                    list.setColor(attr(node, "name"), Color3f(toVector3f(attr(node, "value"))));
                    break;
                case TTransform:
                    checkAttributes(node, {"name"});
                    list.setTransform(attr(node, "name"), transform);
                    break;
                case TTranslate: {
                    checkAttributes(node, {"value"});
                    Vector3f v = toVector3f(attr(node, "value"));
                    float m[16] = {1, 0, 0, v[0], 0, 1, 0, v[1], 0, 0, 1, v[2], 0, 0, 0, 1};
                    transform = fromRows(m) * transform;
                } break;
                case TMatrix: {
                    checkAttributes(node, {"value"});
                    std::vector<std::string> tok = tokenize(attr(node, "value"));
                    if (tok.size() != 16) throw NoriException("Expected 16 values");
                    float m[16];
                    for (int k = 0; k < 16; ++k) m[k] = toFloat(tok[k]);
                    transform = fromRows(m) * transform;
                } break;
                case TScale: {
                    checkAttributes(node, {"value"});
                    Vector3f v = toVector3f(attr(node, "value"));
                    float m[16] = {v[0], 0, 0, 0, 0, v[1], 0, 0, 0, 0, v[2], 0, 0, 0, 0, 1};
                    transform = fromRows(m) * transform;
                } break;
                case TRotate: {
                    checkAttributes(node, {"angle", "axis"});
                    float angle = degToRad(toFloat(attr(node, "angle")));
                    Vector3f axis = toVector3f(attr(node, "axis")).normalized();
                    transform = rotation(angle, axis) * transform;
                } break;
                case TLookAt: {
                    checkAttributes(node, {"origin", "target", "up"});
                    Vector3f origin = toVector3f(attr(node, "origin"));
                    Vector3f target = toVector3f(attr(node, "target"));
                    Vector3f up = toVector3f(attr(node, "up"));
                    Vector3f dir = (target - origin).normalized();
                    Vector3f left = up.normalized().cross(dir).normalized();
                    Vector3f newUp = dir.cross(left).normalized();
                    float m[16] = {left[0], newUp[0], dir[0], origin[0], left[1], newUp[1], dir[1], origin[1],
                                   left[2], newUp[2], dir[2], origin[2], 0,       0,        0,      1};
                    transform = fromRows(m) * transform;
                } break;
                default:
                    throw NoriException("Unhandled element \"%s\"", node.name);
                }
            }
        } catch (const NoriException &e) {
            for (auto *c : children) delete c;
            delete result;
            std::string msg = e.what();
            if (msg.rfind("Error while parsing", 0) == 0) throw;
            throw NoriException("Error while parsing \"%s\": %s (at %s)", filename, msg, where(node.offset));
        }
        return result;
    }
};

std::string dirOf(const std::string &path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string(".") : path.substr(0, p);
}

} // namespace

NoriObject *loadFromXMLString(const std::string &text, const std::string &dir, const std::string &name) {
    XmlDocument doc;
    std::string err;
    size_t off = 0;
    if (!parseXml(text, doc, err, off))
        throw NoriException("Error while parsing \"%s\": %s (at %s)", name, err, xmlPosition(text, off));
    getFileResolver()->prepend(dir);
    struct Pop {
        ~Pop() { getFileResolver()->clear(); }
    } pop;
    Loader L{text, name, Transform()};
    PropertyList list;
    for (auto &n : doc.nodes)
        if (n->type == XmlNode::Element) return L.parseTag(*n, list, TInvalid);
    throw NoriException("Error while parsing \"%s\": no root element", name);
}

NoriObject *loadFromXML(const std::string &filename) {
    std::ifstream is(filename, std::ios::binary);
    if (!is.good()) throw NoriException("Error while parsing \"%s\": unable to open the file", filename);
    std::string text((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
    return loadFromXMLString(text, dirOf(filename), filename);
}

NORI_NAMESPACE_END
