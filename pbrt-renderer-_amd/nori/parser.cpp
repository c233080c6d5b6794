// XML scene loader.  The scene format is Nori's (src/cpu/parser.cpp): its tag
// names, nesting rules, attributes, property conversions, transform
// composition and error texts are the interface kept here; the loader itself
// is a recursive descent over the parsed document (xml.h) in three node kinds.
//
//   object     scene, mesh, bsdf, emitter, medium, phase, camera, integrator,
//              sampler, rfilter, test: creates the plugin registered under its
//              `type` attribute (the scene tag takes none), which must be of
//              the tag's class; its object children are added (addChild, then
//              setParent) in document order and the object is activated.
//   property   boolean, integer, float, string, point, vector, color (name and
//              value attributes) and transform (name only): set on the
//              enclosing object's property list.
//   transform  translate (value), scale (value), rotate (axis, angle in
//   operation  degrees), matrix (16 values, row by row), lookat (origin, target,
//              up), only inside a transform; each one left-multiplies the
//              transform built so far, which starts at the identity.
//
// Placement: the root must be an object; a transform holds only transform
// operations and those appear only inside a transform; every other node needs
// an object as its parent.
#include "parser.h"

#include "xml.h"

#include <fstream>
#include <memory>

NORI_NAMESPACE_BEGIN

namespace {

enum class Kind { Object, Property, Transform, TransformOp };

struct TagInfo {
    Kind kind;
    int id; // object: NoriObject::EClassType; property / operation: a Prop / Op value
};

enum Prop { PBoolean, PInteger, PFloat, PString, PPoint, PVector, PColor };
enum Op { OTranslate, OScale, ORotate, OMatrix, OLookAt };

bool lookupTag(const std::string &name, TagInfo &info) {
    static const std::map<std::string, TagInfo> table = {
        {"scene", {Kind::Object, NoriObject::EScene}},
        {"mesh", {Kind::Object, NoriObject::EMesh}},
        {"bsdf", {Kind::Object, NoriObject::EBSDF}},
        {"emitter", {Kind::Object, NoriObject::EEmitter}},
        {"medium", {Kind::Object, NoriObject::EMedium}},
        {"phase", {Kind::Object, NoriObject::EPhaseFunction}},
        {"camera", {Kind::Object, NoriObject::ECamera}},
        {"integrator", {Kind::Object, NoriObject::EIntegrator}},
        {"sampler", {Kind::Object, NoriObject::ESampler}},
        {"rfilter", {Kind::Object, NoriObject::EReconstructionFilter}},
        {"test", {Kind::Object, NoriObject::ETest}},
        {"boolean", {Kind::Property, PBoolean}},
        {"integer", {Kind::Property, PInteger}},
        {"float", {Kind::Property, PFloat}},
        {"string", {Kind::Property, PString}},
        {"point", {Kind::Property, PPoint}},
        {"vector", {Kind::Property, PVector}},
        {"color", {Kind::Property, PColor}},
        {"transform", {Kind::Transform, 0}},
        {"translate", {Kind::TransformOp, OTranslate}},
        {"scale", {Kind::TransformOp, OScale}},
        {"rotate", {Kind::TransformOp, ORotate}},
        {"matrix", {Kind::TransformOp, OMatrix}},
        {"lookat", {Kind::TransformOp, OLookAt}},
    };
    auto it = table.find(name);
    if (it == table.end()) return false;
    info = it->second;
    return true;
}

// 4x4 from 16 row-major values.
Transform rowMajor(std::initializer_list<float> v) {
    float m[16];
    std::copy(v.begin(), v.end(), m);
    return Transform(m);
}

// Right-handed rotation by `radians` about the unit vector `u`:
// R = cos I + sin [u]x + (1 - cos) u u^T.
Transform axisAngle(const Vector3f &u, float radians) {
    const float c = std::cos(radians), s = std::sin(radians), k = 1.0f - c;
    float m[16] = {0};
    for (int r = 0; r < 3; ++r)
        for (int col = 0; col < 3; ++col) m[4 * r + col] = k * u[r] * u[col] + (r == col ? c : 0.0f);
    // the skew-symmetric part s [u]x
    m[1] -= s * u[2]; m[2] += s * u[1];
    m[4] += s * u[2]; m[6] -= s * u[0];
    m[8] -= s * u[1]; m[9] += s * u[0];
    m[15] = 1.0f;
    return Transform(m);
}

// Camera-style frame at `origin` looking at `target`: columns are the left
// axis, the re-orthogonalised up axis, the view direction and the origin.
Transform lookAtFrame(const Vector3f &origin, const Vector3f &target, const Vector3f &up) {
    const Vector3f fwd = (target - origin).normalized();
    const Vector3f left = up.normalized().cross(fwd).normalized();
    const Vector3f up2 = fwd.cross(left).normalized();
    return rowMajor({left[0], up2[0], fwd[0], origin[0],
                     left[1], up2[1], fwd[1], origin[1],
                     left[2], up2[2], fwd[2], origin[2],
                     0.0f, 0.0f, 0.0f, 1.0f});
}

class SceneLoader {
public:
    SceneLoader(const std::string &text, const std::string &file) : m_text(text), m_file(file) {}

    NoriObject *loadRoot(const XmlNode &node) {
        TagInfo info = classify(node);
        if (info.kind != Kind::Object)
            fail(node, "root element \"" + node.name + "\" must be a Nori object");
        return guarded(node, [&] { return loadObject(node, info); });
    }

private:
    const std::string &m_text, &m_file;

    [[noreturn]] void fail(const XmlNode &node, const std::string &what) const {
        throw NoriException("Error while parsing \"%s\": %s (at %s)", m_file, what, xmlPosition(m_text, node.offset));
    }

    // Exceptions from plugins and conversions get the file and position added
    // once (the innermost node that saw them).
    template <typename F> auto guarded(const XmlNode &node, F &&f) -> decltype(f()) {
        try {
            return f();
        } catch (const NoriException &e) {
            const std::string msg = e.what();
            if (msg.rfind("Error while parsing", 0) == 0) throw;
            throw NoriException("Error while parsing \"%s\": %s (at %s)", m_file, msg, xmlPosition(m_text, node.offset));
        }
    }

    TagInfo classify(const XmlNode &node) const {
        TagInfo info;
        if (!lookupTag(node.name, info)) fail(node, "unexpected tag \"" + node.name + "\"");
        return info;
    }

    // Children that are elements, with comments and declarations skipped.
    template <typename F> void forElements(const XmlNode &node, F &&f) const {
        for (auto &ch : node.children) {
            if (ch->type == XmlNode::Comment || ch->type == XmlNode::Declaration) continue;
            if (ch->type != XmlNode::Element) fail(*ch, "unexpected content");
            f(*ch);
        }
    }

    void checkPlacement(const XmlNode &child, Kind parent, Kind kind) const {
        if ((parent == Kind::Transform) != (kind == Kind::TransformOp))
            fail(child, "transform nodes can only contain transform operations");
        if (parent != Kind::Object && parent != Kind::Transform)
            fail(child, "node \"" + child.name + "\" requires a Nori object as parent");
    }

    // The node's attributes must be exactly `names`.
    void expectAttributes(const XmlNode &node, std::initializer_list<const char *> names) const {
        for (auto &a : node.attributes) {
            bool known = false;
            for (const char *n : names) known = known || a.first == n;
            if (!known)
                throw NoriException("Error while parsing \"%s\": unexpected attribute \"%s\" in \"%s\" at %s", m_file,
                                    a.first, node.name, xmlPosition(m_text, node.offset));
        }
        for (const char *n : names)
            if (!node.attribute(n))
                throw NoriException("Error while parsing \"%s\": missing attribute \"%s\" in \"%s\" at %s", m_file, n,
                                    node.name, xmlPosition(m_text, node.offset));
    }

    static const std::string &value(const XmlNode &node, const char *name) { return *node.attribute(name); }

    NoriObject *loadObject(const XmlNode &node, const TagInfo &info) {
        PropertyList props;
        std::vector<std::unique_ptr<NoriObject>> children;
        forElements(node, [&](const XmlNode &ch) {
            const TagInfo ci = classify(ch);
            checkPlacement(ch, Kind::Object, ci.kind);
            guarded(ch, [&] {
                if (ci.kind == Kind::Object) children.emplace_back(loadObject(ch, ci));
                else if (ci.kind == Kind::Property) loadProperty(ch, ci, props);
                else loadTransform(ch, props);
                return 0;
            });
        });
        return guarded(node, [&] {
            std::string type = "scene";
            if (info.id == NoriObject::EScene) {
                expectAttributes(node, {});
            } else {
                expectAttributes(node, {"type"});
                type = value(node, "type");
            }
            std::unique_ptr<NoriObject> obj(NoriObjectFactory::createInstance(type, props));
            if ((int)obj->getClassType() != info.id)
                throw NoriException("Unexpectedly constructed an object of type <%s> (expected type <%s>): %s",
                                    NoriObject::classTypeName(obj->getClassType()),
                                    NoriObject::classTypeName((NoriObject::EClassType)info.id), obj->toString());
            for (auto &c : children) {
                NoriObject *raw = c.release();
                obj->addChild(raw);
                raw->setParent(obj.get());
            }
            obj->activate();
            return obj.release();
        });
    }

    void loadProperty(const XmlNode &node, const TagInfo &info, PropertyList &props) {
        forElements(node, [&](const XmlNode &ch) { checkPlacement(ch, Kind::Property, classify(ch).kind); });
        expectAttributes(node, {"name", "value"});
        const std::string &name = value(node, "name"), &v = value(node, "value");
        switch (info.id) {
        case PBoolean: props.setBoolean(name, toBool(v)); break;
        case PInteger: props.setInteger(name, toInt(v)); break;
        case PFloat: props.setFloat(name, toFloat(v)); break;
        case PString: props.setString(name, v); break;
        case PPoint: props.setPoint(name, toVector3f(v)); break;
        case PVector: props.setVector(name, toVector3f(v)); break;
        default: props.setColor(name, Color3f(toVector3f(v))); break;
        }
    }

    void loadTransform(const XmlNode &node, PropertyList &props) {
        Transform acc;
        forElements(node, [&](const XmlNode &ch) {
            const TagInfo ci = classify(ch);
            checkPlacement(ch, Kind::Transform, ci.kind);
            acc = guarded(ch, [&] { return operation(ch, ci.id); }) * acc;
        });
        expectAttributes(node, {"name"});
        props.setTransform(value(node, "name"), acc);
    }

    Transform operation(const XmlNode &node, int op) {
        forElements(node, [&](const XmlNode &ch) { checkPlacement(ch, Kind::TransformOp, classify(ch).kind); });
        switch (op) {
        case OTranslate: {
            expectAttributes(node, {"value"});
            const Vector3f t = toVector3f(value(node, "value"));
            return rowMajor({1, 0, 0, t[0], 0, 1, 0, t[1], 0, 0, 1, t[2], 0, 0, 0, 1});
        }
        case OScale: {
            expectAttributes(node, {"value"});
            const Vector3f s = toVector3f(value(node, "value"));
            return rowMajor({s[0], 0, 0, 0, 0, s[1], 0, 0, 0, 0, s[2], 0, 0, 0, 0, 1});
        }
        case ORotate:
            expectAttributes(node, {"angle", "axis"});
            return axisAngle(toVector3f(value(node, "axis")).normalized(), degToRad(toFloat(value(node, "angle"))));
        case OMatrix: {
            expectAttributes(node, {"value"});
            const std::vector<std::string> tok = tokenize(value(node, "value"));
            if (tok.size() != 16) throw NoriException("Expected 16 values");
            float m[16];
            for (int k = 0; k < 16; ++k) m[k] = toFloat(tok[k]);
            return Transform(m);
        }
        default:
            expectAttributes(node, {"origin", "target", "up"});
            return lookAtFrame(toVector3f(value(node, "origin")), toVector3f(value(node, "target")),
                               toVector3f(value(node, "up")));
        }
    }
};

std::string directoryOf(const std::string &path) {
    const size_t p = path.find_last_of('/');
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
    struct ResetResolver {
        ~ResetResolver() { getFileResolver()->clear(); }
    } reset;
    for (auto &n : doc.nodes)
        if (n->type == XmlNode::Element) return SceneLoader(text, name).loadRoot(*n);
    throw NoriException("Error while parsing \"%s\": no root element", name);
}

NoriObject *loadFromXML(const std::string &filename) {
    std::ifstream is(filename, std::ios::binary);
    if (!is.good()) throw NoriException("Error while parsing \"%s\": unable to open the file", filename);
    std::string text((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
    return loadFromXMLString(text, directoryOf(filename), filename);
}

NORI_NAMESPACE_END
