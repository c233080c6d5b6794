// Wavefront OBJ mesh plugin "obj" (semantics of src/cpu/obj.cpp): positions
// transformed by `toWorld`, quads split into two triangles, vertices
// deduplicated by their (position, uv, normal) index triple, normals from the
// file or, without them, angle-weighted face normals; texture coordinates kept
// as written.
#include "mesh.h"

#include <fstream>
#include <unordered_map>

NORI_NAMESPACE_BEGIN

class WavefrontOBJ : public Mesh {
public:
    explicit WavefrontOBJ(const PropertyList &propList) {
        std::string filename = getFileResolver()->resolve(propList.getString("filename"));
        std::ifstream is(filename);
        if (is.fail()) throw NoriException("Unable to open OBJ file \"%s\"!", filename);
        Transform trafo = propList.getTransform("toWorld", Transform());

        std::vector<Vector3f> positions, normals;
        std::vector<Vector2f> texcoords;
        std::vector<uint32_t> indices;
        std::vector<Key> vertices;
        std::unordered_map<Key, uint32_t, KeyHash> vmap;
        std::string line;
        while (std::getline(is, line)) {
            std::istringstream ls(line);
            std::string prefix;
            ls >> prefix;
            if (prefix == "v") {
                Point3f p;
                ls >> p.x() >> p.y() >> p.z();
                p = trafo.transformPoint(p);
                m_bbox.expandBy(p);
                positions.push_back(p);
            } else if (prefix == "vt") {
                Vector2f tc;
                ls >> tc.x() >> tc.y();
                texcoords.push_back(tc);
            } else if (prefix == "vn") {
                Normal3f n;
                ls >> n.x() >> n.y() >> n.z();
                normals.push_back(trafo.transformNormal(n).normalized());
            } else if (prefix == "f") {
                std::string v[4];
                ls >> v[0] >> v[1] >> v[2] >> v[3];
                Key k[6];
                int nv = 3;
                k[0] = Key(v[0]);
                k[1] = Key(v[1]);
                k[2] = Key(v[2]);
                if (!v[3].empty()) {
                    k[3] = Key(v[3]);
                    k[4] = k[0];
                    k[5] = k[2];
                    nv = 6;
                }
                for (int i = 0; i < nv; ++i) {
                    auto it = vmap.find(k[i]);
                    if (it == vmap.end()) {
                        vmap[k[i]] = (uint32_t)vertices.size();
                        indices.push_back((uint32_t)vertices.size());
                        vertices.push_back(k[i]);
                    } else {
                        indices.push_back(it->second);
                    }
                }
            }
        }
        const uint32_t nvtx = (uint32_t)vertices.size(), ntri = (uint32_t)(indices.size() / 3);
        m_F.resize(3, ntri);
        std::copy(indices.begin(), indices.end(), m_F.data());
        m_V.resize(3, nvtx);
        for (uint32_t i = 0; i < nvtx; ++i) {
            if (vertices[i].p == 0 || vertices[i].p > positions.size())
                throw NoriException("OBJ file \"%s\": vertex index %i out of range", filename, vertices[i].p);
            m_V.setCol3(i, positions[vertices[i].p - 1]);
        }
        m_N.resize(3, nvtx);
        if (!normals.empty()) {
            for (uint32_t i = 0; i < nvtx; ++i) {
                if (vertices[i].n == Key::NONE || vertices[i].n == 0 || vertices[i].n > normals.size())
                    throw NoriException("OBJ file \"%s\": missing or out-of-range normal index", filename);
                m_N.setCol3(i, normals[vertices[i].n - 1]);
            }
        } else {
            // angle-weighted face normals
            auto angle = [](const Vector3f &a, const Vector3f &b) {
                float d = a.normalized().dot(b.normalized());
                return std::acos(std::max(-1.0f, std::min(1.0f, d)));
            };
            std::vector<Vector3f> acc(nvtx, Vector3f(0.0f));
            for (uint32_t t = 0; t < ntri; ++t) {
                uint32_t i0 = m_F(0, t), i1 = m_F(1, t), i2 = m_F(2, t);
                Vector3f p0 = m_V.col3(i0), p1 = m_V.col3(i1), p2 = m_V.col3(i2);
                Vector3f fn = (p1 - p0).cross(p2 - p0);
                acc[i0] += fn * angle(p1 - p0, p2 - p0);
                acc[i1] += fn * angle(p0 - p1, p2 - p1);
                acc[i2] += fn * angle(p0 - p2, p1 - p2);
            }
            for (uint32_t i = 0; i < nvtx; ++i) {
                float len = acc[i].norm();
                m_N.setCol3(i, len > 1e-8f ? acc[i] / len : acc[i]);
            }
        }
        if (!texcoords.empty()) {
            m_UV.resize(2, nvtx);
            for (uint32_t i = 0; i < nvtx; ++i) {
                if (vertices[i].uv == Key::NONE || vertices[i].uv == 0 || vertices[i].uv > texcoords.size())
                    throw NoriException("OBJ file \"%s\": missing or out-of-range texture coordinate index", filename);
                m_UV(0, i) = texcoords[vertices[i].uv - 1].x();
                m_UV(1, i) = texcoords[vertices[i].uv - 1].y();
            }
        }
        m_name = filename;
    }

private:
    struct Key {
        static constexpr uint32_t NONE = 0xFFFFFFFFu;
        uint32_t p = NONE, n = NONE, uv = NONE;
        Key() {}
        explicit Key(const std::string &s) {
            std::vector<std::string> t = tokenize(s, "/", true);
            if (t.size() < 1 || t.size() > 3) throw NoriException("Invalid vertex data: \"%s\"", s);
            p = toUInt(t[0]);
            if (t.size() >= 2 && !t[1].empty()) uv = toUInt(t[1]);
            if (t.size() >= 3 && !t[2].empty()) n = toUInt(t[2]);
        }
        bool operator==(const Key &o) const { return p == o.p && n == o.n && uv == o.uv; }
    };
    struct KeyHash {
        size_t operator()(const Key &k) const {
            size_t h = std::hash<uint32_t>()(k.p);
            h = h * 37 + std::hash<uint32_t>()(k.uv);
            return h * 37 + std::hash<uint32_t>()(k.n);
        }
    };
};

NORI_REGISTER_CLASS(WavefrontOBJ, "obj");

NORI_NAMESPACE_END
