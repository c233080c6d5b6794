// Wavefront OBJ mesh plugin "obj" (the plugin name and properties of
// src/cpu/obj.cpp: `filename`, `toWorld`).
//
// Statements read (everything else is skipped):
//   v x y z      position, transformed by toWorld
//   vt u v       texture coordinate, kept as written
//   vn x y z     normal, transformed by the inverse transpose of toWorld and
//                normalised
//   f c0 c1 ...  polygon whose corners are p, p/t, p//n or p/t/n (1-based
//                indices; negative ones count back from the last element read
//                so far), fanned into triangles (c0, ci, ci+1)
//
// Vertex and triangle order: triangles in face order; a vertex is created the
// first time a corner's (p, t, n) triple occurs, so corners that repeat a
// triple share it.  When the file has normals (texture coordinates), every
// corner must name one.  Without normals, a vertex normal is the normalised
// sum, over the triangles at the vertex, of the triangle's edge cross product
// times the triangle's angle at the vertex.
#include "mesh.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <unordered_map>

NORI_NAMESPACE_BEGIN

namespace {

constexpr uint32_t NO_INDEX = 0xFFFFFFFFu;

// 0-based indices of one face corner into the position / texcoord / normal lists.
struct Corner {
    uint32_t p = NO_INDEX, t = NO_INDEX, n = NO_INDEX;
    bool operator==(const Corner &o) const { return p == o.p && t == o.t && n == o.n; }
};

struct CornerHash {
    size_t operator()(const Corner &c) const {
        // splitmix64 finaliser over the packed triple
        uint64_t x = ((uint64_t)c.p << 32) ^ ((uint64_t)c.t << 21) ^ (uint64_t)c.n;
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return (size_t)(x ^ (x >> 31));
    }
};

// Line-by-line reader over the whole file held in memory.
class ObjReader {
public:
    ObjReader(const std::string &text, const std::string &file) : m_text(text), m_file(file) {}

    bool nextLine() {
        if (m_pos >= m_text.size()) return false;
        size_t end = m_text.find('\n', m_pos);
        if (end == std::string::npos) end = m_text.size();
        m_cur = m_pos;
        m_end = end;
        m_pos = end + 1;
        ++m_line;
        return true;
    }
    // Next whitespace-separated token of the current line ("" at its end).
    std::string token() {
        while (m_cur < m_end && isSpace(m_text[m_cur])) ++m_cur;
        size_t b = m_cur;
        while (m_cur < m_end && !isSpace(m_text[m_cur])) ++m_cur;
        return m_text.substr(b, m_cur - b);
    }
    float number() {
        std::string s = token();
        char *e = nullptr;
        float v = std::strtof(s.c_str(), &e);
        if (s.empty() || *e != '\0') fail("expected a number, found \"" + s + "\"");
        return v;
    }
    [[noreturn]] void fail(const std::string &what) const {
        throw NoriException("OBJ file \"%s\", line %i: %s", m_file, (int)m_line, what);
    }

private:
    static bool isSpace(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }
    const std::string &m_text;
    const std::string &m_file;
    size_t m_pos = 0, m_cur = 0, m_end = 0, m_line = 0;
};

// One index of a corner: 1-based, or negative relative to `count`.
uint32_t resolveIndex(const ObjReader &rd, const std::string &s, size_t count, const char *what) {
    char *e = nullptr;
    long v = std::strtol(s.c_str(), &e, 10);
    if (s.empty() || *e != '\0') rd.fail(std::string("bad ") + what + " index \"" + s + "\"");
    long i = v > 0 ? v - 1 : (long)count + v;
    if (v == 0 || i < 0 || i >= (long)count) rd.fail(std::string(what) + " index " + s + " out of range");
    return (uint32_t)i;
}

Corner parseCorner(const ObjReader &rd, const std::string &tok, size_t np, size_t nt, size_t nn) {
    Corner c;
    const size_t s1 = tok.find('/');
    c.p = resolveIndex(rd, tok.substr(0, s1), np, "position");
    if (s1 == std::string::npos) return c;
    const size_t s2 = tok.find('/', s1 + 1);
    const std::string t = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
    if (!t.empty()) c.t = resolveIndex(rd, t, nt, "texture coordinate");
    if (s2 != std::string::npos) {
        const std::string n = tok.substr(s2 + 1);
        if (n.find('/') != std::string::npos) rd.fail("corner \"" + tok + "\" has more than three indices");
        if (!n.empty()) c.n = resolveIndex(rd, n, nn, "normal");
    }
    return c;
}

// Angle between two edge vectors (robust near 0 and pi).
float cornerAngle(const Vector3f &a, const Vector3f &b) { return std::atan2(a.cross(b).norm(), a.dot(b)); }

} // namespace

class WavefrontOBJ : public Mesh {
public:
    explicit WavefrontOBJ(const PropertyList &propList) {
        const std::string filename = getFileResolver()->resolve(propList.getString("filename"));
        std::ifstream is(filename, std::ios::binary);
        if (is.fail()) throw NoriException("Unable to open OBJ file \"%s\"!", filename);
        std::ostringstream buf;
        buf << is.rdbuf();
        const std::string text = buf.str();
        const Transform toWorld = propList.getTransform("toWorld", Transform());

        std::vector<Vector3f> pos, nrm;
        std::vector<Vector2f> uv;
        std::vector<Corner> verts;  // unique corners, in order of first use
        std::vector<uint32_t> tris; // 3 vertex ids per triangle
        std::unordered_map<Corner, uint32_t, CornerHash> ids;
        auto vertexOf = [&](const Corner &c) {
            auto it = ids.emplace(c, (uint32_t)verts.size());
            if (it.second) verts.push_back(c);
            return it.first->second;
        };

        ObjReader rd(text, filename);
        std::vector<uint32_t> poly;
        while (rd.nextLine()) {
            const std::string kw = rd.token();
            if (kw == "v") {
                Point3f p;
                p.x() = rd.number();
                p.y() = rd.number();
                p.z() = rd.number();
                pos.push_back(toWorld.transformPoint(p));
            } else if (kw == "vt") {
                Vector2f t;
                t.x() = rd.number();
                t.y() = rd.number();
                uv.push_back(t);
            } else if (kw == "vn") {
                Normal3f n;
                n.x() = rd.number();
                n.y() = rd.number();
                n.z() = rd.number();
                nrm.push_back(toWorld.transformNormal(n).normalized());
            } else if (kw == "f") {
                poly.clear();
                for (std::string tok = rd.token(); !tok.empty(); tok = rd.token())
                    poly.push_back(vertexOf(parseCorner(rd, tok, pos.size(), uv.size(), nrm.size())));
                if (poly.size() < 3) rd.fail("a face needs at least three corners");
                for (size_t i = 1; i + 1 < poly.size(); ++i) {
                    tris.push_back(poly[0]);
                    tris.push_back(poly[i]);
                    tris.push_back(poly[i + 1]);
                }
            }
        }

        const size_t nv = verts.size(), ntri = tris.size() / 3;
        m_F.resize(3, ntri);
        std::copy(tris.begin(), tris.end(), m_F.data());
        m_V.resize(3, nv);
        for (size_t i = 0; i < nv; ++i) {
            m_V.setCol3(i, pos[verts[i].p]);
            m_bbox.expandBy(Point3f(pos[verts[i].p]));
        }
        if (!uv.empty()) {
            m_UV.resize(2, nv);
            for (size_t i = 0; i < nv; ++i) {
                if (verts[i].t == NO_INDEX)
                    throw NoriException("OBJ file \"%s\": a face corner has no texture coordinate", filename);
                m_UV(0, i) = uv[verts[i].t].x();
                m_UV(1, i) = uv[verts[i].t].y();
            }
        }
        m_N.resize(3, nv);
        if (!nrm.empty()) {
            for (size_t i = 0; i < nv; ++i) {
                if (verts[i].n == NO_INDEX)
                    throw NoriException("OBJ file \"%s\": a face corner has no normal", filename);
                m_N.setCol3(i, nrm[verts[i].n]);
            }
        } else {
            std::vector<Vector3f> sum(nv, Vector3f(0.0f));
            for (size_t t = 0; t < ntri; ++t) {
                const uint32_t v[3] = {m_F(0, t), m_F(1, t), m_F(2, t)};
                const Vector3f p[3] = {m_V.col3(v[0]), m_V.col3(v[1]), m_V.col3(v[2])};
                const Vector3f area = (p[1] - p[0]).cross(p[2] - p[0]);
                for (int k = 0; k < 3; ++k) {
                    const Vector3f a = p[(k + 1) % 3] - p[k], b = p[(k + 2) % 3] - p[k];
                    sum[v[k]] += area * cornerAngle(a, b);
                }
            }
            for (size_t i = 0; i < nv; ++i) {
                const float len = sum[i].norm();
                m_N.setCol3(i, len > 1e-8f ? sum[i] / len : sum[i]);
            }
        }
        m_name = filename;
    }
};

NORI_REGISTER_CLASS(WavefrontOBJ, "obj");

NORI_NAMESPACE_END
