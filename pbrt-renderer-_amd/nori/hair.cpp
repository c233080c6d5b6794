// Hair mesh plugin "hair" (semantics of src/cpu/hair.cpp:21-324): a strand file
// in Yuksel's .hair format (a 128-byte little-endian header "HAIR", strand and
// point counts, flags for the optional per-strand segment counts, points,
// thickness, transparency and colour arrays, default segments / thickness)
// becomes tubes of `sides` vertices per control point.  Each ring lies in the
// plane normal to the fiber tangent (forward difference at the root, backward
// at the tip, central between; after `toWorld`), its vertices at angle
// 2 pi k / sides around the strand with outward normals, the tangent as the
// vertex tangent, u = position along the strand in [0, 1] and v = (h + 1) / 2
// with h = -cos(angle) the offset across the fiber (PBRT 3e, 9.9.1).
// Properties: filename, sides (6), radiusScale (1), maxStrands (all), toWorld.
#include "mesh.h"

#include <cmath>
#include <cstring>
#include <fstream>

NORI_NAMESPACE_BEGIN

namespace {

struct StrandFile {
    std::vector<uint32_t> points_per_strand;
    std::vector<float> points;    // 3 per point
    std::vector<float> thickness; // per point, or empty
    float default_thickness = 1.0f;
};

template <class T> void read_array(std::ifstream &f, std::vector<T> &v, size_t n, const char *what,
                                   const std::string &path) {
    v.resize(n);
    if (n && !f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(n * sizeof(T))))
        throw NoriException("hair file \"%s\": truncated %s array", path, what);
}

StrandFile read_strands(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw NoriException("hair file \"%s\": cannot open", path);
    unsigned char hdr[128];
    if (!f.read(reinterpret_cast<char *>(hdr), sizeof hdr) || std::memcmp(hdr, "HAIR", 4) != 0)
        throw NoriException("hair file \"%s\": missing HAIR signature", path);
    auto u32 = [&](int off) {
        uint32_t v;
        std::memcpy(&v, hdr + off, 4);
        return v;
    };
    auto f32 = [&](int off) {
        float v;
        std::memcpy(&v, hdr + off, 4);
        return v;
    };
    const uint32_t nstrands = u32(4), npoints = u32(8), flags = u32(12), default_segments = u32(16);
    if (!(flags & 2u)) throw NoriException("hair file \"%s\": no points array", path);
    StrandFile s;
    s.default_thickness = f32(20);
    std::vector<uint16_t> segments;
    if (flags & 1u) read_array(f, segments, nstrands, "segments", path);
    read_array(f, s.points, 3 * (size_t)npoints, "points", path);
    if (flags & 4u) read_array(f, s.thickness, npoints, "thickness", path);
    // transparency (flag 8) and colour (flag 16) follow; they are not used
    s.points_per_strand.resize(nstrands);
    uint64_t total = 0;
    for (uint32_t k = 0; k < nstrands; ++k) {
        s.points_per_strand[k] = (flags & 1u ? segments[k] : default_segments) + 1u;
        total += s.points_per_strand[k];
    }
    if (total != npoints)
        throw NoriException("hair file \"%s\": the strands hold %u points, the header says %u", path,
                            (unsigned)total, npoints);
    return s;
}

} // namespace

class HairMesh : public Mesh {
public:
    explicit HairMesh(const PropertyList &propList) {
        const std::string filename = getFileResolver()->resolve(propList.getString("filename"));
        const int sides = propList.getInteger("sides", 6);
        const float radius_scale = propList.getFloat("radiusScale", 1.0f);
        const int max_strands = propList.getInteger("maxStrands", -1);
        const Transform trafo = propList.getTransform("toWorld", Transform());
        if (sides < 3) throw NoriException("hair: \"sides\" must be at least 3");

        StrandFile sf = read_strands(filename);
        size_t nstrands = sf.points_per_strand.size();
        if (max_strands > 0 && (size_t)max_strands < nstrands) nstrands = (size_t)max_strands;

        uint64_t nv = 0, nt = 0;
        for (size_t k = 0; k < nstrands; ++k) {
            const uint32_t np = sf.points_per_strand[k];
            if (np < 2) continue;
            nv += (uint64_t)np * sides;
            nt += (uint64_t)(np - 1) * 2 * sides;
        }
        if (nv >= 0xFFFFFFFFull || nt >= 0xFFFFFFFFull) throw NoriException("hair \"%s\": too many vertices", filename);
        m_V.resize(3, nv);
        m_N.resize(3, nv);
        m_T.resize(3, nv);
        m_UV.resize(2, nv);
        m_F.resize(3, nt);

        uint32_t vi = 0, fi = 0;
        size_t first = 0; // first point of strand k in the file's arrays
        std::vector<Point3f> pw;
        for (size_t k = 0; k < nstrands; first += sf.points_per_strand[k], ++k) {
            const uint32_t np = sf.points_per_strand[k];
            if (np < 2) continue;
            pw.resize(np);
            for (uint32_t j = 0; j < np; ++j) {
                const float *q = &sf.points[3 * (first + j)];
                pw[j] = trafo.transformPoint(Point3f(q[0], q[1], q[2]));
            }
            const uint32_t ring0 = vi;
            for (uint32_t j = 0; j < np; ++j) {
                const Point3f &P = pw[j];
                m_bbox.expandBy(P);
                const Vector3f tangent =
                    (j == 0 ? pw[1] - pw[0] : j == np - 1 ? pw[j] - pw[j - 1] : pw[j + 1] - pw[j - 1]).normalized();
                // ring plane: two unit vectors normal to the tangent
                const Vector3f helper = std::fabs(tangent.x()) < 0.9f ? Vector3f(1, 0, 0) : Vector3f(0, 1, 0);
                const Vector3f b = tangent.cross(helper).normalized();
                const Vector3f n = tangent.cross(b).normalized();
                const float thick = sf.thickness.empty() ? sf.default_thickness : sf.thickness[first + j];
                const float radius = thick * 0.5f * radius_scale;
                const float u = (float)j / (float)(np - 1);
                for (int si = 0; si < sides; ++si) {
                    const float ang = 2.0f * (float)M_PI * (float)si / (float)sides;
                    const Vector3f radial = (b * std::cos(ang) + n * std::sin(ang)).normalized();
                    m_V.setCol3(vi, P + radial * radius);
                    m_N.setCol3(vi, radial);
                    m_T.setCol3(vi, tangent);
                    m_UV(0, vi) = u;
                    m_UV(1, vi) = (1.0f - std::cos(ang)) * 0.5f; // (h + 1) / 2, h = -cos(angle)
                    ++vi;
                }
                if (j > 0) {
                    const uint32_t prev = ring0 + (j - 1) * sides, cur = ring0 + j * sides;
                    for (int si = 0; si < sides; ++si) {
                        const uint32_t nx = (uint32_t)((si + 1) % sides);
                        m_F(0, fi) = prev + si;
                        m_F(1, fi) = prev + nx;
                        m_F(2, fi) = cur + nx;
                        ++fi;
                        m_F(0, fi) = prev + si;
                        m_F(1, fi) = cur + nx;
                        m_F(2, fi) = cur + si;
                        ++fi;
                    }
                }
            }
        }
        m_name = filename;
    }
};

NORI_REGISTER_CLASS(HairMesh, "hair");

NORI_NAMESPACE_END
