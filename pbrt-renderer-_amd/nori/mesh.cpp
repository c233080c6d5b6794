// Mesh plugin base (src/cpu/mesh.cpp semantics) and the discrete area pdf
// (include/dpdf.h semantics).
#include "mesh.h"

#include "bsdf.h"
#include "emitter.h"

NORI_NAMESPACE_BEGIN

float DiscretePDF::normalize() {
    m_sum = m_cdf.back();
    if (m_sum > 0.0f) {
        m_normalization = 1.0f / m_sum;
        for (size_t i = 1; i < m_cdf.size(); ++i) m_cdf[i] *= m_normalization;
        m_cdf.back() = 1.0f;
        m_normalized = true;
    } else {
        m_normalization = 0.0f;
    }
    return m_sum;
}

size_t DiscretePDF::sample(float x) const {
    auto it = std::lower_bound(m_cdf.begin(), m_cdf.end(), x);
    size_t index = (size_t)std::max<ptrdiff_t>(0, it - m_cdf.begin() - 1);
    return std::min(index, m_cdf.size() - 2);
}

size_t DiscretePDF::sampleReuse(float &x) const {
    size_t index = sample(x);
    x = (x - m_cdf[index]) / (m_cdf[index + 1] - m_cdf[index]);
    return index;
}

std::string Intersection::toString() const {
    if (!mesh) return "Intersection[invalid]";
    return format("Intersection[\n  p = %s,\n  t = %f,\n  uv = %s,\n  shFrame = %s,\n  geoFrame = %s,\n  mesh = %s\n]",
                  p.toString(), t, uv.toString(), indent(shFrame.toString()), indent(geoFrame.toString()),
                  mesh->getName());
}

Mesh::~Mesh() {
    delete m_bsdf;
    delete m_emitter;
}

void Mesh::activate() {
    if (!m_bsdf) m_bsdf = static_cast<BSDF *>(NoriObjectFactory::createInstance("diffuse", PropertyList()));
    m_dpdf.clear();
    m_dpdf.reserve(getTriangleCount());
    for (uint32_t i = 0; i < getTriangleCount(); ++i) m_dpdf.append(surfaceArea(i));
    m_dpdf.normalize();
}

float Mesh::surfaceArea(uint32_t index) const {
    Point3f p0 = m_V.col3(m_F(0, index)), p1 = m_V.col3(m_F(1, index)), p2 = m_V.col3(m_F(2, index));
    return 0.5f * (p1 - p0).cross(p2 - p0).norm();
}

void Mesh::samplePosition(const Point2f &sample, Point3f &p, Normal3f &n, float &pdf) const {
    float x = sample.x();
    size_t tri = m_dpdf.sampleReuse(x);
    Point3f p0 = m_V.col3(m_F(0, tri)), p1 = m_V.col3(m_F(1, tri)), p2 = m_V.col3(m_F(2, tri));
    float sr1 = std::sqrt(x), r2 = sample.y();
    p = p0 * (1.0f - sr1) + p1 * (sr1 * (1.0f - r2)) + p2 * (sr1 * r2);
    n = (p1 - p0).cross(p2 - p0).normalized();
    pdf = 1.0f / m_dpdf.getSum();
}

BoundingBox3f Mesh::getBoundingBox(uint32_t index) const {
    BoundingBox3f b(m_V.col3(m_F(0, index)));
    b.expandBy(m_V.col3(m_F(1, index)));
    b.expandBy(m_V.col3(m_F(2, index)));
    return b;
}

Point3f Mesh::getCentroid(uint32_t index) const {
    return (m_V.col3(m_F(0, index)) + m_V.col3(m_F(1, index)) + m_V.col3(m_F(2, index))) * (1.0f / 3.0f);
}

void Mesh::addChild(NoriObject *obj) {
    switch (obj->getClassType()) {
    case EBSDF:
        if (m_bsdf) throw NoriException("Mesh: tried to register multiple BSDF instances!");
        m_bsdf = static_cast<BSDF *>(obj);
        break;
    case EEmitter:
        if (m_emitter) throw NoriException("Mesh: tried to register multiple Emitter instances!");
        m_emitter = static_cast<Emitter *>(obj);
        m_emitter->setMesh(this);
        break;
    default:
        throw NoriException("Mesh::addChild(<%s>) is not supported!", classTypeName(obj->getClassType()));
    }
}

std::string Mesh::toString() const {
    return format("Mesh[\n  name = \"%s\",\n  vertexCount = %i,\n  triangleCount = %i,\n  bsdf = %s,\n  emitter = %s\n]",
                  m_name, m_V.cols(), m_F.cols(), m_bsdf ? indent(m_bsdf->toString()) : std::string("null"),
                  m_emitter ? indent(m_emitter->toString()) : std::string("null"));
}

NORI_NAMESPACE_END
