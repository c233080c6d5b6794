// Triangle meshes, intersections and the discrete area pdf of the plugin
// surface (include/mesh.h, include/dpdf.h of the reference).
#pragma once
#include "object.h"

NORI_NAMESPACE_BEGIN

class BSDF;
class Emitter;
class Mesh;

/// Discrete pdf over a mesh's triangles (include/dpdf.h): append weights, then
/// normalize(); the sum before normalisation is kept.
class DiscretePDF {
public:
    void clear() { m_cdf.assign(1, 0.0f); m_sum = 0.0f; m_normalization = 0.0f; m_normalized = false; }
    void reserve(size_t n) { m_cdf.reserve(n + 1); }
    void append(float v) { m_cdf.push_back(m_cdf.back() + v); }
    size_t size() const { return m_cdf.size() - 1; }
    float getSum() const { return m_sum; }
    float getNormalization() const { return m_normalization; }
    const std::vector<float> &cdf() const { return m_cdf; }
    float normalize();
    /// Index i with cdf[i] <= x < cdf[i+1].
    size_t sample(float x) const;
    /// Same, and x is remapped to [0, 1) inside the chosen interval.
    size_t sampleReuse(float &x) const;

private:
    std::vector<float> m_cdf{0.0f};
    float m_sum = 0.0f, m_normalization = 0.0f;
    bool m_normalized = false;
};

struct Intersection {
    Point3f p;
    float t = 0.0f;
    Point2f uv;
    Frame shFrame, geoFrame;
    const Mesh *mesh = nullptr;
    uint32_t primitive = 0; ///< triangle index within the mesh
    Vector3f toLocal(const Vector3f &d) const { return shFrame.toLocal(d); }
    Vector3f toWorld(const Vector3f &d) const { return shFrame.toWorld(d); }
    std::string toString() const;
};

class Mesh : public NoriObject {
public:
    virtual ~Mesh();
    /// Default BSDF (diffuse) and the area pdf (src/cpu/mesh.cpp:26-40).
    void activate() override;
    uint32_t getTriangleCount() const { return (uint32_t)m_F.cols(); }
    uint32_t getVertexCount() const { return (uint32_t)m_V.cols(); }
    float surfaceArea(uint32_t index) const;
    /// Uniform point on the surface with the GPU path's barycentric mapping
    /// (1-sqrt(r1), sqrt(r1)(1-r2), sqrt(r1) r2; shaders/Emitter.hlsli:82-85).
    void samplePosition(const Point2f &sample, Point3f &p, Normal3f &n, float &pdf) const;
    const BoundingBox3f &getBoundingBox() const { return m_bbox; }
    BoundingBox3f getBoundingBox(uint32_t index) const;
    Point3f getCentroid(uint32_t index) const;
    const MatrixXf &getVertexPositions() const { return m_V; }
    const MatrixXf &getVertexNormals() const { return m_N; }
    const MatrixXf &getVertexTexCoords() const { return m_UV; }
    const MatrixXu &getIndices() const { return m_F; }
    const MatrixXf &getVertexTangents() const { return m_T; }
    bool hasTangents() const { return m_T.cols() > 0; }
    bool isEmitter() const { return m_emitter != nullptr; }
    Emitter *getEmitter() { return m_emitter; }
    const Emitter *getEmitter() const { return m_emitter; }
    const BSDF *getBSDF() const { return m_bsdf; }
    const DiscretePDF &getDiscretePDF() const { return m_dpdf; }
    void addChild(NoriObject *child) override;
    const std::string &getName() const { return m_name; }
    std::string toString() const override;
    EClassType getClassType() const override { return EMesh; }

protected:
    Mesh() {}
    std::string m_name;
    MatrixXf m_V;  ///< 3 x vertices
    MatrixXf m_N;  ///< 3 x vertices, or empty
    MatrixXf m_T;  ///< 3 x vertices fiber tangents (hair), or empty
    MatrixXf m_UV; ///< 2 x vertices, or empty
    MatrixXu m_F;  ///< 3 x triangles
    BSDF *m_bsdf = nullptr;
    Emitter *m_emitter = nullptr;
    BoundingBox3f m_bbox;
    DiscretePDF m_dpdf;
};

NORI_NAMESPACE_END
