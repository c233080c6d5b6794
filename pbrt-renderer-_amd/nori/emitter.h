// Emitter plugin interface (include/emitter.h:11-31 of the reference).
#pragma once
#include "mesh.h"

NORI_NAMESPACE_BEGIN

class Emitter : public NoriObject {
public:
    virtual Color3f getRadiance() const = 0;
    /// Samples a point on the emitter: position, normal and area pdf.
    virtual Color3f sample(const Point2f &sample, Point3f &p, Normal3f &n, float &pdf) const = 0;
    virtual float pdf() const = 0;
    void setMesh(Mesh *mesh) { m_mesh = mesh; }
    const Mesh *getMesh() const { return m_mesh; }
    EClassType getClassType() const override { return EEmitter; }

protected:
    Mesh *m_mesh = nullptr;
};

NORI_NAMESPACE_END
