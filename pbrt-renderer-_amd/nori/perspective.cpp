// "perspective" camera (src/cpu/perspective.cpp:22-178 semantics: horizontal
// fov, near/far clip, thin lens) and the "gaussian" / "box" reconstruction
// filters (src/cpu/rfilter.cpp).  The GPU path's film is box-filtered over the
// pixel (jittered sample positions), which is also how the reference's GPU
// renderer accumulates; the filter object is kept for the plugin surface.
#include "camera.h"

NORI_NAMESPACE_BEGIN

class GaussianFilter : public ReconstructionFilter {
public:
    explicit GaussianFilter(const PropertyList &p) {
        float r = p.getFloat("radius", 2.0f);
        m_radius = Vector2f(r, r);
        m_stddev = p.getFloat("stddev", 0.5f);
        m_alpha = std::exp(-m_stddev * m_stddev * r * r);
    }
    float eval(float x) const override { return std::max(0.0f, std::exp(-m_stddev * m_stddev * x * x) - m_alpha); }
    std::string toString() const override { return format("GaussianFilter[radius=%f, stddev=%f]", m_radius.x(), m_stddev); }

private:
    float m_stddev, m_alpha;
};

class BoxFilter : public ReconstructionFilter {
public:
    explicit BoxFilter(const PropertyList &) { m_radius = Vector2f(0.5f, 0.5f); }
    float eval(float) const override { return 1.0f; }
    std::string toString() const override { return "BoxFilter[]"; }
};

class PerspectiveCamera : public Camera {
public:
    explicit PerspectiveCamera(const PropertyList &p) {
        m_outputSize.x() = p.getInteger("width", 1280);
        m_outputSize.y() = p.getInteger("height", 720);
        m_cameraToWorld = p.getTransform("toWorld", Transform());
        m_fov = p.getFloat("fov", 30.0f);
        m_nearClip = p.getFloat("nearClip", 1e-4f);
        m_farClip = p.getFloat("farClip", 1e4f);
        m_lensRadius = p.getFloat("lensRadius", 0.0f);
        m_focalDistance = p.getFloat("focalDistance", 1.0f);
        if (m_outputSize.x() <= 0 || m_outputSize.y() <= 0)
            throw NoriException("PerspectiveCamera: invalid image size %ix%i", m_outputSize.x(), m_outputSize.y());
    }

    void activate() override {
        if (!m_rfilter)
            m_rfilter = static_cast<ReconstructionFilter *>(NoriObjectFactory::createInstance("gaussian", PropertyList()));
    }
    ~PerspectiveCamera() override { delete m_rfilter; }

    // Sample position (pixels) -> camera-space direction through the z = 1
    // plane: x = (1 - 2 sx) tan(fov/2), y = (1 - 2 sy) tan(fov/2) / aspect.
    Color3f sampleRay(Ray3f &ray, const Point2f &samplePosition, const Point2f &apertureSample) const override {
        float sx = samplePosition.x() / (float)m_outputSize.x(), sy = samplePosition.y() / (float)m_outputSize.y();
        float aspect = (float)m_outputSize.x() / (float)m_outputSize.y();
        float th = std::tan(degToRad(m_fov * 0.5f));
        Vector3f d = Vector3f((1.0f - 2.0f * sx) * th, (1.0f - 2.0f * sy) * th / aspect, 1.0f).normalized();
        Point3f o(0.0f);
        if (m_lensRadius > 0.0f) {
            float r = std::sqrt(apertureSample.x()), phi = 2.0f * M_PI_F * apertureSample.y();
            Point3f lens(m_lensRadius * r * std::cos(phi), m_lensRadius * r * std::sin(phi), 0.0f);
            Point3f focus = d * (m_focalDistance / d.z());
            o = lens;
            d = (focus - lens).normalized();
        }
        float invZ = 1.0f / d.z();
        ray.o = m_cameraToWorld.transformPoint(o);
        ray.d = m_cameraToWorld.transformVector(d);
        ray.mint = m_nearClip * invZ;
        ray.maxt = m_farClip * invZ;
        ray.update();
        return Color3f(1.0f);
    }

    nori_camera getGPUCamera() const override {
        nori_camera c;
        for (int k = 0; k < 16; ++k) c.to_world[k] = m_cameraToWorld.matrix()[k];
        c.fov_deg = m_fov;
        c.width = (uint32_t)m_outputSize.x();
        c.height = (uint32_t)m_outputSize.y();
        c.near_clip = m_nearClip;
        c.far_clip = m_farClip;
        c.lens_radius = m_lensRadius;
        c.focal_distance = m_focalDistance;
        return c;
    }

    float getLensRadius() const override { return m_lensRadius; }
    float getFocalDistance() const override { return m_focalDistance; }

    void addChild(NoriObject *obj) override {
        if (obj->getClassType() != EReconstructionFilter)
            throw NoriException("Camera::addChild(<%s>) is not supported!", classTypeName(obj->getClassType()));
        if (m_rfilter) throw NoriException("Camera: tried to register multiple reconstruction filters!");
        m_rfilter = static_cast<ReconstructionFilter *>(obj);
    }

    std::string toString() const override {
        return format("PerspectiveCamera[\n  cameraToWorld = %s,\n  outputSize = [%i, %i],\n  fov = %f,\n  clip = [%f, %f],"
                      "\n  lensRadius = %f,\n  focalDistance = %f\n]",
                      indent(m_cameraToWorld.toString(), 18), m_outputSize.x(), m_outputSize.y(), m_fov, m_nearClip,
                      m_farClip, m_lensRadius, m_focalDistance);
    }

private:
    Transform m_cameraToWorld;
    float m_fov, m_nearClip, m_farClip, m_lensRadius, m_focalDistance;
};

NORI_REGISTER_CLASS(PerspectiveCamera, "perspective");
NORI_REGISTER_CLASS(GaussianFilter, "gaussian");
NORI_REGISTER_CLASS(BoxFilter, "box");

NORI_NAMESPACE_END
