// Scene container plugin (include/scene.h:16-93 of the reference).
#pragma once
#include "accel.h"
#include "medium.h"

NORI_NAMESPACE_BEGIN

class Integrator;
class Sampler;
class Camera;

class Scene : public NoriObject {
public:
    explicit Scene(const PropertyList &);
    virtual ~Scene();
    const Accel *getAccel() const { return m_accel; }
    const Integrator *getIntegrator() const { return m_integrator; }
    Integrator *getIntegrator() { return m_integrator; }
    const Camera *getCamera() const { return m_camera; }
    const Sampler *getSampler() const { return m_sampler; }
    Sampler *getSampler() { return m_sampler; }
    const std::vector<Mesh *> &getMeshes() const { return m_meshes; }
    const Medium *getMedium() const { return m_medium; }
    bool rayIntersect(const Ray3f &ray, Intersection &its) const { return m_accel->rayIntersect(ray, its, false); }
    bool rayIntersect(const Ray3f &ray) const {
        Intersection its;
        return m_accel->rayIntersect(ray, its, true);
    }
    const BoundingBox3f &getBoundingBox() const { return m_accel->getBoundingBox(); }
    /// Loads the environment map (or the flat 0.5 grey fallback), packs and
    /// uploads the scene, creates the default sampler.
    void activate() override;
    void addChild(NoriObject *obj) override;
    std::string toString() const override;
    EClassType getClassType() const override { return EScene; }
    const std::string &getEnvmap() const { return m_envmap; }
    float getEnvmapScale() const { return m_envmapScale; }
    float getEvCompensation() const { return m_evCompensation; }
    /// Whether the environment came from a file (false: the flat grey fallback).
    bool envmapLoaded() const { return m_envmapLoaded; }

private:
    std::vector<Mesh *> m_meshes;
    Integrator *m_integrator = nullptr;
    Sampler *m_sampler = nullptr;
    Camera *m_camera = nullptr;
    Accel *m_accel = nullptr;
    Medium *m_medium = nullptr;
    std::string m_envmap = "textures/white_furnace.hdr";
    float m_envmapScale = 1.0f;
    float m_evCompensation = 0.0f;
    bool m_envmapLoaded = false;
};

NORI_NAMESPACE_END
