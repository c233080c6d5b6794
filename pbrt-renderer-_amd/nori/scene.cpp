// Scene plugin (src/cpu/scene.cpp semantics) with the environment map of the
// reference's GPU renderer: the `envmap` file (resolved against the scene
// file's directory) scaled by `envmapScale`, or, when it cannot be read, a flat
// grey of 0.5 (src/gpu/DXRApp_Textures.cpp:181-213).
#include "scene.h"

#include "camera.h"
#include "image.h"
#include "integrator.h"
#include "sampler.h"

#include <fstream>
#include <sstream>

NORI_NAMESPACE_BEGIN

Scene::Scene(const PropertyList &p) {
    m_envmap = p.getString("envmap", "textures/white_furnace.hdr");
    m_envmapScale = p.getFloat("envmapScale", 1.0f);
    m_evCompensation = p.getFloat("evCompensation", 0.0f);
    m_accel = new Accel();
}

Scene::~Scene() {
    delete m_accel;
    delete m_sampler;
    delete m_camera;
    delete m_integrator;
    delete m_medium;
}

void Scene::activate() {
    if (!m_camera) throw NoriException("No camera was specified!");
    if (!m_sampler) m_sampler = static_cast<Sampler *>(NoriObjectFactory::createInstance("independent", PropertyList()));
    ImageRGB env;
    std::string path = getFileResolver()->resolve(m_envmap);
    std::ifstream probe(path);
    m_envmapLoaded = false;
    if (probe.good()) {
        probe.close();
        env = readImage(path);
        m_envmapLoaded = true;
    } else {
        cerr << "[envmap] No HDR at '" << path << "' - falling back to flat gray." << endl;
        env.width = env.height = 1;
        env.rgb = {0.5f, 0.5f, 0.5f};
    }
    m_accel->setEnvironment(env.width, env.height, std::move(env.rgb), m_envmapScale);
    m_accel->setMedium(m_medium);
    m_accel->build();
    if (!m_integrator) cerr << "Warning: No integrator was specified." << endl;
}

void Scene::addChild(NoriObject *obj) {
    switch (obj->getClassType()) {
    case EMesh: {
        Mesh *mesh = static_cast<Mesh *>(obj);
        m_accel->addMesh(mesh);
        m_meshes.push_back(mesh);
    } break;
    case EEmitter:
        delete obj; // a free emitter is ignored, as the reference does
        break;
    case ESampler:
        if (m_sampler) throw NoriException("There can only be one sampler per scene!");
        m_sampler = static_cast<Sampler *>(obj);
        break;
    case ECamera:
        if (m_camera) throw NoriException("There can only be one camera per scene!");
        m_camera = static_cast<Camera *>(obj);
        break;
    case EIntegrator:
        if (m_integrator) throw NoriException("There can only be one integrator per scene!");
        m_integrator = static_cast<Integrator *>(obj);
        break;
    case EMedium:
        if (m_medium) throw NoriException("There can only be one medium per scene!");
        m_medium = static_cast<Medium *>(obj);
        break;
    default:
        throw NoriException("Scene::addChild(<%s>) is not supported!", classTypeName(obj->getClassType()));
    }
}

std::string Scene::toString() const {
    // one "name = value" line per member; nested descriptions indented under it
    auto describe = [](const NoriObject *o) { return o ? indent(o->toString(), 2) : std::string("null"); };
    std::ostringstream os;
    os << "Scene[\n"
       << "  integrator = " << describe(m_integrator) << ",\n"
       << "  sampler = " << describe(m_sampler) << ",\n"
       << "  camera = " << describe(m_camera) << ",\n"
       << "  medium = " << describe(m_medium) << ",\n"
       << "  envmap = \"" << m_envmap << "\",\n"
       << "  meshes (" << m_meshes.size() << ") = [";
    for (size_t i = 0; i < m_meshes.size(); ++i)
        os << (i ? ",\n    " : "\n    ") << indent(m_meshes[i]->toString(), 4);
    os << (m_meshes.empty() ? "]\n]" : "\n  ]\n]");
    return os.str();
}

NORI_REGISTER_CLASS(Scene, "scene");

NORI_NAMESPACE_END
