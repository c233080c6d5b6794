// C entry points of libnori_host.so (include/nori_host.h).
#include "../../include/nori_host.h"

#include "camera.h"
#include "emitter.h"
#include "image.h"
#include "integrator.h"
#include "parser.h"
#include "sampler.h"
#include "scene.h"

#include <cstdlib>
#include <cstring>

using namespace nori;

struct nori_host_scene {
    std::unique_ptr<Scene> scene;
};

namespace {

void copyErr(const std::string &m, char *err, int errlen) {
    if (!err || errlen <= 0) return;
    std::strncpy(err, m.c_str(), (size_t)errlen - 1);
    err[errlen - 1] = '\0';
}

template <typename F> int guarded(char *err, int errlen, F &&f) {
    try {
        f();
        return NORI_OK;
    } catch (const NoriException &e) {
        copyErr(e.what(), err, errlen);
        return NORI_E_ARG;
    } catch (const std::exception &e) {
        copyErr(e.what(), err, errlen);
        return NORI_E_STATE;
    }
}

nori_host_scene *wrap(NoriObject *obj) {
    if (!obj || obj->getClassType() != NoriObject::EScene) {
        delete obj;
        throw NoriException("The root element of the scene file is not a <scene>");
    }
    auto *h = new nori_host_scene;
    h->scene.reset(static_cast<Scene *>(obj));
    return h;
}

} // namespace

extern "C" {

int nori_xml_load(const char *filename, nori_host_scene **out, char *err, int errlen) {
    if (!filename || !out) return NORI_E_ARG;
    *out = nullptr;
    return guarded(err, errlen, [&] { *out = wrap(loadFromXML(filename)); });
}

int nori_xml_load_string(const char *text, const char *dir, nori_host_scene **out, char *err, int errlen) {
    if (!text || !out) return NORI_E_ARG;
    *out = nullptr;
    return guarded(err, errlen,
                   [&] { *out = wrap(loadFromXMLString(text, dir ? dir : ".", "<string>")); });
}

void nori_xml_free(nori_host_scene *s) { delete s; }

int nori_xml_info_get(const nori_host_scene *s, nori_xml_info *info) {
    if (!s || !info) return NORI_E_ARG;
    std::memset(info, 0, sizeof(*info));
    const Scene &sc = *s->scene;
    info->width = (uint32_t)sc.getCamera()->getOutputSize().x();
    info->height = (uint32_t)sc.getCamera()->getOutputSize().y();
    info->spp = (uint32_t)sc.getSampler()->getSampleCount();
    info->meshes = (uint32_t)sc.getMeshes().size();
    for (auto *m : sc.getMeshes()) {
        info->triangles += m->getTriangleCount();
        info->vertices += m->getVertexCount();
        info->emitters += m->isEmitter() ? 1 : 0;
    }
    info->has_medium = sc.getMedium() ? 1 : 0;
    info->envmap_loaded = sc.envmapLoaded() ? 1 : 0;
    if (sc.getIntegrator()) {
        std::string n = sc.getIntegrator()->toString();
        std::strncpy(info->integrator, n.c_str(), sizeof(info->integrator) - 1);
    }
    return NORI_OK;
}

int nori_xml_desc(const nori_host_scene *s, nori_scene_desc *desc, nori_camera *camera) {
    if (!s) return NORI_E_ARG;
    if (desc) *desc = s->scene->getAccel()->desc();
    if (camera) *camera = s->scene->getCamera()->getGPUCamera();
    return NORI_OK;
}

int nori_xml_render(nori_host_scene *s, float *rgb, char *err, int errlen) {
    if (!s || !rgb) return NORI_E_ARG;
    return guarded(err, errlen, [&] {
        Scene &sc = *s->scene;
        if (!sc.getIntegrator()) throw NoriException("The scene has no integrator");
        sc.getIntegrator()->preprocess(&sc);
        std::vector<float> img;
        sc.getIntegrator()->render(&sc, img);
        std::memcpy(rgb, img.data(), img.size() * sizeof(float));
    });
}

int nori_xml_li(nori_host_scene *s, const nori_ray *rays, int n, uint64_t seed, float *rgb, char *err, int errlen) {
    if (!s || n < 0 || (n > 0 && (!rays || !rgb))) return NORI_E_ARG;
    return guarded(err, errlen, [&] {
        Scene &sc = *s->scene;
        if (!sc.getIntegrator()) throw NoriException("The scene has no integrator");
        sc.getIntegrator()->preprocess(&sc);
        std::unique_ptr<Sampler> sampler = sc.getSampler()->clone();
        ImageBlock block(Vector2i((int)(seed & 0x7fffffff), (int)(seed >> 32)), Vector2i(1, 1));
        sampler->prepare(block);
        std::vector<Ray3f> rs((size_t)n);
        for (int i = 0; i < n; ++i) {
            const nori_ray &r = rays[i];
            rs[(size_t)i] = Ray3f(Point3f(r.o[0], r.o[1], r.o[2]), Vector3f(r.d[0], r.d[1], r.d[2]), r.mint, r.maxt);
        }
        std::vector<Color3f> L((size_t)n);
        sc.getIntegrator()->LiBatch(&sc, sampler.get(), rs.data(), rs.size(), L.data());
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c) rgb[3 * i + c] = L[(size_t)i][c];
    });
}

int nori_image_write(const char *filename, const float *rgb, uint32_t width, uint32_t height, char *err, int errlen) {
    if (!filename || !rgb || width == 0 || height == 0) return NORI_E_ARG;
    return guarded(err, errlen, [&] {
        ImageRGB img;
        img.width = width;
        img.height = height;
        img.rgb.assign(rgb, rgb + (size_t)width * height * 3);
        writeImage(filename, img);
    });
}

int nori_image_read(const char *filename, float **rgb, uint32_t *width, uint32_t *height, char *err, int errlen) {
    if (!filename || !rgb || !width || !height) return NORI_E_ARG;
    *rgb = nullptr;
    return guarded(err, errlen, [&] {
        ImageRGB img = readImage(filename);
        float *buf = (float *)std::malloc(img.rgb.size() * sizeof(float));
        if (!buf) throw NoriException("out of memory");
        std::memcpy(buf, img.rgb.data(), img.rgb.size() * sizeof(float));
        *rgb = buf;
        *width = img.width;
        *height = img.height;
    });
}

void nori_image_free(float *rgb) { std::free(rgb); }

int nori_plugin_names(char *buf, int len) {
    if (!buf || len <= 0) return NORI_E_ARG;
    std::string s;
    for (auto &n : NoriObjectFactory::registeredNames()) s += (s.empty() ? "" : ",") + n;
    copyErr(s, buf, len);
    return (int)s.size() < len ? NORI_OK : NORI_E_LIMIT;
}

} // extern "C"
