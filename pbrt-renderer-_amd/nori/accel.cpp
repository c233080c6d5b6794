// GPU-backed Accel: packing as src/gpu/DXRApp_Scene.cpp:172-433 does it (one
// material record per mesh, global indices, UV V flipped except on hair meshes,
// zero tangents on non-hair meshes), upload through nori_gpu_upload_scene, and
// ray queries through nori_gpu_trace.
#include "accel.h"

#include "bsdf.h"
#include "emitter.h"
#include "gpu.h"
#include "image.h"
#include "medium.h"

#include <cstring>
#include <fstream>
#include <unordered_map>

NORI_NAMESPACE_BEGIN

Accel::~Accel() {
    if (m_ctx) nori_gpu_destroy(m_ctx);
    clear();
}

void Accel::clear() {
    for (auto *m : m_meshes) delete m;
    m_meshes.clear();
    m_bbox = BoundingBox3f();
}

void Accel::addMesh(Mesh *mesh) {
    m_meshes.push_back(mesh);
    m_bbox.expandBy(mesh->getBoundingBox());
}

void Accel::setEnvironment(uint32_t width, uint32_t height, std::vector<float> rgb, float scale) {
    m_envW = width;
    m_envH = height;
    m_env = std::move(rgb);
    m_envScale = scale;
}

namespace {
// VOL1 density file (src/gpu/DXRApp_Scene.cpp:583-615): "VOL1", W, H, D
// (uint32), a bounding box (6 floats, not used: the medium's bounds place the
// grid), then W * H * D floats, x fastest.
void readVol1(const std::string &path, uint32_t &W, uint32_t &H, uint32_t &D, std::vector<float> &data) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw NoriException("Unable to open volume file \"%s\"", path);
    char magic[4];
    f.read(magic, 4);
    if (!f || std::memcmp(magic, "VOL1", 4) != 0) throw NoriException("\"%s\" is not a VOL1 volume", path);
    uint32_t dims[3];
    float box[6];
    f.read(reinterpret_cast<char *>(dims), 12);
    f.read(reinterpret_cast<char *>(box), 24);
    W = dims[0];
    H = dims[1];
    D = dims[2];
    if (!f || W == 0 || H == 0 || D == 0 || (uint64_t)W * H * D > (1ull << 31))
        throw NoriException("\"%s\": bad VOL1 dimensions", path);
    data.resize((size_t)W * H * D);
    f.read(reinterpret_cast<char *>(data.data()), (std::streamsize)(data.size() * 4));
    if (!f) throw NoriException("\"%s\": truncated VOL1 data", path);
}
} // namespace

void Accel::setMedium(const Medium *medium) {
    m_vols.clear();
    m_volGrid.clear();
    if (!medium) return;
    BoundingBox3f box = medium->hasExplicitBounds() ? medium->getExplicitBounds() : m_bbox;
    Color3f sa = medium->getSigmaARGB(), ss = medium->getSigmaSRGB();
    nori_volume v;
    std::memset(&v, 0, sizeof(v));
    for (int k = 0; k < 3; ++k) {
        v.vmin[k] = box.min[k];
        v.vmax[k] = box.max[k];
        v.sigma_a[k] = sa[k];
        v.sigma_s[k] = ss[k];
    }
    v.g = medium->getPhaseG();
    v.heterogeneous = medium->isHeterogeneous() ? 1u : 0u;
    v.brick = 4;
    if (medium->isHeterogeneous() && !medium->getVolumePath().empty()) {
        readVol1(getFileResolver()->resolve(medium->getVolumePath()), v.width, v.height, v.depth, m_volGrid);
        v.density = m_volGrid.data();
    }
    m_vols.push_back(v);
}

void Accel::build() {
    m_pos.clear(); m_nrm.clear(); m_uv.clear(); m_tan.clear(); m_idx.clear();
    m_meshTriOffset.clear(); m_meshTriCount.clear(); m_mats.clear(); m_triOffset.clear();
    m_texPixels.clear(); m_texDesc.clear();
    // Texture files of the materials (src/gpu/DXRApp_Textures.cpp:452-486):
    // resolved against the scene's directory, one texture per resolved path
    // (the first use decides sRGB: albedo maps are sRGB, the others linear); a
    // file that does not load leaves the map unset.
    std::unordered_map<std::string, uint32_t> texCache;
    auto loadTexture = [&](const std::string &rel, bool srgb) -> uint32_t {
        if (rel.empty()) return NORI_NO_TEXTURE;
        const std::string key = getFileResolver()->resolve(rel);
        auto it = texCache.find(key);
        if (it != texCache.end()) return it->second;
        uint32_t w = 0, h = 0;
        std::vector<uint8_t> px;
        std::string err;
        if (!loadTextureRGBA8(key, w, h, px, err)) {
            cerr << "[texture] Failed to load: " << key << " (" << err << ")" << endl;
            return NORI_NO_TEXTURE;
        }
        const uint32_t idx = (uint32_t)m_texDesc.size();
        m_texPixels.push_back(std::move(px));
        m_texDesc.push_back(nori_texture{w, h, srgb ? 1u : 0u, nullptr});
        texCache[key] = idx;
        return idx;
    };
    bool anyNormals = false, anyUV = false, anyTan = false;
    for (auto *m : m_meshes) {
        anyNormals |= m->getVertexNormals().cols() > 0;
        anyUV |= m->getVertexTexCoords().cols() > 0;
        anyTan |= m->hasTangents();
    }
    uint32_t v0 = 0, t0 = 0;
    for (auto *m : m_meshes) {
        const MatrixXf &V = m->getVertexPositions(), &N = m->getVertexNormals(), &UV = m->getVertexTexCoords(),
                       &T = m->getVertexTangents();
        const MatrixXu &F = m->getIndices();
        const uint32_t nv = m->getVertexCount(), nt = m->getTriangleCount();
        for (uint32_t i = 0; i < nv; ++i) {
            m_pos.insert(m_pos.end(), {V(0, i), V(1, i), V(2, i)});
            if (anyNormals) {
                if (N.cols() == nv) m_nrm.insert(m_nrm.end(), {N(0, i), N(1, i), N(2, i)});
                else m_nrm.insert(m_nrm.end(), {0.0f, 0.0f, 0.0f});
            }
            if (anyUV) {
                if (UV.cols() == nv) m_uv.insert(m_uv.end(), {UV(0, i), m->hasTangents() ? UV(1, i) : 1.0f - UV(1, i)});
                else m_uv.insert(m_uv.end(), {0.0f, 0.0f});
            }
            if (anyTan) {
                if (T.cols() == nv) m_tan.insert(m_tan.end(), {T(0, i), T(1, i), T(2, i)});
                else m_tan.insert(m_tan.end(), {0.0f, 0.0f, 0.0f});
            }
        }
        for (uint32_t t = 0; t < nt; ++t)
            for (int k = 0; k < 3; ++k) {
                uint32_t vi = F(k, t);
                if (vi >= nv) throw NoriException("Mesh \"%s\": vertex index out of range", m->getName());
                m_idx.push_back(v0 + vi);
            }
        const BSDFGPUData gd = m->getBSDF()->getGPUData();
        nori_material mat = toGPUMaterial(gd);
        mat.albedoTexIndex = loadTexture(gd.albedoTexture, true);
        mat.normalTexIndex = loadTexture(gd.normalTexture, false);
        mat.roughnessTexIndex = loadTexture(gd.roughnessTexture, false);
        mat.metallicTexIndex = loadTexture(gd.metallicTexture, false);
        mat.specularTexIndex = loadTexture(gd.specularTexture, false);
        mat.subsurfaceTexIndex = loadTexture(gd.subsurfaceTexture, false);
        mat.alphaTexIndex = loadTexture(gd.alphaTexture, false);
        if (m->isEmitter()) {
            Color3f L = m->getEmitter()->getRadiance();
            mat.isEmitter = 1;
            for (int k = 0; k < 3; ++k) mat.radiance[k] = L[k];
        }
        mat.vertexOffset = v0;
        mat.vertexCount = nv;
        m_mats.push_back(mat);
        m_triOffset.push_back(t0);
        m_meshTriOffset.push_back(t0);
        m_meshTriCount.push_back(nt);
        v0 += nv;
        t0 += nt;
    }
    std::memset(&m_desc, 0, sizeof(m_desc));
    m_desc.num_vertices = v0;
    m_desc.num_triangles = t0;
    m_desc.num_meshes = (uint32_t)m_meshes.size();
    m_desc.positions = m_pos.data();
    m_desc.normals = anyNormals ? m_nrm.data() : nullptr;
    m_desc.uvs = anyUV ? m_uv.data() : nullptr;
    m_desc.tangents = anyTan ? m_tan.data() : nullptr;
    m_desc.indices = m_idx.data();
    m_desc.mesh_tri_offset = m_meshTriOffset.data();
    m_desc.mesh_tri_count = m_meshTriCount.data();
    m_desc.materials = m_mats.data();
    m_desc.env_width = m_env.empty() ? 0 : m_envW;
    m_desc.env_height = m_env.empty() ? 0 : m_envH;
    m_desc.env_rgb = m_env.empty() ? nullptr : m_env.data();
    m_desc.env_scale = m_envScale;
    for (size_t k = 0; k < m_texDesc.size(); ++k) m_texDesc[k].rgba = m_texPixels[k].data();
    m_desc.num_textures = (uint32_t)m_texDesc.size();
    m_desc.textures = m_texDesc.empty() ? nullptr : m_texDesc.data();
    m_desc.num_volumes = (uint32_t)m_vols.size();
    m_desc.volumes = m_vols.empty() ? nullptr : m_vols.data();
    m_uploaded = false;
}

nori_ctx *Accel::context() const {
    if (m_desc.num_triangles == 0) throw NoriException("Accel: the scene has no triangles");
    if (!m_ctx) {
        int rc = nori_gpu_create(&m_ctx, 0);
        if (rc != NORI_OK) {
            m_ctx = nullptr;
            throw NoriException("No GPU available for the Nori GPU path (nori_gpu_create returned %i)", rc);
        }
    }
    if (!m_uploaded) {
        gpuCheck(nori_gpu_upload_scene(m_ctx, &m_desc), "nori_gpu_upload_scene", m_ctx);
        m_uploaded = true;
    }
    return m_ctx;
}

void Accel::fillIntersection(const Ray3f &ray, const nori_hit &h, Intersection &its) const {
    // mesh of the global triangle id
    auto it = std::upper_bound(m_triOffset.begin(), m_triOffset.end(), h.prim);
    size_t mi = (size_t)(it - m_triOffset.begin()) - 1;
    const Mesh *mesh = m_meshes[mi];
    uint32_t f = h.prim - m_triOffset[mi];
    its.mesh = mesh;
    its.primitive = f;
    its.t = h.t;
    its.p = ray(h.t);
    const MatrixXu &F = mesh->getIndices();
    uint32_t i0 = F(0, f), i1 = F(1, f), i2 = F(2, f);
    float b0 = 1.0f - h.u - h.v, b1 = h.u, b2 = h.v;
    const MatrixXf &V = mesh->getVertexPositions(), &N = mesh->getVertexNormals(), &UV = mesh->getVertexTexCoords();
    Point3f p0 = V.col3(i0), p1 = V.col3(i1), p2 = V.col3(i2);
    if (UV.cols() > 0)
        its.uv = Point2f(b0 * UV(0, i0) + b1 * UV(0, i1) + b2 * UV(0, i2), b0 * UV(1, i0) + b1 * UV(1, i1) + b2 * UV(1, i2));
    else
        its.uv = Point2f(b1, b2);
    its.geoFrame = Frame((p1 - p0).cross(p2 - p0).normalized());
    if (N.cols() > 0) {
        Vector3f n = N.col3(i0) * b0 + N.col3(i1) * b1 + N.col3(i2) * b2;
        its.shFrame = n.squaredNorm() > 1e-16f ? Frame(n.normalized()) : its.geoFrame;
    } else {
        its.shFrame = its.geoFrame;
    }
}

bool Accel::rayIntersect(const Ray3f &ray, Intersection &its, bool shadowRay) const {
    std::vector<Ray3f> rays{ray};
    std::vector<Intersection> out;
    std::vector<uint8_t> hit;
    rayIntersect(rays, out, hit, shadowRay);
    if (hit[0] && !shadowRay) its = out[0];
    return hit[0] != 0;
}

void Accel::rayIntersect(const std::vector<Ray3f> &rays, std::vector<Intersection> &its, std::vector<uint8_t> &hits,
                         bool shadowRay) const {
    const size_t n = rays.size();
    hits.assign(n, 0);
    its.assign(shadowRay ? 0 : n, Intersection());
    if (n == 0 || m_desc.num_triangles == 0) return;
    nori_ctx *ctx = context();
    std::vector<nori_ray> rr(n);
    for (size_t i = 0; i < n; ++i) {
        const Ray3f &r = rays[i];
        // adaptive epsilon (src/cpu/accel.cpp:435-438): a ray left at the default
        // mint = Epsilon gets max(Epsilon, Epsilon * max|o|)
        float mint = r.mint;
        if (mint == Epsilon) {
            float mo = std::max(std::fabs(r.o[0]), std::max(std::fabs(r.o[1]), std::fabs(r.o[2])));
            mint = std::max(mint, mint * mo);
        }
        rr[i] = nori_ray{{r.o[0], r.o[1], r.o[2]}, mint, {r.d[0], r.d[1], r.d[2]}, r.maxt};
    }
    std::vector<nori_hit> hh(n);
    gpuCheck(nori_gpu_trace(ctx, rr.data(), (int)n, hh.data(), shadowRay ? 1 : 0), "nori_gpu_trace", ctx);
    for (size_t i = 0; i < n; ++i) {
        hits[i] = hh[i].prim != 0xFFFFFFFFu;
        if (hits[i] && !shadowRay) fillIntersection(rays[i], hh[i], its[i]);
    }
}

NORI_NAMESPACE_END
