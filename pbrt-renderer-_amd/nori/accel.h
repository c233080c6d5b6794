// Acceleration structure of the plugin surface (include/accel.h:62-83 of the
// reference): addMesh / build / rayIntersect.  build() packs every mesh into the
// flat SoA layout of the GPU path; the first GPU use uploads it (the library
// builds the BVH8 there); rayIntersect traces on the GPU.
#pragma once
#include "../../include/nori_gpu.h"
#include "mesh.h"

NORI_NAMESPACE_BEGIN

class Accel {
public:
    Accel() {}
    ~Accel();
    Accel(const Accel &) = delete;
    Accel &operator=(const Accel &) = delete;

    void addMesh(Mesh *mesh);
    /// Equirectangular RGB environment (width x height x 3 floats, row 0 = +y),
    /// uploaded with the geometry; empty: none.
    void setEnvironment(uint32_t width, uint32_t height, std::vector<float> rgb, float scale);
    /// The scene's participating medium (src/gpu/DXRApp_Scene.cpp:452-492 and
    /// src/gpu/DXRApp_Textures.cpp:495-518): its box is the explicit bounds or
    /// the scene's bounding box; a heterogeneous medium with a `volume` file
    /// reads its VOL1 density grid (majorant bricks of 4 voxels).
    void setMedium(const class Medium *medium);
    void build();
    void clear();
    const BoundingBox3f &getBoundingBox() const { return m_bbox; }

    /// Closest hit (shadowRay = false) or any hit (true), on the GPU.
    bool rayIntersect(const Ray3f &ray, Intersection &its, bool shadowRay) const;
    /// Batch form: hits[i] = whether ray i hit; its[i] filled for closest hits.
    void rayIntersect(const std::vector<Ray3f> &rays, std::vector<Intersection> &its, std::vector<uint8_t> &hits,
                      bool shadowRay) const;

    /// The GPU context holding this scene (created and uploaded on first use;
    /// throws NoriException without a GPU).
    nori_ctx *context() const;
    const std::vector<Mesh *> &meshes() const { return m_meshes; }
    /// The flat scene uploaded to the GPU (valid after build()).
    const nori_scene_desc &desc() const { return m_desc; }

private:
    void fillIntersection(const Ray3f &ray, const nori_hit &h, Intersection &its) const;

    std::vector<Mesh *> m_meshes;
    std::vector<uint32_t> m_triOffset; // first global triangle of each mesh
    BoundingBox3f m_bbox;
    mutable nori_ctx *m_ctx = nullptr;
    mutable bool m_uploaded = false;
    // packed arrays (kept alive for desc())
    std::vector<float> m_pos, m_nrm, m_uv, m_tan, m_env;
    std::vector<uint32_t> m_idx, m_meshTriOffset, m_meshTriCount;
    std::vector<nori_material> m_mats;
    std::vector<std::vector<uint8_t>> m_texPixels; // decoded RGBA8 texture files
    std::vector<nori_texture> m_texDesc;
    std::vector<nori_volume> m_vols;
    std::vector<float> m_volGrid;
    uint32_t m_envW = 0, m_envH = 0;
    float m_envScale = 1.0f;
    nori_scene_desc m_desc{};
};

NORI_NAMESPACE_END
