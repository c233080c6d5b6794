// Host-side scene container and procedural generators.
#pragma once
#include "../../include/nori_gpu.h"
#include "nori_math.h"
#include <vector>

namespace ng {

struct SceneData {
    std::vector<float> pos, nrm, uv, tan;
    std::vector<uint32_t> idx, mesh_tri_offset, mesh_tri_count;
    std::vector<nori_material> mats;
    uint32_t env_w = 0, env_h = 0;
    float env_scale = 1.0f;
    std::vector<float> env_rgb;
    std::vector<nori_volume> vols;           // density pointers into grids
    std::vector<std::vector<float>> grids;   // one dense grid per heterogeneous volume
    std::vector<std::vector<uint8_t>> tex_data; // RGBA8 texels of each texture
    std::vector<nori_texture> texs;             // descriptors into tex_data (set by finalize)
    std::vector<uint32_t> tex_w, tex_h, tex_srgb;
    nori_camera camera{};

    void finalize();
    void fill_desc(nori_scene_desc &d) const;
};

void icosphere_mesh(uint32_t level, std::vector<float> &pos, std::vector<uint32_t> &tri);
SceneData generate_scene(const nori_scene_options &opt);

} // namespace ng
