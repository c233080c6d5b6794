// CPU oracle — test infrastructure only.
//
// The reference has no CPU integrator (nothing registers an Integrator; its path
// integrator exists only as the HLSL RayGen megakernel, shaders/Shaders.hlsl:25-651).
// This is a CPU restatement of RayGen as the `path_mis` integrator, run over
// 32x32 image blocks (NORI_BLOCK_SIZE, include/block.h:17) on a std::thread pool.
//
// Sampling conventions chosen here rather than taken from the HLSL (it draws from an
// Owen-scrambled Sobol' stream; shaders/RNG.hlsli:80-98):
//   * one pcg32 stream per pixel, seed(render_seed ^ pixel, PCG32_DEFAULT_STREAM),
//     sample s starting at advance(s << 20);
//   * camera: Nori's perspective camera (src/cpu/perspective.cpp:50-123) with a
//     box-filtered jittered pixel sample;
//   * envmap next-event estimation draws from the alias table (one 32-bit integer,
//     a coin, two jitters) or, with env_cdf=1, from the reference 2D CDF;
//   * a scene without an envmap contributes no environment light.
#pragma once
#include "orc_bsdf.h"
#include "orc_env.h"
#include "orc_geom.h"
#include "orc_hair.h"
#include "orc_pcg32.h"
#include "orc_volume.h"
#include "orc_sobol.h"
#include "orc_texture.h"

namespace orc {

struct Camera {
    float to_world[16]; // row-major camera-to-world
    float fov_deg;      // horizontal
    uint32_t width, height;
    float near_clip, far_clip, lens_radius, focal_distance;
};

struct SceneData {
    uint32_t nv = 0, nt = 0, nmesh = 0;
    std::vector<float> pos, nrm;      // nv*3 (nrm may be empty)
    std::vector<float> uv, tan;       // nv*2, nv*3 (may be empty)
    std::vector<uint32_t> idx;        // nt*3, global vertex indices
    std::vector<uint32_t> tri_mesh;   // nt
    std::vector<MatRec> mats;         // nmesh
    std::vector<uint32_t> mesh_tri_offset, mesh_tri_count;
    Bvh2 bvh;
    Envmap env;
    std::vector<uint32_t> emitters;            // mesh ids with isEmitter
    std::vector<std::vector<float>> emit_cdf;  // per emitter: Nori DiscretePDF cdf (tris+1)
    std::vector<float> emit_area;              // per emitter: total area
    std::vector<Volume> vols;                  // participating media
    bool has_pass = false;                     // mirror/dielectric meshes present (shadow rays pass them)
    std::vector<Texture> tex;                  // scene textures (mip chains)
    bool has_alpha = false;                    // some mesh has an alpha texture

    void finalize(); // BVH + emitter tables from the raw arrays
    V3f vtx(uint32_t i) const { return V3f(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]); }
};

struct RenderParams {
    uint32_t spp, sample_offset, max_depth, rr_depth;
    uint64_t seed;
    int env_cdf;   // 1: sample the envmap with the reference CDF instead of the alias table
    int nthreads;
    int sampler;   // 0: pcg32 per pixel; 1: Owen-scrambled Sobol' (shaders/RNG.hlsli)
};

struct RenderStats {
    uint64_t extension_rays = 0, shadow_rays = 0, samples = 0;
};

// Denoiser features of one sample (shaders/Shaders.hlsl:99-106, 630-638):
// albedo and normal at the first non-delta interaction (the envmap radiance
// and a zero normal for a miss, white and the shading normal for an emitter).
struct Aov {
    V3f albedo{0.0f}, normal{0.0f};
    bool done = false;
    void capture(V3f a, V3f n) {
        if (done) return;
        albedo = a;
        normal = n;
        done = true;
    }
};

void camera_ray(const Camera &cam, float px, float py, float ax, float ay, Ray &ray);
V3f path_mis_li(const SceneData &s, const Camera &cam, Ray ray, Pcg32 &rng, const RenderParams &p, RenderStats &st);
// accum: RGBA per pixel (radiance sum, sample count); aov_albedo / aov_normal
// (optional): the same for the denoiser features.
void render(const SceneData &s, const Camera &cam, const RenderParams &p, float *accum, RenderStats *stats,
            float *aov_albedo = nullptr, float *aov_normal = nullptr);

} // namespace orc
