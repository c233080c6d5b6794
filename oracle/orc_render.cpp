// CPU oracle — test infrastructure only.  See orc_render.h.
#include "orc_render.h"
#include <atomic>
#include <cstring>
#include <thread>

namespace orc {

static const float PI_F = 3.14159265358979323846f;

void SceneData::finalize() {
    bvh.build(pos.data(), idx.data(), nt);
    has_pass = false;
    has_alpha = false;
    for (uint32_t m = 0; m < nmesh; ++m) {
        has_pass = has_pass || mats[m].type == 1 || mats[m].type == 2;
        has_alpha = has_alpha || (mats[m].alphaTex != 0xFFFFFFFFu && !tex.empty());
    }
    emitters.clear(); emit_cdf.clear(); emit_area.clear();
    for (uint32_t m = 0; m < nmesh; ++m) {
        if (!mats[m].isEmitter) continue;
        // include/dpdf.h append/normalize over triangle areas (src/cpu/mesh.cpp:32-48)
        std::vector<float> cdf(1, 0.0f);
        for (uint32_t k = 0; k < mesh_tri_count[m]; ++k) {
            uint32_t f = mesh_tri_offset[m] + k;
            V3f p0 = vtx(idx[3 * f]), p1 = vtx(idx[3 * f + 1]), p2 = vtx(idx[3 * f + 2]);
            float a = 0.5f * length(cross(p1 - p0, p2 - p0));
            cdf.push_back(cdf.back() + a);
        }
        float sum = cdf.back();
        if (sum > 0) {
            float norm = 1.0f / sum;
            for (size_t i = 1; i < cdf.size(); ++i) cdf[i] *= norm;
            cdf.back() = 1.0f;
        }
        emitters.push_back(m);
        emit_cdf.push_back(cdf);
        emit_area.push_back(sum);
    }
}

// Nori's perspective camera, src/cpu/perspective.cpp:50-123, with the
// sample-to-camera matrix written in closed form: a sample (sx, sy) in [0,1]^2
// maps to the camera-space direction ((1-2sx)tan(fov/2), (1-2sy)tan(fov/2)/aspect, 1).
void camera_ray(const Camera &cam, float px, float py, float ax, float ay, Ray &ray) {
    float sx = px * (1.0f / (float)cam.width), sy = py * (1.0f / (float)cam.height);
    float aspect = (float)cam.width / (float)cam.height;
    float th = std::tan(cam.fov_deg * (PI_F / 180.0f) * 0.5f);
    V3f d = normalize(V3f((1.0f - 2.0f * sx) * th, (1.0f - 2.0f * sy) * th / aspect, 1.0f));
    V3f o(0.0f);
    if (cam.lens_radius > 0.0f) {
        float r = std::sqrt(ax), phi = 2.0f * PI_F * ay; // squareToUniformDisk, src/cpu/warp.cpp:42-47
        V3f pl(cam.lens_radius * (r * std::cos(phi)), cam.lens_radius * (r * std::sin(phi)), 0.0f);
        float ft = cam.focal_distance / d.z;
        V3f pf = d * ft;
        o = pl;
        d = normalize(pf - pl);
    }
    float invz = 1.0f / d.z;
    const float *M = cam.to_world;
    ray.o = V3f(M[0] * o.x + M[1] * o.y + M[2] * o.z + M[3], M[4] * o.x + M[5] * o.y + M[6] * o.z + M[7],
                M[8] * o.x + M[9] * o.y + M[10] * o.z + M[11]);
    ray.d = V3f(M[0] * d.x + M[1] * d.y + M[2] * d.z, M[4] * d.x + M[5] * d.y + M[6] * d.z,
                M[8] * d.x + M[9] * d.y + M[10] * d.z);
    ray.mint = cam.near_clip * invz;
    ray.maxt = cam.far_clip * invz;
    ray.drcp = V3f(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
}

static inline Ray mk_ray(V3f o, V3f d, float mint, float maxt) {
    Ray r;
    r.o = o; r.d = d; r.mint = mint; r.maxt = maxt;
    r.drcp = V3f(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    return r;
}

// Alpha test of a candidate (shaders/Shaders.hlsl:674-724): a = red of the
// texture at the candidate's UV (level 0) for radiance rays; for shadow rays
// alpha when below 1, else red; ignored below 0.01, else accepted when
// a^2 >= xi with xi = pcg_hash(seed ^ pcg_hash(prim)) * 2^-32 (the ray's seed
// and the triangle only, independent of the traversal order).
static bool alpha_accept(const SceneData &s, uint32_t prim, float u, float v, uint32_t seed, bool shadow) {
    const MatRec &m = s.mats[s.tri_mesh[prim]];
    if (m.alphaTex == 0xFFFFFFFFu || s.uv.empty()) return true;
    const uint32_t i0 = s.idx[3 * prim], i1 = s.idx[3 * prim + 1], i2 = s.idx[3 * prim + 2];
    const float w = 1.0f - u - v;
    const float tu = s.uv[2 * i0] * w + s.uv[2 * i1] * u + s.uv[2 * i2] * v;
    const float tv = s.uv[2 * i0 + 1] * w + s.uv[2 * i1 + 1] * u + s.uv[2 * i2 + 1] * v;
    const Rgba c = sample(s.tex[m.alphaTex], tu, tv, 0.0f);
    const float a = shadow ? (c.a < 1.0f ? c.a : c.r) : c.r;
    if (a < 0.01f) return false;
    const float xi = (float)pcg_hash(seed ^ pcg_hash(prim)) * 2.3283064365386963e-10f;
    return a * a >= xi;
}
// Seeds: a radiance ray's from its path stream's state words, a shadow ray's
// from its seed word (alpha_seed_shadow).
static uint32_t alpha_seed(const Pcg32 &r) {
    return pcg_hash((uint32_t)r.state ^ pcg_hash((uint32_t)(r.state >> 32)));
}
static uint32_t alpha_seed(const SobolRng &r) { return pcg_hash(r.index ^ pcg_hash(r.dim)); }
static uint32_t alpha_seed_shadow(uint32_t word) { return pcg_hash(word ^ 0x2545F491u); }
// Seed word of a surface NEE shadow ray: a transmittance seed drawn from the
// stream in scenes with media, else the stream's state hash (no draw).
template <class Rng> static uint32_t shadow_word(const SceneData &s, Rng &rng) {
    return s.vols.empty() ? alpha_seed(rng) : rng.next_uint();
}

// Texture maps at a hit (shaders/Shaders.hlsl:232-315) with the level of
// detail from one view pixel's UV footprint at the segment's length
// (shaders/GeometryUtils.hlsli:120-154; dims 0: level 0).
static float apply_textures(const SceneData &s, uint32_t dims_max, uint32_t f, float bu, float bv, float hit_t,
                           const MatRec &mr, V3f &N, MatT<float> &m) {
    const uint32_t i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
    const float w = 1.0f - bu - bv;
    const float tu = s.uv[2 * i0] * w + s.uv[2 * i1] * bu + s.uv[2 * i2] * bv;
    const float tv = s.uv[2 * i0 + 1] * w + s.uv[2 * i1 + 1] * bu + s.uv[2 * i2 + 1] * bv;
    const V3f p0 = s.vtx(i0), p1 = s.vtx(i1), p2 = s.vtx(i2);
    const float u0 = s.uv[2 * i0], v0 = s.uv[2 * i0 + 1], u1 = s.uv[2 * i1], v1 = s.uv[2 * i1 + 1];
    const float u2 = s.uv[2 * i2], v2 = s.uv[2 * i2 + 1];
    const V3f e1 = p1 - p0, e2 = p2 - p0;
    const float world_area = 0.5f * length(cross(e1, e2));
    const float uv_area = 0.5f * std::fabs((u1 - u0) * (v2 - v0) - (u2 - u0) * (v1 - v0));
    float foot = 0.0f;
    if (world_area >= 1e-12f && uv_area >= 1e-12f && dims_max > 0)
        foot = hit_t * (1.0f / (float)dims_max) * std::sqrt(uv_area / world_area);
    auto look = [&](uint32_t ti) { return sample(s.tex[ti], tu, tv, lod_for(s.tex[ti], foot)); };
    if (mr.albedoTex != 0xFFFFFFFFu) {
        const Rgba c = look(mr.albedoTex);
        m.albedo = V3f(c.r, c.g, c.b);
    }
    if (mr.normalTex != 0xFFFFFFFFu) {
        const float du1 = u1 - u0, dv1 = v1 - v0, du2 = u2 - u0, dv2 = v2 - v0;
        const float det = du1 * dv2 - du2 * dv1;
        V3f T, B;
        if (std::fabs(det) > 1e-8f) {
            const float inv = 1.0f / det;
            T = normalize((e1 * dv2 - e2 * dv1) * inv);
            T = normalize(T - N * dot(N, T));
            B = cross(N, T);
        } else {
            build_onb(N, T, B);
        }
        const Rgba c = look(mr.normalTex);
        N = normalize(T * (c.r * 2.0f - 1.0f) + B * (c.g * 2.0f - 1.0f) + N * (c.b * 2.0f - 1.0f));
    }
    if (mr.type != MAT_HAIR && mr.roughnessTex != 0xFFFFFFFFu) {
        const float r = look(mr.roughnessTex).r;
        m.alpha = r;
        m.roughness = r;
    }
    if (mr.specularTex != 0xFFFFFFFFu) m.specular = look(mr.specularTex).r * 0.5f;
    if (mr.subsurfaceTex != 0xFFFFFFFFu) m.subsurface = look(mr.subsurfaceTex).r;
    return foot;
}

// shaders/Emitter.hlsli:17-31
static uint32_t cdf_sample(const std::vector<float> &c, uint32_t n, float u) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (c[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    return (uint32_t)std::max(0, (int)lo - 1);
}

// Shadow ray: 0 if occluded, else the Fresnel transmission through the mirror
// and dielectric meshes it crosses (shaders/RayQueryTrace.hlsli:96-124; those
// meshes and the alpha-tested ones are the non-opaque instances, every other
// mesh occludes).  `word` seeds the alpha tests: the ray's transmittance seed
// in scenes with media, else the path stream's state hash where the ray was
// made (shadow_word).
static float shadow_pass_factor(const SceneData &s, V3f o, V3f d, float maxt, RenderStats &st, uint32_t word,
                                float mint = 0.0f) {
    ++st.shadow_rays;
    if (!s.has_pass && !s.has_alpha)
        return s.bvh.trace(mk_ray(o, d, mint, maxt), true).prim != 0xFFFFFFFFu ? 0.0f : 1.0f;
    const uint32_t seed = s.has_alpha ? alpha_seed_shadow(word) : 0u;
    return shadow_filter(s.bvh, mk_ray(o, d, mint, maxt), [&](uint32_t f, float u, float v, float cosi) {
        const MatRec &m = s.mats[s.tri_mesh[f]];
        if (s.has_alpha && m.alphaTex != 0xFFFFFFFFu) return alpha_accept(s, f, u, v, seed, true) ? -1.0f : 1.0f;
        if (m.type == MAT_MIRROR || m.type == MAT_DIELECTRIC)
            return 1.0f - fresnel_dielectric(cosi, std::max(m.extIOR, 1e-6f), m.intIOR);
        return -1.0f;
    });
}

// Transmittance of an unoccluded shadow ray through the scene's volumes, with
// the ray's own stream seed(seed, PCG32_DEFAULT_STREAM).
static V3f shadow_transmittance(const SceneData &s, V3f o, V3f d, float tmax, uint32_t seed) {
    if (s.vols.empty()) return V3f(1.0f);
    Pcg32 r;
    r.seed(seed, PCG32_DEFAULT_STREAM);
    return multi_volume_transmittance(s.vols, o, d, tmax, r);
}

// Area-light point sample (shaders/Emitter.hlsli:43-89): emitter, triangle,
// barycentrics (1 - sqrt r1, sqrt r1 (1 - r2), sqrt r1 r2).
struct LightSample {
    V3f p, n, Le;
    float pdfA;
};
template <class Rng> static LightSample sample_area_light(const SceneData &s, Rng &rng) {
    const uint32_t E = (uint32_t)s.emitters.size();
    uint32_t pick = std::min((uint32_t)(rng.next_float() * (float)E), E - 1);
    uint32_t em = s.emitters[pick];
    uint32_t ntri = s.mesh_tri_count[em];
    uint32_t k = std::min(cdf_sample(s.emit_cdf[pick], ntri + 1, rng.next_float()), ntri - 1);
    uint32_t ef = s.mesh_tri_offset[em] + k;
    V3f p0 = s.vtx(s.idx[3 * ef]), p1 = s.vtx(s.idx[3 * ef + 1]), p2 = s.vtx(s.idx[3 * ef + 2]);
    float r1 = rng.next_float(), r2 = rng.next_float();
    float sr1 = std::sqrt(r1);
    LightSample ls;
    ls.p = (1.0f - sr1) * p0 + sr1 * (1.0f - r2) * p1 + sr1 * r2 * p2;
    ls.n = normalize(cross(p1 - p0, p2 - p0));
    ls.pdfA = 1.0f / (s.emit_area[pick] * (float)E);
    ls.Le = V3f(s.mats[em].radiance[0], s.mats[em].radiance[1], s.mats[em].radiance[2]);
    return ls;
}

// Volume NEE toward the area lights at a scatter point (shaders/Emitter.hlsli:334-385).
template <class Rng>
static V3f volume_nee_area(const SceneData &s, V3f P, V3f wo, const Volume &v, Rng &rng, RenderStats &st) {
    if (volume_nee_skip(v, P) || s.emitters.empty()) return V3f(0.0f);
    LightSample ls = sample_area_light(s, rng);
    uint32_t tr_seed = rng.next_uint();
    V3f tl = ls.p - P;
    float dist = length(tl);
    V3f wi = tl / dist;
    float cl = dot(ls.n, -wi);
    if (cl <= 0.0f) return V3f(0.0f);
    const float pass = shadow_pass_factor(s, P, wi, dist - 0.001f, st, tr_seed, 0.001f);
    if (pass == 0.0f) return V3f(0.0f);
    float fp = hg_eval(dot(wo, wi), v.g);
    float pe = ls.pdfA * dist * dist / cl;
    float w = balance_heuristic(pe, fp);
    V3f tr = shadow_transmittance(s, P, wi, dist, tr_seed);
    return ls.Le * fp / std::max(pe, 1e-20f) * w * tr * pass;
}

// Volume NEE toward the envmap (shaders/Emitter.hlsli:387-434).
template <class Rng>
static V3f volume_nee_env(const SceneData &s, V3f P, V3f wo, const Volume &v, Rng &rng, const RenderParams &p,
                          RenderStats &st) {
    if (volume_nee_skip(v, P) || !s.env.valid()) return V3f(0.0f);
    V3f dir, Le;
    float pe;
    if (p.env_cdf) {
        float u1 = rng.next_float(), u2 = rng.next_float();
        s.env.sample_cdf(u1, u2, dir, Le, pe);
    } else {
        uint32_t r = rng.next_uint();
        float coin = rng.next_float(), jx = rng.next_float(), jy = rng.next_float();
        s.env.sample_alias(r, coin, jx, jy, dir, Le, pe);
    }
    uint32_t tr_seed = rng.next_uint();
    if (pe <= 0.0f) return V3f(0.0f);
    const float pass = shadow_pass_factor(s, P, dir, 1e20f, st, tr_seed, 0.001f);
    if (pass == 0.0f) return V3f(0.0f);
    float fp = hg_eval(dot(wo, dir), v.g);
    float w = balance_heuristic(pe, fp);
    V3f tr = shadow_transmittance(s, P, dir, 1e20f, tr_seed);
    return Le * fp / std::max(pe, 1e-20f) * w * tr * pass;
}

// Interpolated shading normal of triangle f at (u, v), the geometric normal
// where the mesh has none or they cancel (shaders/GeometryUtils.hlsli:59-80).
static V3f interp_normal(const SceneData &s, uint32_t f, float u, float v) {
    const uint32_t i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
    if (!s.nrm.empty()) {
        V3f n0(s.nrm[3 * i0], s.nrm[3 * i0 + 1], s.nrm[3 * i0 + 2]);
        V3f n1(s.nrm[3 * i1], s.nrm[3 * i1 + 1], s.nrm[3 * i1 + 2]);
        V3f n2(s.nrm[3 * i2], s.nrm[3 * i2 + 1], s.nrm[3 * i2 + 2]);
        V3f n = n0 * (1.0f - u - v) + n1 * u + n2 * v;
        if (dot(n, n) >= 1e-8f) return normalize(n);
    }
    return normalize(cross(s.bvh.e1[f], s.bvh.e2[f]));
}

// Random walk through the interior of a subsurface mesh
// (shaders/Subsurface.hlsli:48-141): free flights -ln(xi)/sigma_t, closest hits
// from tmin 1e-4 up to the flight length, Fresnel reflection (and total
// internal reflection) or refraction out at the mesh's own boundary, a ray
// that reaches another mesh ends the walk, Henyey-Greenstein scattering that
// multiplies the weight by the per-channel albedo, Russian roulette after the
// fifth step, at most 256 steps.
template <class Rng>
static bool sss_walk(const SceneData &s, V3f pos, V3f dir, uint32_t mesh, float sigma_t, V3f alb, float g,
                     float int_ior, float ext_ior, Rng &rng, V3f &exit_p, V3f &exit_d, V3f &exit_n, V3f &w) {
    w = V3f(1.0f);
    for (int step = 0; step < 256; ++step) {
        const float dist = -std::log(std::max(rng.next_float(), 1e-10f)) / sigma_t;
        const Ray r = mk_ray(pos, dir, 1e-4f, dist);
        Hit h;
        if (s.has_alpha) {
            const uint32_t seed = alpha_seed(rng);
            h = trace_accept(s.bvh, r, false,
                             [&](uint32_t f, float u, float v) { return alpha_accept(s, f, u, v, seed, false); });
        } else {
            h = s.bvh.trace(r, false);
        }
        if (h.prim != 0xFFFFFFFFu) {
            if (s.tri_mesh[h.prim] != mesh) {
                w = V3f(0.0f);
                return false;
            }
            const V3f bp = pos + dir * h.t;
            V3f nb = interp_normal(s, h.prim, h.u, h.v);
            if (dot(nb, dir) < 0.0f) nb = -nb;
            const float fr = fresnel_dielectric(dot(dir, nb), int_ior, ext_ior);
            if (rng.next_float() < fr) {
                dir = normalize(reflect(dir, nb));
                pos = bp - nb * 1e-4f;
                continue;
            }
            const V3f out = refract(dir, -nb, int_ior / ext_ior);
            if (dot(out, out) < 1e-8f) {
                dir = normalize(reflect(dir, nb));
                pos = bp - nb * 1e-4f;
                continue;
            }
            exit_p = bp;
            exit_d = normalize(out);
            exit_n = nb;
            return true;
        }
        pos = pos + dir * dist;
        w = w * alb;
        const float x1 = rng.next_float(), x2 = rng.next_float();
        float ppdf;
        dir = hg_sample(dir, g, x1, x2, ppdf);
        if (step > 4) {
            const float q = std::min(std::max(maxc(w), 0.05f), 0.95f);
            if (rng.next_float() >= q) {
                w = V3f(0.0f);
                return false;
            }
            w = w / q;
        }
    }
    w = V3f(0.0f);
    return false;
}

// Next-event estimation at a subsurface entry with the boundary's reflection
// lobe: one area-light and one envmap sample, balance heuristic
// (shaders/Emitter.hlsli:43-153, 275-326).
template <class Rng>
static V3f sss_nee(const SceneData &s, V3f P, V3f N, V3f Ng, V3f Tg, V3f B, V3f wi, const MatT<float> &m, Rng &rng,
                   const RenderParams &p, RenderStats &st) {
    V3f L(0.0f);
    const V3f so = offset_ray_origin(P, Ng, Ng);
    if (!s.emitters.empty()) {
        LightSample ls = sample_area_light(s, rng);
        uint32_t tr_seed = shadow_word(s, rng);
        V3f tl = ls.p - P;
        float dist = length(tl);
        V3f ww = tl / dist;
        float ct = dot(N, ww), cl = dot(ls.n, -ww);
        float pass = 0.0f;
        if (ct > 0.0f && cl > 0.0f && (pass = shadow_pass_factor(s, so, ww, dist - 0.001f, st, tr_seed)) > 0.0f) {
            V3f wol = to_local(ww, Tg, B, N);
            V3f fv = material_eval(wi, wol, m);
            float pb = material_pdf(wi, wol, m);
            float pe = ls.pdfA * dist * dist / cl;
            float w = balance_heuristic(pe, pb);
            V3f tr = shadow_transmittance(s, so, ww, dist, tr_seed);
            L += ls.Le * fv * std::fabs(ct) / std::max(pe, 1e-20f) * w * tr * pass;
        }
    }
    if (s.env.valid()) {
        V3f dir, Le;
        float pe;
        if (p.env_cdf) {
            float u1 = rng.next_float(), u2 = rng.next_float();
            s.env.sample_cdf(u1, u2, dir, Le, pe);
        } else {
            uint32_t r = rng.next_uint();
            float coin = rng.next_float(), jx = rng.next_float(), jy = rng.next_float();
            s.env.sample_alias(r, coin, jx, jy, dir, Le, pe);
        }
        uint32_t tr_seed = shadow_word(s, rng);
        float ct = dot(N, dir);
        float pass = 0.0f;
        if (pe > 0.0f && ct > 0.0f && (pass = shadow_pass_factor(s, so, dir, 1e20f, st, tr_seed)) > 0.0f) {
            V3f wol = to_local(dir, Tg, B, N);
            V3f fv = material_eval(wi, wol, m);
            float pb = material_pdf(wi, wol, m);
            float w = balance_heuristic(pe, pb);
            V3f tr = shadow_transmittance(s, so, dir, 1e20f, tr_seed);
            L += Le * fv * std::fabs(ct) / std::max(pe, 1e-20f) * w * tr * pass;
        }
    }
    return L;
}

// RayGen body, shaders/Shaders.hlsl:55-627, for the surface materials of the
// scene (diffuse, mirror, dielectric, microfacet, Disney).
template <class Rng>
static V3f path_li(const SceneData &s, const Camera &cam, Ray ray, Rng &rng, const RenderParams &p, RenderStats &st,
                   Aov *aov = nullptr) {
    Aov aov_unused;
    Aov &av = aov ? *aov : aov_unused;
    V3f L(0.0f), T(1.0f);
    float last_pdf = 0.0f, eta = 1.0f;
    const uint32_t E = (uint32_t)s.emitters.size();
    for (uint32_t bounce = 0; bounce < p.max_depth; ++bounce) {
        ++st.extension_rays;
        Hit h;
        if (s.has_alpha) {
            const uint32_t seed = alpha_seed(rng);
            h = trace_accept(s.bvh, ray, false,
                             [&](uint32_t f, float u, float v) { return alpha_accept(s, f, u, v, seed, false); });
        } else {
            h = s.bvh.trace(ray, false);
        }
        // Participating media between the ray origin and the surface
        // (shaders/Shaders.hlsl:133-180).
        bool vol_scattered = false;
        if (!s.vols.empty()) {
            float tsurf = h.prim != 0xFFFFFFFFu ? h.t : 1e20f;
            V3f vw(1.0f);
            float ts;
            uint32_t vi;
            bool sc = sample_volume_free_flight(s.vols, ray.o, ray.d, tsurf, rng, vw, ts, vi);
            T = T * vw;
            if (sc) {
                vol_scattered = true;
                const Volume &v = s.vols[vi];
                V3f sig = v.sigma_t();
                T = T * V3f(v.sigma_s.x / std::max(sig.x, 1e-20f), v.sigma_s.y / std::max(sig.y, 1e-20f),
                            v.sigma_s.z / std::max(sig.z, 1e-20f));
                V3f P = ray.o + ray.d * ts;
                L += T * volume_nee_area(s, P, ray.d, v, rng, st);
                L += T * volume_nee_env(s, P, ray.d, v, rng, p, st);
                float xi1 = rng.next_float(), xi2 = rng.next_float(), ppdf;
                V3f nd = hg_sample(ray.d, v.g, xi1, xi2, ppdf);
                ray = mk_ray(P, nd, 0.0f, 1e20f);
                last_pdf = std::max(ppdf, 1e-20f);
            }
        }
        if (vol_scattered) {
            // scatter event: on to Russian roulette
        } else if (h.prim == 0xFFFFFFFFu) {
            av.capture(s.env.valid() ? s.env.eval(ray.d) : V3f(0.0f), V3f(0.0f));
            if (s.env.valid()) {
                V3f env = s.env.eval(ray.d);
                if (last_pdf == 0.0f) L += T * env;
                else L += T * env * balance_heuristic(last_pdf, s.env.pdf(ray.d));
            }
            break;
        }
        if (!vol_scattered) { // surface interaction
        V3f P = ray.o + ray.d * h.t;
        uint32_t f = h.prim, mesh = s.tri_mesh[f];
        const MatRec &mr = s.mats[mesh];
        MatT<float> m = MatT<float>::from(mr);
        V3f Ng = normalize(cross(s.bvh.e1[f], s.bvh.e2[f]));
        V3f N = Ng;
        if (!s.nrm.empty()) {
            uint32_t i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
            V3f n0(s.nrm[3 * i0], s.nrm[3 * i0 + 1], s.nrm[3 * i0 + 2]);
            V3f n1(s.nrm[3 * i1], s.nrm[3 * i1 + 1], s.nrm[3 * i1 + 2]);
            V3f n2(s.nrm[3 * i2], s.nrm[3 * i2 + 1], s.nrm[3 * i2 + 2]);
            float w0 = 1.0f - h.u - h.v;
            V3f n = n0 * w0 + n1 * h.u + n2 * h.v;
            float l2 = dot(n, n);
            N = l2 < 1e-8f ? Ng : normalize(n);
        }
        bool back = dot(Ng, ray.d) > 0.0f;
        if (back) Ng = -Ng;
        if (dot(N, Ng) < 0.0f) N = -N;
        float foot = 0.0f;
        if (!s.tex.empty() && !s.uv.empty() &&
            (mr.albedoTex != 0xFFFFFFFFu || mr.normalTex != 0xFFFFFFFFu || mr.roughnessTex != 0xFFFFFFFFu ||
             mr.specularTex != 0xFFFFFFFFu || mr.subsurfaceTex != 0xFFFFFFFFu))
            foot = apply_textures(s, std::max(cam.width, cam.height), f, h.u, h.v, h.t, mr, N, m);

        if (mr.isEmitter) {
            av.capture(V3f(1.0f), N);
            V3f Le(mr.radiance[0], mr.radiance[1], mr.radiance[2]);
            if (last_pdf == 0.0f) {
                L += T * Le;
            } else {
                // shaders/Emitter.hlsli:91-103
                V3f dv = P - ray.o;
                float d2 = dot(dv, dv), pe = 0.0f;
                if (d2 >= 1e-12f) {
                    float dist = std::sqrt(d2);
                    float cl = std::fabs(dot(N, -dv / dist));
                    if (cl >= 1e-8f) {
                        uint32_t k = 0;
                        while (s.emitters[k] != mesh) ++k;
                        pe = (1.0f / (s.emit_area[k] * (float)E)) * d2 / cl;
                    }
                }
                L += T * Le * balance_heuristic(last_pdf, pe);
            }
            break;
        }

        V3f Tg, B;
        float hh = 0.0f;
        if (m.type == MAT_HAIR) {
            // shaders/Shaders.hlsl:340-387: fiber tangent frame, h from UV
            uint32_t i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
            float w0 = 1.0f - h.u - h.v;
            V3f ht(0.0f);
            if (!s.tan.empty()) {
                V3f t0(s.tan[3 * i0], s.tan[3 * i0 + 1], s.tan[3 * i0 + 2]);
                V3f t1(s.tan[3 * i1], s.tan[3 * i1 + 1], s.tan[3 * i1 + 2]);
                V3f t2(s.tan[3 * i2], s.tan[3 * i2 + 1], s.tan[3 * i2 + 2]);
                V3f t = t0 * w0 + t1 * h.u + t2 * h.v;
                float l2 = dot(t, t);
                ht = l2 > 1e-8f ? t / std::sqrt(l2) : V3f(0.0f);
            }
            float uvx = 0.0f, uvy = 0.0f;
            if (!s.uv.empty()) {
                uvx = s.uv[2 * i0] * w0 + s.uv[2 * i1] * h.u + s.uv[2 * i2] * h.v;
                uvy = s.uv[2 * i0 + 1] * w0 + s.uv[2 * i1 + 1] * h.u + s.uv[2 * i2 + 1] * h.v;
            }
            if (dot(ht, ht) > 1e-8f) {
                ht = normalize(ht);
                hh = uvy * 2.0f - 1.0f;
            } else {
                // hair card: fiber axis from dP/dV, h across the card width
                V3f p0 = s.vtx(i0), p1 = s.vtx(i1), p2 = s.vtx(i2);
                float du1 = 0, dv1 = 0, du2 = 0, dv2 = 0;
                if (!s.uv.empty()) {
                    du1 = s.uv[2 * i1] - s.uv[2 * i0]; dv1 = s.uv[2 * i1 + 1] - s.uv[2 * i0 + 1];
                    du2 = s.uv[2 * i2] - s.uv[2 * i0]; dv2 = s.uv[2 * i2 + 1] - s.uv[2 * i0 + 1];
                }
                float det = du1 * dv2 - du2 * dv1;
                if (std::fabs(det) > 1e-8f) ht = normalize((-du1 * (p1 - p0) + du2 * (p2 - p0)) * (1.0f / det));
                else build_onb(N, ht, B);
                hh = uvx * 2.0f - 1.0f;
            }
            Tg = normalize(ht - N * dot(ht, N));
            B = cross(N, Tg);
        } else {
            build_onb(N, Tg, B);
        }
        {
            V3f V = -ray.d;
            float ndv = dot(N, V);
            if (ndv <= 0.0f && dot(V, Ng) > 0.0f) {
                N = normalize(N + V * (1e-3f - ndv));
                build_onb(N, Tg, B);
            }
        }
        V3f wi = to_local(-ray.d, Tg, B, N);

        if (m.type == MAT_MIRROR) {
            ray = mk_ray(P + Ng * 0.001f, reflect(ray.d, N), 0.0f, 1e20f);
            last_pdf = 0.0f;
        } else if (m.type == MAT_DIELECTRIC) {
            V3f I = ray.d;
            float etaI = back ? m.intIOR : m.extIOR, etaT = back ? m.extIOR : m.intIOR;
            float Fr = fresnel_dielectric(dot(-I, N), etaI, etaT);
            V3f nd;
            bool refr = false;
            if (rng.next_float() < Fr) {
                nd = reflect(I, N);
            } else {
                nd = refract(I, N, etaI / etaT);
                if (dot(nd, nd) < 0.001f) nd = reflect(I, N);
                else refr = true;
            }
            V3f on = dot(nd, Ng) > 0.0f ? Ng : -Ng;
            ray = mk_ray(P + on * 0.001f, nd, 0.0f, 1e20f);
            last_pdf = 0.0f;
            if (refr) eta *= etaI / etaT;
        } else if (m.type == MAT_SSS) {
            // Random-walk subsurface entry, shaders/Shaders.hlsl:513-612.
            av.capture(m.albedo, N);
            const V3f I = ray.d;
            m.roughness *= 1.0f + (0.65f - 1.0f) * std::min(std::max(m.specular * 2.0f, 0.0f), 1.0f);
            const float rough = m.roughness, aR = std::max(rough * rough, 1e-4f);
            V3f wmL(0.0f, 0.0f, 1.0f);
            if (rough > 1e-3f) {
                float u0 = rng.next_float(), u1 = rng.next_float();
                wmL = beckmann_sample(u0, u1, aR);
            }
            V3f wm = normalize(to_world(wmL, Tg, B, N));
            float cos_m = dot(-I, wm);
            if (cos_m <= 0.0f) {
                wm = N;
                wmL = V3f(0.0f, 0.0f, 1.0f);
                cos_m = dot(-I, N);
            }
            const float g1i = rough > 1e-3f ? smith_g1(wi, wmL, aR) : 1.0f;
            const float Fr = fresnel_dielectric(cos_m, m.extIOR, m.intIOR);
            L += T * sss_nee(s, P, N, Ng, Tg, B, wi, m, rng, p, st);
            if (rng.next_float() < Fr) {
                const V3f refl = reflect(I, wm);
                const V3f woL = to_local(refl, Tg, B, N);
                if (woL.z <= 0.0f) break;
                const float wgt = rough > 1e-3f ? g1i * smith_g1(woL, wmL, aR) * cos_m / std::max(wi.z * wmL.z, 1e-6f)
                                                : 1.0f;
                ray = mk_ray(P + Ng * 0.001f, refl, 0.0f, 1e20f);
                T = T * wgt;
                last_pdf = rough > 1e-3f ? Fr * beckmann_d_cos(wmL, aR) / std::max(4.0f * cos_m, 1e-6f) : 0.0f;
            } else {
                const V3f refr = refract(I, wm, m.extIOR / m.intIOR);
                if (dot(refr, refr) < 1e-8f) {
                    ray = mk_ray(P + Ng * 0.001f, reflect(I, wm), 0.0f, 1e20f);
                    last_pdf = 0.0f;
                } else {
                    V3f a_blur = m.albedo, detail(1.0f);
                    if (mr.albedoTex != 0xFFFFFFFFu && !s.tex.empty() && !s.uv.empty()) {
                        const uint32_t i0 = s.idx[3 * f], i1 = s.idx[3 * f + 1], i2 = s.idx[3 * f + 2];
                        const float w0 = 1.0f - h.u - h.v;
                        const float tu = s.uv[2 * i0] * w0 + s.uv[2 * i1] * h.u + s.uv[2 * i2] * h.v;
                        const float tv = s.uv[2 * i0 + 1] * w0 + s.uv[2 * i1 + 1] * h.u + s.uv[2 * i2 + 1] * h.v;
                        const Texture &tx = s.tex[mr.albedoTex];
                        const Rgba c = sample(tx, tu, tv, lod_for(tx, foot) + 4.0f);
                        a_blur = V3f(c.r, c.g, c.b);
                        auto cl = [](float a, float b) { return std::min(std::max(a / std::max(b, 1e-3f), 0.5f), 2.0f); };
                        detail = V3f(cl(m.albedo.x, a_blur.x), cl(m.albedo.y, a_blur.y), cl(m.albedo.z, a_blur.z));
                    }
                    const V3f A = a_blur * V3f(mr.sheen, mr.sheenTint, mr.clearcoat);
                    const float sigma_t = 1.0f / std::max(m.subsurface, 1e-4f);
                    const V3f A2 = A * A, A3 = A2 * A;
                    const V3f ex = A * -5.09406f + A2 * 2.61188f - A3 * 4.31805f;
                    const V3f alb(1.0f - std::exp(ex.x), 1.0f - std::exp(ex.y), 1.0f - std::exp(ex.z));
                    V3f ep, ed, en, tw;
                    if (!sss_walk(s, P - Ng * 0.001f, normalize(refr), mesh, sigma_t, alb, mr.anisotropic, m.intIOR,
                                  m.extIOR, rng, ep, ed, en, tw))
                        break;
                    T = T * tw * detail;
                    ray = mk_ray(offset_ray_origin(ep, en, en), ed, 0.0f, 1e20f);
                    last_pdf = 0.0f;
                }
            }
        } else if (m.type == MAT_DIFFUSE || m.type == MAT_MICROFACET || m.type == MAT_DISNEY || m.type == MAT_HAIR) {
            av.capture(m.albedo, N);
            const bool hair = m.type == MAT_HAIR;
            auto side = [&](V3f w) { return hair && dot(w, Ng) < 0.0f ? -Ng : Ng; }; // shaders/Emitter.hlsli:122-126
            V3f so = offset_ray_origin(P, Ng, Ng);
            // Area light NEE, shaders/Emitter.hlsli:43-89, 105-153
            if (E > 0) {
                LightSample ls = sample_area_light(s, rng);
                uint32_t tr_seed = shadow_word(s, rng);
                V3f tl = ls.p - P;
                float dist = length(tl);
                V3f ww = tl / dist;
                float ct = dot(N, ww), cl = dot(ls.n, -ww);
                V3f sso = hair ? offset_ray_origin(P, side(ww), side(ww)) : so;
                float pass = 0.0f;
                if ((hair || ct > 0.0f) && cl > 0.0f && (pass = shadow_pass_factor(s, sso, ww, dist - 0.001f, st, tr_seed)) > 0.0f) {
                    V3f wol = to_local(ww, Tg, B, N);
                    V3f fv = material_eval_h(wi, wol, m, hh);
                    float pb = material_pdf_h(wi, wol, m, hh);
                    float pe = ls.pdfA * dist * dist / cl;
                    float w = balance_heuristic(pe, pb);
                    V3f tr = shadow_transmittance(s, sso, ww, dist, tr_seed);
                    L += T * (ls.Le * fv * std::fabs(ct) / std::max(pe, 1e-20f) * w) * tr * pass;
                }
            }
            // Envmap NEE, shaders/Emitter.hlsli:275-326
            if (s.env.valid()) {
                V3f dir, Le;
                float pe;
                if (p.env_cdf) {
                    float u1 = rng.next_float(), u2 = rng.next_float();
                    s.env.sample_cdf(u1, u2, dir, Le, pe);
                } else {
                    uint32_t r = rng.next_uint();
                    float coin = rng.next_float(), jx = rng.next_float(), jy = rng.next_float();
                    s.env.sample_alias(r, coin, jx, jy, dir, Le, pe);
                }
                uint32_t tr_seed = shadow_word(s, rng);
                float ct = dot(N, dir);
                V3f sso = hair ? offset_ray_origin(P, side(dir), side(dir)) : so;
                float pass = 0.0f;
                if (pe > 0.0f && (hair || ct > 0.0f) && (pass = shadow_pass_factor(s, sso, dir, 1e20f, st, tr_seed)) > 0.0f) {
                    V3f wol = to_local(dir, Tg, B, N);
                    V3f fv = material_eval_h(wi, wol, m, hh);
                    float pb = material_pdf_h(wi, wol, m, hh);
                    float w = balance_heuristic(pe, pb);
                    V3f tr = shadow_transmittance(s, sso, dir, 1e20f, tr_seed);
                    L += T * (Le * fv * std::fabs(ct) / std::max(pe, 1e-20f) * w) * tr * pass;
                }
            }
            float u[4];
            int nd = material_sample_dims_h(m.type);
            for (int i = 0; i < nd; ++i) u[i] = rng.next_float();
            V3f wo;
            float pdf;
            V3f wgt = material_sample_h(wi, u, m, hh, wo, pdf);
            if (pdf <= 0.0f || (wgt.x == 0.0f && wgt.y == 0.0f && wgt.z == 0.0f)) break;
            V3f ww = to_world(wo, Tg, B, N);
            ray = mk_ray(offset_ray_origin(P, side(ww), side(ww)), ww, 0.0f, 1e20f);
            T *= wgt;
            last_pdf = std::max(pdf, 1e-20f);
        } else {
            break;
        }
        } // surface interaction
        // Russian roulette, shaders/Shaders.hlsl:620-626
        if (bounce >= p.rr_depth) {
            float q = std::min(maxc(T) * eta * eta, 0.95f);
            if (rng.next_float() >= q) break;
            T /= q;
        }
    }
    if (!std::isfinite(L.x) || !std::isfinite(L.y) || !std::isfinite(L.z)) L = V3f(0.0f);
    return L;
}

V3f path_mis_li(const SceneData &s, const Camera &cam, Ray ray, Pcg32 &rng, const RenderParams &p, RenderStats &st) {
    return path_li(s, cam, ray, rng, p, st);
}

// One camera sample of pixel (x, y) with generator `rng` (pixel jitter, lens,
// then the path).
template <class Rng>
static V3f camera_sample(const SceneData &s, const Camera &cam, uint32_t x, uint32_t y, Rng &rng,
                         const RenderParams &p, RenderStats &st, Aov *aov) {
    float jx = rng.next_float(), jy = rng.next_float();
    float ax = 0.5f, ay = 0.5f;
    if (cam.lens_radius > 0.0f) { ax = rng.next_float(); ay = rng.next_float(); }
    Ray ray;
    camera_ray(cam, (float)x + jx, (float)y + jy, ax, ay, ray);
    return path_li(s, cam, ray, rng, p, st, aov);
}

void render(const SceneData &s, const Camera &cam, const RenderParams &p, float *accum, RenderStats *stats,
            float *aov_albedo, float *aov_normal) {
    const uint32_t W = cam.width, H = cam.height, BS = 32;
    const uint32_t bx = (W + BS - 1) / BS, by = (H + BS - 1) / BS, nblocks = bx * by;
    std::atomic<uint32_t> next(0);
    int nth = p.nthreads > 0 ? p.nthreads : (int)std::thread::hardware_concurrency();
    std::vector<RenderStats> per(nth);
    auto work = [&](int tid) {
        RenderStats &st = per[tid];
        while (true) {
            uint32_t b = next.fetch_add(1);
            if (b >= nblocks) break;
            uint32_t x0 = (b % bx) * BS, y0 = (b / bx) * BS;
            for (uint32_t y = y0; y < std::min(y0 + BS, H); ++y)
                for (uint32_t x = x0; x < std::min(x0 + BS, W); ++x) {
                    uint32_t pix = y * W + x;
                    float *a = accum + 4 * (size_t)pix;
                    for (uint32_t si = 0; si < p.spp; ++si) {
                        V3f L;
                        Aov av;
                        if (p.sampler == 1) {
                            SobolRng rng(pix, p.sample_offset + si);
                            L = camera_sample(s, cam, x, y, rng, p, st, &av);
                        } else {
                            Pcg32 rng = sample_stream(p.seed, pix, p.sample_offset + si);
                            L = camera_sample(s, cam, x, y, rng, p, st, &av);
                        }
                        a[0] += L.x; a[1] += L.y; a[2] += L.z; a[3] += 1.0f;
                        if (aov_albedo) {
                            float *q = aov_albedo + 4 * (size_t)pix;
                            q[0] += av.albedo.x; q[1] += av.albedo.y; q[2] += av.albedo.z; q[3] += 1.0f;
                        }
                        if (aov_normal) {
                            float *q = aov_normal + 4 * (size_t)pix;
                            q[0] += av.normal.x; q[1] += av.normal.y; q[2] += av.normal.z; q[3] += 1.0f;
                        }
                        ++st.samples;
                    }
                }
        }
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) th.emplace_back(work, t);
    for (auto &t : th) t.join();
    if (stats) {
        *stats = RenderStats{};
        for (auto &q : per) {
            stats->extension_rays += q.extension_rays;
            stats->shadow_rays += q.shadow_rays;
            stats->samples += q.samples;
        }
    }
}

} // namespace orc
