// Seeded procedural scenes for the benchmark configurations, produced in the
// flat layout of nori_scene_desc (one material record per mesh, global vertex
// indices).  The generator stream is pcg32.seed(0x5EED0000 + config, 1) unless
// the caller gives a seed.
//
//   config 1  Cornell box: floor, ceiling, back, red left and green right wall
//             (2 triangles each), two white icospheres, a ceiling quad light with
//             radiance (17, 12, 4); pinhole camera with a 40 degree horizontal fov.
//   config 2  grid x grid Disney icospheres on a Lambertian ground quad, lit by a
//             sky-gradient envmap with a 0.5 degree sun disc of radiance 5e4 in
//             a seeded direction.
//   config 3  hair: strands rooted uniformly on the upper hemisphere of a
//             radius-10 scalp, each a polyline whose direction random-walks
//             (Gaussian perturbation plus a pull downwards), tessellated as
//             ribbons of half-width 0.04 that face the camera (2 triangles per
//             segment).  Per strand one hair material (beta_m = beta_n = 0.3,
//             alpha = 2 degrees, eta = 1.55).  Variant 0: eumelanin U[0.3, 8]
//             and pheomelanin U[0, 0.5] per strand, a diffuse head and three
//             quad lights; variant 1 (white furnace): colour 1 (sigma_a = 0),
//             hair only, under a uniform unit envmap.
//   config 4  heterogeneous medium: a grid^3 density (default 256) of seeded
//             4-octave value-noise fBm, thresholded at 0.9 of its range,
//             clamped >= 0 and scaled so that max sigma_t = 20 (albedo 0.9,
//             sigma_s = 18, sigma_a = 2); Henyey-Greenstein g = 0, 0.3 or 0.8
//             (variant 0, 1, 2); majorants on 16^3-voxel bricks; in an open
//             diffuse box (floor and three walls) under the config-2 envmap.
#include "scene_gen.h"
#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>

namespace ng {

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) { s = pcg_seed(seed, 1); }
    float next() {
        uint64_t inc = (1ull << 1u) | 1u;
        uint32_t u = (pcg_next(s, inc) >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &u, 4);
        return f - 1.0f;
    }
    // Standard normal pair (Box-Muller).
    void gauss2(float &a, float &b) {
        float u1 = 1.0f - next(), u2 = next();
        float r = std::sqrt(-2.0f * std::log(u1));
        a = r * std::cos(2.0f * PI * u2);
        b = r * std::sin(2.0f * PI * u2);
    }
};

// Default record: the BSDFGPUData defaults of include/bsdf.h:46-89.
nori_material default_material() {
    nori_material m;
    std::memset(&m, 0, sizeof(m));
    m.type = NORI_MAT_DIFFUSE;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.5f;
    m.intIOR = 1.5046f;
    m.extIOR = 1.000277f;
    m.alpha = 0.1f;
    m.albedoTexIndex = m.normalTexIndex = m.roughnessTexIndex = m.metallicTexIndex = NORI_NO_TEXTURE;
    m.specularTexIndex = m.subsurfaceTexIndex = m.alphaTexIndex = NORI_NO_TEXTURE;
    m.roughness = 0.5f;
    m.specular = 0.5f;
    m.sheenTint = 0.5f;
    m.clearcoatGloss = 1.0f;
    m.betaN = 0.3f;
    return m;
}

struct Builder {
    SceneData &s;
    bool attrs = false; // scene carries per-vertex UVs and tangents (zero where a mesh has none)
    uint32_t begin_mesh() { return (uint32_t)s.mesh_tri_offset.size(); }
    void add_mesh(const std::vector<float> &pos, const std::vector<float> &nrm, const std::vector<uint32_t> &tri,
                  const nori_material &mat, const std::vector<float> *uv = nullptr,
                  const std::vector<float> *tan = nullptr) {
        uint32_t v0 = (uint32_t)(s.pos.size() / 3), t0 = (uint32_t)(s.idx.size() / 3);
        s.pos.insert(s.pos.end(), pos.begin(), pos.end());
        s.nrm.insert(s.nrm.end(), nrm.begin(), nrm.end());
        if (attrs) {
            if (uv) s.uv.insert(s.uv.end(), uv->begin(), uv->end());
            else s.uv.resize(s.uv.size() + pos.size() / 3 * 2, 0.0f);
            if (tan) s.tan.insert(s.tan.end(), tan->begin(), tan->end());
            else s.tan.resize(s.tan.size() + pos.size(), 0.0f);
        }
        for (uint32_t i : tri) s.idx.push_back(i + v0);
        nori_material m = mat;
        m.vertexOffset = v0;
        m.vertexCount = (uint32_t)(pos.size() / 3);
        m.indexOffset = 3 * t0;
        m.indexCount = (uint32_t)tri.size();
        s.mesh_tri_offset.push_back(t0);
        s.mesh_tri_count.push_back((uint32_t)(tri.size() / 3));
        s.mats.push_back(m);
    }
    // Axis-aligned quad p, p+a, p+a+b, p+b; triangle normal = normalize(a x b).
    void quad(f3 p, f3 a, f3 b, const nori_material &mat) {
        f3 n = normalize(cross(a, b));
        f3 q[4] = {p, p + a, p + a + b, p + b};
        std::vector<float> pos, nrm;
        for (auto &v : q) {
            pos.insert(pos.end(), {v.x, v.y, v.z});
            nrm.insert(nrm.end(), {n.x, n.y, n.z});
        }
        add_mesh(pos, nrm, {0, 1, 2, 0, 2, 3}, mat);
    }
    void icosphere(f3 c, float r, uint32_t level, const nori_material &mat) {
        std::vector<float> pos, nrm;
        std::vector<uint32_t> tri;
        icosphere_mesh(level, pos, tri);
        nrm = pos;
        for (size_t i = 0; i < pos.size(); i += 3) {
            pos[i] = c.x + r * pos[i];
            pos[i + 1] = c.y + r * pos[i + 1];
            pos[i + 2] = c.z + r * pos[i + 2];
        }
        add_mesh(pos, nrm, tri, mat);
    }
};

// Quad light of side `size` centred at c, emitting towards `target`.
void quad_light(Builder &b, f3 c, f3 target, float size, f3 radiance) {
    nori_material m = default_material();
    m.isEmitter = 1;
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 0.0f;
    m.radiance[0] = radiance.x; m.radiance[1] = radiance.y; m.radiance[2] = radiance.z;
    f3 n = normalize(target - c);
    f3 up = std::fabs(n.y) < 0.9f ? mk(0, 1, 0) : mk(1, 0, 0);
    f3 a = normalize(cross(up, n)), bb = cross(n, a); // a x bb = n
    b.quad(c - (a + bb) * (0.5f * size), a * size, bb * size, m);
}

// Chiang et al. 2016 eq. 9 inverted: the colour whose sigma_a the colour-to-sigma_a
// fit returns is exp(-g(beta_n) sqrt(sigma_a)).
float hair_color_from_sigma_a(float sa, float bn) {
    float g = 5.969f - 0.215f * bn + 2.532f * bn * bn - 10.73f * bn * bn * bn + 5.574f * bn * bn * bn * bn +
              0.245f * bn * bn * bn * bn * bn;
    return std::exp(-g * std::sqrt(sa));
}

void hair_scene(Builder &b, Rng &rng, const nori_scene_options &opt, f3 eye) {
    const uint32_t ns = opt.strands ? opt.strands : 50000, nseg = opt.subdiv ? opt.subdiv : 16;
    const bool furnace = opt.variant == 1;
    const float R = 10.0f, len = 14.0f, half_w = 0.04f, seg = len / (float)nseg;
    b.attrs = true;
    std::vector<f3> P(nseg + 1), T(nseg + 1);
    std::vector<float> pos, nrm, uv, tan;
    std::vector<uint32_t> tri;
    for (uint32_t i = 0; i < nseg; ++i) {
        uint32_t a0 = 2 * i, b0 = a0 + 1, a1 = a0 + 2, b1 = a0 + 3;
        tri.insert(tri.end(), {a0, b0, a1, b0, b1, a1});
    }
    for (uint32_t k = 0; k < ns; ++k) {
        // root uniform on the upper hemisphere (y uniform gives uniform area)
        float y = rng.next(), phi = 2.0f * PI * rng.next();
        float r = std::sqrt(std::max(0.0f, 1.0f - y * y));
        f3 n = mk(r * std::cos(phi), y, r * std::sin(phi));
        P[0] = n * (R * 1.001f);
        f3 d = n;
        for (uint32_t i = 1; i <= nseg; ++i) {
            float g0, g1, g2, g3;
            rng.gauss2(g0, g1);
            rng.gauss2(g2, g3);
            d = normalize(d + mk(0.0f, -0.12f, 0.0f) + mk(g0, g1, g2) * 0.1f);
            f3 p = P[i - 1] + d * seg;
            float l = length(p);
            if (l < R * 1.02f) p = p * (R * 1.02f / l); // stay outside the head
            P[i] = p;
        }
        for (uint32_t i = 0; i <= nseg; ++i) {
            f3 t = P[std::min(i + 1, nseg)] - P[i > 0 ? i - 1 : 0];
            T[i] = normalize(t);
        }
        pos.clear(); nrm.clear(); uv.clear(); tan.clear();
        for (uint32_t i = 0; i <= nseg; ++i) {
            f3 v = eye - P[i];
            f3 side = cross(T[i], v);
            float sl = length(side);
            side = sl > 1e-12f ? side * (1.0f / sl) : normalize(cross(T[i], mk(0.3f, 0.5f, 0.81f)));
            f3 nn = normalize(cross(side, T[i]));
            if (dot(nn, v) < 0.0f) nn = nn * -1.0f;
            f3 pa = P[i] - side * half_w, pb = P[i] + side * half_w;
            float u = (float)i / (float)nseg;
            pos.insert(pos.end(), {pa.x, pa.y, pa.z, pb.x, pb.y, pb.z});
            nrm.insert(nrm.end(), {nn.x, nn.y, nn.z, nn.x, nn.y, nn.z});
            uv.insert(uv.end(), {u, 0.0f, u, 1.0f}); // h = 2 v - 1 across the ribbon
            tan.insert(tan.end(), {T[i].x, T[i].y, T[i].z, T[i].x, T[i].y, T[i].z});
        }
        nori_material m = default_material();
        m.type = NORI_MAT_HAIR;
        m.roughness = 0.3f; // beta_m
        m.betaN = 0.3f;
        m.alpha = 2.0f;
        m.intIOR = 1.55f;
        if (furnace) {
            m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
        } else {
            // melanin concentrations -> sigma_a (PBRT 3e §9.9.7) -> colour
            float ce = 0.3f + 7.7f * rng.next(), cp = 0.5f * rng.next();
            const float eu[3] = {0.419f, 0.697f, 1.37f}, ph[3] = {0.187f, 0.4f, 1.05f};
            for (int c = 0; c < 3; ++c) m.albedo[c] = hair_color_from_sigma_a(ce * eu[c] + cp * ph[c], m.betaN);
        }
        b.add_mesh(pos, nrm, tri, m, &uv, &tan);
    }
    if (!furnace) {
        nori_material skin = default_material();
        skin.albedo[0] = 0.55f; skin.albedo[1] = 0.42f; skin.albedo[2] = 0.36f;
        b.icosphere(mk(0, 0, 0), R, 5, skin);
        quad_light(b, mk(18, 22, 26), mk(0, 0, 0), 8.0f, mk(6.0f, 5.6f, 5.0f));   // key
        quad_light(b, mk(-28, 4, 16), mk(0, 0, 0), 10.0f, mk(1.5f, 1.7f, 2.0f));  // fill
        quad_light(b, mk(0, 14, -30), mk(0, 0, 0), 8.0f, mk(9.0f, 9.0f, 9.0f));   // rim
    }
}

// Sky gradient with a 0.5-degree sun disc of radiance 5e4 at a seeded
// elevation (20-60 degrees) and azimuth; 2048 x 1024 by default.
void sky_sun_env(SceneData &s, Rng &rng, const nori_scene_options &opt) {
    uint32_t W = opt.env_width ? opt.env_width : 2048, H = opt.env_height ? opt.env_height : 1024;
    float elev = (20.0f + 40.0f * rng.next()) * (PI / 180.0f), azim = 2.0f * PI * rng.next();
    f3 sun = mk(std::cos(elev) * std::cos(azim), std::sin(elev), std::cos(elev) * std::sin(azim));
    float cos_disc = std::cos(0.5f * PI / 180.0f);
    s.env_w = W; s.env_h = H; s.env_scale = 1.0f;
    s.env_rgb.resize((size_t)W * H * 3);
    for (uint32_t y = 0; y < H; ++y) {
        float theta = PI * ((float)y + 0.5f) / (float)H;
        float ct = std::cos(theta), st = std::sin(theta);
        f3 sky;
        if (ct > 0.0f) sky = lerp3(mk(0.85f, 0.9f, 1.0f), mk(0.25f, 0.45f, 0.95f), std::sqrt(ct));
        else sky = lerp3(mk(0.85f, 0.9f, 1.0f), mk(0.2f, 0.18f, 0.15f), std::sqrt(-ct));
        for (uint32_t x = 0; x < W; ++x) {
            float phi = 2.0f * PI * ((float)x + 0.5f) / (float)W;
            f3 d = mk(st * std::cos(phi), ct, st * std::sin(phi));
            f3 c = dot(d, sun) >= cos_disc ? mk(5e4f) : sky;
            float *o = &s.env_rgb[((size_t)y * W + x) * 3];
            o[0] = c.x; o[1] = c.y; o[2] = c.z;
        }
    }
}

// 32-bit PCG hash (Jarzynski and Olano 2020, "Hash Functions for GPU Rendering").
uint32_t hash32(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// Value noise: trilinear interpolation of per-lattice-point values in [0, 1)
// from a hash of the lattice coordinates and the seed.
float lattice(int x, int y, int z, uint32_t seed) {
    uint32_t h = hash32((uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ (uint32_t)z * 83492791u ^ seed);
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}
float value_noise(float x, float y, float z, uint32_t seed) {
    float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    float tx = x - fx, ty = y - fy, tz = z - fz;
    auto sm = [](float t) { return t * t * (3.0f - 2.0f * t); };
    tx = sm(tx); ty = sm(ty); tz = sm(tz);
    float c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int k = 0; k < 2; ++k) c[a][b][k] = lattice(ix + a, iy + b, iz + k, seed);
    float x00 = c[0][0][0] + (c[1][0][0] - c[0][0][0]) * tx, x10 = c[0][1][0] + (c[1][1][0] - c[0][1][0]) * tx;
    float x01 = c[0][0][1] + (c[1][0][1] - c[0][0][1]) * tx, x11 = c[0][1][1] + (c[1][1][1] - c[0][1][1]) * tx;
    float y0 = x00 + (x10 - x00) * ty, y1 = x01 + (x11 - x01) * ty;
    return y0 + (y1 - y0) * tz;
}

void medium_scene(SceneData &s, Builder &b, Rng &rng, const nori_scene_options &opt) {
    const uint32_t N = opt.grid ? opt.grid : 256;
    const uint32_t seed = hash32((uint32_t)(rng.next() * 16777216.0f));
    std::vector<float> grid((size_t)N * N * N);
    for (uint32_t z = 0; z < N; ++z)
        for (uint32_t y = 0; y < N; ++y)
            for (uint32_t x = 0; x < N; ++x) {
                float px = ((float)x + 0.5f) / (float)N, py = ((float)y + 0.5f) / (float)N, pz = ((float)z + 0.5f) / (float)N;
                float f = 0.0f, amp = 0.5f, freq = 4.0f;
                for (int o = 0; o < 4; ++o) {
                    f += amp * value_noise(px * freq, py * freq, pz * freq, seed + (uint32_t)o * 101u);
                    amp *= 0.5f;
                    freq *= 2.0f;
                }
                // soft falloff towards the box faces keeps the cloud inside
                float e = std::min(std::min(std::min(px, 1 - px), std::min(py, 1 - py)), std::min(pz, 1 - pz));
                grid[(size_t)x + (size_t)N * ((size_t)y + (size_t)N * z)] = f * std::min(1.0f, e * 8.0f);
            }
    float hi = 0.0f;
    for (float v : grid) hi = std::max(hi, v);
    const float thr = 0.9f * hi * 0.5f;
    float top = 0.0f;
    for (float &v : grid) {
        v = std::max(0.0f, v - thr);
        top = std::max(top, v);
    }
    for (float &v : grid) v = top > 0.0f ? v / top : 0.0f;
    s.grids.push_back(std::move(grid));
    nori_volume v;
    std::memset(&v, 0, sizeof(v));
    const float lo[3] = {-1, 0, -1}, up[3] = {1, 2, 1};
    for (int k = 0; k < 3; ++k) {
        v.vmin[k] = lo[k];
        v.vmax[k] = up[k];
        v.sigma_s[k] = 18.0f;
        v.sigma_a[k] = 2.0f;
    }
    const float gs[3] = {0.0f, 0.3f, 0.8f};
    v.g = gs[std::min(opt.variant, 2u)];
    v.heterogeneous = 1;
    v.width = v.height = v.depth = N;
    v.density = s.grids.back().data();
    v.brick = 16;
    s.vols.push_back(v);
    nori_material wall = default_material();
    wall.albedo[0] = wall.albedo[1] = wall.albedo[2] = 0.6f;
    b.quad(mk(-3, 0, 3), mk(6, 0, 0), mk(0, 0, -6), wall);  // floor, normal +y
    b.quad(mk(-3, 0, -3), mk(6, 0, 0), mk(0, 3, 0), wall);  // back wall, normal +z
    b.quad(mk(-3, 0, 3), mk(0, 0, -6), mk(0, 3, 0), wall);  // left wall, normal +x
    b.quad(mk(3, 0, -3), mk(0, 0, 6), mk(0, 3, 0), wall);   // right wall, normal -x
}

void lookat(f3 eye, f3 target, f3 up, float *M) {
    f3 dir = normalize(target - eye);
    f3 left = normalize(cross(up, dir));
    f3 nup = cross(dir, left);
    float m[16] = {left.x, nup.x, dir.x, eye.x, left.y, nup.y, dir.y, eye.y,
                   left.z, nup.z, dir.z, eye.z, 0, 0, 0, 1};
    std::memcpy(M, m, sizeof(m));
}

} // namespace

// Unit icosphere: the 20 faces of the icosahedron, each split into 4 per level
// with shared, renormalised edge midpoints.  level L gives 20*4^L triangles.
void icosphere_mesh(uint32_t level, std::vector<float> &pos, std::vector<uint32_t> &tri) {
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    std::vector<f3> v = {mk(-1, t, 0), mk(1, t, 0),  mk(-1, -t, 0), mk(1, -t, 0), mk(0, -1, t),  mk(0, 1, t),
                         mk(0, -1, -t), mk(0, 1, -t), mk(t, 0, -1),  mk(t, 0, 1),  mk(-t, 0, -1), mk(-t, 0, 1)};
    for (auto &p : v) p = normalize(p);
    std::vector<uint32_t> f = {0, 11, 5, 0, 5,  1, 0, 1, 7, 0, 7,  10, 0, 10, 11, 1, 5, 9, 5, 11,
                               4, 11, 10, 2, 10, 7, 6, 7, 1, 8, 3,  9, 4, 3, 4,  2, 3, 2, 6, 3,
                               6, 8, 3,  8, 9,  4, 9, 5, 2, 4, 11, 6, 2, 10, 8,  6, 7, 9, 8, 1};
    for (uint32_t l = 0; l < level; ++l) {
        std::map<uint64_t, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            f3 m = normalize((v[a] + v[b]) * 0.5f);
            v.push_back(m);
            uint32_t id = (uint32_t)v.size() - 1;
            mid[key] = id;
            return id;
        };
        std::vector<uint32_t> nf;
        nf.reserve(f.size() * 4);
        for (size_t i = 0; i < f.size(); i += 3) {
            uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
            uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
        }
        f.swap(nf);
    }
    pos.clear();
    for (auto &p : v) pos.insert(pos.end(), {p.x, p.y, p.z});
    tri = f;
}

void SceneData::finalize() {
    // Surface areas (float sum of 0.5*|e1 x e2|, include/dpdf.h append order) and
    // the uniform emitter selection probability (shaders/Emitter.hlsli:51,87).
    uint32_t ne = 0;
    for (auto &m : mats) ne += m.isEmitter ? 1 : 0;
    for (size_t k = 0; k < mats.size(); ++k) {
        float sum = 0.0f;
        for (uint32_t i = 0; i < mesh_tri_count[k]; ++i) {
            uint32_t f = mesh_tri_offset[k] + i;
            const float *a = &pos[3 * idx[3 * f]], *b = &pos[3 * idx[3 * f + 1]], *c = &pos[3 * idx[3 * f + 2]];
            f3 p0 = mk(a[0], a[1], a[2]), p1 = mk(b[0], b[1], b[2]), p2 = mk(c[0], c[1], c[2]);
            sum += 0.5f * length(cross(p1 - p0, p2 - p0));
        }
        mats[k].surfaceArea = sum;
        mats[k].emitterSelectionProb = mats[k].isEmitter && ne ? 1.0f / (float)ne : 0.0f;
    }
    texs.clear();
    for (size_t i = 0; i < tex_data.size(); ++i) texs.push_back(nori_texture{tex_w[i], tex_h[i], tex_srgb[i], tex_data[i].data()});
}

// ------------------------------------------------------------ config 5
// Procedural textures: an sRGB albedo pattern (two seeded colours mixed by
// checker, stripes or value noise) and a tangent-space normal map from a sum
// of sines height field; one alpha texture of vertical strands for the hair
// cards.  Rows are filled in parallel.
template <class F> void parallel_rows(uint32_t n, F f) {
    const uint32_t nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    std::atomic<uint32_t> next(0);
    for (uint32_t t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (uint32_t y; (y = next.fetch_add(1)) < n;) f(y);
        });
    for (auto &t : th) t.join();
}

void add_texture(SceneData &s, uint32_t w, uint32_t h, bool srgb, std::vector<uint8_t> &&rgba) {
    s.tex_data.push_back(std::move(rgba));
    s.tex_w.push_back(w);
    s.tex_h.push_back(h);
    s.tex_srgb.push_back(srgb ? 1u : 0u);
}

uint8_t to_byte(float v) { return (uint8_t)std::min(255.0f, std::max(0.0f, v * 255.0f + 0.5f)); }

void albedo_texture(SceneData &s, uint32_t n, uint32_t seed) {
    Rng r(0xA1BED000ull + seed);
    const f3 ca = mk(r.next(), r.next(), r.next()), cb = mk(r.next(), r.next(), r.next());
    const uint32_t kind = seed % 3, freq = 2 + (uint32_t)(r.next() * 14.0f);
    std::vector<uint8_t> t((size_t)n * n * 4);
    parallel_rows(n, [&](uint32_t y) {
        for (uint32_t x = 0; x < n; ++x) {
            const float u = ((float)x + 0.5f) / (float)n, v = ((float)y + 0.5f) / (float)n;
            float m;
            if (kind == 0) m = (((uint32_t)(u * freq) + (uint32_t)(v * freq)) & 1u) ? 1.0f : 0.0f;
            else if (kind == 1) m = 0.5f + 0.5f * std::sin(2.0f * PI * (float)freq * (u + 0.3f * v));
            else m = value_noise(u * freq, v * freq, 0.5f, seed);
            const f3 c = ca * (1.0f - m) + cb * m;
            uint8_t *o = &t[((size_t)y * n + x) * 4];
            o[0] = to_byte(c.x); o[1] = to_byte(c.y); o[2] = to_byte(c.z); o[3] = 255;
        }
    });
    add_texture(s, n, n, true, std::move(t));
}

void normal_texture(SceneData &s, uint32_t n, uint32_t seed) {
    Rng r(0x0B0A7000ull + seed);
    const float fu = 1.0f + std::floor(r.next() * 8.0f), fv = 1.0f + std::floor(r.next() * 8.0f), amp = 0.02f + 0.05f * r.next();
    std::vector<uint8_t> t((size_t)n * n * 4);
    parallel_rows(n, [&](uint32_t y) {
        for (uint32_t x = 0; x < n; ++x) {
            const float u = ((float)x + 0.5f) / (float)n, v = ((float)y + 0.5f) / (float)n;
            // height h = amp (sin(2 pi fu u) + sin(2 pi fv v)); normal ~ (-dh/du, -dh/dv, 1)
            const float du = amp * 2.0f * PI * fu * std::cos(2.0f * PI * fu * u);
            const float dv = amp * 2.0f * PI * fv * std::cos(2.0f * PI * fv * v);
            const f3 nn = normalize(mk(-du, -dv, 1.0f));
            uint8_t *o = &t[((size_t)y * n + x) * 4];
            o[0] = to_byte(nn.x * 0.5f + 0.5f); o[1] = to_byte(nn.y * 0.5f + 0.5f); o[2] = to_byte(nn.z * 0.5f + 0.5f);
            o[3] = 255;
        }
    });
    add_texture(s, n, n, false, std::move(t));
}

void strands_alpha_texture(SceneData &s, uint32_t n) {
    std::vector<uint8_t> t((size_t)n * n * 4);
    parallel_rows(n, [&](uint32_t y) {
        for (uint32_t x = 0; x < n; ++x) {
            const float u = ((float)x + 0.5f) / (float)n, v = ((float)y + 0.5f) / (float)n;
            // eight strands across the card, thinning towards the tip (v = 1)
            const float w = 0.5f * (1.0f - 0.7f * v);
            const float c = std::fabs(std::fmod(u * 8.0f, 1.0f) - 0.5f) / std::max(w, 1e-3f);
            const float a = std::max(0.0f, 1.0f - c * c);
            uint8_t *o = &t[((size_t)y * n + x) * 4];
            o[0] = to_byte(a); o[1] = to_byte(a); o[2] = to_byte(a); o[3] = to_byte(a);
        }
    });
    add_texture(s, n, n, false, std::move(t));
}

float terrain_height(float x, float z) {
    float h = 0.0f, amp = 2.0f, f = 0.04f;
    for (int o = 0; o < 4; ++o, amp *= 0.5f, f *= 2.0f) h += amp * (value_noise(x * f, z * f, 3.7f, 0x7E44u + o) - 0.5f);
    return h;
}

// Batched large scene: a heightfield terrain, displaced icospheres, alpha-tested
// hair cards, 128 textures (64 albedo + 64 normal maps) plus the cards' alpha.
void batched_scene(SceneData &s, Builder &b, Rng &rng, const nori_scene_options &opt) {
    const uint32_t N = opt.grid ? opt.grid : 2048, level = opt.subdiv ? opt.subdiv : 4;
    const uint32_t nsph = opt.meshes ? opt.meshes : 199, ncards = opt.strands ? opt.strands : 20000;
    const uint32_t tsz = opt.tex_size ? opt.tex_size : 1024;
    for (uint32_t i = 0; i < 64; ++i) albedo_texture(s, tsz, i);
    for (uint32_t i = 0; i < 64; ++i) normal_texture(s, tsz, i);
    strands_alpha_texture(s, tsz);
    const uint32_t ALPHA_TEX = 128;
    // terrain: N x N cells over [-50, 50]^2, uv = world / 10 (tiling textures)
    {
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> tri;
        const float S = 100.0f, h = S / (float)N;
        for (uint32_t j = 0; j <= N; ++j)
            for (uint32_t i = 0; i <= N; ++i) {
                const float x = -50.0f + h * (float)i, z = -50.0f + h * (float)j;
                const float y = terrain_height(x, z);
                const float e = 0.5f * h;
                const f3 nn = normalize(mk(terrain_height(x - e, z) - terrain_height(x + e, z), 2.0f * e,
                                           terrain_height(x, z - e) - terrain_height(x, z + e)));
                pos.insert(pos.end(), {x, y, z});
                nrm.insert(nrm.end(), {nn.x, nn.y, nn.z});
                uv.insert(uv.end(), {x / 10.0f, z / 10.0f});
            }
        for (uint32_t j = 0; j < N; ++j)
            for (uint32_t i = 0; i < N; ++i) {
                const uint32_t a = j * (N + 1) + i, bb = a + 1, c = a + N + 1, d = c + 1;
                tri.insert(tri.end(), {a, c, bb, bb, c, d});
            }
        nori_material m = default_material();
        m.albedoTexIndex = 0;
        m.normalTexIndex = 64;
        b.add_mesh(pos, nrm, tri, m, &uv);
    }
    // displaced icospheres with spherical uv: the first 36 of every 199 one
    // subdivision level finer (at the defaults 36 x 20,480 + 163 x 5,120
    // triangles, so the scene holds 10,000,448 with the terrain and the cards)
    std::vector<float> ipos, fpos;
    std::vector<uint32_t> itri, ftri;
    icosphere_mesh(level, ipos, itri);
    icosphere_mesh(level + 1, fpos, ftri);
    const uint32_t nfine = (nsph * 36 + 198) / 199;
    for (uint32_t k = 0; k < nsph; ++k) {
        const std::vector<float> &sp = k < nfine ? fpos : ipos;
        const std::vector<uint32_t> &st = k < nfine ? ftri : itri;
        const float r = 0.5f + 1.5f * rng.next();
        const float x = -40.0f + 80.0f * rng.next(), z = -40.0f + 80.0f * rng.next();
        const f3 c = mk(x, terrain_height(x, z) + r * 0.8f, z);
        const uint32_t ns = 0x51u + k;
        std::vector<float> pos(sp.size()), nrm(sp.size()), uv(sp.size() / 3 * 2);
        for (size_t v = 0; v < sp.size(); v += 3) {
            const f3 d = mk(sp[v], sp[v + 1], sp[v + 2]);
            const float rr = r * (0.9f + 0.2f * value_noise(d.x * 3.0f + 7.0f, d.y * 3.0f, d.z * 3.0f, ns));
            const f3 p = c + d * rr;
            pos[v] = p.x; pos[v + 1] = p.y; pos[v + 2] = p.z;
            nrm[v] = d.x; nrm[v + 1] = d.y; nrm[v + 2] = d.z;
            uv[v / 3 * 2] = 0.5f + std::atan2(d.z, d.x) / (2.0f * PI);
            uv[v / 3 * 2 + 1] = std::acos(std::max(-1.0f, std::min(1.0f, d.y))) / PI;
        }
        nori_material m = default_material();
        const float pick = rng.next();
        const uint32_t t = 1 + k % 63;
        if (pick < 0.4f) {
            m.type = NORI_MAT_DISNEY;
            m.metallic = rng.next();
            m.roughness = 0.1f + 0.9f * rng.next();
            m.clearcoat = rng.next();
            m.sheen = rng.next();
            m.albedoTexIndex = t;
            m.normalTexIndex = 64 + t;
        } else if (pick < 0.8f) {
            m.albedoTexIndex = t;
            m.normalTexIndex = 64 + t;
        } else {
            m.type = NORI_MAT_DIELECTRIC;
            m.intIOR = 1.5f;
            m.extIOR = 1.0f;
        }
        b.add_mesh(pos, nrm, st, m, &uv);
    }
    // hair cards: 0.08 x 0.8 quads standing on the terrain near the camera's
    // focus distance, alpha-tested with the strands texture
    {
        std::vector<float> pos, nrm, uv;
        std::vector<uint32_t> tri;
        for (uint32_t k = 0; k < ncards; ++k) {
            const float x = -6.0f + 12.0f * rng.next(), z = -22.0f + 6.0f * rng.next(), yaw = 2.0f * PI * rng.next();
            const f3 base = mk(x, terrain_height(x, z), z), a = mk(std::cos(yaw), 0.0f, std::sin(yaw)) * 0.08f;
            const f3 up = mk(0.15f * (rng.next() - 0.5f), 0.8f, 0.15f * (rng.next() - 0.5f));
            const f3 n = normalize(cross(a, up));
            const uint32_t v0 = (uint32_t)(pos.size() / 3);
            const f3 q[4] = {base - a * 0.5f, base + a * 0.5f, base + a * 0.5f + up, base - a * 0.5f + up};
            const float quv[8] = {0, 0, 1, 0, 1, 1, 0, 1};
            for (int i = 0; i < 4; ++i) {
                pos.insert(pos.end(), {q[i].x, q[i].y, q[i].z});
                nrm.insert(nrm.end(), {n.x, n.y, n.z});
                uv.insert(uv.end(), {quv[2 * i], quv[2 * i + 1]});
            }
            tri.insert(tri.end(), {v0, v0 + 1, v0 + 2, v0, v0 + 2, v0 + 3});
        }
        nori_material m = default_material();
        m.type = NORI_MAT_DISNEY;
        m.albedo[0] = 0.35f; m.albedo[1] = 0.22f; m.albedo[2] = 0.12f;
        m.roughness = 0.5f;
        m.alphaTexIndex = ALPHA_TEX;
        b.add_mesh(pos, nrm, tri, m, &uv);
    }
}

SceneData generate_scene(const nori_scene_options &opt) {
    SceneData s;
    Builder b{s};
    const int cfg = opt.config;
    Rng rng(opt.seed ? opt.seed : 0x5EED0000ull + (uint64_t)cfg);
    if (cfg == 1) {
        uint32_t level = opt.subdiv ? opt.subdiv : 4;
        nori_material white = default_material(), red = white, green = white, light = white;
        white.albedo[0] = white.albedo[1] = white.albedo[2] = 0.73f;
        red.albedo[0] = 0.65f; red.albedo[1] = 0.05f; red.albedo[2] = 0.05f;
        green.albedo[0] = 0.12f; green.albedo[1] = 0.45f; green.albedo[2] = 0.15f;
        b.quad(mk(-1, 0, 1), mk(2, 0, 0), mk(0, 0, -2), white);  // floor, normal +y
        b.quad(mk(-1, 2, -1), mk(2, 0, 0), mk(0, 0, 2), white);  // ceiling, normal -y
        b.quad(mk(-1, 0, -1), mk(2, 0, 0), mk(0, 2, 0), white);  // back wall, normal +z
        b.quad(mk(-1, 0, 1), mk(0, 0, -2), mk(0, 2, 0), red);    // left wall, normal +x
        b.quad(mk(1, 0, -1), mk(0, 0, 2), mk(0, 2, 0), green);   // right wall, normal -x
        float r1 = 0.33f + 0.04f * rng.next(), r2 = 0.33f + 0.04f * rng.next();
        b.icosphere(mk(-0.42f, r1, -0.30f + 0.1f * rng.next()), r1, level, white);
        b.icosphere(mk(0.42f, r2, 0.25f + 0.1f * rng.next()), r2, level, white);
        light.isEmitter = 1;
        light.albedo[0] = light.albedo[1] = light.albedo[2] = 0.0f;
        light.radiance[0] = 17.0f; light.radiance[1] = 12.0f; light.radiance[2] = 4.0f;
        b.quad(mk(-0.25f, 1.999f, -0.25f), mk(0.5f, 0, 0), mk(0, 0, 0.5f), light); // normal -y
        s.camera.fov_deg = 40.0f;
        lookat(mk(0, 1, 3.7f), mk(0, 1, 0), mk(0, 1, 0), s.camera.to_world);
        s.camera.width = opt.width ? opt.width : 256;
        s.camera.height = opt.height ? opt.height : 256;
    } else if (cfg == 2) {
        uint32_t level = opt.subdiv ? opt.subdiv : 6, g = opt.grid ? opt.grid : 8;
        nori_material ground = default_material();
        ground.albedo[0] = ground.albedo[1] = ground.albedo[2] = 0.5f;
        b.quad(mk(-20, 0, 20), mk(40, 0, 0), mk(0, 0, -40), ground);
        for (uint32_t i = 0; i < g; ++i)
            for (uint32_t j = 0; j < g; ++j) {
                nori_material m = default_material();
                m.type = NORI_MAT_DISNEY;
                m.albedo[0] = rng.next(); m.albedo[1] = rng.next(); m.albedo[2] = rng.next();
                m.metallic = rng.next();
                m.specular = rng.next();
                m.sheen = rng.next();
                m.clearcoat = rng.next();
                m.roughness = 0.05f + 0.95f * rng.next();
                m.anisotropic = 0.8f * rng.next();
                float x = ((float)i - 0.5f * (float)(g - 1)) * 1.0f, z = ((float)j - 0.5f * (float)(g - 1)) * 1.0f;
                b.icosphere(mk(x, 0.4f, z), 0.4f, level, m);
            }
        sky_sun_env(s, rng, opt);
        s.camera.fov_deg = 50.0f;
        lookat(mk(0, 4.5f, 9.5f), mk(0, 0.3f, 0), mk(0, 1, 0), s.camera.to_world);
        s.camera.width = opt.width ? opt.width : 512;
        s.camera.height = opt.height ? opt.height : 512;
    } else if (cfg == 3) {
        const f3 eye = mk(0, 0, 45), target = mk(0, -1, 0);
        hair_scene(b, rng, opt, eye);
        if (opt.variant == 1) {
            uint32_t W = opt.env_width ? opt.env_width : 64, H = opt.env_height ? opt.env_height : 32;
            s.env_w = W; s.env_h = H; s.env_scale = 1.0f;
            s.env_rgb.assign((size_t)W * H * 3, 1.0f);
        }
        s.camera.fov_deg = 45.0f;
        lookat(eye, target, mk(0, 1, 0), s.camera.to_world);
        s.camera.width = opt.width ? opt.width : 512;
        s.camera.height = opt.height ? opt.height : 512;
    } else if (cfg == 4) {
        medium_scene(s, b, rng, opt);
        sky_sun_env(s, rng, opt);
        s.camera.fov_deg = 45.0f;
        lookat(mk(0, 1.4f, 5.0f), mk(0, 0.9f, 0), mk(0, 1, 0), s.camera.to_world);
        s.camera.width = opt.width ? opt.width : 512;
        s.camera.height = opt.height ? opt.height : 512;
    } else if (cfg == 5) {
        b.attrs = true;
        batched_scene(s, b, rng, opt);
        sky_sun_env(s, rng, opt);
        s.camera.fov_deg = 40.0f;
        lookat(mk(0, 2.0f, -24.5f), mk(0, 0.8f, -19.5f), mk(0, 1, 0), s.camera.to_world);
        s.camera.width = opt.width ? opt.width : 512;
        s.camera.height = opt.height ? opt.height : 512;
    } else {
        throw std::invalid_argument("unknown scene config");
    }
    s.camera.near_clip = 1e-4f;
    s.camera.far_clip = 1e4f;
    // config 5: thin lens, aperture radius 0.05 focused at 5
    s.camera.lens_radius = cfg == 5 ? 0.05f : 0.0f;
    s.camera.focal_distance = cfg == 5 ? 5.0f : 1.0f;
    s.finalize();
    return s;
}

void SceneData::fill_desc(nori_scene_desc &d) const {
    std::memset(&d, 0, sizeof(d));
    d.num_vertices = (uint32_t)(pos.size() / 3);
    d.num_triangles = (uint32_t)(idx.size() / 3);
    d.num_meshes = (uint32_t)mats.size();
    d.positions = pos.data();
    d.normals = nrm.empty() ? nullptr : nrm.data();
    d.uvs = uv.empty() ? nullptr : uv.data();
    d.tangents = tan.empty() ? nullptr : tan.data();
    d.indices = idx.data();
    d.mesh_tri_offset = mesh_tri_offset.data();
    d.mesh_tri_count = mesh_tri_count.data();
    d.materials = mats.data();
    d.env_width = env_w;
    d.env_height = env_h;
    d.env_rgb = env_rgb.empty() ? nullptr : env_rgb.data();
    d.env_scale = env_scale;
    d.num_volumes = (uint32_t)vols.size();
    d.volumes = vols.empty() ? nullptr : vols.data();
    d.num_textures = (uint32_t)texs.size();
    d.textures = texs.empty() ? nullptr : texs.data();
}

} // namespace ng
