// CPU oracle — test infrastructure only.
//
// Participating media as shaders/Volume.hlsl evaluates them (FP32):
//   * ray/AABB slabs (:14-27), Henyey-Greenstein eval/sample (:31-65);
//   * density: trilinear lookup of a dense W x H x D grid over the volume's box,
//     clamp addressing, texel centres at (i + 0.5) / W (the Texture3D sample of
//     :95-102); homogeneous volumes have density 1;
//   * local majorant per brick (:119-164) and null-collision free flight with a
//     hero channel (max sigma_t) (:236-319), ratio tracking (:326-372), segment
//     enumeration (:187-222), the multi-volume entry points (:381-454);
//   * volume NEE at scatter points (shaders/Emitter.hlsli:334-434), including
//     its skip when minDist * min sigma_t > 8.
//
// Conventions chosen here (the reference leaves them open):
//   * the brick majorant is the maximum density over every voxel the trilinear
//     lookup can touch inside the brick (the brick's voxels plus a one-voxel
//     apron), so density / majorant <= 1 everywhere; the reference takes the
//     maximum over the brick's own voxels only, where interpolation can exceed
//     it and the clamped real-collision probability biases the estimate;
//   * the brick edge is a configurable number of voxels (16 by default);
//   * free flight and ratio tracking walk the brick grid as a 3D DDA: the brick
//     index is carried across faces (the current brick +-1 along the axis the
//     ray leaves by) and each exit distance is taken from the ray origin.  The
//     reference re-derives the brick from the rounded position after stepping
//     exactly onto a face (with a 1e-5 minimum step), so which brick the next
//     step sees depends on the last bit of that position: two FP32
//     implementations that round the position differently (a fused
//     multiply-add) then diverge at brick faces;
//   * the transmittance of a next-event shadow ray is estimated with its own
//     pcg32 stream seed(s, PCG32_DEFAULT_STREAM), s being one 32-bit draw from the
//     path's stream taken right after the light sample (only when the scene has
//     volumes), and only for shadow rays that reach the light.
#pragma once
#include "orc_math.h"
#include "orc_pcg32.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace orc {

struct Volume {
    V3f vmin, vmax, sigma_a, sigma_s;
    float g = 0.0f;
    bool hetero = false;
    uint32_t W = 0, H = 0, D = 0; // density grid (empty: density 1)
    std::vector<float> density;   // x fastest
    uint32_t mW = 0, mH = 0, mD = 0, brick = 16;
    std::vector<float> majorant;  // per brick, x fastest

    V3f sigma_t() const { return sigma_a + sigma_s; }
    float sigma_t_max() const { V3f s = sigma_t(); return std::max(s.x, std::max(s.y, s.z)); }
    bool has_grid() const { return hetero && !density.empty(); }

    void build_majorant() {
        if (!has_grid()) return;
        const uint32_t B = brick ? brick : 16;
        mW = (W + B - 1) / B;
        mH = (H + B - 1) / B;
        mD = (D + B - 1) / B;
        majorant.assign((size_t)mW * mH * mD, 0.0f);
        auto range = [](uint32_t k, uint32_t m, uint32_t n, int &lo, int &hi) {
            // voxels a trilinear lookup touches for normalised coordinates in
            // [k/m, (k+1)/m]: floor(u*n - 0.5) .. floor(u*n - 0.5) + 1, clamped
            double a = (double)k / m * n - 0.5, b = (double)(k + 1) / m * n - 0.5;
            lo = std::max(0, (int)std::floor(a));
            hi = std::min((int)n - 1, (int)std::floor(b) + 1);
        };
        for (uint32_t bz = 0; bz < mD; ++bz)
            for (uint32_t by = 0; by < mH; ++by)
                for (uint32_t bx = 0; bx < mW; ++bx) {
                    int x0, x1, y0, y1, z0, z1;
                    range(bx, mW, W, x0, x1);
                    range(by, mH, H, y0, y1);
                    range(bz, mD, D, z0, z1);
                    float m = 0.0f;
                    for (int z = z0; z <= z1; ++z)
                        for (int y = y0; y <= y1; ++y)
                            for (int x = x0; x <= x1; ++x)
                                m = std::max(m, density[(size_t)x + (size_t)W * ((size_t)y + (size_t)H * z)]);
                    majorant[(size_t)bx + (size_t)mW * ((size_t)by + (size_t)mH * bz)] = m;
                }
    }

    float density_at(V3f p) const {
        if (!has_grid()) return 1.0f;
        V3f size = V3f(std::max(vmax.x - vmin.x, 1e-6f), std::max(vmax.y - vmin.y, 1e-6f), std::max(vmax.z - vmin.z, 1e-6f));
        V3f uvw = V3f((p.x - vmin.x) / size.x, (p.y - vmin.y) / size.y, (p.z - vmin.z) / size.z);
        float u = uvw.x * (float)W - 0.5f, v = uvw.y * (float)H - 0.5f, w = uvw.z * (float)D - 0.5f;
        float fu = std::floor(u), fv = std::floor(v), fw = std::floor(w);
        float tx = u - fu, ty = v - fv, tz = w - fw;
        int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw;
        auto at = [&](int x, int y, int z) {
            x = std::min(std::max(x, 0), (int)W - 1);
            y = std::min(std::max(y, 0), (int)H - 1);
            z = std::min(std::max(z, 0), (int)D - 1);
            return density[(size_t)x + (size_t)W * ((size_t)y + (size_t)H * z)];
        };
        float c00 = at(x0, y0, z0) * (1 - tx) + at(x0 + 1, y0, z0) * tx;
        float c10 = at(x0, y0 + 1, z0) * (1 - tx) + at(x0 + 1, y0 + 1, z0) * tx;
        float c01 = at(x0, y0, z0 + 1) * (1 - tx) + at(x0 + 1, y0, z0 + 1) * tx;
        float c11 = at(x0, y0 + 1, z0 + 1) * (1 - tx) + at(x0 + 1, y0 + 1, z0 + 1) * tx;
        float c0 = c00 * (1 - ty) + c10 * ty, c1 = c01 * (1 - ty) + c11 * ty;
        return c0 * (1 - tz) + c1 * tz;
    }

    // shaders/Volume.hlsl:119-164
    void local_majorant(V3f pos, V3f dir, float &mu, float &bd) const {
        const float gmu = sigma_t_max();
        if (!has_grid()) {
            mu = gmu;
            bd = 1e20f;
            return;
        }
        V3f size = V3f(std::max(vmax.x - vmin.x, 1e-6f), std::max(vmax.y - vmin.y, 1e-6f), std::max(vmax.z - vmin.z, 1e-6f));
        const float mres[3] = {(float)mW, (float)mH, (float)mD};
        const float sz[3] = {size.x, size.y, size.z}, lo[3] = {vmin.x, vmin.y, vmin.z};
        const float p[3] = {pos.x, pos.y, pos.z}, d[3] = {dir.x, dir.y, dir.z};
        int bc[3];
        for (int i = 0; i < 3; ++i) {
            float bf = (p[i] - lo[i]) / sz[i] * mres[i];
            bc[i] = std::min(std::max((int)bf, 0), (int)mres[i] - 1);
        }
        mu = majorant[(size_t)bc[0] + (size_t)mW * ((size_t)bc[1] + (size_t)mH * bc[2])] * gmu;
        float dist[3];
        for (int i = 0; i < 3; ++i) {
            if (std::fabs(d[i]) < 1e-20f) {
                dist[i] = 1e20f;
            } else {
                float edge = (float)bc[i] + (d[i] > 0.0f ? 1.0f : 0.0f);
                float world = lo[i] + edge * (sz[i] / mres[i]);
                dist[i] = std::max((world - p[i]) / d[i], 1e-5f);
            }
        }
        bd = std::min(dist[0], std::min(dist[1], dist[2]));
    }
};

// A ray's walk through the majorant bricks (3D DDA).  Without a grid the whole
// box is one brick with the global majorant.
struct BrickWalk {
    int bc[3] = {0, 0, 0};
    bool grid = false;

    void start(const Volume &v, V3f o, V3f d, float t) {
        grid = v.has_grid();
        if (!grid) return;
        const V3f p = o + d * t;
        const float pp[3] = {p.x, p.y, p.z}, lo[3] = {v.vmin.x, v.vmin.y, v.vmin.z};
        const float sz[3] = {std::max(v.vmax.x - v.vmin.x, 1e-6f), std::max(v.vmax.y - v.vmin.y, 1e-6f),
                             std::max(v.vmax.z - v.vmin.z, 1e-6f)};
        const int m[3] = {(int)v.mW, (int)v.mH, (int)v.mD};
        for (int i = 0; i < 3; ++i) bc[i] = std::min(std::max((int)((pp[i] - lo[i]) / sz[i] * (float)m[i]), 0), m[i] - 1);
    }
    float majorant(const Volume &v) const {
        const float gmu = v.sigma_t_max();
        return grid ? v.majorant[(size_t)bc[0] + (size_t)v.mW * ((size_t)bc[1] + (size_t)v.mH * bc[2])] * gmu : gmu;
    }
    // ray parameter where the ray leaves the current brick, and the axis it leaves by
    float exit(const Volume &v, V3f o, V3f d, int &axis) const {
        axis = -1;
        if (!grid) return 1e20f;
        const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, lo[3] = {v.vmin.x, v.vmin.y, v.vmin.z};
        const float sz[3] = {std::max(v.vmax.x - v.vmin.x, 1e-6f), std::max(v.vmax.y - v.vmin.y, 1e-6f),
                             std::max(v.vmax.z - v.vmin.z, 1e-6f)};
        const float m[3] = {(float)v.mW, (float)v.mH, (float)v.mD};
        float best = 1e20f;
        for (int i = 0; i < 3; ++i) {
            if (std::fabs(dd[i]) < 1e-20f) continue;
            const float edge = lo[i] + (float)(bc[i] + (dd[i] > 0.0f ? 1 : 0)) * (sz[i] / m[i]);
            const float ti = (edge - oo[i]) / dd[i];
            if (ti < best) { best = ti; axis = i; }
        }
        return best;
    }
    // into the neighbouring brick across `axis`; false once the walk leaves the grid
    bool advance(const Volume &v, V3f d, int axis) {
        if (axis < 0) return false;
        const float dd[3] = {d.x, d.y, d.z};
        const int m[3] = {(int)v.mW, (int)v.mH, (int)v.mD};
        bc[axis] += dd[axis] > 0.0f ? 1 : -1;
        return bc[axis] >= 0 && bc[axis] < m[axis];
    }
};

// shaders/Volume.hlsl:14-27
inline bool ray_aabb(V3f o, V3f d, V3f bmin, V3f bmax, float &tn, float &tf) {
    V3f inv(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    V3f t0((bmin.x - o.x) * inv.x, (bmin.y - o.y) * inv.y, (bmin.z - o.z) * inv.z);
    V3f t1((bmax.x - o.x) * inv.x, (bmax.y - o.y) * inv.y, (bmax.z - o.z) * inv.z);
    tn = std::max(std::max(std::min(t0.x, t1.x), std::min(t0.y, t1.y)), std::min(t0.z, t1.z));
    tf = std::min(std::min(std::max(t0.x, t1.x), std::max(t0.y, t1.y)), std::max(t0.z, t1.z));
    return tf >= std::max(tn, 0.0f);
}

// shaders/Volume.hlsl:31-36
inline float hg_eval(float cos_theta, float g) {
    float g2 = g * g, den = 1.0f + g2 - 2.0f * g * cos_theta;
    return (1.0f / (4.0f * 3.14159265358979323846f)) * (1.0f - g2) / (den * std::sqrt(std::max(den, 1e-8f)));
}

// shaders/Volume.hlsl:38-65 with BuildONB (shaders/GeometryUtils.hlsli:160-165)
inline V3f hg_sample(V3f wo, float g, float xi1, float xi2, float &pdf) {
    float ct;
    if (std::fabs(g) < 1e-3f) {
        ct = 1.0f - 2.0f * xi1;
    } else {
        float sqr = (1.0f - g * g) / (1.0f + g - 2.0f * g * xi1);
        ct = (1.0f + g * g - sqr * sqr) / (2.0f * g);
        ct = std::min(std::max(ct, -1.0f), 1.0f);
    }
    float st = std::sqrt(std::max(0.0f, 1.0f - ct * ct)), phi = 2.0f * 3.14159265358979323846f * xi2;
    V3f up = std::fabs(wo.y) < 0.999f ? V3f(0, 1, 0) : V3f(1, 0, 0);
    V3f T = normalize(cross(up, wo)), B = cross(wo, T);
    V3f wi = T * (st * std::cos(phi)) + B * (st * std::sin(phi)) + wo * ct;
    pdf = hg_eval(ct, g);
    return normalize(wi);
}

struct VolSegment {
    uint32_t vi;
    float tn, tf;
};

// shaders/Volume.hlsl:187-222 (at most 8 segments, sorted by entry distance)
inline uint32_t enumerate_segments(const std::vector<Volume> &vols, V3f o, V3f d, float tmax, VolSegment seg[8]) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < vols.size() && n < 8; ++i) {
        float tn, tf;
        if (!ray_aabb(o, d, vols[i].vmin, vols[i].vmax, tn, tf)) continue;
        tn = std::max(tn, 0.0f);
        tf = std::min(tf, tmax);
        if (tf <= tn) continue;
        uint32_t k = n;
        for (; k > 0; --k) {
            if (seg[k - 1].tn <= tn) break;
            seg[k] = seg[k - 1];
        }
        seg[k] = VolSegment{i, tn, tf};
        ++n;
    }
    return n;
}

// shaders/Volume.hlsl:236-319, the bricks walked as a DDA (BrickWalk).  One
// uniform per tentative collision, a second for the real/null decision.
template <class Rng>
inline bool null_collision_free_flight(const Volume &v, V3f o, V3f d, float tn, float tf, Rng &rng, V3f &weight,
                                       float &ts) {
    const float gmu = v.sigma_t_max();
    const V3f st = v.sigma_t();
    float t = tn;
    BrickWalk w;
    w.start(v, o, d, t);
    for (uint32_t step = 0; step < 512; ++step) {
        const float mu = w.majorant(v);
        int axis;
        const float texit = w.exit(v, o, d, axis);
        if (mu < 1e-8f) {
            t = std::max(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) { ts = tf; return false; }
            continue;
        }
        float xi = std::max(rng.next_float(), 1e-10f);
        float dt = -std::log(xi) / mu;
        if (t + dt > texit) {
            t = std::max(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) { ts = tf; return false; }
            continue;
        }
        t += dt;
        if (t >= tf) { ts = tf; return false; }
        float dens = v.density_at(o + d * t);
        float p_real = std::min(std::max((dens * gmu) / std::max(mu, 1e-20f), 0.0f), 1.0f);
        if (p_real >= 1.0f - 1e-6f || rng.next_float() < p_real) {
            float inv = 1.0f / std::max(gmu, 1e-20f);
            weight = weight * V3f(st.x * inv, st.y * inv, st.z * inv);
            ts = t;
            return true;
        }
        float den = std::max(mu * (1.0f - p_real), 1e-20f);
        weight = weight * V3f((mu - dens * st.x) / den, (mu - dens * st.y) / den, (mu - dens * st.z) / den);
    }
    ts = tf;
    return false;
}

// shaders/Volume.hlsl:326-372, the bricks walked as a DDA (BrickWalk).
template <class Rng>
inline V3f ratio_tracking(const Volume &v, V3f o, V3f d, float tn, float tf, Rng &rng) {
    const V3f st = v.sigma_t();
    V3f tr(1.0f);
    float t = tn;
    BrickWalk w;
    w.start(v, o, d, t);
    for (uint32_t step = 0; step < 512; ++step) {
        const float mu = w.majorant(v);
        int axis;
        const float texit = w.exit(v, o, d, axis);
        if (mu < 1e-8f) {
            t = std::max(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) return tr;
            continue;
        }
        float xi = std::max(rng.next_float(), 1e-10f);
        float dt = -std::log(xi) / mu;
        if (t + dt > texit) {
            t = std::max(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) return tr;
            continue;
        }
        t += dt;
        if (t >= tf) return tr;
        float dens = v.density_at(o + d * t);
        float inv = 1.0f / std::max(mu, 1e-20f);
        tr = tr * V3f((mu - dens * st.x) * inv, (mu - dens * st.y) * inv, (mu - dens * st.z) * inv);
        if (std::max(tr.x, std::max(tr.y, tr.z)) < 1e-6f) return V3f(0.0f);
    }
    return tr;
}

// shaders/Volume.hlsl:381-415
template <class Rng>
inline bool sample_volume_free_flight(const std::vector<Volume> &vols, V3f o, V3f d, float tmax, Rng &rng,
                                      V3f &weight, float &ts, uint32_t &vi) {
    ts = tmax;
    vi = 0;
    if (vols.empty()) return false;
    VolSegment seg[8];
    uint32_t n = enumerate_segments(vols, o, d, tmax, seg);
    for (uint32_t k = 0; k < n; ++k) {
        const Volume &v = vols[seg[k].vi];
        if (v.sigma_t_max() < 1e-8f) continue;
        float th;
        if (null_collision_free_flight(v, o, d, seg[k].tn, seg[k].tf, rng, weight, th)) {
            ts = th;
            vi = seg[k].vi;
            return true;
        }
    }
    return false;
}

// shaders/Volume.hlsl:419-454
template <class Rng>
inline V3f multi_volume_transmittance(const std::vector<Volume> &vols, V3f o, V3f d, float tmax, Rng &rng) {
    if (vols.empty()) return V3f(1.0f);
    VolSegment seg[8];
    uint32_t n = enumerate_segments(vols, o, d, tmax, seg);
    V3f tr(1.0f);
    for (uint32_t k = 0; k < n; ++k) {
        const Volume &v = vols[seg[k].vi];
        V3f st = v.sigma_t();
        if (std::max(st.x, std::max(st.y, st.z)) < 1e-8f) continue;
        V3f s;
        if (v.hetero) {
            s = ratio_tracking(v, o, d, seg[k].tn, seg[k].tf, rng);
        } else {
            float len = seg[k].tf - seg[k].tn;
            s = V3f(std::exp(-st.x * len), std::exp(-st.y * len), std::exp(-st.z * len));
        }
        tr = tr * s;
        if (std::max(tr.x, std::max(tr.y, tr.z)) < 1e-6f) return V3f(0.0f);
    }
    return tr;
}

// The NEE skip of shaders/Emitter.hlsli:336-345 / 389-398.
inline bool volume_nee_skip(const Volume &v, V3f p) {
    V3f st = v.sigma_t();
    float stmin = std::min(st.x, std::min(st.y, st.z));
    float md = std::min(std::min(std::min(p.x - v.vmin.x, p.y - v.vmin.y), std::min(p.z - v.vmin.z, v.vmax.x - p.x)),
                        std::min(v.vmax.y - p.y, v.vmax.z - p.z));
    return md * stmin > 8.0f;
}

} // namespace orc
