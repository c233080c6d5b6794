// CPU oracle — test infrastructure only.
//
// Ray/triangle intersection and Nori's BVH2, restated.
//   Möller–Trumbore          src/cpu/mesh.cpp:51-89
//   slab test                include/bbox.h:360-392
//   binned + serial SAH build src/cpu/accel.cpp:55-323 (run serially here)
//   stack traversal          src/cpu/accel.cpp:429-488
//
// Canonical hit (shared with the GPU path): the smallest t over all triangles
// accepted by the Möller–Trumbore test with mint <= t <= maxt; equal t is broken
// by the smallest global primitive index.  Nori's own leaf loop keeps the last
// hit at equal t, which depends on its (TBB-ordered) leaf order; the explicit
// tie-break makes the result independent of the acceleration structure.
#pragma once
#include "orc_math.h"
#include <vector>

namespace orc {

struct Ray {
    V3f o, d, drcp;
    float mint, maxt;
};

struct Hit {
    float t, u, v;
    uint32_t prim; // 0xFFFFFFFF = miss
};

// src/cpu/mesh.cpp:51-89, operating on (p0, e1 = p1-p0, e2 = p2-p0).
inline bool intersect_tri(const V3f &p0, const V3f &e1, const V3f &e2, const Ray &r, float &u, float &v, float &t) {
    V3f pvec = cross(r.d, e2);
    float det = dot(e1, pvec);
    if (det > -1e-8f && det < 1e-8f) return false;
    float inv_det = 1.0f / det;
    V3f tvec = r.o - p0;
    u = dot(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    V3f qvec = cross(tvec, e1);
    v = dot(r.d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qvec) * inv_det;
    return t >= r.mint && t <= r.maxt;
}

struct BBox {
    V3f mn, mx;
    BBox() : mn(3.4028235e38f), mx(-3.4028235e38f) {}
    void expand(const V3f &p) {
        mn = V3f(std::min(mn.x, p.x), std::min(mn.y, p.y), std::min(mn.z, p.z));
        mx = V3f(std::max(mx.x, p.x), std::max(mx.y, p.y), std::max(mx.z, p.z));
    }
    void expand(const BBox &b) {
        mn = V3f(std::min(mn.x, b.mn.x), std::min(mn.y, b.mn.y), std::min(mn.z, b.mn.z));
        mx = V3f(std::max(mx.x, b.mx.x), std::max(mx.y, b.mx.y), std::max(mx.z, b.mx.z));
    }
    float area() const {
        V3f d = mx - mn;
        float r = 0.0f;
        for (int i = 0; i < 3; ++i) {
            float term = 1.0f;
            for (int j = 0; j < 3; ++j)
                if (i != j) term *= d[j];
            r += term;
        }
        return r * 2.0f;
    }
    int largest_axis() const {
        V3f e = mx - mn;
        if (e.x >= e.y && e.x >= e.z) return 0;
        if (e.y >= e.x && e.y >= e.z) return 1;
        return 2;
    }
    // include/bbox.h:360-392.  With robust = true the far bound is widened by
    // (1 + 2*gamma_3) (PBRT 3e §3.9.2): Nori's exact comparisons can cull a box
    // whose entry point equals the current hit distance to the last ulp, and then
    // miss an equal-t hit with a smaller primitive id.  The widening only affects
    // which boxes are visited, never which triangles are accepted.
    bool hit(const Ray &r, bool robust = true) const {
        const float g = robust ? 1.0f + 2.0f * 3.5762791e-07f : 1.0f;
        float nearT = -INFINITY, farT = INFINITY;
        for (int i = 0; i < 3; ++i) {
            float o = r.o[i], lo = mn[i], hi = mx[i];
            if (r.d[i] == 0) {
                if (o < lo || o > hi) return false;
            } else {
                float t1 = (lo - o) * r.drcp[i], t2 = (hi - o) * r.drcp[i];
                if (t1 > t2) std::swap(t1, t2);
                nearT = std::max(t1, nearT);
                farT = std::min(t2, farT);
                if (!(nearT <= farT * g)) return false;
            }
        }
        return r.mint <= farT * g && nearT <= r.maxt * g;
    }
};

struct Bvh2 {
    struct Node {
        BBox box;
        uint32_t right_or_start; // inner: right child; leaf: first index
        uint32_t size;           // leaf: triangle count
        uint8_t leaf, used;
    };
    std::vector<V3f> p0, e1, e2; // per-triangle precomputation
    std::vector<V3f> centroid;
    std::vector<BBox> tbox;
    std::vector<Node> nodes;
    std::vector<uint32_t> index;

    void build(const float *pos, const uint32_t *idx, uint32_t nt);
    // robust = false uses Nori's exact slab comparisons.
    Hit trace(const Ray &ray, bool any_hit, uint64_t *nodes_visited = nullptr, uint64_t *tris_tested = nullptr,
              bool robust = true) const;
};

// Shadow ray of the renderer (shaders/RayQueryTrace.hlsli:96-124): 0 when a
// triangle whose mesh does not pass (ior_of(prim).x <= 0) lies in [mint, maxt];
// otherwise the product over the pass-through triangles in the interval of
// 1 - fresnel(|d . Ng|, ior.x, ior.y).  Every triangle in the interval is
// visited, so each crossing counts once.
template <class IorOf, class Fresnel>
float shadow_pass(const Bvh2 &b, const Ray &ray, IorOf ior_of, Fresnel fresnel) {
    if (b.nodes.empty() || b.index.empty() || ray.maxt < ray.mint) return 1.0f;
    float tr = 1.0f;
    uint32_t stack[64], sp = 0, ni = 0;
    while (true) {
        const Bvh2::Node &n = b.nodes[ni];
        if (!n.box.hit(ray, true)) {
            if (sp == 0) break;
            ni = stack[--sp];
            continue;
        }
        if (!n.leaf) {
            stack[sp++] = n.right_or_start;
            ++ni;
            continue;
        }
        for (uint32_t i = n.right_or_start, e = n.right_or_start + n.size; i < e; ++i) {
            const uint32_t f = b.index[i];
            float u, v, t;
            if (!intersect_tri(b.p0[f], b.e1[f], b.e2[f], ray, u, v, t)) continue;
            const auto ior = ior_of(f);
            if (!(ior.first > 0.0f)) return 0.0f;
            const V3f ng = normalize(cross(b.e1[f], b.e2[f]));
            tr *= 1.0f - fresnel(std::fabs(dot(ray.d, ng)), ior.first, ior.second);
        }
        if (sp == 0) break;
        ni = stack[--sp];
    }
    return tr;
}

// Closest hit so far: smaller t, equal t resolved to the smaller primitive id.
inline void take_hit(Hit &best, float t, float u, float v, uint32_t prim) {
    if (t < best.t || (t == best.t && prim < best.prim)) {
        best.t = t; best.u = u; best.v = v; best.prim = prim;
    }
}

// Closest or any hit where accept(prim, u, v) decides whether a candidate
// that passes the triangle test counts (the reference's any-hit programs:
// alpha test).  Ties in t go to the smaller primitive id, as in trace().
template <class Accept> Hit trace_accept(const Bvh2 &b, const Ray &ray_in, bool any_hit, Accept accept) {
    Hit best{INFINITY, 0, 0, 0xFFFFFFFFu};
    if (b.nodes.empty() || b.index.empty() || ray_in.maxt < ray_in.mint) return best;
    Ray ray = ray_in;
    uint32_t stack[64], sp = 0, ni = 0;
    while (true) {
        const Bvh2::Node &n = b.nodes[ni];
        if (!n.box.hit(ray, true)) {
            if (sp == 0) break;
            ni = stack[--sp];
            continue;
        }
        if (!n.leaf) {
            stack[sp++] = n.right_or_start;
            ++ni;
            continue;
        }
        for (uint32_t i = n.right_or_start, e = n.right_or_start + n.size; i < e; ++i) {
            const uint32_t f = b.index[i];
            float u, v, t;
            if (!intersect_tri(b.p0[f], b.e1[f], b.e2[f], ray, u, v, t) || !accept(f, u, v)) continue;
            take_hit(best, t, u, v, f);
            if (any_hit) return best;
            ray.maxt = best.t;
        }
        if (sp == 0) break;
        ni = stack[--sp];
    }
    return best;
}

// Shadow ray through an any-hit filter: factor(prim, u, v, |d . Ng|) < 0 means
// the candidate occludes (result 0); otherwise the transmission is multiplied
// by it (1: the candidate is ignored).  Every triangle in [mint, maxt] is
// visited once.
template <class Factor> float shadow_filter(const Bvh2 &b, const Ray &ray, Factor factor) {
    if (b.nodes.empty() || b.index.empty() || ray.maxt < ray.mint) return 1.0f;
    float tr = 1.0f;
    uint32_t stack[64], sp = 0, ni = 0;
    while (true) {
        const Bvh2::Node &n = b.nodes[ni];
        if (!n.box.hit(ray, true)) {
            if (sp == 0) break;
            ni = stack[--sp];
            continue;
        }
        if (!n.leaf) {
            stack[sp++] = n.right_or_start;
            ++ni;
            continue;
        }
        for (uint32_t i = n.right_or_start, e = n.right_or_start + n.size; i < e; ++i) {
            const uint32_t f = b.index[i];
            float u, v, t;
            if (!intersect_tri(b.p0[f], b.e1[f], b.e2[f], ray, u, v, t)) continue;
            const V3f ng = normalize(cross(b.e1[f], b.e2[f]));
            const float k = factor(f, u, v, std::fabs(dot(ray.d, ng)));
            if (k < 0.0f) return 0.0f;
            tr *= k;
        }
        if (sp == 0) break;
        ni = stack[--sp];
    }
    return tr;
}

// Exhaustive test over every triangle (ground truth for small scenes).
Hit brute_trace(const Bvh2 &tris, const Ray &ray, bool any_hit);

inline Ray make_ray(const float *r8) {
    Ray r;
    r.o = V3f(r8[0], r8[1], r8[2]);
    r.mint = r8[3];
    r.d = V3f(r8[4], r8[5], r8[6]);
    r.maxt = r8[7];
    r.drcp = V3f(1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z);
    return r;
}

} // namespace orc
