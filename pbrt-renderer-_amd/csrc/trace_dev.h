// Device-side traversal of the compressed BVH8 (bvh8.h), one ray per lane of a
// wave64.  Replaces the DXR hardware traversal behind TraceClosestInline /
// TraceShadowInline (shaders/RayQueryTrace.hlsli:36-129) and serves Nori's
// Accel::rayIntersect semantics (src/cpu/accel.cpp:429-488):
//
//   * triangle test: Möller–Trumbore of src/cpu/mesh.cpp:51-89 on the stored
//     (p0, e1, e2), evaluated without contraction so the result equals the CPU
//     evaluation bit for bit;
//   * closest hit: smallest t in [tmin, tmax], equal t resolved to the smaller
//     global primitive id (independent of traversal order);
//   * box test: the distance to each decoded plane fl(q sc + p) as one fma
//     q (sc rcp) + (p - o) rcp, moved outward by a slack bounding its rounding
//     errors (visit_node; rays with a zero direction component use the
//     two-step decode), and the current hit distance widened by
//     (1 + 2*gamma_3) (Pharr, Jakob, Humphreys, PBRT 3e §3.9.2), so no box that
//     holds an accepted hit, or an equal-t hit with a smaller id, is ever culled;
//   * stack: the node groups (child base, hit bits | inner mask) of Ylitie et
//     al. 2017, kept in LDS, one column per lane; descending into one child
//     pushes the rest of its group, so the entries belong to strictly
//     increasing tree levels that have inner children: a tree of D levels
//     needs D - 1 entries;
//   * schedule (POSTPONE, closest-hit queries): either a lane tests all
//     triangles of a node's hit leaves right after visiting it, or, as Ylitie
//     et al. 2017 (section 4) postpone them, a wave-wide ballot counts the
//     lanes that hold triangles: while fewer than a fifth of the active lanes
//     do and a lane still has child nodes to visit, its triangle group goes
//     onto its stack and it keeps descending, so lanes testing triangles do
//     not hold up the lanes walking nodes; otherwise it tests one triangle per
//     step.  A lane keeps at most one postponed group on its stack (sharing it
//     with node groups: a node group always has child hit bits in its top
//     byte, a triangle group never does), so the stack needs one entry more.
//     Occlusion queries do not postpone: an early
//     occluder ends them.
#pragma once
#include "nori_math.h"

namespace ng {

struct TraceOut {
    float t, u, v;
    uint32_t prim; // 0xFFFFFFFF: no hit
};

__device__ __forceinline__ float dot_exact(f3 a, f3 b) {
#pragma clang fp contract(off)
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ f3 cross_exact(f3 a, f3 b) {
#pragma clang fp contract(off)
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ f3 sub_exact(f3 a, f3 b) {
#pragma clang fp contract(off)
    return f3{a.x - b.x, a.y - b.y, a.z - b.z};
}

// Möller–Trumbore, src/cpu/mesh.cpp:51-89.
__device__ __forceinline__ bool mt_test(f3 o, f3 d, f3 p0, f3 e1, f3 e2, float tmin, float tmax, float &t, float &u,
                                        float &v) {
#pragma clang fp contract(off)
    f3 pvec = cross_exact(d, e2);
    float det = dot_exact(e1, pvec);
    if (det > -1e-8f && det < 1e-8f) return false;
    float inv_det = 1.0f / det;
    f3 tvec = sub_exact(o, p0);
    u = dot_exact(tvec, pvec) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    f3 qvec = cross_exact(tvec, e1);
    v = dot_exact(d, qvec) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot_exact(e2, qvec) * inv_det;
    return t >= tmin && t <= tmax;
}

__device__ __forceinline__ float byte_f(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

// Candidate filter that accepts every triangle (opaque geometry).
struct AcceptAll {
    __device__ __forceinline__ bool operator()(uint32_t, uint32_t, uint32_t, float, float, f3, f3) const {
        return true;
    }
};

// One ray's traversal as a resumable state machine: start() sets it up,
// each step() visits one node (its child boxes, then the triangles among its
// hit children) or pops the stack, and returns true once the ray is finished
// (res final).  ANY = occlusion query (finishes on the first accepted hit).
// stk: this lane's LDS column (entry i at stk[i * stride]).
// accept(prim, mesh, flags, u, v, e1, e2) decides whether a candidate that passes the
// triangle test counts as a hit (the any-hit programs of shaders/Shaders.hlsl:
// 674-735 and shaders/RayQueryTrace.hlsli:43-124: alpha test, glass and
// mirrors on shadow rays); a rejected candidate leaves the search as it was.
// Every triangle is tested at most once per ray.
template <bool ANY, bool COUNT, bool POSTPONE = false> struct Bvh8Walker {
    f3 o, d, rcp;
    float tmin;
    uint32_t octinv;
    bool nx, ny, nz;
    bool fused; // every |rcp| below 1e28: the one-fma box test with its slack (else the two-step decode)
    TraceOut res;
    uint2 group;
    uint2 tg; // POSTPONE: pending triangles (first triangle, hit mask in bits 0..23); y = 0: none
    int sp;
    bool postponed; // POSTPONE: a triangle group of this ray is on the stack
    uint32_t cnt_nodes, cnt_tris; // summed over every ray this walker ran

    __device__ __forceinline__ void start(f3 o_, f3 d_, float tmin_, float tmax) {
        o = o_;
        d = d_;
        tmin = tmin_;
        res = TraceOut{tmax, 0.0f, 0.0f, 0xFFFFFFFFu};
        rcp = f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        nx = rcp.x < 0.0f;
        ny = rcp.y < 0.0f;
        nz = rcp.z < 0.0f;
        // a zero (or denormal-small) direction component makes its slab test a
        // containment test of the origin, which the fused form cannot express
        fused = fabsf(rcp.x) < 1e28f && fabsf(rcp.y) < 1e28f && fabsf(rcp.z) < 1e28f;
        octinv = 7u - ((nx ? 4u : 0u) | (ny ? 2u : 0u) | (nz ? 1u : 0u));
        group = make_uint2(0u, (1u << (24u + octinv)) | 1u);
        tg = make_uint2(0u, 0u);
        sp = 0;
        postponed = false;
    }

    // Hit bits of a node's children (bits 0..23 leaf triangles, 24..31 inner
    // children by octant-sorted slot).  w: the near/far plane words per axis
    // for this ray's octant (x lo, x hi, far x lo, far x hi, then y, z).
    //
    // FUSED: slab distances as one fma per plane, t = q A + B with A = sc rcp
    // (exact: sc is a power of two) and B = (p - o) rcp, moved outward by a
    // per-axis slack E that bounds every rounding between this form and the
    // exact distance to the decoded plane fl(q sc + p): the plane's own
    // rounding (u |p + q sc| |rcp|), those of p - o, of the product and of rcp
    // (3u |B| and u q |A|), the final fma and the shifted constants
    // (u (255 |A| + |B| + E) and u (|B| + E)).  E = 16u (|B| + 255 |A| +
    // |p rcp|) covers their sum about three times over, so a near distance is
    // never later and a far one never earlier than exact, and no box that the
    // exact test enters is culled.  Otherwise (a ray with a zero direction
    // component, |rcp| >= 1e28): the two-step decode (X - o) rcp with the far
    // distance widened by (1 + 2 gamma_3), where an infinite rcp keeps the slab
    // a containment test of the origin.
    template <bool FUSED>
    __device__ __forceinline__ uint32_t child_hits(const uint4 &b, const uint32_t *w, f3 p, f3 sc) const {
        const float far_scale = 1.0f + 2.0f * 3.5762791e-07f; // 1 + 2*gamma_3, gamma_3 = 3u/(1-3u), u = 2^-24
        const float u16 = 16.0f * 5.9604645e-08f;
        f3 A, Bn, Bf;
        if (FUSED) {
            A = f3{sc.x * rcp.x, sc.y * rcp.y, sc.z * rcp.z};
            const f3 B = f3{(p.x - o.x) * rcp.x, (p.y - o.y) * rcp.y, (p.z - o.z) * rcp.z};
            const f3 E = f3{u16 * (fabsf(B.x) + __builtin_fmaf(255.0f, fabsf(A.x), fabsf(p.x * rcp.x))),
                            u16 * (fabsf(B.y) + __builtin_fmaf(255.0f, fabsf(A.y), fabsf(p.y * rcp.y))),
                            u16 * (fabsf(B.z) + __builtin_fmaf(255.0f, fabsf(A.z), fabsf(p.z * rcp.z)))};
            Bn = f3{B.x - E.x, B.y - E.y, B.z - E.z};
            Bf = f3{B.x + E.x, B.y + E.y, B.z + E.z};
        }
        // widened: a box entered at exactly res.t may hold an equal-t hit with a
        // smaller primitive id
        const float tcap = res.t * far_scale;
        uint32_t hitmask = 0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            uint32_t meta = ((s < 4 ? b.z : b.w) >> (8 * (s & 3))) & 0xFFu;
            if (meta == 0u) continue;
            const int k = s & 3, h = s < 4 ? 0 : 1;
            float qnx = byte_f(w[0 + h], k), qfx = byte_f(w[2 + h], k);
            float qny = byte_f(w[4 + h], k), qfy = byte_f(w[6 + h], k);
            float qnz = byte_f(w[8 + h], k), qfz = byte_f(w[10 + h], k);
            float tn, tf;
            if (FUSED) {
                float tnx = __builtin_fmaf(qnx, A.x, Bn.x), tfx = __builtin_fmaf(qfx, A.x, Bf.x);
                float tny = __builtin_fmaf(qny, A.y, Bn.y), tfy = __builtin_fmaf(qfy, A.y, Bf.y);
                float tnz = __builtin_fmaf(qnz, A.z, Bn.z), tfz = __builtin_fmaf(qfz, A.z, Bf.z);
                tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
                tf = fminf(fminf(fminf(tfx, tfy), tfz), tcap);
            } else {
                float tnx = (__builtin_fmaf(qnx, sc.x, p.x) - o.x) * rcp.x;
                float tfx = (__builtin_fmaf(qfx, sc.x, p.x) - o.x) * rcp.x;
                float tny = (__builtin_fmaf(qny, sc.y, p.y) - o.y) * rcp.y;
                float tfy = (__builtin_fmaf(qfy, sc.y, p.y) - o.y) * rcp.y;
                float tnz = (__builtin_fmaf(qnz, sc.z, p.z) - o.z) * rcp.z;
                float tfz = (__builtin_fmaf(qfz, sc.z, p.z) - o.z) * rcp.z;
                tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
                tf = fminf(fminf(fminf(tfx, tfy), tfz), res.t) * far_scale;
            }
            if (tn <= tf) {
                uint32_t is_inner = (meta & 0x18u) == 0x18u;
                uint32_t bitidx = is_inner ? ((meta ^ octinv) & 0x1Fu) : (meta & 0x1Fu);
                hitmask |= (meta >> 5) << bitidx;
            }
        }
        return hitmask;
    }

    // Visits the next hit child node of `group` (which has one): the rest of
    // the group goes onto the stack, `group` becomes the node's hit inner
    // children and the return value its hit leaf triangles (bits 0..23, from
    // triangle *tri_base).
    __device__ __forceinline__ uint32_t visit_node(const uint4 *__restrict__ nodes, uint2 *stk, uint32_t stride,
                                                   uint32_t &tri_base) {
        uint32_t bit = 31u - __builtin_clz(group.y & 0xFF000000u);
        group.y &= ~(1u << bit);
        uint32_t slot = (bit - 24u) ^ octinv;
        uint32_t node_idx = group.x + __builtin_popcount((group.y & 0xFFu) & ((1u << slot) - 1u));
        if (group.y & 0xFF000000u) {
            stk[sp * stride] = group;
            ++sp;
        }
        const uint4 *np = nodes + 5 * (size_t)node_idx;
        uint4 a = np[0], b = np[1], c = np[2], dd = np[3], e = np[4];
        if (COUNT) ++cnt_nodes;
        f3 p = f3{__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z)};
        f3 sc = f3{__uint_as_float(((a.w) & 0xFFu) << 23), __uint_as_float(((a.w >> 8) & 0xFFu) << 23),
                   __uint_as_float(((a.w >> 16) & 0xFFu) << 23)};
        uint32_t imask = a.w >> 24;
        // near/far plane words per axis, chosen by the ray octant
        uint32_t nxl = nx ? dd.z : c.x, nxh = nx ? dd.w : c.y, fxl = nx ? c.x : dd.z, fxh = nx ? c.y : dd.w;
        uint32_t nyl = ny ? e.x : c.z, nyh = ny ? e.y : c.w, fyl = ny ? c.z : e.x, fyh = ny ? c.w : e.y;
        uint32_t nzl = nz ? e.z : dd.x, nzh = nz ? e.w : dd.y, fzl = nz ? dd.x : e.z, fzh = nz ? dd.y : e.w;
        const uint32_t w[12] = {nxl, nxh, fxl, fxh, nyl, nyh, fyl, fyh, nzl, nzh, fzl, fzh};
        const uint32_t hitmask = fused ? child_hits<true>(b, w, p, sc) : child_hits<false>(b, w, p, sc);
        group = make_uint2(b.x, (hitmask & 0xFF000000u) | imask);
        tri_base = b.y;
        return hitmask & 0x00FFFFFFu;
    }

    // Tests triangle tri_base + tb; true when an occlusion query is answered.
    template <class Accept>
    __device__ __forceinline__ bool test_triangle(const float4 *__restrict__ tris, uint32_t idx, Accept &accept) {
        const float4 *tp = tris + 3 * (size_t)idx;
        float4 t0 = tp[0], t1 = tp[1], t2 = tp[2];
        if (COUNT) ++cnt_tris;
        float t, u, v;
        if (mt_test(o, d, f3{t0.x, t0.y, t0.z}, f3{t1.x, t1.y, t1.z}, f3{t2.x, t2.y, t2.z}, tmin, res.t, t, u, v)) {
            uint32_t prim = __float_as_uint(t0.w);
            if (!accept(prim, __float_as_uint(t1.w), __float_as_uint(t2.w), u, v, f3{t1.x, t1.y, t1.z},
                        f3{t2.x, t2.y, t2.z}))
                return false;
            if (t < res.t || prim < res.prim) {
                res.t = t;
                res.u = u;
                res.v = v;
                res.prim = prim;
                if (ANY) return true;
            }
        }
        return false;
    }

    template <class Accept>
    __device__ __forceinline__ bool step(const uint4 *__restrict__ nodes, const float4 *__restrict__ tris, uint2 *stk,
                                         uint32_t stride, Accept &accept) {
        if (!POSTPONE || ANY) {
            if (group.y & 0xFF000000u) {
                uint32_t base;
                uint32_t trimask = visit_node(nodes, stk, stride, base);
                while (trimask) {
                    uint32_t tb = __builtin_ctz(trimask);
                    trimask &= trimask - 1u;
                    if (test_triangle(tris, base + tb, accept)) return true;
                }
            }
            if (!(group.y & 0xFF000000u)) {
                if (sp == 0) return true;
                --sp;
                group = stk[sp * stride];
            }
            return false;
        }
        // postponing schedule: a node visit or a pop, then at most one triangle
        if (tg.y == 0u) {
            if (group.y & 0xFF000000u) {
                uint32_t base;
                const uint32_t m = visit_node(nodes, stk, stride, base);
                tg = make_uint2(base, m);
            } else {
                if (sp == 0) return true;
                --sp;
                const uint2 e = stk[sp * stride];
                if (e.y & 0xFF000000u) {
                    group = e;
                } else {
                    tg = e;
                    postponed = false;
                }
            }
        }
        const bool want = tg.y != 0u;
        const uint64_t wants = __ballot(want), active = __ballot(1);
        if (want) {
            if (!postponed && (group.y & 0xFF000000u) && 5 * __popcll(wants) < __popcll(active)) {
                stk[sp * stride] = tg; // postponed: few lanes of the wave would test triangles now
                ++sp;
                tg.y = 0u;
                postponed = true;
            } else {
                const uint32_t tb = __builtin_ctz(tg.y);
                tg.y &= tg.y - 1u;
                if (test_triangle(tris, tg.x + tb, accept)) return true;
            }
        }
        return false;
    }
};

// Whole traversal of one ray (counters, optional: nodes and triangles visited).
template <bool ANY, bool COUNT, bool POSTPONE = false, class Accept = AcceptAll>
__device__ __forceinline__ TraceOut traverse_bvh8(const uint4 *__restrict__ nodes, const float4 *__restrict__ tris, f3 o,
                                                  f3 d, float tmin, float tmax, uint2 *stk, uint32_t stride,
                                                  uint32_t *nnodes = nullptr, uint32_t *ntris = nullptr,
                                                  Accept accept = Accept()) {
    Bvh8Walker<ANY, COUNT, POSTPONE> w;
    w.cnt_nodes = w.cnt_tris = 0;
    w.start(o, d, tmin, tmax);
    while (!w.step(nodes, tris, stk, stride, accept)) {
    }
    if (COUNT) {
        *nnodes = w.cnt_nodes;
        *ntris = w.cnt_tris;
    }
    return w.res;
}

} // namespace ng
