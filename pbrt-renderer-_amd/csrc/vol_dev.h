// Participating media on the device (shaders/Volume.hlsl): null-collision free
// flight with a hero channel over per-brick majorants, ratio-tracking
// transmittance, Henyey-Greenstein phase function, and the NEE skip of
// shaders/Emitter.hlsli:336-345.
//
// Layout: one VolumeDev per medium (at most NORI_MAX_VOLUMES); the density grid
// is W x H x D floats (x fastest) looked up trilinearly with clamping (texel
// centres at (i + 0.5) / W); the majorant grid holds, per brick, the maximum
// density over every voxel a lookup inside the brick can touch (the brick's
// voxels plus a one-voxel apron), so density / majorant <= 1.
//
// Floating-point contraction is off in these functions, as in the CPU
// restatement, so both evaluate the same roundings; positions along a ray are
// formed by vol_point, not by the (contracted) f3 operators.
//
// Free flight and ratio tracking walk the majorant bricks as a 3D DDA: the
// brick index is carried across faces (the current brick +-1 along the axis the
// ray leaves by) and each exit distance is measured from the ray origin.  The
// reference (shaders/Volume.hlsl:119-164, 236-372) instead steps exactly onto a
// face and re-derives the brick from the rounded position (with a 1e-5 minimum
// step), so the brick seen next hangs on the position's last bit.
#pragma once
#include "nori_math.h"

namespace ng {

struct VolumeDev {
    f3 vmin, vmax, sigma_a, sigma_s;
    float g;
    uint32_t hetero;
    const float *density; // null: density 1
    uint32_t W, H, D;
    const float *majorant; // null: the global majorant everywhere
    uint32_t mW, mH, mD;
};

NG_HD f3 vol_sigma_t(const VolumeDev &v) { return v.sigma_a + v.sigma_s; }
NG_HD float vol_sigma_t_max(const VolumeDev &v) { return maxc(vol_sigma_t(v)); }

// o + d t, rounded per component without contraction
NG_HD f3 vol_point(f3 o, f3 d, float t) {
#pragma clang fp contract(off)
    return mk(o.x + d.x * t, o.y + d.y * t, o.z + d.z * t);
}

NG_HD float vol_density(const VolumeDev &v, f3 p) {
#pragma clang fp contract(off)
    if (!v.hetero || !v.density) return 1.0f;
    f3 size = mk(fmaxf(v.vmax.x - v.vmin.x, 1e-6f), fmaxf(v.vmax.y - v.vmin.y, 1e-6f), fmaxf(v.vmax.z - v.vmin.z, 1e-6f));
    float u = (p.x - v.vmin.x) / size.x * (float)v.W - 0.5f;
    float w1 = (p.y - v.vmin.y) / size.y * (float)v.H - 0.5f;
    float w2 = (p.z - v.vmin.z) / size.z * (float)v.D - 0.5f;
    float fu = floorf(u), fv = floorf(w1), fw = floorf(w2);
    float tx = u - fu, ty = w1 - fv, tz = w2 - fw;
    int x0 = (int)fu, y0 = (int)fv, z0 = (int)fw;
    const int W = (int)v.W, H = (int)v.H, D = (int)v.D;
    int xa = min(max(x0, 0), W - 1), xb = min(max(x0 + 1, 0), W - 1);
    int ya = min(max(y0, 0), H - 1), yb = min(max(y0 + 1, 0), H - 1);
    int za = min(max(z0, 0), D - 1), zb = min(max(z0 + 1, 0), D - 1);
    auto at = [&](int x, int y, int z) { return v.density[(size_t)x + (size_t)W * ((size_t)y + (size_t)H * z)]; };
    float c00 = at(xa, ya, za) * (1 - tx) + at(xb, ya, za) * tx;
    float c10 = at(xa, yb, za) * (1 - tx) + at(xb, yb, za) * tx;
    float c01 = at(xa, ya, zb) * (1 - tx) + at(xb, ya, zb) * tx;
    float c11 = at(xa, yb, zb) * (1 - tx) + at(xb, yb, zb) * tx;
    float c0 = c00 * (1 - ty) + c10 * ty, c1 = c01 * (1 - ty) + c11 * ty;
    return c0 * (1 - tz) + c1 * tz;
}

// shaders/Volume.hlsl:119-164
NG_HD void vol_local_majorant(const VolumeDev &v, f3 pos, f3 dir, float &mu, float &bd) {
#pragma clang fp contract(off)
    const float gmu = vol_sigma_t_max(v);
    if (!v.hetero || !v.majorant) {
        mu = gmu;
        bd = 1e20f;
        return;
    }
    f3 size = mk(fmaxf(v.vmax.x - v.vmin.x, 1e-6f), fmaxf(v.vmax.y - v.vmin.y, 1e-6f), fmaxf(v.vmax.z - v.vmin.z, 1e-6f));
    const float mres[3] = {(float)v.mW, (float)v.mH, (float)v.mD};
    const float sz[3] = {size.x, size.y, size.z}, lo[3] = {v.vmin.x, v.vmin.y, v.vmin.z};
    const float p[3] = {pos.x, pos.y, pos.z}, d[3] = {dir.x, dir.y, dir.z};
    int bc[3];
    for (int i = 0; i < 3; ++i) {
        float bf = (p[i] - lo[i]) / sz[i] * mres[i];
        bc[i] = min(max((int)bf, 0), (int)mres[i] - 1);
    }
    mu = v.majorant[(size_t)bc[0] + (size_t)v.mW * ((size_t)bc[1] + (size_t)v.mH * bc[2])] * gmu;
    float best = 1e20f;
    for (int i = 0; i < 3; ++i) {
        float di = 1e20f;
        if (fabsf(d[i]) >= 1e-20f) {
            float edge = (float)bc[i] + (d[i] > 0.0f ? 1.0f : 0.0f);
            float world = lo[i] + edge * (sz[i] / mres[i]);
            di = fmaxf((world - p[i]) / d[i], 1e-5f);
        }
        best = fminf(best, di);
    }
    bd = best;
}

// A ray's walk through the majorant bricks (3D DDA); without a grid the whole
// box is one brick with the global majorant.
struct BrickWalk {
    int bc[3];
    bool grid;

    NG_HD void start(const VolumeDev &v, f3 o, f3 d, float t) {
#pragma clang fp contract(off)
        grid = v.hetero && v.majorant;
        bc[0] = bc[1] = bc[2] = 0;
        if (!grid) return;
        const f3 p = vol_point(o, d, t);
        const float pp[3] = {p.x, p.y, p.z}, lo[3] = {v.vmin.x, v.vmin.y, v.vmin.z};
        const float sz[3] = {fmaxf(v.vmax.x - v.vmin.x, 1e-6f), fmaxf(v.vmax.y - v.vmin.y, 1e-6f),
                             fmaxf(v.vmax.z - v.vmin.z, 1e-6f)};
        const int m[3] = {(int)v.mW, (int)v.mH, (int)v.mD};
        for (int i = 0; i < 3; ++i) bc[i] = min(max((int)((pp[i] - lo[i]) / sz[i] * (float)m[i]), 0), m[i] - 1);
    }
    NG_HD float majorant(const VolumeDev &v) const {
        const float gmu = vol_sigma_t_max(v);
        return grid ? v.majorant[(size_t)bc[0] + (size_t)v.mW * ((size_t)bc[1] + (size_t)v.mH * bc[2])] * gmu : gmu;
    }
    // ray parameter where the ray leaves the current brick, and the axis it leaves by
    NG_HD float exit(const VolumeDev &v, f3 o, f3 d, int &axis) const {
#pragma clang fp contract(off)
        axis = -1;
        if (!grid) return 1e20f;
        const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, lo[3] = {v.vmin.x, v.vmin.y, v.vmin.z};
        const float sz[3] = {fmaxf(v.vmax.x - v.vmin.x, 1e-6f), fmaxf(v.vmax.y - v.vmin.y, 1e-6f),
                             fmaxf(v.vmax.z - v.vmin.z, 1e-6f)};
        const float m[3] = {(float)v.mW, (float)v.mH, (float)v.mD};
        float best = 1e20f;
        for (int i = 0; i < 3; ++i) {
            if (fabsf(dd[i]) < 1e-20f) continue;
            const float edge = lo[i] + (float)(bc[i] + (dd[i] > 0.0f ? 1 : 0)) * (sz[i] / m[i]);
            const float ti = (edge - oo[i]) / dd[i];
            if (ti < best) { best = ti; axis = i; }
        }
        return best;
    }
    // into the neighbouring brick across `axis`; false once the walk leaves the
    // grid (static indices only, so the walker stays in registers)
    NG_HD bool advance(const VolumeDev &v, f3 d, int axis) {
        if (axis == 0) {
            bc[0] += d.x > 0.0f ? 1 : -1;
            return bc[0] >= 0 && bc[0] < (int)v.mW;
        }
        if (axis == 1) {
            bc[1] += d.y > 0.0f ? 1 : -1;
            return bc[1] >= 0 && bc[1] < (int)v.mH;
        }
        if (axis == 2) {
            bc[2] += d.z > 0.0f ? 1 : -1;
            return bc[2] >= 0 && bc[2] < (int)v.mD;
        }
        return false;
    }
};

// shaders/Volume.hlsl:14-27
NG_HD bool vol_ray_aabb(f3 o, f3 d, f3 bmin, f3 bmax, float &tn, float &tf) {
#pragma clang fp contract(off)
    f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    f3 t0 = mk((bmin.x - o.x) * inv.x, (bmin.y - o.y) * inv.y, (bmin.z - o.z) * inv.z);
    f3 t1 = mk((bmax.x - o.x) * inv.x, (bmax.y - o.y) * inv.y, (bmax.z - o.z) * inv.z);
    tn = fmaxf(fmaxf(fminf(t0.x, t1.x), fminf(t0.y, t1.y)), fminf(t0.z, t1.z));
    tf = fminf(fminf(fmaxf(t0.x, t1.x), fmaxf(t0.y, t1.y)), fmaxf(t0.z, t1.z));
    return tf >= fmaxf(tn, 0.0f);
}

// shaders/Volume.hlsl:31-65
NG_HD float hg_eval(float ct, float g) {
#pragma clang fp contract(off)
    float g2 = g * g, den = 1.0f + g2 - 2.0f * g * ct;
    return (1.0f / (4.0f * PI)) * (1.0f - g2) / (den * sqrtf(fmaxf(den, 1e-8f)));
}
NG_HD f3 hg_sample(f3 wo, float g, float xi1, float xi2, float &pdf) {
#pragma clang fp contract(off)
    float ct;
    if (fabsf(g) < 1e-3f) {
        ct = 1.0f - 2.0f * xi1;
    } else {
        float sqr = (1.0f - g * g) / (1.0f + g - 2.0f * g * xi1);
        ct = fminf(fmaxf((1.0f + g * g - sqr * sqr) / (2.0f * g), -1.0f), 1.0f);
    }
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct)), phi = 2.0f * PI * xi2;
    f3 T, B;
    build_onb(wo, T, B);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    f3 wi = T * (st * cp) + B * (st * sp) + wo * ct;
    pdf = hg_eval(ct, g);
    return normalize(wi);
}

struct VolSeg {
    uint32_t vi;
    float tn, tf;
};

// The segment [tn, tf] of ray (o, d) inside one medium's box, clipped to
// [0, tmax] (one step of vol_segments); false when empty.
NG_HD bool vol_segment(const VolumeDev &v, f3 o, f3 d, float tmax, float &tn, float &tf) {
    if (!vol_ray_aabb(o, d, v.vmin, v.vmax, tn, tf)) return false;
    tn = fmaxf(tn, 0.0f);
    tf = fminf(tf, tmax);
    return tf > tn;
}

// shaders/Volume.hlsl:187-222
NG_HD uint32_t vol_segments(const VolumeDev *vols, uint32_t nv, f3 o, f3 d, float tmax, VolSeg seg[8]) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < nv && n < 8; ++i) {
        float tn, tf;
        if (!vol_segment(vols[i], o, d, tmax, tn, tf)) continue;
        uint32_t k = n;
        for (; k > 0; --k) {
            if (seg[k - 1].tn <= tn) break;
            seg[k] = seg[k - 1];
        }
        seg[k] = VolSeg{i, tn, tf};
        ++n;
    }
    return n;
}

// shaders/Volume.hlsl:236-319, the bricks walked as a DDA (BrickWalk).  One
// uniform per tentative collision, a second for the real/null decision.
template <class R>
NG_HD bool vol_free_flight(const VolumeDev &v, f3 o, f3 d, float tn, float tf, R &rng, f3 &weight, float &ts) {
#pragma clang fp contract(off)
    const float gmu = vol_sigma_t_max(v);
    const f3 st = vol_sigma_t(v);
    float t = tn;
    BrickWalk w;
    w.start(v, o, d, t);
    for (uint32_t step = 0; step < 512; ++step) {
        const float mu = w.majorant(v);
        int axis;
        const float texit = w.exit(v, o, d, axis);
        if (mu < 1e-8f) {
            t = fmaxf(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) { ts = tf; return false; }
            continue;
        }
        float xi = fmaxf(rng_float(rng), 1e-10f);
        float dt = -logf(xi) / mu;
        if (t + dt > texit) {
            t = fmaxf(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) { ts = tf; return false; }
            continue;
        }
        t += dt;
        if (t >= tf) { ts = tf; return false; }
        float dens = vol_density(v, vol_point(o, d, t));
        float p_real = fminf(fmaxf((dens * gmu) / fmaxf(mu, 1e-20f), 0.0f), 1.0f);
        if (p_real >= 1.0f - 1e-6f || rng_float(rng) < p_real) {
            float inv = 1.0f / fmaxf(gmu, 1e-20f);
            weight = mk(weight.x * (st.x * inv), weight.y * (st.y * inv), weight.z * (st.z * inv));
            ts = t;
            return true;
        }
        float den = fmaxf(mu * (1.0f - p_real), 1e-20f);
        weight = mk(weight.x * ((mu - dens * st.x) / den), weight.y * ((mu - dens * st.y) / den),
                    weight.z * ((mu - dens * st.z) / den));
    }
    ts = tf;
    return false;
}

// shaders/Volume.hlsl:326-372, the bricks walked as a DDA (BrickWalk).
template <class R>
NG_HD f3 vol_ratio_tracking(const VolumeDev &v, f3 o, f3 d, float tn, float tf, R &rng) {
#pragma clang fp contract(off)
    const f3 st = vol_sigma_t(v);
    f3 tr = mk(1.0f);
    float t = tn;
    BrickWalk w;
    w.start(v, o, d, t);
    for (uint32_t step = 0; step < 512; ++step) {
        const float mu = w.majorant(v);
        int axis;
        const float texit = w.exit(v, o, d, axis);
        if (mu < 1e-8f) {
            t = fmaxf(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) return tr;
            continue;
        }
        float xi = fmaxf(rng_float(rng), 1e-10f);
        float dt = -logf(xi) / mu;
        if (t + dt > texit) {
            t = fmaxf(t, texit);
            if (t >= tf || !w.advance(v, d, axis)) return tr;
            continue;
        }
        t += dt;
        if (t >= tf) return tr;
        float dens = vol_density(v, vol_point(o, d, t));
        float inv = 1.0f / fmaxf(mu, 1e-20f);
        tr = mk(tr.x * ((mu - dens * st.x) * inv), tr.y * ((mu - dens * st.y) * inv), tr.z * ((mu - dens * st.z) * inv));
        if (maxc(tr) < 1e-6f) return mk(0.0f);
    }
    return tr;
}

// shaders/Volume.hlsl:381-415
template <class R>
NG_HD bool vol_sample_free_flight(const VolumeDev *vols, uint32_t nv, f3 o, f3 d, float tmax, R &rng,
                                  f3 &weight, float &ts, uint32_t &vi) {
    ts = tmax;
    vi = 0;
    if (nv == 0) return false;
    if (nv == 1) {
        // one medium: its segment directly (the segment list would be a
        // dynamically indexed array, kept in scratch memory)
        float tn, tf, th;
        if (!vol_segment(vols[0], o, d, tmax, tn, tf) || vol_sigma_t_max(vols[0]) < 1e-8f) return false;
        if (!vol_free_flight(vols[0], o, d, tn, tf, rng, weight, th)) return false;
        ts = th;
        return true;
    }
    VolSeg seg[8];
    uint32_t n = vol_segments(vols, nv, o, d, tmax, seg);
    for (uint32_t k = 0; k < n; ++k) {
        const VolumeDev &v = vols[seg[k].vi];
        if (vol_sigma_t_max(v) < 1e-8f) continue;
        float th;
        if (vol_free_flight(v, o, d, seg[k].tn, seg[k].tf, rng, weight, th)) {
            ts = th;
            vi = seg[k].vi;
            return true;
        }
    }
    return false;
}

// shaders/Volume.hlsl:419-454
template <class R>
NG_HD f3 vol_transmittance(const VolumeDev *vols, uint32_t nv, f3 o, f3 d, float tmax, R &rng) {
#pragma clang fp contract(off)
    if (nv == 0) return mk(1.0f);
    if (nv == 1) {
        // one medium: its segment directly, as in vol_sample_free_flight
        float tn, tf;
        const VolumeDev &v = vols[0];
        const f3 st = vol_sigma_t(v);
        if (!vol_segment(v, o, d, tmax, tn, tf) || maxc(st) < 1e-8f) return mk(1.0f);
        const f3 tr = v.hetero ? vol_ratio_tracking(v, o, d, tn, tf, rng)
                               : mk(expf(-st.x * (tf - tn)), expf(-st.y * (tf - tn)), expf(-st.z * (tf - tn)));
        return maxc(tr) < 1e-6f ? mk(0.0f) : tr;
    }
    VolSeg seg[8];
    uint32_t n = vol_segments(vols, nv, o, d, tmax, seg);
    f3 tr = mk(1.0f);
    for (uint32_t k = 0; k < n; ++k) {
        const VolumeDev &v = vols[seg[k].vi];
        f3 st = vol_sigma_t(v);
        if (maxc(st) < 1e-8f) continue;
        f3 s;
        if (v.hetero) {
            s = vol_ratio_tracking(v, o, d, seg[k].tn, seg[k].tf, rng);
        } else {
            float len = seg[k].tf - seg[k].tn;
            s = mk(expf(-st.x * len), expf(-st.y * len), expf(-st.z * len));
        }
        tr = tr * s;
        if (maxc(tr) < 1e-6f) return mk(0.0f);
    }
    return tr;
}

// shaders/Emitter.hlsli:336-345
NG_HD bool vol_nee_skip(const VolumeDev &v, f3 p) {
#pragma clang fp contract(off)
    f3 st = vol_sigma_t(v);
    float stmin = fminf(st.x, fminf(st.y, st.z));
    float md = fminf(fminf(fminf(p.x - v.vmin.x, p.y - v.vmin.y), fminf(p.z - v.vmin.z, v.vmax.x - p.x)),
                     fminf(v.vmax.y - p.y, v.vmax.z - p.z));
    return md * stmin > 8.0f;
}

} // namespace ng
