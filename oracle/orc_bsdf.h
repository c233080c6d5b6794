// CPU oracle — test infrastructure only.
//
// BSDFs of the reference's GPU path restated in templated C++ (T = float for the
// FP32 render oracle, T = double for the FP64 BSDF oracle).  Local frame, z up;
// `wi` is the direction towards the viewer and `wo` the direction towards the
// light, as in shaders/Material.hlsli:18-194.
//
//   diffuse     shaders/Material.hlsli:20-25, 64-69, 109-121
//   microfacet  shaders/Microfacet.hlsli:13-125 (Beckmann, rational Smith G1)
//   disney      shaders/Disney.hlsli:10-275, shaders/Material.hlsli:31-39, 75-84, 127-183
#pragma once
#include "orc_math.h"

namespace orc {

enum { MAT_DIFFUSE = 0, MAT_MIRROR = 1, MAT_DIELECTRIC = 2, MAT_MICROFACET = 3, MAT_DISNEY = 4, MAT_HAIR = 5, MAT_SSS = 6 };

// The 144-byte material record of the reference (src/gpu/DXRApp.h:72-111 ==
// shaders/Common.hlsli:55-89), field for field.
struct MatRec {
    uint32_t type;
    float albedo[3];
    float intIOR, extIOR, alpha;
    uint32_t isEmitter;
    float radiance[3];
    uint32_t indexOffset, vertexOffset, indexCount, vertexCount;
    float surfaceArea;
    uint32_t emitterCdfOffset;
    uint32_t albedoTex, normalTex, roughnessTex, metallicTex, specularTex, subsurfaceTex, alphaTex;
    float roughness, metallic, specular, specularTint, sheen, sheenTint, subsurface, clearcoat, clearcoatGloss, anisotropic;
    float betaN;
    float emitterSelectionProb;
};
static_assert(sizeof(MatRec) == 144, "material record must be 144 bytes");

template <typename T> struct MatT {
    int type;
    V3<T> albedo;
    T intIOR, extIOR, alpha;
    T roughness, metallic, specular, specularTint, sheen, sheenTint, subsurface, clearcoat, clearcoatGloss;
    T betaN;
    static MatT from(const MatRec &m) {
        MatT r;
        r.type = (int)m.type;
        r.albedo = V3<T>(T(m.albedo[0]), T(m.albedo[1]), T(m.albedo[2]));
        r.intIOR = m.intIOR; r.extIOR = m.extIOR; r.alpha = m.alpha;
        r.roughness = m.roughness; r.metallic = m.metallic; r.specular = m.specular;
        r.specularTint = m.specularTint; r.sheen = m.sheen; r.sheenTint = m.sheenTint;
        r.subsurface = m.subsurface; r.clearcoat = m.clearcoat; r.clearcoatGloss = m.clearcoatGloss;
        r.betaN = m.betaN;
        return r;
    }
};

template <typename T> inline T pow5(T x) { return std::pow(x, T(5)); }

// ---------------------------------------------------------------- microfacet
template <typename T> T beckmann_d(V3<T> wh, T a) {
    if (wh.z <= T(0)) return T(0);
    T c2 = wh.z * wh.z, t2 = (T(1) - c2) / c2, c3 = c2 * wh.z, a2 = a * a;
    return std::exp(-t2 / a2) / (Consts<T>::PI * a2 * c3);
}
template <typename T> T beckmann_d_cos(V3<T> wh, T a) {
    if (wh.z <= T(0)) return T(0);
    T c2 = wh.z * wh.z, t2 = (T(1) - c2) / c2, a2 = a * a;
    return std::exp(-t2 / a2) / (Consts<T>::PI * a2 * c2);
}
template <typename T> V3<T> beckmann_sample(T u0, T u1, T a) {
    T phi = T(2) * Consts<T>::PI * u1;
    T t2 = -a * a * std::log(std::max(T(1e-8), T(1) - u0));
    T c = T(1) / std::sqrt(T(1) + t2);
    T s = std::sqrt(std::max(T(0), T(1) - c * c));
    return {s * std::cos(phi), s * std::sin(phi), c};
}
template <typename T> T smith_g1(V3<T> v, V3<T> wh, T a) {
    T dvw = dot(v, wh), dvn = v.z;
    if (dvw / dvn <= T(0)) return T(0);
    T c2 = v.z * v.z;
    if (c2 < T(1e-10)) return T(0);
    T tant = std::sqrt(std::max(T(0), T(1) - c2)) / std::fabs(v.z);
    if (tant < T(1e-10)) return T(1);
    T b = T(1) / (a * tant);
    if (b >= T(1.6)) return T(1);
    T b2 = b * b;
    return (T(3.535) * b + T(2.181) * b2) / (T(1) + T(2.276) * b + T(2.577) * b2);
}
template <typename T> V3<T> microfacet_eval(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    V3<T> kd = m.albedo;
    T ks = T(1) - maxc(kd);
    V3<T> wh = normalize(wi + wo);
    T D = beckmann_d(wh, m.alpha);
    T F = fresnel_dielectric(dot(wh, wi), m.extIOR, m.intIOR);
    T G = smith_g1(wi, wh, m.alpha) * smith_g1(wo, wh, m.alpha);
    T den = T(4) * wi.z * wo.z;
    return kd * Consts<T>::INV_PI + V3<T>(T(1)) * (ks * D * F * G / std::max(den, T(1e-10)));
}
template <typename T> T microfacet_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return T(0);
    T ks = T(1) - maxc(m.albedo);
    V3<T> wh = normalize(wi + wo);
    T dc = beckmann_d_cos(wh, m.alpha);
    T jh = T(1) / std::max(T(4) * dot(wh, wo), T(1e-10));
    return ks * dc * jh + (T(1) - ks) * wo.z * Consts<T>::INV_PI;
}

// ---------------------------------------------------------------- Disney
template <typename T> V3<T> disney_burley(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    V3<T> wh = normalize(wi + wo);
    T cd = std::fabs(dot(wo, wh));
    T fd90 = T(0.5) + T(2) * m.roughness * cd * cd;
    T fv = T(1) + (fd90 - T(1)) * pow5(T(1) - wi.z);
    T fl = T(1) + (fd90 - T(1)) * pow5(T(1) - wo.z);
    return m.albedo * Consts<T>::INV_PI * fv * fl;
}
template <typename T> V3<T> disney_hk(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    V3<T> wh = normalize(wi + wo);
    T cd = std::fabs(dot(wo, wh));
    T f90 = m.roughness * cd * cd;
    T fv = T(1) + (f90 - T(1)) * pow5(T(1) - wi.z);
    T fl = T(1) + (f90 - T(1)) * pow5(T(1) - wo.z);
    T cs = std::max(wi.z + wo.z, T(1e-4));
    T ss = T(1.25) * (fv * fl * (T(1) / cs - T(0.5)) + T(0.5));
    return m.albedo * Consts<T>::INV_PI * ss;
}
template <typename T> V3<T> disney_diffuse_lobe(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    return lerp3(disney_burley(wi, wo, m), disney_hk(wi, wo, m), m.subsurface);
}
template <typename T> T ggx_d(T nh, T a) {
    if (nh <= T(0)) return T(0);
    T a2 = a * a;
    T den = nh * nh * (a2 - T(1)) + T(1);
    return a2 / std::max(Consts<T>::PI * den * den, T(1e-10));
}
template <typename T> T ggx_g1(T nv, T a) {
    if (nv <= T(0)) return T(0);
    T a2 = a * a;
    T den = nv + std::sqrt(a2 + (T(1) - a2) * nv * nv);
    return T(2) * nv / std::max(den, T(1e-10));
}
template <typename T> V3<T> ggx_sample_wh(T u0, T u1, T a) {
    T a2 = a * a;
    T c2 = (T(1) - u0) / std::max(T(1) + (a2 - T(1)) * u0, T(1e-10));
    T c = std::sqrt(saturate(c2));
    T s = std::sqrt(std::max(T(0), T(1) - c2));
    T phi = T(2) * Consts<T>::PI * u1;
    return {s * std::cos(phi), s * std::sin(phi), c};
}
template <typename T> T ggx_pdf_wo(V3<T> wi, V3<T> wo, T a) {
    V3<T> wh = normalize(wi + wo);
    if (wh.z <= T(0)) return T(0);
    return ggx_d(wh.z, a) * wh.z / std::max(T(4) * std::fabs(dot(wh, wo)), T(1e-10));
}
template <typename T> V3<T> disney_spec_f0(const MatT<T> &m) {
    V3<T> base = m.albedo;
    T lum = luminance(base);
    V3<T> ctint = lum > T(0) ? base / lum : V3<T>(T(1));
    V3<T> cspec0 = T(0.08) * m.specular * lerp3(V3<T>(T(1)), ctint, m.specularTint);
    return lerp3(cspec0, base, m.metallic);
}
template <typename T> V3<T> disney_spec_eval(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    V3<T> wh = normalize(wi + wo);
    if (wh.z <= T(0)) return V3<T>(T(0));
    T a = std::max(m.roughness * m.roughness, T(1e-4));
    T vh = std::fabs(dot(wi, wh));
    T D = ggx_d(wh.z, a);
    V3<T> f0 = disney_spec_f0(m);
    T t = pow5(T(1) - vh);
    V3<T> F = f0 + (V3<T>(T(1)) - f0) * t;
    T G = ggx_g1(wi.z, a) * ggx_g1(wo.z, a);
    return F * (D * G / std::max(T(4) * wi.z * wo.z, T(1e-10)));
}
template <typename T> T disney_spec_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return T(0);
    return ggx_pdf_wo(wi, wo, std::max(m.roughness * m.roughness, T(1e-4)));
}
template <typename T> V3<T> disney_sheen(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (m.sheen <= T(0) || wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    V3<T> wh = normalize(wi + wo);
    T cd = std::fabs(dot(wo, wh));
    V3<T> base = m.albedo;
    T lum = luminance(base);
    V3<T> ctint = lum > T(0) ? base / lum : V3<T>(T(1));
    V3<T> cs = lerp3(V3<T>(T(1)), ctint, m.sheenTint);
    return cs * (m.sheen * pow5(T(1) - cd));
}
template <typename T> T gtr1_d(T nh, T a) {
    if (nh <= T(0)) return T(0);
    T a2 = a * a;
    if (a2 >= T(0.9999)) return Consts<T>::INV_PI;
    T t = T(1) + (a2 - T(1)) * nh * nh;
    return (a2 - T(1)) / (Consts<T>::PI * std::log(a2) * t);
}
template <typename T> V3<T> gtr1_sample_wh(T u0, T u1, T a) {
    T a2 = a * a;
    T c2 = a2 >= T(0.9999) ? T(1) - u0 : (T(1) - std::pow(a2, T(1) - u0)) / (T(1) - a2);
    T c = std::sqrt(saturate(c2));
    T s = std::sqrt(std::max(T(0), T(1) - c2));
    T phi = T(2) * Consts<T>::PI * u1;
    return {s * std::cos(phi), s * std::sin(phi), c};
}
template <typename T> T gtr1_pdf_wo(V3<T> wi, V3<T> wo, T a) {
    V3<T> wh = normalize(wi + wo);
    if (wh.z <= T(0)) return T(0);
    return gtr1_d(wh.z, a) * wh.z / std::max(T(4) * std::fabs(dot(wh, wo)), T(1e-10));
}
template <typename T> T disney_cc_alpha(const MatT<T> &m) { return lerp1(T(0.1), T(0.001), m.clearcoatGloss); }
template <typename T> T disney_cc_eval(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (m.clearcoat <= T(0) || wi.z <= T(0) || wo.z <= T(0)) return T(0);
    V3<T> wh = normalize(wi + wo);
    if (wh.z <= T(0)) return T(0);
    T D = gtr1_d(wh.z, disney_cc_alpha(m));
    T vh = std::fabs(dot(wi, wh));
    T F = T(0.04) + (T(1) - T(0.04)) * pow5(T(1) - vh);
    T G = ggx_g1(wi.z, T(0.25)) * ggx_g1(wo.z, T(0.25));
    return T(0.25) * m.clearcoat * D * F * G / std::max(T(4) * wi.z * wo.z, T(1e-10));
}
template <typename T> T disney_cc_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (m.clearcoat <= T(0) || wi.z <= T(0) || wo.z <= T(0)) return T(0);
    return gtr1_pdf_wo(wi, wo, disney_cc_alpha(m));
}
template <typename T> void disney_lobe_probs(const MatT<T> &m, T &pd, T &ps, T &pc) {
    T wd = T(1) - m.metallic, ws = T(1), wc = T(0.25) * m.clearcoat;
    T tot = wd + ws + wc;
    T inv = tot > T(0) ? T(1) / tot : T(0);
    pd = wd * inv; ps = ws * inv; pc = wc * inv;
}
template <typename T> V3<T> disney_eval(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    V3<T> fd = disney_diffuse_lobe(wi, wo, m);
    V3<T> fs = disney_spec_eval(wi, wo, m);
    V3<T> fsh = disney_sheen(wi, wo, m);
    T fc = disney_cc_eval(wi, wo, m);
    return (fd + fsh) * (T(1) - m.metallic) + fs + V3<T>(fc);
}
template <typename T> T disney_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return T(0);
    T pd, ps, pc;
    disney_lobe_probs(m, pd, ps, pc);
    return pd * (wo.z * Consts<T>::INV_PI) + ps * disney_spec_pdf(wi, wo, m) + pc * disney_cc_pdf(wi, wo, m);
}

// ---------------------------------------------------------------- dispatch
// Boundary reflection lobe of the random-walk subsurface material:
// Beckmann with alpha = max(roughness^2, 1e-4), dielectric Fresnel
// (shaders/Material.hlsli:48-57, 92-100).
template <typename T> V3<T> sss_eval(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
    const T a = std::max(m.roughness * m.roughness, T(1e-4));
    const V3<T> wh = normalize(wi + wo);
    const T Fr = fresnel_dielectric(dot(wh, wi), m.extIOR, m.intIOR);
    const T G = smith_g1(wi, wh, a) * smith_g1(wo, wh, a);
    return V3<T>(beckmann_d(wh, a) * Fr * G / std::max(T(4) * wi.z * wo.z, T(1e-10)));
}
template <typename T> T sss_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    if (wi.z <= T(0) || wo.z <= T(0)) return T(0);
    const T a = std::max(m.roughness * m.roughness, T(1e-4));
    const V3<T> wh = normalize(wi + wo);
    const T Fr = fresnel_dielectric(dot(wh, wi), m.extIOR, m.intIOR);
    return Fr * beckmann_d_cos(wh, a) / std::max(T(4) * dot(wh, wo), T(1e-10));
}

template <typename T> V3<T> material_eval(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    switch (m.type) {
    case MAT_DIFFUSE:
        if (wi.z <= T(0) || wo.z <= T(0)) return V3<T>(T(0));
        return m.albedo * Consts<T>::INV_PI;
    case MAT_MICROFACET: return microfacet_eval(wi, wo, m);
    case MAT_DISNEY: return disney_eval(wi, wo, m);
    case MAT_SSS: return sss_eval(wi, wo, m);
    default: return V3<T>(T(0));
    }
}
template <typename T> T material_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m) {
    switch (m.type) {
    case MAT_DIFFUSE:
        if (wi.z <= T(0) || wo.z <= T(0)) return T(0);
        return wo.z * Consts<T>::INV_PI;
    case MAT_MICROFACET: return microfacet_pdf(wi, wo, m);
    case MAT_DISNEY: return disney_pdf(wi, wo, m);
    case MAT_SSS: return sss_pdf(wi, wo, m);
    default: return T(0);
    }
}
// Number of uniform numbers material_sample consumes for type `t` (diffuse 2,
// microfacet 2, Disney 3: one lobe choice then a 2D sample).
inline int material_sample_dims(int t) { return t == MAT_DISNEY ? 3 : 2; }

// Returns f*cos/pdf; writes wo and pdf.  u[] holds material_sample_dims() numbers.
template <typename T> V3<T> material_sample(V3<T> wi, const T *u, const MatT<T> &m, V3<T> &wo, T &pdf) {
    wo = V3<T>(0, 0, 1);
    pdf = T(0);
    if (m.type == MAT_DIFFUSE) {
        if (wi.z <= T(0)) return V3<T>(T(0));
        wo = cosine_sample_hemisphere(u[0], u[1]);
        pdf = wo.z * Consts<T>::INV_PI;
        return m.albedo;
    }
    if (m.type == MAT_MICROFACET) {
        if (wi.z <= T(0)) return V3<T>(T(0));
        T ks = T(1) - maxc(m.albedo);
        T ux = u[0], uy = u[1];
        if (ux < ks) {
            ux /= ks;
            V3<T> wh = beckmann_sample(ux, uy, m.alpha);
            wo = T(2) * dot(wh, wi) * wh - wi;
        } else {
            ux = (ux - ks) / (T(1) - ks);
            wo = cosine_sample_hemisphere(ux, uy);
        }
        if (wo.z <= T(0)) return V3<T>(T(0));
        pdf = microfacet_pdf(wi, wo, m);
        if (pdf <= T(0)) return V3<T>(T(0));
        return microfacet_eval(wi, wo, m) * wo.z / pdf;
    }
    if (m.type == MAT_DISNEY) {
        if (wi.z <= T(0)) return V3<T>(T(0));
        T pd, ps, pc;
        disney_lobe_probs(m, pd, ps, pc);
        if (u[0] < pd) {
            wo = cosine_sample_hemisphere(u[1], u[2]);
        } else if (u[0] < pd + ps) {
            V3<T> wh = ggx_sample_wh(u[1], u[2], std::max(m.roughness * m.roughness, T(1e-4)));
            wo = T(2) * dot(wh, wi) * wh - wi;
        } else {
            V3<T> wh = gtr1_sample_wh(u[1], u[2], disney_cc_alpha(m));
            wo = T(2) * dot(wh, wi) * wh - wi;
        }
        if (wo.z <= T(0)) return V3<T>(T(0));
        pdf = pd * (wo.z * Consts<T>::INV_PI) + ps * disney_spec_pdf(wi, wo, m) + pc * disney_cc_pdf(wi, wo, m);
        if (pdf <= T(0)) return V3<T>(T(0));
        return disney_eval(wi, wo, m) * wo.z / pdf;
    }
    return V3<T>(T(0));
}

} // namespace orc
