// CPU oracle — test infrastructure only.
//
// Chiang et al. 2016 hair BCSDF as the reference's GPU path evaluates it,
// shaders/Hair.hlsli:26-432 (PBRT 4th ed. §9.9), templated on the float type.
// Hair frame: x = fiber tangent, z = ribbon/tube normal; sin(theta) = w.x,
// phi = atan2(w.z, w.y).  `wi` is the viewing direction, `wo` the light
// direction; h in [-1, 1] is the offset across the fiber.
#pragma once
#include "orc_bsdf.h"

namespace orc {

// I0(x) = sum_k (x^2/4)^k / (k!)^2 summed to convergence (26 terms: at x = 12
// term 25 is 2e-16 of the sum), each term from the last by (x^2/4)/k^2.  The
// reference (shaders/Hair.hlsli:26-40) stops at 10 terms, 2 % short at x = 12.
template <typename T> T hair_i0(T x) {
    const T q = T(0.25) * x * x;
    T val = T(1), t = T(1);
    for (int k = 1; k < 26; ++k) {
        t *= q / T(k * k);
        val += t;
    }
    return val;
}
// log I0(x); above 12 its asymptotic expansion x - log(2 pi x)/2 + 1/(8x) +
// 1/(16x^2).  The reference (shaders/Hair.hlsli:43-48) halves the 1/(8x) term.
template <typename T> T hair_log_i0(T x) {
    if (x > T(12)) {
        const T r = T(1) / x;
        return x - T(0.5) * std::log(T(2) * Consts<T>::PI * x) + r * (T(0.125) + T(0.0625) * r);
    }
    return std::log(hair_i0(x));
}
template <typename T> T hair_mp(T cos_i, T cos_o, T sin_i, T sin_o, T v) { // shaders/Hair.hlsli:55-66
    T a = cos_i * cos_o / v, b = sin_i * sin_o / v;
    if (v <= T(0.1)) return std::exp(hair_log_i0(a) - b - T(1) / v + T(0.6931) + std::log(T(1) / (T(2) * v)));
    return (std::exp(-b) * hair_i0(a)) / (std::sinh(T(1) / v) * T(2) * v);
}
template <typename T> T hair_logistic(T x, T s) {
    T ex = std::exp(-std::fabs(x) / s);
    return ex / (s * (T(1) + ex) * (T(1) + ex));
}
template <typename T> T hair_logistic_cdf(T x, T s) { return T(1) / (T(1) + std::exp(-x / s)); }
template <typename T> T hair_trimmed_logistic(T x, T s, T a, T b) {
    T den = hair_logistic_cdf(b, s) - hair_logistic_cdf(a, s);
    return den > T(1e-10) ? hair_logistic(x, s) / den : T(0);
}
template <typename T> T hair_sample_trimmed_logistic(T u, T s, T a, T b) { // :90-97
    T ca = hair_logistic_cdf(a, s), cb = hair_logistic_cdf(b, s);
    T t = ca + u * (cb - ca);
    t = std::min(std::max(t, T(1e-6)), T(1) - T(1e-6));
    return -s * std::log(T(1) / t - T(1));
}
template <typename T> T hair_phi(int p, T go, T gt) { return T(2) * T(p) * gt - T(2) * go + T(p) * Consts<T>::PI; }
template <typename T> T hair_np(T phi, int p, T s, T go, T gt) { // :108-117
    const T PI = Consts<T>::PI;
    T dphi = phi - hair_phi(p, go, gt);
    while (dphi > PI) dphi -= T(2) * PI;
    while (dphi < -PI) dphi += T(2) * PI;
    return hair_trimmed_logistic(dphi, s, -PI, PI);
}
// shaders/Hair.hlsli:125-133 (Chiang et al. 2016 eq. 9)
template <typename T> V3<T> hair_sigma_a_from_color(V3<T> c, T bn) {
    T den = T(5.969) - T(0.215) * bn + T(2.532) * bn * bn - T(10.73) * bn * bn * bn + T(5.574) * bn * bn * bn * bn +
            T(0.245) * bn * bn * bn * bn * bn;
    auto one = [&](T x) { return x > T(1e-4) ? std::pow(std::log(x) / den, T(2)) : T(100); };
    return V3<T>(one(c.x), one(c.y), one(c.z));
}
template <typename T> struct HairParams {
    T v[4], s, sin2k[3], cos2k[3];
    V3<T> sigma_a;
};
template <typename T> HairParams<T> hair_params(const MatT<T> &m) { // :145-177
    HairParams<T> p;
    T bm = m.roughness;
    T v0 = T(0.726) * bm + T(0.812) * bm * bm + T(3.7) * std::pow(bm, T(20));
    v0 = v0 * v0;
    p.v[0] = v0;
    p.v[1] = T(0.25) * v0;
    p.v[2] = T(4) * v0;
    p.v[3] = p.v[2];
    T bn = m.betaN;
    p.s = T(0.626657069) * (T(0.265) * bn + T(1.194) * bn * bn + T(5.372) * std::pow(bn, T(22)));
    T ar = m.alpha * Consts<T>::PI / T(180);
    p.sin2k[0] = std::sin(ar);
    p.cos2k[0] = std::sqrt(std::max(T(0), T(1) - p.sin2k[0] * p.sin2k[0]));
    for (int i = 1; i < 3; ++i) {
        p.sin2k[i] = T(2) * p.cos2k[i - 1] * p.sin2k[i - 1];
        p.cos2k[i] = p.cos2k[i - 1] * p.cos2k[i - 1] - p.sin2k[i - 1] * p.sin2k[i - 1];
    }
    p.sigma_a = hair_sigma_a_from_color(m.albedo, bn);
    return p;
}
template <typename T> void hair_ap(T cos_o, T sin_o, T h, T eta, V3<T> sa, V3<T> ap[4]) { // :182-208
    T etap = std::sqrt(std::max(T(0), eta * eta - sin_o * sin_o)) / std::max(cos_o, T(1e-6));
    T sgt = std::min(std::max(h / etap, T(-1)), T(1));
    T cgt = std::sqrt(std::max(T(0), T(1) - sgt * sgt));
    T stt = sin_o / eta;
    T ctt = std::sqrt(std::max(T(0), T(1) - stt * stt));
    T len = T(2) * cgt / std::max(ctt, T(1e-6));
    V3<T> Tr(std::exp(-sa.x * len), std::exp(-sa.y * len), std::exp(-sa.z * len));
    T cgo = std::sqrt(std::max(T(0), T(1) - h * h));
    T f = fresnel_dielectric(cos_o * cgo, T(1), eta);
    ap[0] = V3<T>(f);
    ap[1] = Tr * ((T(1) - f) * (T(1) - f));
    ap[2] = ap[1] * Tr * f;
    V3<T> tf = Tr * f;
    V3<T> den(std::max(T(1) - tf.x, T(1e-10)), std::max(T(1) - tf.y, T(1e-10)), std::max(T(1) - tf.z, T(1e-10)));
    ap[3] = ap[2] * f * Tr / den;
}
template <typename T> void hair_tilt(int p, T sin_o, T cos_o, const HairParams<T> &hp, T &sp, T &cp) { // :214-239
    if (p == 0) {
        sp = sin_o * hp.cos2k[1] - cos_o * hp.sin2k[1];
        cp = cos_o * hp.cos2k[1] + sin_o * hp.sin2k[1];
    } else if (p == 1) {
        sp = sin_o * hp.cos2k[0] + cos_o * hp.sin2k[0];
        cp = cos_o * hp.cos2k[0] - sin_o * hp.sin2k[0];
    } else if (p == 2) {
        sp = sin_o * hp.cos2k[2] + cos_o * hp.sin2k[2];
        cp = cos_o * hp.cos2k[2] - sin_o * hp.sin2k[2];
    } else {
        sp = sin_o;
        cp = cos_o;
    }
    cp = std::fabs(cp);
}
template <typename T> struct HairGeom {
    T sin_o, cos_o, phi_o, go, gt, eta;
};
template <typename T> HairGeom<T> hair_geom(V3<T> wi, T h, T eta) {
    HairGeom<T> g;
    g.sin_o = wi.x;
    g.cos_o = std::sqrt(std::max(T(0), T(1) - g.sin_o * g.sin_o));
    g.phi_o = std::atan2(wi.z, wi.y);
    g.go = std::asin(std::min(std::max(h, T(-1)), T(1)));
    g.eta = eta;
    T etap = std::sqrt(std::max(T(0), eta * eta - g.sin_o * g.sin_o)) / std::max(g.cos_o, T(1e-6));
    g.gt = std::asin(std::min(std::max(h / etap, T(-1)), T(1)));
    return g;
}
// shaders/Hair.hlsli:248-300
template <typename T> V3<T> hair_eval(V3<T> wi, V3<T> wo, const MatT<T> &m, T h) {
    HairGeom<T> g = hair_geom(wi, h, m.intIOR);
    T sin_i = wo.x, cos_i = std::sqrt(std::max(T(0), T(1) - sin_i * sin_i));
    T phi = std::atan2(wo.z, wo.y) - g.phi_o;
    HairParams<T> hp = hair_params(m);
    V3<T> ap[4];
    hair_ap(g.cos_o, g.sin_o, h, g.eta, hp.sigma_a, ap);
    V3<T> f(T(0));
    for (int p = 0; p < 3; ++p) {
        T sp, cp;
        hair_tilt(p, g.sin_o, g.cos_o, hp, sp, cp);
        f += ap[p] * (hair_mp(cos_i, cp, sin_i, sp, hp.v[p]) * hair_np(phi, p, hp.s, g.go, g.gt));
    }
    f += ap[3] * (hair_mp(cos_i, g.cos_o, sin_i, g.sin_o, hp.v[3]) / (T(2) * Consts<T>::PI));
    T ac = std::fabs(wo.z);
    if (ac > T(1e-6)) f = f / ac;
    return f;
}
// shaders/Hair.hlsli:303-348
template <typename T> T hair_pdf(V3<T> wi, V3<T> wo, const MatT<T> &m, T h) {
    HairGeom<T> g = hair_geom(wi, h, m.intIOR);
    T sin_i = wo.x, cos_i = std::sqrt(std::max(T(0), T(1) - sin_i * sin_i));
    T phi = std::atan2(wo.z, wo.y) - g.phi_o;
    HairParams<T> hp = hair_params(m);
    V3<T> ap[4];
    hair_ap(g.cos_o, g.sin_o, h, g.eta, hp.sigma_a, ap);
    T lum[4], sum = 0;
    for (int i = 0; i < 4; ++i) { lum[i] = luminance(ap[i]); sum += lum[i]; }
    T inv = sum > T(0) ? T(1) / sum : T(0);
    T pdf = 0;
    for (int p = 0; p < 3; ++p) {
        T sp, cp;
        hair_tilt(p, g.sin_o, g.cos_o, hp, sp, cp);
        pdf += hair_mp(cos_i, cp, sin_i, sp, hp.v[p]) * lum[p] * inv * hair_np(phi, p, hp.s, g.go, g.gt);
    }
    pdf += hair_mp(cos_i, g.cos_o, sin_i, g.sin_o, hp.v[3]) * lum[3] * inv * (T(1) / (T(2) * Consts<T>::PI));
    return pdf;
}
// shaders/Hair.hlsli:351-432; u[0] lobe, u[1], u[2] longitudinal, u[3] azimuthal.
template <typename T> V3<T> hair_sample(V3<T> wi, const T *u, const MatT<T> &m, T h, V3<T> &wo, T &pdf) {
    HairGeom<T> g = hair_geom(wi, h, m.intIOR);
    HairParams<T> hp = hair_params(m);
    V3<T> ap[4];
    hair_ap(g.cos_o, g.sin_o, h, g.eta, hp.sigma_a, ap);
    T apdf[4], sum = 0;
    for (int i = 0; i < 4; ++i) { apdf[i] = luminance(ap[i]); sum += apdf[i]; }
    T inv = sum > T(0) ? T(1) / sum : T(0);
    for (int i = 0; i < 4; ++i) apdf[i] *= inv;
    T u0 = u[0];
    int p = 0;
    T cdf = apdf[0];
    while (p < 3 && u0 > cdf) {
        ++p;
        cdf += apdf[p];
    }
    T sp, cp;
    hair_tilt(p, g.sin_o, g.cos_o, hp, sp, cp);
    T u1 = u[1], u2 = u[2];
    T ct = T(1) + hp.v[p] * std::log(std::max(u1, T(1e-5)) + (T(1) - u1) * std::exp(T(-2) / hp.v[p]));
    T st = std::sqrt(std::max(T(0), T(1) - ct * ct));
    T cphi = std::cos(T(2) * Consts<T>::PI * u2);
    T sin_i = -ct * sp + st * cphi * cp;
    T cos_i = std::sqrt(std::max(T(0), T(1) - sin_i * sin_i));
    T dphi = p < 3 ? hair_phi(p, g.go, g.gt) + hair_sample_trimmed_logistic(u[3], hp.s, -Consts<T>::PI, Consts<T>::PI)
                   : T(2) * Consts<T>::PI * u[3];
    T phi_i = g.phi_o + dphi;
    wo = V3<T>(sin_i, cos_i * std::cos(phi_i), cos_i * std::sin(phi_i));
    pdf = hair_pdf(wi, wo, m, h);
    if (pdf <= T(0)) return V3<T>(T(0));
    return hair_eval(wi, wo, m, h) * (std::fabs(wo.z) / pdf);
}

// Eumelanin/pheomelanin concentrations -> sigma_a (PBRT 3e §9.9.7; the
// reference only maps colour to sigma_a) and the colour that the reference's
// colour-to-sigma_a fit maps back to that sigma_a exactly.
template <typename T> V3<T> hair_sigma_a_from_melanin(T ce, T cp) {
    return V3<T>(ce * T(0.419) + cp * T(0.187), ce * T(0.697) + cp * T(0.4), ce * T(1.37) + cp * T(1.05));
}

// Material dispatch including hair (the h argument is ignored by the others).
template <typename T> V3<T> material_eval_h(V3<T> wi, V3<T> wo, const MatT<T> &m, T h) {
    return m.type == MAT_HAIR ? hair_eval(wi, wo, m, h) : material_eval(wi, wo, m);
}
template <typename T> T material_pdf_h(V3<T> wi, V3<T> wo, const MatT<T> &m, T h) {
    return m.type == MAT_HAIR ? hair_pdf(wi, wo, m, h) : material_pdf(wi, wo, m);
}
inline int material_sample_dims_h(int t) { return t == MAT_HAIR ? 4 : material_sample_dims(t); }
template <typename T> V3<T> material_sample_h(V3<T> wi, const T *u, const MatT<T> &m, T h, V3<T> &wo, T &pdf) {
    if (m.type == MAT_HAIR) {
        wo = V3<T>(0, 0, 1);
        pdf = T(0);
        return hair_sample(wi, u, m, h, wo, pdf);
    }
    return material_sample(wi, u, m, wo, pdf);
}

} // namespace orc
