// BSDFs of the GPU path (host/device), local frame with z = shading normal;
// `wi` points to the viewer and `wo` to the light (shaders/Material.hlsli:18-194).
//
//   diffuse     shaders/Material.hlsli:20-25, 64-69, 109-121
//   microfacet  shaders/Microfacet.hlsli:13-125
//   Disney      shaders/Disney.hlsli:10-275 and shaders/Material.hlsli:31-39, 75-84, 127-183
//
// Algebraically identical rewrites used for FP32 accuracy:
//   * GGX/GTR1 denominators 1 + (a^2-1)cos^2 are evaluated as
//     sin^2 + a^2 cos^2 with sin^2 = x^2 + y^2 of the half vector, avoiding the
//     cancellation near the specular peak;
//   * pow(x, 5) is x^4 * x.
#pragma once
#include "../../include/nori_gpu.h"
#include "nori_math.h"

namespace ng {

// Fields of the 144-byte record that shading reads.
struct Mat {
    uint32_t type;
    f3 albedo;
    float intIOR, extIOR, alpha;
    float roughness, metallic, specular, specularTint, sheen, sheenTint, subsurface, clearcoat, clearcoatGloss;
    float betaN;
};

NG_HD Mat load_mat(const nori_material &r) {
    Mat m;
    m.type = r.type;
    m.albedo = mk(r.albedo[0], r.albedo[1], r.albedo[2]);
    m.intIOR = r.intIOR; m.extIOR = r.extIOR; m.alpha = r.alpha;
    m.roughness = r.roughness; m.metallic = r.metallic; m.specular = r.specular; m.specularTint = r.specularTint;
    m.sheen = r.sheen; m.sheenTint = r.sheenTint; m.subsurface = r.subsurface; m.clearcoat = r.clearcoat;
    m.clearcoatGloss = r.clearcoatGloss; m.betaN = r.betaN;
    return m;
}

NG_HD float pow5(float x) {
    float x2 = x * x;
    return x2 * x2 * x;
}

// ------------------------------------------------------------- microfacet
NG_HD float beckmann_d(f3 wh, float a) {
    if (wh.z <= 0.0f) return 0.0f;
    float c2 = wh.z * wh.z, s2 = wh.x * wh.x + wh.y * wh.y, a2 = a * a;
    return expf(-(s2 / c2) / a2) / (PI * a2 * c2 * wh.z);
}
NG_HD float beckmann_d_cos(f3 wh, float a) {
    if (wh.z <= 0.0f) return 0.0f;
    float c2 = wh.z * wh.z, s2 = wh.x * wh.x + wh.y * wh.y, a2 = a * a;
    return expf(-(s2 / c2) / a2) / (PI * a2 * c2);
}
NG_HD f3 beckmann_sample(float u0, float u1, float a) {
    float phi = 2.0f * PI * u1;
    float t2 = -a * a * logf(fmaxf(1e-8f, 1.0f - u0));
    float c = 1.0f / sqrtf(1.0f + t2);
    float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    float sp, cp;
    sincosf(phi, &sp, &cp);
    return mk(s * cp, s * sp, c);
}
NG_HD float smith_g1(f3 v, f3 wh, float a) {
    float dvw = dot(v, wh), dvn = v.z;
    if (dvw / dvn <= 0.0f) return 0.0f;
    float c2 = v.z * v.z;
    if (c2 < 1e-10f) return 0.0f;
    float tant = sqrtf(fmaxf(0.0f, 1.0f - c2)) / fabsf(v.z);
    if (tant < 1e-10f) return 1.0f;
    float b = 1.0f / (a * tant);
    if (b >= 1.6f) return 1.0f;
    float b2 = b * b;
    return (3.535f * b + 2.181f * b2) / (1.0f + 2.276f * b + 2.577f * b2);
}
NG_HD f3 microfacet_eval(f3 wi, f3 wo, const Mat &m) {
    if (wi.z <= 0.0f || wo.z <= 0.0f) return mk(0.0f);
    float ks = 1.0f - maxc(m.albedo);
    f3 wh = normalize(wi + wo);
    float D = beckmann_d(wh, m.alpha);
    float F = fresnel_dielectric(dot(wh, wi), m.extIOR, m.intIOR);
    float G = smith_g1(wi, wh, m.alpha) * smith_g1(wo, wh, m.alpha);
    return m.albedo * INV_PI + mk(ks * D * F * G / fmaxf(4.0f * wi.z * wo.z, 1e-10f));
}
NG_HD float microfacet_pdf(f3 wi, f3 wo, const Mat &m) {
    if (wi.z <= 0.0f || wo.z <= 0.0f) return 0.0f;
    float ks = 1.0f - maxc(m.albedo);
    f3 wh = normalize(wi + wo);
    return ks * beckmann_d_cos(wh, m.alpha) / fmaxf(4.0f * dot(wh, wo), 1e-10f) + (1.0f - ks) * wo.z * INV_PI;
}

// ------------------------------------------------------------- Disney
struct DisneyPre {
    f3 f0, csheen;
    float a_spec, a_cc, pd, ps, pc;
};
NG_HD DisneyPre disney_pre(const Mat &m) {
    DisneyPre q;
    float lum = luminance(m.albedo);
    f3 ctint = lum > 0.0f ? m.albedo * (1.0f / lum) : mk(1.0f);
    f3 cspec0 = (0.08f * m.specular) * lerp3(mk(1.0f), ctint, m.specularTint);
    q.f0 = lerp3(cspec0, m.albedo, m.metallic);
    q.csheen = lerp3(mk(1.0f), ctint, m.sheenTint);
    q.a_spec = fmaxf(m.roughness * m.roughness, 1e-4f);
    q.a_cc = 0.1f + (0.001f - 0.1f) * m.clearcoatGloss;
    float wd = 1.0f - m.metallic, wc = 0.25f * m.clearcoat, tot = wd + 1.0f + wc;
    float inv = tot > 0.0f ? 1.0f / tot : 0.0f;
    q.pd = wd * inv;
    q.ps = inv;
    q.pc = wc * inv;
    return q;
}
// GGX D with the denominator sin^2 + a^2 cos^2 (== 1 + (a^2-1) cos^2).
NG_HD float ggx_d(f3 wh, float a) {
    if (wh.z <= 0.0f) return 0.0f;
    float a2 = a * a;
    float den = (wh.x * wh.x + wh.y * wh.y) + a2 * (wh.z * wh.z);
    return a2 / fmaxf(PI * den * den, 1e-10f);
}
NG_HD float ggx_g1(float nv, float a) {
    if (nv <= 0.0f) return 0.0f;
    float a2 = a * a;
    return 2.0f * nv / fmaxf(nv + sqrtf(a2 + (1.0f - a2) * nv * nv), 1e-10f);
}
NG_HD f3 ggx_sample_wh(float u0, float u1, float a) {
    float a2 = a * a;
    float c2 = (1.0f - u0) / fmaxf(1.0f + (a2 - 1.0f) * u0, 1e-10f);
    float c = sqrtf(saturatef_(c2)), s = sqrtf(fmaxf(0.0f, 1.0f - c2));
    float sp, cp;
    sincosf(2.0f * PI * u1, &sp, &cp);
    return mk(s * cp, s * sp, c);
}
NG_HD float gtr1_d(f3 wh, float a) {
    if (wh.z <= 0.0f) return 0.0f;
    float a2 = a * a;
    if (a2 >= 0.9999f) return INV_PI;
    float t = (wh.x * wh.x + wh.y * wh.y) + a2 * (wh.z * wh.z);
    return (a2 - 1.0f) / (PI * logf(a2) * t);
}
NG_HD f3 gtr1_sample_wh(float u0, float u1, float a) {
    float a2 = a * a;
    float c2 = a2 >= 0.9999f ? 1.0f - u0 : (1.0f - powf(a2, 1.0f - u0)) / (1.0f - a2);
    float c = sqrtf(saturatef_(c2)), s = sqrtf(fmaxf(0.0f, 1.0f - c2));
    float sp, cp;
    sincosf(2.0f * PI * u1, &sp, &cp);
    return mk(s * cp, s * sp, c);
}

// f and pdf together (the sample, NEE and MIS paths all need both).
NG_HD void disney_eval_pdf(f3 wi, f3 wo, const Mat &m, const DisneyPre &q, f3 &f, float &pdf) {
    f = mk(0.0f);
    pdf = 0.0f;
    if (wi.z <= 0.0f || wo.z <= 0.0f) return;
    f3 wh = normalize(wi + wo);
    float cd = fabsf(dot(wo, wh));
    float fwi = pow5(1.0f - wi.z), fwo = pow5(1.0f - wo.z);
    // Burley diffuse and the Hanrahan-Krueger blend (shaders/Disney.hlsli:10-53)
    float fd90 = 0.5f + 2.0f * m.roughness * cd * cd;
    float burley = (1.0f + (fd90 - 1.0f) * fwi) * (1.0f + (fd90 - 1.0f) * fwo);
    float fss90 = m.roughness * cd * cd;
    float fss = (1.0f + (fss90 - 1.0f) * fwi) * (1.0f + (fss90 - 1.0f) * fwo);
    float ss = 1.25f * (fss * (1.0f / fmaxf(wi.z + wo.z, 1e-4f) - 0.5f) + 0.5f);
    f3 fdiff = m.albedo * (INV_PI * (burley + (ss - burley) * m.subsurface));
    f3 fsheen = mk(0.0f);
    if (m.sheen > 0.0f) fsheen = q.csheen * (m.sheen * pow5(1.0f - cd));
    f3 fspec = mk(0.0f), fcc = mk(0.0f);
    float pspec = 0.0f, pcc = 0.0f;
    if (wh.z > 0.0f) {
        float vh = fabsf(dot(wi, wh));
        float fh = pow5(1.0f - vh);
        float D = ggx_d(wh, q.a_spec);
        float G = ggx_g1(wi.z, q.a_spec) * ggx_g1(wo.z, q.a_spec);
        f3 F = q.f0 + (mk(1.0f) - q.f0) * fh;
        float inv4 = 1.0f / fmaxf(4.0f * wi.z * wo.z, 1e-10f);
        fspec = F * (D * G * inv4);
        float jac = wh.z / fmaxf(4.0f * fabsf(dot(wh, wo)), 1e-10f);
        pspec = D * jac;
        if (m.clearcoat > 0.0f) {
            float Dc = gtr1_d(wh, q.a_cc);
            float Fc = 0.04f + 0.96f * fh;
            float Gc = ggx_g1(wi.z, 0.25f) * ggx_g1(wo.z, 0.25f);
            fcc = mk(0.25f * m.clearcoat * Dc * Fc * Gc * inv4);
            pcc = Dc * jac;
        }
    }
    f = (fdiff + fsheen) * (1.0f - m.metallic) + fspec + fcc;
    pdf = q.pd * (wo.z * INV_PI) + q.ps * pspec + q.pc * pcc;
}

// ------------------------------------------------------------- hair
// Chiang et al. 2016 hair BCSDF (PBRT 4e §9.9) as shaders/Hair.hlsli:26-432
// evaluates it.  Frame: x = fiber tangent, z = surface normal; sin(theta) = w.x,
// phi = atan2(w.z, w.y); h in [-1, 1] is the offset across the fiber.  Material
// fields: roughness = beta_m, betaN = beta_n, alpha = scale tilt in degrees,
// intIOR = eta, albedo = fiber colour (mapped to sigma_a by Chiang eq. 9).
// I0(x) = sum_k (x^2/4)^k / (k!)^2 for x <= 12, summed to convergence: term k
// from term k-1 by the factor (x^2/4)/k^2 (no overflowing powers or
// factorials); at x = 12 term 25 is 2e-16 of the sum.  The reference
// (shaders/Hair.hlsli:26-40) stops at 10 terms, 2 % short of I0 at x = 12.
NG_HD float hair_i0(float x) {
    const float q = 0.25f * x * x;
    float val = 1.0f, t = 1.0f;
    for (int k = 1; k < 26; ++k) {
        t *= q / (float)(k * k);
        val += t;
    }
    return val;
}
// log I0(x); above 12 its asymptotic expansion x - log(2 pi x)/2 + 1/(8x) +
// 1/(16x^2) (remainder under 4e-5).  The reference (:43-48) writes the 1/(8x)
// term halved, 1/(16x) short.
NG_HD float hair_log_i0(float x) {
    if (x > 12.0f) {
        const float r = 1.0f / x;
        return x - 0.5f * logf(2.0f * PI * x) + r * (0.125f + 0.0625f * r);
    }
    return logf(hair_i0(x));
}
NG_HD float hair_mp(float cos_i, float cos_o, float sin_i, float sin_o, float v) { // :55-66
    float a = cos_i * cos_o / v, b = sin_i * sin_o / v;
    if (v <= 0.1f) return expf(hair_log_i0(a) - b - 1.0f / v + 0.6931f + logf(1.0f / (2.0f * v)));
    return (expf(-b) * hair_i0(a)) / (sinhf(1.0f / v) * 2.0f * v);
}
NG_HD float hair_logistic(float x, float s) {
    float ex = expf(-fabsf(x) / s);
    return ex / (s * (1.0f + ex) * (1.0f + ex));
}
NG_HD float hair_logistic_cdf(float x, float s) { return 1.0f / (1.0f + expf(-x / s)); }
// Logistic trimmed to [-pi, pi].
NG_HD float hair_trimmed_logistic(float x, float s) {
    float den = hair_logistic_cdf(PI, s) - hair_logistic_cdf(-PI, s);
    return den > 1e-10f ? hair_logistic(x, s) / den : 0.0f;
}
NG_HD float hair_sample_trimmed_logistic(float u, float s) { // :90-97
    float ca = hair_logistic_cdf(-PI, s), cb = hair_logistic_cdf(PI, s);
    float t = fminf(fmaxf(ca + u * (cb - ca), 1e-6f), 1.0f - 1e-6f);
    return -s * logf(1.0f / t - 1.0f);
}
NG_HD float hair_phi(int p, float go, float gt) { return 2.0f * (float)p * gt - 2.0f * go + (float)p * PI; }
NG_HD float hair_np(float phi, int p, float s, float go, float gt) { // :108-117
    float dphi = phi - hair_phi(p, go, gt);
    while (dphi > PI) dphi -= 2.0f * PI;
    while (dphi < -PI) dphi += 2.0f * PI;
    return hair_trimmed_logistic(dphi, s);
}
NG_HD float hair_sa1(float c, float den) { // :125-133
    if (c <= 1e-4f) return 100.0f;
    float l = logf(c) / den;
    return l * l;
}
struct HairPre {
    float v[4], s, sin2k[3], cos2k[3];
    f3 sa;
};
NG_HD HairPre hair_pre(const Mat &m) { // :145-177
    HairPre p;
    float bm = m.roughness;
    float v0 = 0.726f * bm + 0.812f * bm * bm + 3.7f * powf(bm, 20.0f);
    v0 = v0 * v0;
    p.v[0] = v0;
    p.v[1] = 0.25f * v0;
    p.v[2] = 4.0f * v0;
    p.v[3] = p.v[2];
    float bn = m.betaN;
    p.s = 0.626657069f * (0.265f * bn + 1.194f * bn * bn + 5.372f * powf(bn, 22.0f));
    float ar = m.alpha * (PI / 180.0f);
    p.sin2k[0] = sinf(ar);
    p.cos2k[0] = sqrtf(fmaxf(0.0f, 1.0f - p.sin2k[0] * p.sin2k[0]));
    for (int i = 1; i < 3; ++i) {
        p.sin2k[i] = 2.0f * p.cos2k[i - 1] * p.sin2k[i - 1];
        p.cos2k[i] = p.cos2k[i - 1] * p.cos2k[i - 1] - p.sin2k[i - 1] * p.sin2k[i - 1];
    }
    float b2 = bn * bn, b3 = b2 * bn, b4 = b3 * bn, b5 = b4 * bn;
    float den = 5.969f - 0.215f * bn + 2.532f * b2 - 10.73f * b3 + 5.574f * b4 + 0.245f * b5;
    p.sa = mk(hair_sa1(m.albedo.x, den), hair_sa1(m.albedo.y, den), hair_sa1(m.albedo.z, den));
    return p;
}
// Attenuation per lobe (:182-208).
NG_HD void hair_ap(float cos_o, float sin_o, float h, float eta, f3 sa, f3 ap[4]) {
    float etap = sqrtf(fmaxf(0.0f, eta * eta - sin_o * sin_o)) / fmaxf(cos_o, 1e-6f);
    float sgt = fminf(fmaxf(h / etap, -1.0f), 1.0f);
    float cgt = sqrtf(fmaxf(0.0f, 1.0f - sgt * sgt));
    float stt = sin_o / eta;
    float ctt = sqrtf(fmaxf(0.0f, 1.0f - stt * stt));
    float len = 2.0f * cgt / fmaxf(ctt, 1e-6f);
    f3 Tr = mk(expf(-sa.x * len), expf(-sa.y * len), expf(-sa.z * len));
    float cgo = sqrtf(fmaxf(0.0f, 1.0f - h * h));
    float f = fresnel_dielectric(cos_o * cgo, 1.0f, eta);
    ap[0] = mk(f);
    ap[1] = Tr * ((1.0f - f) * (1.0f - f));
    ap[2] = ap[1] * Tr * f;
    f3 tf = Tr * f;
    f3 den = mk(fmaxf(1.0f - tf.x, 1e-10f), fmaxf(1.0f - tf.y, 1e-10f), fmaxf(1.0f - tf.z, 1e-10f));
    f3 a3 = ap[2] * f * Tr;
    ap[3] = mk(a3.x / den.x, a3.y / den.y, a3.z / den.z);
}
// Scale-tilted outgoing angle of lobe p (:214-239).
NG_HD void hair_tilt(int p, float sin_o, float cos_o, const HairPre &hp, float &sp, float &cp) {
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
    cp = fabsf(cp);
}
struct HairGeom {
    float sin_o, cos_o, phi_o, go, gt;
};
NG_HD HairGeom hair_geom(f3 wi, float h, float eta) {
    HairGeom g;
    g.sin_o = wi.x;
    g.cos_o = sqrtf(fmaxf(0.0f, 1.0f - g.sin_o * g.sin_o));
    g.phi_o = atan2f(wi.z, wi.y);
    g.go = asinf(fminf(fmaxf(h, -1.0f), 1.0f));
    float etap = sqrtf(fmaxf(0.0f, eta * eta - g.sin_o * g.sin_o)) / fmaxf(g.cos_o, 1e-6f);
    g.gt = asinf(fminf(fmaxf(h / etap, -1.0f), 1.0f));
    return g;
}
// f (divided by |wo.z|, :248-300) and the lobe-mixture pdf (:303-348) together.
NG_HD void hair_eval_pdf(f3 wi, f3 wo, const Mat &m, float h, const HairGeom &g, const HairPre &hp, const f3 ap[4],
                         f3 &f, float &pdf) {
    float sin_i = wo.x, cos_i = sqrtf(fmaxf(0.0f, 1.0f - sin_i * sin_i));
    float phi = atan2f(wo.z, wo.y) - g.phi_o;
    float lum[4], sum = 0.0f;
    for (int i = 0; i < 4; ++i) {
        lum[i] = luminance(ap[i]);
        sum += lum[i];
    }
    float inv = sum > 0.0f ? 1.0f / sum : 0.0f;
    f = mk(0.0f);
    pdf = 0.0f;
    for (int p = 0; p < 3; ++p) {
        float sp, cp;
        hair_tilt(p, g.sin_o, g.cos_o, hp, sp, cp);
        float mn = hair_mp(cos_i, cp, sin_i, sp, hp.v[p]) * hair_np(phi, p, hp.s, g.go, g.gt);
        f += ap[p] * mn;
        pdf += mn * (lum[p] * inv);
    }
    float m3 = hair_mp(cos_i, g.cos_o, sin_i, g.sin_o, hp.v[3]) * (1.0f / (2.0f * PI));
    f += ap[3] * m3;
    pdf += m3 * (lum[3] * inv);
    float ac = fabsf(wo.z);
    if (ac > 1e-6f) f = f * (1.0f / ac);
}
NG_HD void hair_eval_pdf(f3 wi, f3 wo, const Mat &m, float h, f3 &f, float &pdf) {
    HairGeom g = hair_geom(wi, h, m.intIOR);
    HairPre hp = hair_pre(m);
    f3 ap[4];
    hair_ap(g.cos_o, g.sin_o, h, m.intIOR, hp.sa, ap);
    hair_eval_pdf(wi, wo, m, h, g, hp, ap, f, pdf);
}
// :351-432; u[0] picks the lobe, u[1], u[2] the longitudinal angle, u[3] the azimuth.
NG_HD f3 hair_sample(f3 wi, const float *u, const Mat &m, float h, f3 &wo, float &pdf) {
    HairGeom g = hair_geom(wi, h, m.intIOR);
    HairPre hp = hair_pre(m);
    f3 ap[4];
    hair_ap(g.cos_o, g.sin_o, h, m.intIOR, hp.sa, ap);
    float apdf[4], sum = 0.0f;
    for (int i = 0; i < 4; ++i) {
        apdf[i] = luminance(ap[i]);
        sum += apdf[i];
    }
    float inv = sum > 0.0f ? 1.0f / sum : 0.0f;
    int p = 0;
    float cdf = apdf[0] * inv;
    while (p < 3 && u[0] > cdf) {
        ++p;
        cdf += apdf[p] * inv;
    }
    float sp, cp;
    hair_tilt(p, g.sin_o, g.cos_o, hp, sp, cp);
    float vp = hp.v[p];
    float ct = 1.0f + vp * logf(fmaxf(u[1], 1e-5f) + (1.0f - u[1]) * expf(-2.0f / vp));
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float cphi = cosf(2.0f * PI * u[2]);
    float sin_i = -ct * sp + st * cphi * cp;
    float cos_i = sqrtf(fmaxf(0.0f, 1.0f - sin_i * sin_i));
    float dphi = p < 3 ? hair_phi(p, g.go, g.gt) + hair_sample_trimmed_logistic(u[3], hp.s) : 2.0f * PI * u[3];
    float phi_i = g.phi_o + dphi;
    float sphi, cphi_i;
    sincosf(phi_i, &sphi, &cphi_i);
    wo = mk(sin_i, cos_i * cphi_i, cos_i * sphi);
    f3 f;
    hair_eval_pdf(wi, wo, m, h, g, hp, ap, f, pdf);
    if (pdf <= 0.0f) return mk(0.0f);
    return f * (fabsf(wo.z) / pdf);
}

// ------------------------------------------------------------- dispatch
NG_HD void material_eval_pdf(f3 wi, f3 wo, const Mat &m, f3 &f, float &pdf) {
    if (m.type == NORI_MAT_DISNEY) {
        DisneyPre q = disney_pre(m);
        disney_eval_pdf(wi, wo, m, q, f, pdf);
    } else if (m.type == NORI_MAT_DIFFUSE) {
        if (wi.z <= 0.0f || wo.z <= 0.0f) { f = mk(0.0f); pdf = 0.0f; return; }
        f = m.albedo * INV_PI;
        pdf = wo.z * INV_PI;
    } else if (m.type == NORI_MAT_MICROFACET) {
        f = microfacet_eval(wi, wo, m);
        pdf = microfacet_pdf(wi, wo, m);
    } else if (m.type == NORI_MAT_SUBSURFACE) {
        // the boundary's Beckmann reflection lobe (shaders/Material.hlsli:48-57, 92-100)
        if (wi.z <= 0.0f || wo.z <= 0.0f) { f = mk(0.0f); pdf = 0.0f; return; }
        const float a = fmaxf(m.roughness * m.roughness, 1e-4f);
        const f3 wh = normalize(wi + wo);
        const float Fr = fresnel_dielectric(dot(wh, wi), m.extIOR, m.intIOR);
        const float G = smith_g1(wi, wh, a) * smith_g1(wo, wh, a);
        f = mk(beckmann_d(wh, a) * Fr * G / fmaxf(4.0f * wi.z * wo.z, 1e-10f));
        pdf = Fr * beckmann_d_cos(wh, a) / fmaxf(4.0f * dot(wh, wo), 1e-10f);
    } else {
        f = mk(0.0f);
        pdf = 0.0f;
    }
}
// Dispatch including hair (h is ignored by the other materials).  HAIR =
// false compiles the hair BCSDF out, for callers that never see hair (its
// per-lobe tables would otherwise cost registers and scratch everywhere).
template <bool HAIR = true>
NG_HD void material_eval_pdf(f3 wi, f3 wo, const Mat &m, float h, f3 &f, float &pdf) {
    if (HAIR && m.type == NORI_MAT_HAIR) hair_eval_pdf(wi, wo, m, h, f, pdf);
    else material_eval_pdf(wi, wo, m, f, pdf);
}
NG_HD int material_sample_dims(uint32_t t) { return t == NORI_MAT_HAIR ? 4 : t == NORI_MAT_DISNEY ? 3 : 2; }

// Returns f*|cos|/pdf; u holds material_sample_dims(type) uniforms.
template <bool HAIR = true>
NG_HD f3 material_sample(f3 wi, const float *u, const Mat &m, float h, f3 &wo, float &pdf) {
    wo = mk(0, 0, 1);
    pdf = 0.0f;
    if (HAIR && m.type == NORI_MAT_HAIR) return hair_sample(wi, u, m, h, wo, pdf);
    if (wi.z <= 0.0f) return mk(0.0f);
    if (m.type == NORI_MAT_DIFFUSE) {
        wo = cosine_sample_hemisphere(u[0], u[1]);
        pdf = wo.z * INV_PI;
        return m.albedo;
    }
    if (m.type == NORI_MAT_MICROFACET) {
        float ks = 1.0f - maxc(m.albedo), ux = u[0];
        if (ux < ks) {
            f3 wh = beckmann_sample(ux / ks, u[1], m.alpha);
            wo = 2.0f * dot(wh, wi) * wh - wi;
        } else {
            wo = cosine_sample_hemisphere((ux - ks) / (1.0f - ks), u[1]);
        }
        if (wo.z <= 0.0f) return mk(0.0f);
        pdf = microfacet_pdf(wi, wo, m);
        if (pdf <= 0.0f) return mk(0.0f);
        return microfacet_eval(wi, wo, m) * (wo.z / pdf);
    }
    if (m.type == NORI_MAT_DISNEY) {
        DisneyPre q = disney_pre(m);
        if (u[0] < q.pd) {
            wo = cosine_sample_hemisphere(u[1], u[2]);
        } else if (u[0] < q.pd + q.ps) {
            f3 wh = ggx_sample_wh(u[1], u[2], q.a_spec);
            wo = 2.0f * dot(wh, wi) * wh - wi;
        } else {
            f3 wh = gtr1_sample_wh(u[1], u[2], q.a_cc);
            wo = 2.0f * dot(wh, wi) * wh - wi;
        }
        if (wo.z <= 0.0f) return mk(0.0f);
        f3 f;
        disney_eval_pdf(wi, wo, m, q, f, pdf);
        if (pdf <= 0.0f) return mk(0.0f);
        return f * (wo.z / pdf);
    }
    return mk(0.0f);
}

} // namespace ng
