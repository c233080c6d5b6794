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
    } else {
        f = mk(0.0f);
        pdf = 0.0f;
    }
}
NG_HD int material_sample_dims(uint32_t t) { return t == NORI_MAT_DISNEY ? 3 : 2; }

// Returns f*cos/pdf; u holds material_sample_dims(type) uniforms.
NG_HD f3 material_sample(f3 wi, const float *u, const Mat &m, f3 &wo, float &pdf) {
    wo = mk(0, 0, 1);
    pdf = 0.0f;
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
