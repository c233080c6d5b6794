"""Closed-form pins of the oracle's Beckmann microfacet, Disney lobes and thin-lens
camera: values at configurations where the published formulas reduce to
numbers computed here by hand, integrals that the distributions must satisfy,
and sampling densities derived from the samplers' inverse transforms.

Beckmann follows shaders/Microfacet.hlsli:13-125.  Its BeckmannD is
exp(-tan^2/a^2) / (pi a^2 cos^3): the solid-angle density of the half vectors
BeckmannSample draws (the textbook D times cos theta_h, as Nori's
Warp::squareToBeckmannPdf, src/cpu/warp.cpp:101-113), and BeckmannDCosTheta
carries one more cos theta_h.  The reference's GPU eval and pdf therefore both
equal the textbook ones (src/cpu/microfacet.cpp:37-88) times cos theta_h, and
its pdf integrates to E[cos theta_h] < 1; the oracle keeps that, and these
tests pin it against the sampler's own density.

Disney follows shaders/Disney.hlsli:10-275 (Burley 2012, "Physically-Based
Shading at Disney", section 5); the thin lens src/cpu/perspective.cpp:101-112.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", "pbrt-renderer-_amd"))
import nori_gpu as ng  # noqa: E402
import orc  # noqa: E402

# Rec. 709 luminance weights (shaders/GeometryUtils.hlsli:225-228)
LUM = np.array([0.212671, 0.715160, 0.072169])


def _dir(theta, phi=0.0):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def _f32(v):
    return np.asarray(v, np.float32)


def _mat(kind, n=1, **kw):
    m = ng.default_materials(n)
    m["type"] = kind
    for k, v in kw.items():
        m[k] = v
    return m


def _eval(m, wi, wo):
    f, _, _ = orc.bsdf_batch(0, m, _f32(np.atleast_2d(wi)), _f32(np.atleast_2d(wo)))
    return f


def _pdf(m, wi, wo):
    _, p, _ = orc.bsdf_batch(1, m, _f32(np.atleast_2d(wi)), _f32(np.atleast_2d(wo)))
    return p


# ------------------------------------------------------------------ Beckmann
@pytest.mark.parametrize("a", [0.05, 0.2, 0.5, 1.0])
def test_beckmann_d_cdf(a):
    # BeckmannSample sets tan^2 = -a^2 log(1 - u): P(tan^2 theta <= t) = 1 - exp(-t/a^2).
    # BeckmannD must be that distribution's solid-angle density:
    # integral_0^theta D 2 pi sin = 1 - exp(-tan^2 theta / a^2) (a cos^4 slip gives 1/cos more mass).
    a = np.float32(a).astype(np.float64)
    th = np.linspace(0, math.pi / 2 * 0.999, 200001)
    wh = np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], -1)
    d = orc.microfacet_fn(0, wh, wh, a)
    cum = np.concatenate([[0], np.cumsum(0.5 * (d[1:] * np.sin(th[1:]) + d[:-1] * np.sin(th[:-1])) * np.diff(th))])
    cum *= 2 * math.pi
    expect = 1 - np.exp(-np.tan(th) ** 2 / a ** 2)
    assert np.max(np.abs(cum - expect)) < 1e-5
    # BeckmannDCosTheta carries one more cos theta_h
    dc = orc.microfacet_fn(1, wh, wh, a)
    normal = d > 1e-290  # FP64 normal range
    assert np.allclose(dc[normal], (d * wh[:, 2])[normal], rtol=1e-12, atol=0)
    # at theta = 30 deg: exp(-(1/3)/a^2) / (pi a^2 (3/4)^(3/2))
    w30 = _dir(math.pi / 6)
    v30 = orc.microfacet_fn(0, w30, w30, a)[0]
    assert abs(v30 - math.exp(-1 / 3 / a ** 2) / (math.pi * a ** 2 * 0.75 ** 1.5)) <= 1e-12 * max(v30, 1e-300)
    assert orc.microfacet_fn(0, _dir(0), [0, 0, -1.0], a)[0] == 0.0


def test_smith_g1_against_exact_beckmann_masking():
    # Walter et al. 2007, eq. 27: the rational form approximates the exact Smith
    # G1 = 2 / (1 + erf(b) + exp(-b^2) / (b sqrt(pi))), b = 1/(a tan theta), with
    # relative error below 0.35 %, and equals 1 for b >= 1.6.
    a = 0.3
    for b in np.concatenate([np.linspace(0.02, 1.59, 400), [1.61, 2.0, 5.0]]):
        th = math.atan(1.0 / (a * b))
        v = _dir(th)
        g = orc.microfacet_fn(2, v, [0, 0, 1.0], a)[0]
        exact = 2.0 / (1 + math.erf(b) + math.exp(-b * b) / (b * math.sqrt(math.pi)))
        if b >= 1.6:
            assert g == 1.0
            assert exact > 0.998
        else:
            assert abs(g - exact) <= 0.0035 * exact, (b, g, exact)
    # tabulated: b = 1 -> (3.535 + 2.181) / (1 + 2.276 + 2.577) (the published coefficients)
    th = math.atan(1.0 / a)
    assert abs(orc.microfacet_fn(2, _dir(th), [0, 0, 1.0], a)[0] - 5.716 / 5.853) < 1e-12
    # v below the microsurface (v.wh and v.n of opposite sign) is masked
    assert orc.microfacet_fn(2, _dir(1.2), _dir(-1.4), a)[0] == 0.0
    # normal incidence: tan theta = 0 -> 1
    assert orc.microfacet_fn(2, _dir(0), _dir(0), a)[0] == 1.0


@pytest.mark.parametrize("kd,a,eta", [(0.2, 0.1, 1.5), (0.5, 0.3, 1.33), (0.0, 0.05, 1.7)])
def test_microfacet_normal_incidence(kd, a, eta):
    # wi = wo = n: wh = n, D = 1/(pi a^2), G = 1, F = ((eta - 1)/(eta + 1))^2, ks = 1 - kd:
    # f = kd/pi + ks F / (4 pi a^2); pdf = ks / (4 pi a^2) + (1 - ks)/pi.
    m = _mat(ng.MAT_MICROFACET, albedo=kd, alpha=a, intIOR=eta, extIOR=1.0)
    n = [0, 0, 1.0]
    a64 = float(np.float32(a))
    e64 = float(np.float32(eta))
    F = ((e64 - 1) / (e64 + 1)) ** 2
    kd64 = float(np.float32(kd))
    ks = 1 - kd64
    f = _eval(m, n, n)[0]
    assert np.allclose(f, kd64 / math.pi + ks * F / (4 * math.pi * a64 ** 2), rtol=1e-9)
    p = _pdf(m, n, n)[0]
    assert abs(p - (ks / (4 * math.pi * a64 ** 2) + (1 - ks) / math.pi)) <= 1e-9 * p


def _beckmann_sampling_density(wi, wo, a, ks):
    """Density of the directions MicrofacetSample draws, derived from its warps:
    half vectors with solid-angle density exp(-tan^2/a^2)/(pi a^2 cos^3)
    (d/dtheta of 1 - exp(-tan^2 theta/a^2), over 2 pi sin theta), reflected
    (Jacobian 1/(4 |wh.wo|)); the diffuse share cos theta_o / pi."""
    h = wi + wo
    h = h / np.linalg.norm(h, axis=-1, keepdims=True)
    c = h[..., 2]
    t2 = (1 - c * c) / (c * c)
    dh = np.where(c > 0, np.exp(-t2 / a ** 2) / (math.pi * a ** 2 * c ** 3), 0.0)
    q = ks * dh / (4 * np.abs(np.sum(h * wo, -1))) + (1 - ks) * wo[..., 2] / math.pi
    return np.where(wo[..., 2] > 0, q, 0.0), c


@pytest.mark.parametrize("kd,a", [(0.0, 0.3), (0.4, 0.2), (0.1, 0.6)])
def test_microfacet_sampling_density_and_pdf(kd, a):
    rng = np.random.default_rng(11)
    wi = _dir(0.7, 0.3)
    a32 = float(np.float32(a))
    ks = 1 - float(np.float32(kd))
    # chi-square of MicrofacetSample draws against the derived density on a (cos, phi) grid
    cells, nt, nphi = 12, 480, 480
    ct = (np.arange(nt) + 0.5) / nt
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    C, P = np.meshgrid(ct, ph, indexing="ij")
    S = np.sqrt(1 - C * C)
    grid = np.stack([S * np.cos(P), S * np.sin(P), C], -1).reshape(-1, 3)
    q, _ = _beckmann_sampling_density(wi[None], grid, a32, ks)
    mass = q.reshape(nt, nphi) * (1.0 / nt) * (2 * np.pi / nphi)
    cell_mass = mass.reshape(cells, nt // cells, cells, nphi // cells).sum(axis=(1, 3))
    n = 400000
    m = _mat(ng.MAT_MICROFACET, n, albedo=kd, alpha=a, intIOR=1.5, extIOR=1.0)
    w, spdf, wo = orc.bsdf_batch(2, m, np.tile(_f32(wi), (n, 1)), u=rng.random((n, 3)).astype(np.float32))
    up = spdf > 0
    assert abs(cell_mass.sum() - up.mean()) < 4 * math.sqrt(0.25 / n) + 2e-3
    ci = np.minimum((wo[up, 2] * cells).astype(int), cells - 1)
    phi = np.mod(np.arctan2(wo[up, 1], wo[up, 0]), 2 * np.pi)
    cj = np.minimum((phi / (2 * np.pi) * cells).astype(int), cells - 1)
    obs = np.bincount(ci * cells + cj, minlength=cells * cells)
    exp = cell_mass.reshape(-1) * n
    keep = exp > 20
    chi2 = (((obs - exp) ** 2) / exp)[keep].sum()
    assert chi2 < 1.5 * keep.sum() + 40, chi2
    # the oracle's pdf is that density with the specular share times cos theta_h
    wos = wo[up][:2000]
    qq, ch = _beckmann_sampling_density(np.tile(wi, (len(wos), 1)), wos, a32, ks)
    diff = (1 - ks) * wos[:, 2] / math.pi
    mm = _mat(ng.MAT_MICROFACET, len(wos), albedo=kd, alpha=a, intIOR=1.5, extIOR=1.0)
    p = _pdf(mm, np.tile(wi, (len(wos), 1)), wos)
    assert np.allclose(p, (qq - diff) * ch + diff, rtol=2e-5)
    # and the sample weight is eval * cos / pdf at the sampled direction
    f = _eval(mm, np.tile(wi, (len(wos), 1)), wos)
    assert np.allclose(w[up][:2000], f * wos[:, 2:3] / p[:, None], rtol=2e-5)


# ------------------------------------------------------------------ Disney
def _disney(n=1, **kw):
    base = dict(albedo=0.5, roughness=0.5, metallic=0.0, specular=0.0, specularTint=0.0, sheen=0.0, sheenTint=0.0,
                subsurface=0.0, clearcoat=0.0, clearcoatGloss=0.0)
    base.update(kw)
    return _mat(ng.MAT_DISNEY, n, **base)


BASE = np.array([0.8, 0.4, 0.2])


def _f64(x):
    return np.asarray(x, np.float32).astype(np.float64)


def test_disney_burley_retro_reflection():
    # wi = wo at cos theta = c: cos theta_d = 1, F_D90 = 0.5 + 2 r, each Fresnel
    # factor 1 + (F_D90 - 1)(1 - c)^5 (Burley 2012, eq. 3).  r = 0.25 makes F_D90 = 1.
    for c, r in ((0.2, 0.8), (0.5, 0.25), (0.9, 0.05), (0.05, 1.0)):
        w = _dir(math.acos(c), 0.4)
        m = _disney(albedo=BASE, roughness=r)
        f = _eval(m, w, w)[0]
        c32 = float(np.float32(w[2]))
        fd90 = 0.5 + 2 * float(np.float32(r))
        fs = 1 + (fd90 - 1) * (1 - c32) ** 5
        assert np.allclose(f, _f64(BASE) / math.pi * fs * fs, rtol=1e-6), (c, r)


def test_disney_hanrahan_krueger_blend():
    # subsurface = s blends Burley with the Hanrahan-Krueger lobe (Burley 2012,
    # section 5.3): F_ss90 = r cos^2 theta_d, F_ss = 1 + (F_ss90 - 1)(1 - c)^5,
    # f_ss = 1.25 (F_ss(l) F_ss(v) (1/(cos l + cos v) - 0.5) + 0.5).
    # wi = wo at c = 0.5, r = 0.25: Burley = 1; F_ss = 1 - 0.75/32; HK = 1.25 (F_ss^2 / 2 + 0.5).
    w = _dir(math.pi / 3)
    fss = 1 - 0.75 / 32
    hk = 1.25 * (fss * fss * 0.5 + 0.5)
    for s in (0.0, 0.5, 1.0):
        f = _eval(_disney(albedo=BASE, roughness=0.25, subsurface=s), w, w)[0]
        assert np.allclose(f, _f64(BASE) / math.pi * ((1 - s) * 1.0 + s * hk), rtol=1e-6), s
    # r = 1, s = 1 at any c: F_ss90 = 1, so f_ss = 1.25 / (2c) = 0.625 / c
    for th in (0.3, 1.0, 1.4):
        w = _dir(th, 1.1)
        c = float(np.float32(w[2]))
        f = _eval(_disney(albedo=BASE, roughness=1.0, subsurface=1.0), w, w)[0]
        assert np.allclose(f, _f64(BASE) / math.pi * 0.625 / c, rtol=1e-6), th


def test_disney_sheen():
    tint = BASE / (LUM @ BASE)
    # cos theta_d = 1 (wi = wo): (1 - cos theta_d)^5 = 0, no sheen whatever its weight
    w = _dir(0.9, 0.2)
    f0 = _eval(_disney(albedo=BASE), w, w)[0]
    f1 = _eval(_disney(albedo=BASE, sheen=1.0, sheenTint=0.7), w, w)[0]
    assert np.array_equal(f0, f1)
    # wi, wo mirrored about n at 60 deg: wh = n, cos theta_d = 1/2, sheen term
    # = sheen (1/2)^5 lerp(1, base/lum(base), sheenTint), scaled by (1 - metallic)
    wi, wo = _dir(math.pi / 3, 0.0), _dir(math.pi / 3, math.pi)
    for sh, st, met in ((1.0, 0.0, 0.0), (0.6, 1.0, 0.0), (0.8, 0.5, 0.3)):
        kw = dict(albedo=BASE, roughness=0.5, metallic=met)
        d = _eval(_disney(sheen=sh, sheenTint=st, **kw), wi, wo)[0] - _eval(_disney(**kw), wi, wo)[0]
        cd = float(abs(np.dot(_f32(wo).astype(np.float64), [0, 0, 1.0])))
        expect = (1 - float(np.float32(met))) * float(np.float32(sh)) * (1 - cd) ** 5 * (
            1 + (tint - 1) * float(np.float32(st)))
        assert np.allclose(d, expect, rtol=1e-5, atol=1e-9), (sh, st, met, d, expect)


def test_disney_specular_tint_f0_at_normal_incidence():
    # wi = wo = n: cos theta_d = 1 so Burley = base/pi and sheen = 0; GGX D = 1/(pi a^2),
    # G = 1, F = F0 = lerp(0.08 specular lerp(1, base/lum, specularTint), base, metallic).
    n = [0, 0, 1.0]
    tint = BASE / (LUM @ BASE)
    for spec, st, met, r in ((0.5, 0.0, 0.0, 0.4), (1.0, 1.0, 0.0, 0.3), (0.7, 0.5, 0.0, 0.6), (0.5, 0.5, 1.0, 0.5),
                             (0.3, 0.2, 0.4, 0.7)):
        m = _disney(albedo=BASE, specular=spec, specularTint=st, metallic=met, roughness=r)
        f = _eval(m, n, n)[0]
        a = max(float(np.float32(r)) ** 2, 1e-4)
        spec64, st64, met64 = (float(np.float32(x)) for x in (spec, st, met))
        f0 = 0.08 * spec64 * (1 + (tint - 1) * st64)
        f0 = f0 + (_f64(BASE) - f0) * met64
        expect = (1 - met64) * _f64(BASE) / math.pi + f0 / (4 * math.pi * a * a)
        assert np.allclose(f, expect, rtol=1e-5), (spec, st, met, f, expect)


def test_disney_metallic_scales_diffuse_and_sheen():
    # Away from the specular peak (roughness 0.01 -> a = 1e-4, wh 51 deg off the
    # normal: GGX D ~ 1e-8), eval = (1 - metallic)(diffuse + sheen) to ~1e-6.
    wi, wo = _dir(math.pi / 3, 0.0), _dir(math.pi / 3, math.pi / 2)
    kw = dict(albedo=BASE, roughness=0.01, sheen=0.7, sheenTint=0.4, subsurface=0.3)
    f0 = _eval(_disney(metallic=0.0, **kw), wi, wo)[0]
    for met in (0.25, 0.5, 0.9):
        f = _eval(_disney(metallic=met, **kw), wi, wo)[0]
        assert np.allclose(f, (1 - float(np.float32(met))) * f0, rtol=1e-5), met


def test_disney_clearcoat_and_lobe_probabilities_at_normal_incidence():
    # GTR1 (Burley 2012, appendix B): D = (a^2 - 1) / (pi log(a^2) (1 + (a^2 - 1) cos^2)),
    # a = lerp(0.1, 0.001, gloss); at wh = n D = (a^2 - 1)/(pi log(a^2) a^2).
    # Clearcoat term = 0.25 cc D F G / (4 cos cos) with F(1) = 0.04 and G = 1 at n.
    n = [0, 0, 1.0]
    for cc, gloss in ((1.0, 0.0), (0.5, 0.7), (0.8, 1.0)):
        kw = dict(albedo=BASE, roughness=0.5, metallic=0.3)
        d = _eval(_disney(clearcoat=cc, clearcoatGloss=gloss, **kw), n, n)[0] - _eval(_disney(**kw), n, n)[0]
        g64 = float(np.float32(gloss))
        a = 0.1 + (0.001 - 0.1) * g64
        dn = (a * a - 1) / (math.pi * math.log(a * a) * a * a)
        assert np.allclose(d, 0.25 * float(np.float32(cc)) * dn * 0.04 / 4, rtol=1e-5), (cc, gloss)
        # pdf: lobe choice in proportion (1 - metallic, 1, 0.25 cc); each lobe's
        # pdf at wo = wi = n is D(n) cos / (4 cos) = D(n)/4 (diffuse: 1/pi)
        met, r = 0.3, 0.5
        m = _disney(clearcoat=cc, clearcoatGloss=gloss, **kw)
        p = _pdf(m, n, n)[0]
        wd, ws, wc = 1 - float(np.float32(met)), 1.0, 0.25 * float(np.float32(cc))
        ag = float(np.float32(r)) ** 2
        expect = (wd / math.pi + ws / (4 * math.pi * ag * ag) + wc * dn / 4) / (wd + ws + wc)
        assert abs(p - expect) <= 1e-5 * expect, (cc, gloss, p, expect)


def test_gtr1_normalised():
    # integral of D_GTR1 cos over the hemisphere = 1, through the clearcoat-only
    # pdf at normal incidence: wi = n, wo = reflect about wh, pdf = D(wh) cos_h/(4 wh.wo).
    for gloss in (0.0, 0.5, 1.0):
        m = _disney(albedo=1.0, metallic=1.0, clearcoat=1.0, clearcoatGloss=gloss, roughness=1.0)
        # mixture pdf = (ps * GGX + pc * GTR1) with ps = 1/1.25, pc = 0.25/1.25 (metallic 1)
        th = np.linspace(0, math.pi / 2, 20001)[:-1] + math.pi / 4 / 20000
        wo = np.stack([np.sin(th), np.zeros_like(th), np.cos(th)], -1)
        p = _pdf(np.repeat(m, len(th)), np.tile([0, 0, 1.0], (len(th), 1)), wo)
        total = np.sum(p * np.sin(th)) * (th[1] - th[0]) * 2 * math.pi
        # both lobes integrate to 1 over the reflections above the horizon up to the
        # mass of half vectors beyond 45 deg, which GGX at a = 1 leaves out: 1/2 of its mass
        m2 = _disney(albedo=1.0, metallic=1.0, clearcoat=0.0, roughness=1.0)
        p2 = _pdf(np.repeat(m2, len(th)), np.tile([0, 0, 1.0], (len(th), 1)), wo)
        ggx_mass = np.sum(p2 * np.sin(th)) * (th[1] - th[0]) * 2 * math.pi
        assert abs(ggx_mass - 0.5) < 1e-3
        a = 0.1 + (0.001 - 0.1) * float(np.float32(gloss))
        # GTR1 mass of half vectors within 45 deg: closed form of the GTR1 CDF,
        # P(cos^2 >= 1/2) = 1 - log(1 + (a^2 - 1)/2) / log(a^2) ... = log((1 + a^2)/2)/log(a^2) complement
        gtr1_mass = 1 - math.log((1 + a * a) / 2) / math.log(a * a)
        assert abs(total - (0.5 / 1.25 + 0.25 * gtr1_mass / 1.25)) < 2e-3, (gloss, total)


# ------------------------------------------------------------------ thin lens
def test_thin_lens_focus_and_disk():
    # src/cpu/perspective.cpp:101-112: the lens point is R sqrt(u) (cos 2 pi v,
    # sin 2 pi v, 0) in camera space and every ray of one image position passes
    # through the pinhole ray's point at camera-space depth focalDistance.
    R, F = 0.05, 5.0
    # eye near the world origin: FP32 origins then carry ~3e-8 of rounding
    eye, target, up = np.array([0.1, 0.2, -0.3]), np.array([0.5, 1.0, 4.0]), np.array([0.0, 1.0, 0.0])
    cam = ng.lookat_camera(eye, target, up, 40.0, 64, 48, lens_radius=R, focal_distance=F)
    pin = ng.lookat_camera(eye, target, up, 40.0, 64, 48)
    M = np.array(cam.to_world, np.float64).reshape(4, 4)
    ex, ey, ez = M[:3, 0], M[:3, 1], M[:3, 2]
    rng = np.random.default_rng(5)
    for px, py in ((32.0, 24.0), (0.5, 0.5), (60.3, 7.9)):
        lens = rng.random((500, 2))
        smp = np.c_[np.full(500, px), np.full(500, py), lens].astype(np.float32)
        rays = orc.camera_rays(cam, smp).astype(np.float64)
        o, d = rays[:, :3], rays[:, 4:7]  # ray record {o, mint, d, maxt}
        rel = o - eye
        # origins in the lens plane, at radius R sqrt(u) and angle 2 pi v
        assert np.all(np.abs(rel @ ez) < 1e-6)
        assert np.allclose(np.sum(rel * rel, 1), R * R * lens[:, 0].astype(np.float32), rtol=1e-4, atol=1e-8)
        ang = np.arctan2(rel @ ey, rel @ ex)
        dang = np.angle(np.exp(1j * (ang - 2 * np.pi * lens[:, 1].astype(np.float32))))
        big = np.sum(rel * rel, 1) > 1e-2 * R * R
        assert np.all(np.abs(dang[big]) < 1e-4)
        # all rays meet the pinhole ray at depth F
        t = (F - rel @ ez) / (d @ ez)
        pts = o + t[:, None] * d
        pr = orc.camera_rays(pin, np.array([[px, py, 0.5, 0.5]], np.float32)).astype(np.float64)[0]
        tp = F / (pr[4:7] @ ez)
        p0 = pr[:3] + tp * pr[4:7]
        assert np.max(np.linalg.norm(pts - p0, axis=1)) < 2e-5 * F
    # lens points uniform on the disk: ring shares proportional to ring areas
    lens = rng.random((200000, 2))
    smp = np.c_[np.full(len(lens), 32.0), np.full(len(lens), 24.0), lens].astype(np.float32)
    rays = orc.camera_rays(cam, smp).astype(np.float64)
    r2 = np.sum((rays[:, :3] - eye) ** 2, 1) / (R * R)
    counts = np.histogram(r2, bins=10, range=(0, 1))[0]
    assert np.all(np.abs(counts / len(lens) - 0.1) < 4 * math.sqrt(0.09 / len(lens)))
