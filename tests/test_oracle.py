"""CPU checks that pin the oracle before it is trusted as the GPU path's checker.

Each test compares the oracle against something independent of its own code:
published known answers, closed forms, brute force, or invariants that a
plausible mistake (dropped term, wrong index, transposed operand) would break.
"""
import math

import numpy as np
import pytest

import nori_gpu as ng
import orc

MISS = 0xFFFFFFFF


# ------------------------------------------------------------------ pcg32
def test_pcg32_known_answer():
    # Output of the PCG reference demo for seed(42, 54) (O'Neill, pcg-c-basic pcg32-demo).
    got = orc.pcg32(42, 54, 6)
    assert [int(x) for x in got] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


@pytest.mark.parametrize("k", [1, 7, 1000, 123457])
def test_pcg32_advance_equals_skipping(k):
    full = orc.pcg32(0xDEADBEEF, 0xda3e39cb94b95bdb, k + 16)
    jumped = orc.pcg32(0xDEADBEEF, 0xda3e39cb94b95bdb, 16, advance=k)
    assert np.array_equal(full[k:], jumped)


def test_sample_windows_follow_pixel_stream():
    # sample s of a pixel = the pixel's single stream advanced by s * 2^20 draws
    seed, pixel = 5, 1234
    base = orc.pcg32(seed ^ pixel, 0xda3e39cb94b95bdb, 8, advance=3 << 20)
    f = orc.sample_floats(seed, pixel, 3, 8)
    expect = ((base >> 9) | 0x3f800000).view(np.float32) - 1.0
    assert np.array_equal(f, expect)
    assert f.min() >= 0.0 and f.max() < 1.0


# ------------------------------------------------------------------ triangles / BVH
def _tri_scene(p0, p1, p2):
    mats = ng.default_materials(1)
    return ng.HostScene(np.array([p0, p1, p2], np.float32), np.array([[0, 1, 2]]), [0], [1], mats)


def test_moller_trumbore_closed_form():
    # Triangle in the plane z = 2; a ray from the origin along +z hits at t = 2 with
    # barycentrics solving o + t d = p0 + u e1 + v e2.
    s = orc.OracleScene(_tri_scene([-1, -1, 2], [3, -1, 2], [-1, 3, 2]))
    rays = np.array([[0, 0, 0, 1e-4, 0, 0, 1, 1e4],        # interior: u = v = 0.25
                     [2.9, -0.9, 0, 1e-4, 0, 0, 1, 1e4],   # near the p1 corner
                     [2, 2, 0, 1e-4, 0, 0, 1, 1e4],        # outside (u + v > 1)
                     [0, 0, 0, 1e-4, 0, 0, 1, 1.5],        # beyond maxt
                     [0, 0, 3, 1e-4, 0, 0, 1, 1e4],        # triangle behind the origin
                     [0, 0, 0, 1e-4, 1, 0, 0, 1e4]], np.float32)  # parallel: det = 0
    t, prim, u, v = s.trace(rays)
    assert prim[0] == 0 and t[0] == 2.0 and u[0] == 0.25 and v[0] == 0.25
    assert prim[1] == 0 and abs(u[1] - 0.975) < 1e-6 and abs(v[1] - 0.025) < 1e-6
    assert list(prim[2:]) == [MISS] * 4


def test_bvh2_matches_brute_force(cornell):
    s = orc.OracleScene(cornell)
    rng = np.random.default_rng(7)
    n = 20000
    o = rng.uniform([-0.95, 0.05, -0.95], [0.95, 1.95, 3.0], (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.c_[o, np.full(n, 1e-4), d, np.full(n, 1e4)].astype(np.float32)
    a = s.trace(rays)
    b = s.trace(rays, brute=True)
    assert np.array_equal(a[1], b[1])
    hit = a[1] != MISS
    assert hit.mean() > 0.5
    for x, y in zip(a, b):
        assert np.array_equal(x[hit], y[hit])
    # occlusion query agrees with the closest-hit query
    anyh = s.trace(rays, any_hit=True)[1] != MISS
    assert np.array_equal(anyh, hit)


def test_robust_bvh2_matches_brute_force_on_ties(disney_small):
    # Camera rays over dense icospheres hit shared edges at equal t; the widened
    # box test keeps the canonical (smallest id) hit that brute force finds.
    s = orc.OracleScene(disney_small)
    cam = disney_small.camera
    cam.width, cam.height = 96, 96
    rng = np.random.default_rng(8)
    n = 96 * 96
    smp = np.c_[rng.random(n) * 96, rng.random(n) * 96, rng.random(n), rng.random(n)].astype(np.float32)
    rays = orc.camera_rays(cam, smp)
    a, b = s.trace(rays), s.trace(rays, brute=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_closest_hit_tie_break_smallest_prim():
    # Two coincident triangles: the canonical hit is the smaller primitive id.
    mats = ng.default_materials(1)
    pos = np.array([[-1, -1, 2], [3, -1, 2], [-1, 3, 2]] * 2, np.float32)
    sc = ng.HostScene(pos, np.array([[3, 4, 5], [0, 1, 2]]), [0], [2], mats)
    s = orc.OracleScene(sc)
    t, prim, _, _ = s.trace(np.array([[0, 0, 0, 1e-4, 0, 0, 1, 1e4]], np.float32))
    assert prim[0] == 0 and t[0] == 2.0


# ------------------------------------------------------------------ camera
def test_camera_closed_form(cornell):
    cam = cornell.camera
    cam.width, cam.height = 64, 48
    smp = np.array([[32, 24, 0.5, 0.5], [0, 24, 0.5, 0.5], [64, 24, 0.5, 0.5], [32, 0, 0.5, 0.5]], np.float32)
    r = orc.camera_rays(cam, smp)
    d = r[:, 4:7] / np.linalg.norm(r[:, 4:7], axis=1, keepdims=True)
    fwd = np.array([0, 0, -1.0])
    assert np.allclose(d[0], fwd, atol=1e-6)
    half = math.radians(cam.fov_deg) / 2  # Nori's fov is horizontal (src/cpu/perspective.cpp:34-35)
    assert abs(math.acos(d[1] @ fwd) - half) < 1e-5 and abs(math.acos(d[2] @ fwd) - half) < 1e-5
    assert d[1][0] < 0 < d[2][0]  # image x grows to the right (world +x for this camera)
    vhalf = math.atan(math.tan(half) * 48 / 64)
    assert abs(math.acos(d[3] @ fwd) - vhalf) < 1e-5 and d[3][1] > 0  # row 0 is the top
    assert np.allclose(r[:, :3], [0, 1, 3.7], atol=1e-6)


# ------------------------------------------------------------------ envmap
@pytest.fixture(scope="module")
def env_scene(disney_small):
    return orc.OracleScene(disney_small)


def test_env_cdf_tables(env_scene, disney_small):
    marg, cond, ppix, prob, alias = env_scene.env_tables()
    H, W = disney_small.env_rgb.shape[:2]
    assert marg[0] == 0 and marg[-1] == 1.0 and np.all(np.diff(marg) >= 0)
    # rows are scaled by 1/rowSum in float (src/gpu/DXRApp_Textures.cpp:324-329): the last entry is 1 to an ulp
    assert np.all(cond[:, 0] == 0) and np.allclose(cond[:, -1], 1.0, atol=2.5e-7) and np.all(np.diff(cond, axis=1) >= 0)
    # luminance * sin(theta) weights reproduce the per-texel probability
    lum = disney_small.env_rgb @ np.array([0.2126, 0.7152, 0.0722])
    w = lum * np.sin(np.pi * (np.arange(H) + 0.5) / H)[:, None]
    assert np.allclose(ppix.reshape(H, W), w / w.sum(), rtol=2e-4, atol=1e-12)


def test_alias_table_reproduces_distribution():
    rng = np.random.default_rng(3)
    w = rng.exponential(size=5000).astype(np.float32)
    w[::17] = 0.0
    prob, alias = orc.alias_build(w)
    n = len(w)
    recon = (prob.astype(np.float64) + np.bincount(alias, weights=1.0 - prob.astype(np.float64), minlength=n)) / n
    target = w.astype(np.float64) / w.astype(np.float64).sum()
    assert np.abs(recon - target).max() < 1e-6 / n * 10
    assert np.all(recon[::17] < 1e-9)


def test_env_pdf_integrates_to_one(env_scene):
    # midpoint quadrature over (theta, phi): sum pdf * sin(theta) dtheta dphi
    nt, nphi = 512, 1024
    th = (np.arange(nt) + 0.5) / nt * np.pi
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    _, pdf = env_scene.env_eval(dirs)
    integral = (pdf.reshape(nt, nphi) * np.sin(T)).sum() * (np.pi / nt) * (2 * np.pi / nphi)
    assert abs(integral - 1.0) < 2e-3


def test_env_samplers_match_pdf(env_scene):
    rng = np.random.default_rng(11)
    n = 20000
    r = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    coin, jx, jy = (rng.random(n).astype(np.float32) * 0.98 + 0.01 for _ in range(3))
    for mode in (0, 1):
        d, L, pdf, _ = env_scene.env_sample(r, coin, jx, jy, mode=mode)
        L2, pdf2 = env_scene.env_eval(d)
        ok = pdf > 0
        assert np.allclose(pdf[ok], pdf2[ok], rtol=1e-4)
        assert np.allclose(L[ok], L2[ok], rtol=1e-4, atol=1e-5)
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)


def test_env_alias_draw_frequencies(env_scene):
    # chi-square of alias texel draws against ppix (coarsened to 64 bins by rank)
    _, _, ppix, _, _ = env_scene.env_tables()
    rng = np.random.default_rng(5)
    n = 400000
    r = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    coin = rng.random(n).astype(np.float32)
    _, _, _, tex = env_scene.env_sample(r, coin, coin, coin, mode=0)
    order = np.argsort(ppix)
    groups = np.array_split(order, 64)
    exp = np.array([ppix[g].sum() for g in groups]) * n / ppix.sum()
    cnt = np.bincount(tex, minlength=len(ppix))
    obs = np.array([cnt[g].sum() for g in groups])
    keep = exp > 5
    chi2 = (((obs - exp) ** 2) / exp)[keep].sum()
    assert chi2 < 2.0 * keep.sum() + 30


# ------------------------------------------------------------------ BSDFs
def _mats(n, rng, **fixed):
    m = ng.default_materials(n)
    m["type"] = ng.MAT_DISNEY
    m["albedo"] = rng.random((n, 3))
    for k in ("metallic", "specular", "sheen", "clearcoat", "specularTint", "sheenTint", "clearcoatGloss",
              "subsurface"):
        m[k] = rng.random(n)
    m["roughness"] = 0.05 + 0.95 * rng.random(n)
    for k, v in fixed.items():
        m[k] = v
    return m


def _hemi(n, rng):
    v = rng.normal(size=(n, 3))
    v[:, 2] = np.abs(v[:, 2]) + 1e-3
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


def test_lambert_closed_form():
    rng = np.random.default_rng(1)
    n = 1000
    m = ng.default_materials(n)
    m["albedo"] = rng.random((n, 3))
    wi, wo = _hemi(n, rng), _hemi(n, rng)
    f, _, _ = orc.bsdf_batch(0, m, wi, wo)
    _, pdf, _ = orc.bsdf_batch(1, m, wi, wo)
    assert np.allclose(f, m["albedo"] / np.pi, rtol=1e-7)
    assert np.allclose(pdf, wo[:, 2] / np.pi, rtol=1e-7)


def test_disney_ggx_normal_incidence_closed_form():
    # metallic = 1, white base, no clearcoat: f = D*G/(4 cos cos) with F = 1; at
    # wi = wo = n: D = 1/(pi a^2), G = 1, so f = 1/(4 pi a^2), a = max(r^2, 1e-4).
    r = np.array([0.1, 0.3, 0.5, 0.9, 1.0])
    m = ng.default_materials(len(r))
    m["type"], m["albedo"], m["metallic"], m["clearcoat"], m["roughness"] = ng.MAT_DISNEY, 1.0, 1.0, 0.0, r
    z = np.tile([0, 0, 1.0], (len(r), 1)).astype(np.float32)
    f, _, _ = orc.bsdf_batch(0, m, z, z)
    a = np.maximum(r.astype(np.float32).astype(np.float64) ** 2, 1e-4)
    assert np.allclose(f[:, 0], 1.0 / (4 * np.pi * a * a), rtol=1e-6)
    # pdf at normal incidence: pSpec * D * cos/(4 cos) = 0.5 * 1/(4 pi a^2) (pDiffuse = 0, pSpec = 1 with cc = 0)
    _, pdf, _ = orc.bsdf_batch(1, m, z, z)
    assert np.allclose(pdf, 1.0 / (4 * np.pi * a * a), rtol=1e-6)


def test_disney_pdf_is_the_sampling_density():
    # The pdf must be the density the sampler actually draws from: per cell of a
    # (cos theta, phi) grid over the upper hemisphere, the quadrature of pdf()
    # must predict the share of sample() draws landing there (chi-square), and
    # the total must equal the share of draws above the horizon.
    rng = np.random.default_rng(2)
    nt, nphi, cells = 480, 480, 12
    ct = (np.arange(nt) + 0.5) / nt
    ph = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    C, P = np.meshgrid(ct, ph, indexing="ij")
    S = np.sqrt(1 - C * C)
    grid = np.stack([S * np.cos(P), S * np.sin(P), C], -1).reshape(-1, 3).astype(np.float32)
    for k in range(3):
        m1 = _mats(1, rng, roughness=0.35 + 0.2 * k)
        wi = _hemi(1, rng)
        _, pdf, _ = orc.bsdf_batch(1, np.repeat(m1, len(grid)), np.repeat(wi, len(grid), 0), grid)
        mass = pdf.reshape(nt, nphi) * (1.0 / nt) * (2 * np.pi / nphi)
        cell_mass = mass.reshape(cells, nt // cells, cells, nphi // cells).sum(axis=(1, 3))
        n = 200000
        w, spdf, wo = orc.bsdf_batch(2, np.repeat(m1, n), np.repeat(wi, n, 0), u=rng.random((n, 3)).astype(np.float32))
        up = spdf > 0
        assert abs(cell_mass.sum() - up.mean()) < 4 * np.sqrt(0.25 / n) + 2e-3
        ci = np.minimum((wo[up, 2] * cells).astype(int), cells - 1)
        phi = np.mod(np.arctan2(wo[up, 1], wo[up, 0]), 2 * np.pi)
        cj = np.minimum((phi / (2 * np.pi) * cells).astype(int), cells - 1)
        obs = np.bincount(ci * cells + cj, minlength=cells * cells)
        exp = cell_mass.reshape(-1) * n
        keep = exp > 20
        chi2 = (((obs - exp) ** 2) / exp)[keep].sum()
        assert chi2 < 1.5 * keep.sum() + 40, (k, chi2, keep.sum())


def test_disney_sample_consistent_with_eval_pdf():
    rng = np.random.default_rng(4)
    n = 20000
    m = _mats(n, rng)
    wi = _hemi(n, rng)
    u = rng.random((n, 3)).astype(np.float32)
    w, pdf, wo = orc.bsdf_batch(2, m, wi, u=u)
    ok = pdf > 0
    assert ok.mean() > 0.8
    wo32 = wo.astype(np.float32)
    f, _, _ = orc.bsdf_batch(0, m, wi, wo32)
    _, p2, _ = orc.bsdf_batch(1, m, wi, wo32)
    # re-evaluated at the float32-rounded direction, so agreement is to that rounding
    rel = np.abs(p2[ok] - pdf[ok]) / pdf[ok]
    assert np.quantile(rel, 0.95) < 1e-5 and rel.max() < 1e-2
    wr = np.abs(w[ok] - f[ok] * wo32[ok, 2:3] / p2[ok, None]) / np.maximum(w[ok], 1e-12)
    assert np.quantile(wr, 0.95) < 1e-5 and wr.max() < 1e-2
    assert np.all(f >= 0) and np.all(w >= 0)


def test_fp32_oracle_close_to_fp64():
    rng = np.random.default_rng(6)
    n = 20000
    m = _mats(n, rng, roughness=np.clip(rng.random(n), 0.2, 1))
    wi, wo = _hemi(n, rng), _hemi(n, rng)
    f64, _, _ = orc.bsdf_batch(0, m, wi, wo, precision=64)
    f32, _, _ = orc.bsdf_batch(0, m, wi, wo, precision=32)
    rel = np.abs(f32 - f64) / np.maximum(np.abs(f64), 1e-6)
    assert np.quantile(rel, 0.99) < 1e-4


# ------------------------------------------------------------------ integrator
def test_white_furnace_lambert_sphere():
    # A convex Lambertian sphere of albedo rho under a uniform unit envmap reflects
    # exactly rho (every scattered ray escapes); MIS of envmap NEE and BSDF
    # sampling must keep that estimate unbiased.
    pos, tri = [], []
    import ctypes  # noqa: F401  (keeps the binding's structures importable in isolation)
    sc1 = ng.generate_scene(1, subdiv=3)
    sphere_mesh = 5  # first icosphere of the Cornell generator
    o, c = int(sc1.mesh_tri_offset[sphere_mesh]), int(sc1.mesh_tri_count[sphere_mesh])
    tri = sc1.indices[o:o + c]
    used = np.unique(tri)
    remap = np.zeros(tri.max() + 1, np.uint32)
    remap[used] = np.arange(len(used))
    pos = sc1.positions[used]
    nrm = sc1.normals[used]
    ctr = pos.mean(0)
    rho = 0.6
    m = ng.default_materials(1)
    m["albedo"] = rho
    env = np.ones((32, 64, 3), np.float32)
    scene = ng.HostScene(pos, remap[tri], [0], [c], m, normals=nrm, env_rgb=env)
    cam = ng.lookat_camera(ctr + np.array([0, 0, 2.5]), ctr, [0, 1, 0], 20.0, 32, 32)
    acc, _ = orc.OracleScene(scene).render(cam, 64, 4)
    img = acc[:, :3] / acc[:, 3:]
    centre = img.reshape(32, 32, 3)[10:22, 10:22].reshape(-1, 3)
    assert abs(centre.mean() - rho) < 0.01
    # the background is the unit environment, seen directly
    assert np.allclose(img.reshape(32, 32, 3)[0, 0], 1.0, atol=1e-6)


def test_rel_mse_definition():
    a = np.array([1.0, 2.0, 0.0])
    b = np.array([1.5, 2.0, 0.1])
    expect = np.mean(np.array([0.25 / (2.25 + 1e-2), 0.0, 0.01 / (0.01 + 1e-2)]))
    assert abs(orc.rel_mse(a, b) - expect) < 1e-15
