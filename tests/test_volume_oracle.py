"""CPU checks that pin the oracle's participating media (shaders/Volume.hlsl
restated): Henyey-Greenstein normalisation, sampling density and mean cosine;
free-flight and ratio-tracking estimators against closed-form transmittance of
analytic densities; conservative brick majorants; and the Beer-Lambert and
white-furnace renders."""
import math

import numpy as np
import pytest
from scipy import integrate, stats

import nori_gpu as ng
import orc


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.3, 0.8, 0.95])
def test_hg_normalised(g):
    val = integrate.quad(lambda mu: 2 * math.pi * orc.hg_eval(mu, g)[0], -1, 1, limit=200, epsabs=1e-10)[0]
    assert abs(val - 1.0) < 2e-5


@pytest.mark.parametrize("g", [-0.4, 0.0, 0.3, 0.8])
def test_hg_sampling(g):
    # mean cosine of HG is g; the sampled cosine follows the HG marginal
    # 2 pi HG(mu) on [-1, 1]; the returned pdf is HG at the sampled cosine
    rng = np.random.default_rng(7)
    n = 200_000
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    d, pdf = orc.hg_sample(wo, g, rng.random(n), rng.random(n))
    mu = (d * wo).sum(1)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-5)
    assert abs(mu.mean() - g) < 4 * mu.std() / math.sqrt(n) + 1e-4
    assert np.allclose(pdf, orc.hg_eval(mu, g), rtol=1e-3, atol=1e-6)
    edges = np.linspace(-1, 1, 41)
    cdf = np.array([integrate.quad(lambda m: 2 * math.pi * orc.hg_eval(m, g)[0], -1, e)[0] for e in edges])
    expect = np.diff(cdf) * n
    obs = np.histogram(mu, edges)[0]
    keep = expect > 5
    chi2 = ((obs[keep] - expect[keep]) ** 2 / expect[keep]).sum()
    assert stats.chi2.sf(chi2, keep.sum() - 1) > 1e-4


def _box_scene(media, env=1.0):
    # a far, black triangle keeps the BVH non-empty; the rest is the medium
    pos = np.array([[100, 100, 100], [101, 100, 100], [100, 101, 100]], np.float32)
    mats = ng.default_materials(1)
    mats["albedo"] = 0.0
    env_rgb = np.full((8, 16, 3), env, np.float32)
    return ng.HostScene(pos, np.array([[0, 1, 2]], np.uint32), np.array([0], np.uint32), np.array([1], np.uint32),
                        mats, env_rgb=env_rgb, media=media)


def _ramp_grid(n):
    # density (i + 0.5) / n along x: trilinear lookup returns d(u) = u on
    # [0.5/n, 1 - 0.5/n] and clamps to the end values outside
    x = (np.arange(n, dtype=np.float32) + 0.5) / n
    return np.broadcast_to(x[None, None, :], (n, n, n)).copy()


def _ramp_optical_depth(u0, u1, n, sigma_t, size):
    lo, hi = 0.5 / n, 1 - 0.5 / n
    f = lambda u: lo if u < lo else (hi if u > hi else u)
    return sigma_t * size * integrate.quad(f, u0, u1, points=[lo, hi])[0]


def test_ratio_tracking_matches_analytic_transmittance():
    n = 16
    m = ng.MediumSpec([0, 0, 0], [2, 2, 2], sigma_a=[0.5, 1.0, 1.5], sigma_s=[0.5, 1.0, 1.5], density=_ramp_grid(n),
                      brick=4)
    o = orc.OracleScene(_box_scene([m]))
    N = 40_000
    orig = np.tile([[0.0, 1.0, 1.0]], (N, 1))
    d = np.tile([[1.0, 0.0, 0.0]], (N, 1))
    tr, _, _ = o.volume_query(1, 0, orig, d, 0.1, 1.7, seeds=np.arange(N, dtype=np.uint64) + 11)
    for c, st in enumerate([1.0, 2.0, 3.0]):
        expect = math.exp(-_ramp_optical_depth(0.05, 0.85, n, st, 2.0))
        se = tr[:, c].std() / math.sqrt(N)
        assert abs(tr[:, c].mean() - expect) < 4 * se + 1e-4, (c, tr[:, c].mean(), expect)


def test_free_flight_pass_probability_and_weights():
    # Null-collision free flight with the hero channel (max sigma_t): a path
    # passes the segment with the hero transmittance; grey media keep weight 1
    # on passing, scatter weights sigma_t,c / sigma_t,max.
    n = 16
    grid = _ramp_grid(n)
    m = ng.MediumSpec([0, 0, 0], [2, 2, 2], sigma_a=0.2, sigma_s=[0.8, 0.8, 0.8], density=grid, brick=4)
    o = orc.OracleScene(_box_scene([m]))
    N = 40_000
    orig = np.tile([[0.0, 1.0, 1.0]], (N, 1))
    d = np.tile([[1.0, 0.0, 0.0]], (N, 1))
    w, t, sc = o.volume_query(0, 0, orig, d, 0.2, 1.9, seeds=np.arange(N, dtype=np.uint64) * 3 + 1)
    p_pass = math.exp(-_ramp_optical_depth(0.1, 0.95, n, 1.0, 2.0))
    frac = 1 - sc.mean()
    assert abs(frac - p_pass) < 4 * math.sqrt(p_pass * (1 - p_pass) / N)
    assert np.allclose(w[sc == 0], 1.0, atol=1e-6)
    assert np.allclose(w[sc == 1], 1.0, atol=1e-6)  # grey: sigma_t,c / sigma_t,max = 1
    assert ((t[sc == 1] > 0.2) & (t[sc == 1] < 1.9)).all()


def test_chromatic_free_flight_is_unbiased():
    # E[weight * 1(pass)] = per-channel transmittance even though distances are
    # sampled with the hero channel only
    m = ng.MediumSpec([0, 0, 0], [2, 2, 2], sigma_a=[0.1, 0.6, 1.2], sigma_s=[0.4, 0.4, 0.4], density=_ramp_grid(8),
                      brick=2)
    o = orc.OracleScene(_box_scene([m]))
    N = 60_000
    orig = np.tile([[0.0, 1.0, 1.0]], (N, 1))
    d = np.tile([[1.0, 0.0, 0.0]], (N, 1))
    w, t, sc = o.volume_query(0, 0, orig, d, 0.0, 2.0, seeds=np.arange(N, dtype=np.uint64) + 5)
    passed = w * (sc == 0)[:, None]
    for c, st in enumerate([0.5, 1.0, 1.6]):
        expect = math.exp(-_ramp_optical_depth(0.0, 1.0, 8, st, 2.0))
        se = passed[:, c].std() / math.sqrt(N)
        assert abs(passed[:, c].mean() - expect) < 4 * se + 1e-4, (c, passed[:, c].mean(), expect)


def test_majorant_bounds_density():
    rng = np.random.default_rng(3)
    dens = rng.random((20, 24, 28)).astype(np.float32) ** 3
    m = ng.MediumSpec([-1, 0, 2], [1, 3, 4], sigma_a=1.0, sigma_s=2.0, density=dens, brick=5)
    o = orc.OracleScene(_box_scene([m]))
    N = 100_000
    p = rng.uniform([-1, 0, 2], [1, 3, 4], (N, 3)).astype(np.float32)
    dirs = rng.normal(size=(N, 3))
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    _, dv, _ = o.volume_query(2, 0, p)
    mb, _, _ = o.volume_query(3, 0, p, dirs)
    assert (dv * 3.0 <= mb[:, 0] * (1 + 1e-6) + 1e-7).all()
    # the majorant is tight to within the brick and its one-voxel apron
    assert np.median(mb[:, 0] / 3.0) < 1.0
    # the boundary distance is positive and no longer than the brick diagonal
    # (bricks of 2/6 x 3/5 x 2/4 world units: 28 x 24 x 20 voxels in bricks of 5)
    diag = math.sqrt((2 / 6) ** 2 + (3 / 5) ** 2 + (2 / 4) ** 2)
    assert (mb[:, 1] > 0).all() and (mb[:, 1] <= diag + 1e-4).all()


def test_beer_lambert_render():
    # pure absorber between the camera and a unit envmap: every pixel sees
    # exp(-sigma_a * L) = exp(-2) = 0.1353
    m = ng.MediumSpec([-5, -5, -0.5], [5, 5, 0.5], sigma_a=2.0, sigma_s=0.0)
    sc = _box_scene([m])
    cam = ng.lookat_camera([0, 0, -3], [0, 0, 3], [0, 1, 0], 10, 16, 16)
    img, _ = orc.OracleScene(sc).render(cam, 4096, 8, seed=1)
    L = img[:, :3] / img[:, 3:4]
    # camera rays cross the slab at up to 5 degrees: path length 1/cos
    ys, xs = np.mgrid[0:16, 0:16]
    th = math.tan(math.radians(5))
    dx, dy = (1 - 2 * (xs + 0.5) / 16) * th, (1 - 2 * (ys + 0.5) / 16) * th
    expect = np.exp(-2.0 * np.sqrt(1 + dx ** 2 + dy ** 2)).reshape(-1)
    assert abs(L.mean() - expect.mean()) < 3e-3, (L.mean(), expect.mean())


def test_volume_white_furnace_render():
    # non-absorbing heterogeneous medium under a unit envmap: every path ends
    # in the envmap with weight 1 on average, so every pixel converges to 1
    rng = np.random.default_rng(5)
    dens = rng.random((8, 8, 8)).astype(np.float32)
    m = ng.MediumSpec([-1, -1, -1], [1, 1, 1], sigma_a=0.0, sigma_s=[2.0, 2.0, 2.0], g=0.3, density=dens, brick=4)
    sc = _box_scene([m])
    cam = ng.lookat_camera([0, 0, -4], [0, 0, 0], [0, 1, 0], 35, 12, 12)
    img, _ = orc.OracleScene(sc).render(cam, 2048, 64, seed=2)
    L = img[:, :3] / img[:, 3:4]
    assert abs(L.mean() - 1.0) < 1e-2, L.mean()
    assert np.abs(L - 1.0).max() < 0.15
