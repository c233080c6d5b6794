"""Parity at the workloads' real sizes, where the oracle finishes in seconds to a
minute on the host: the Cornell box at 256^2 x 64 spp (same samples as the
oracle), closest hits on the full 5,242,882-triangle Disney grid (bit-exact
{prim, t, u, v} on 2^20 camera rays and 2^18 random rays), alias draws on its
2048x1024 envmap, the 50,000-strand hair white furnace, and the full batched
scene (10,000,448 triangles, 128 1024^2 textures, 20,000 alpha-tested cards,
thin lens): bit-exact closest hits on 2^20 camera rays and same-sample renders
of two of its sixteen views."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from test_gpu_parity import MISS, _camera_rays, _check_hits, _random_rays

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_cornell_full_size_same_samples(gpu):
    # 256 x 256 pixels, 64 spp, depth 8, RR from 3: the workload as specified
    sc = ng.generate_scene(1)
    cam = sc.camera
    assert (cam.width, cam.height) == (256, 256)
    _say("case: cornell 256x256 64spp depth 8, GPU and oracle")
    gpu.upload(sc)
    gpu.reset_film()
    g = gpu.render(ng.make_view(cam, 64, 8), seed=11)
    o, _ = orc.OracleScene(sc).render(cam, 64, 8, seed=11)
    assert np.all(g[:, 3] == 64) and np.array_equal(g[:, 3], o[:, 3])
    gi, oi = g[:, :3] / 64, o[:, :3] / 64
    e = orc.rel_mse(gi, oi)
    close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-5, axis=1)
    _say(f"  relMSE {e:.3e}, pixels within 1e-3: {close.mean():.5f}")
    assert e < 1e-5
    assert close.mean() > 0.97


@pytest.fixture(scope="module")
def disney_full():
    sc = ng.generate_scene(2)
    assert len(sc.indices) == 5_242_882 and sc.env_rgb.shape[:2] == (1024, 2048)
    return sc


def test_disney_grid_full_scene_hits_bit_exact(gpu, disney_full):
    gpu.upload(disney_full)
    _say("case: oracle BVH2 over 5,242,882 triangles")
    o = orc.OracleScene(disney_full)
    for rays, label in ((_camera_rays(disney_full, 1024, 1024, 21), "camera 1024x1024"),
                        (_random_rays([-6, 0.01, -6], [6, 1.5, 6], 1 << 18, 22), "random 2^18")):
        _say(f"case: full disney grid trace {label}")
        gt, gp, gu, gv = gpu.trace(rays)
        ot, op, ou, ov = o.trace(rays)
        bad = np.flatnonzero(gp != op)
        assert len(bad) == 0, f"{label}: {len(bad)} primitive mismatches, first {bad[:5]}"
        hit = op != MISS
        for a, b in ((gt, ot), (gu, ou), (gv, ov)):
            assert np.array_equal(a[hit].view(np.uint32), b[hit].view(np.uint32)), label
        _, ap, _, _ = gpu.trace(rays, any_hit=True)
        assert np.array_equal(ap != MISS, hit), label
        _say(f"  {label}: {hit.mean():.4f} hit, all bit-exact")


def test_envmap_alias_full_size_bit_exact(gpu, disney_full):
    gpu.upload(disney_full)
    o = orc.OracleScene(disney_full)
    rng = np.random.default_rng(23)
    n = 1 << 20
    r = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    coin, jx, jy = (rng.random(n).astype(np.float32) for _ in range(3))
    _say("case: alias draws on the 2048x1024 sky+sun map")
    gd, gL, gpdf, gtex = gpu.envmap_sample(r, coin, jx, jy)
    od, oL, opdf, otex = o.env_sample(r, coin, jx, jy, mode=0)
    assert np.array_equal(gtex, otex)
    assert np.allclose(gd, od, atol=2e-6)
    assert np.allclose(gpdf, opdf, rtol=2e-5)
    # the sun disc (radiance 5e4, 0.5 deg) is drawn in proportion to its share of the pdf
    sun = gL.max(1) > 1e3
    _say(f"  sun texels drawn: {sun.mean():.4f}")
    assert 0.0 < sun.mean() < 1.0


def test_hair_full_furnace(gpu):
    # 50,000 strands x 16 segments as camera-facing ribbons, colour 1 (sigma_a = 0)
    # under a uniform unit envmap: every hair pixel converges to 1 when no path is
    # cut short.  Paths in this dense hair ball scatter dozens of times, so the
    # workload's max depth of 32 truncates a fifth of the energy (reported, not a
    # BCSDF property); energy conservation is asserted at depth 256.
    sc = ng.generate_scene(3, variant=1, width=128, height=128)
    _say(f"case: hair furnace, {len(sc.indices)} triangles, 128x128")
    assert len(sc.indices) >= 1_600_000
    from test_hair_oracle import _first_hits
    hit = _first_hits(orc.OracleScene(sc), sc.camera)
    gpu.upload(sc)
    res = {}
    for depth, spp in ((32, 256), (256, 2048)):
        gpu.reset_film()
        img = gpu.render(ng.make_view(sc.camera, spp, depth), seed=31)
        L = img[:, :3].mean(1) / img[:, 3]
        m, se = float(L[hit].mean()), float(L[hit].std() / np.sqrt(hit.sum()))
        res[depth] = (m, se)
        _say(f"  depth {depth}, {spp} spp: hair pixels {hit.sum()} mean {m:.6f} (standard error {se:.2e})")
    m, se = res[256]
    assert abs(m - 1.0) < max(1e-3, 4 * se), res
    assert res[32][0] < m


def test_batched_full_scene_same_samples(gpu):
    # The headline workload's scene at its full size.  The oracle (Nori's BVH2
    # over 10 M triangles, built in about 20 s on the host) renders two of the
    # sixteen views at a few spp: the thin lens, the 1024^2 albedo and normal
    # maps, the alpha-tested cards and the glass of the real scene.
    sc = ng.generate_scene(5)
    _say(f"case: batched scene, {len(sc.indices)} triangles, {len(sc.textures)} textures")
    assert len(sc.indices) == 10_000_448
    gpu.upload(sc)
    oracle = orc.OracleScene(sc)
    rays = _camera_rays(sc, 1024, 1024, 5)
    _check_hits(gpu, oracle, rays, "batched scene camera rays")
    for (w, h), spp in (((256, 256), 4), ((480, 270), 2)):
        cam = ng.lookat_camera_like(sc.camera, w, h)
        _say(f"case: batched scene view {w}x{h}, {spp} spp, depth 12, same samples")
        gpu.reset_film()
        g = gpu.render(ng.make_view(cam, spp, 12), seed=21)
        o, _ = oracle.render(cam, spp, 12, seed=21)
        assert np.array_equal(g[:, 3], o[:, 3])
        gi, oi = g[:, :3] / spp, o[:, :3] / spp
        e = orc.rel_mse(gi, oi)
        close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-4, axis=1).mean()
        _say(f"  relMSE {e:.3e}, pixels within 1e-3: {close:.4f}")
        assert np.isfinite(gi).all()
        assert e < 1e-3 and close > 0.97
