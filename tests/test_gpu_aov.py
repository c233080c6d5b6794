"""Denoiser feature films on the GPU (nori_gpu_aov): the exact cases of
aov_scenes.py, and the same samples as the oracle on the Cornell box with a
glass and a mirror sphere (looked through) and on the Disney grid; the films
accumulate like the radiance film and a render without the option leaves the
entry point in error."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from aov_scenes import quad_camera, quad_scene, sky_camera

pytestmark = pytest.mark.gpu
ALB = np.array([0.2, 0.4, 0.6], np.float32)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture
def aov_gpu(gpu):
    gpu.set_option("aov", 1)
    yield gpu
    gpu.set_option("aov", 0)


@pytest.mark.parametrize("kind,cam,albedo,normal", [
    ("diffuse", quad_camera, ALB, [0, 0, 1]), ("emitter", quad_camera, [1, 1, 1], [0, 0, 1]),
    ("diffuse", sky_camera, [0.7, 0.7, 0.7], [0, 0, 0]), ("mirror", quad_camera, ALB, [0, 1, 0])])
def test_exact_cases(aov_gpu, kind, cam, albedo, normal):
    gpu = aov_gpu
    gpu.upload(quad_scene(kind))
    gpu.reset_film()
    _say(f"case: aov {kind} {cam.__name__}")
    gpu.render(ng.make_view(cam(), 4, 4), seed=1)
    a, n = gpu.aov(ng.AOV_ALBEDO), gpu.aov(ng.AOV_NORMAL)
    assert np.all(a[:, 3] == 4) and np.all(n[:, 3] == 4)
    assert np.allclose(a[:, :3] / 4, albedo, atol=1e-5)
    assert np.allclose(n[:, :3] / 4, normal, atol=1e-5)


@pytest.mark.parametrize("config", [1, 2])
def test_same_samples_as_oracle(aov_gpu, config):
    gpu = aov_gpu
    sc = ng.generate_scene(1, subdiv=3) if config == 1 else ng.generate_scene(2, subdiv=3, grid=4, env_width=256,
                                                                              env_height=128)
    if config == 1:  # a glass and a mirror sphere: features seen through them
        sc.materials[5]["type"] = ng.MAT_DIELECTRIC
        sc.materials[5]["intIOR"], sc.materials[5]["extIOR"] = 1.5, 1.0
        sc.materials[6]["type"] = ng.MAT_MIRROR
    gpu.upload(sc)
    cam = ng.lookat_camera_like(sc.camera, 48, 48)
    gpu.reset_film()
    _say(f"case: aov config {config}, 48x48 x 8 spp vs oracle")
    g = gpu.render(ng.make_view(cam, 8, 8), seed=4)
    ga, gn = gpu.aov(ng.AOV_ALBEDO), gpu.aov(ng.AOV_NORMAL)
    o, _, oa, on = orc.OracleScene(sc).render(cam, 8, 8, seed=4, aov=True)
    for name, x, y in (("albedo", ga, oa), ("normal", gn, on)):
        assert np.array_equal(x[:, 3], y[:, 3])
        bad = np.abs(x[:, :3] - y[:, :3]) > 1e-4 * np.maximum(np.abs(y[:, :3]), 1.0)
        _say(f"  {name}: {bad.any(axis=1).sum()} of {len(x)} pixels differ")
        # a glass Fresnel coin can flip after FP32 rounding of the refracted direction
        assert bad.any(axis=1).mean() < 0.01
    # the radiance film is unchanged by the option
    assert orc.rel_mse(g[:, :3] / 8, o[:, :3] / 8) < 1e-4


def test_accumulates_and_requires_the_option(gpu, cornell):
    gpu.upload(cornell)
    cam = ng.lookat_camera_like(cornell.camera, 16, 16)
    gpu.set_option("aov", 0)
    gpu.reset_film()
    gpu.render(ng.make_view(cam, 2, 4), seed=1)
    with pytest.raises(RuntimeError, match="aov"):
        gpu.aov(ng.AOV_ALBEDO)
    gpu.set_option("aov", 1)
    try:
        gpu.reset_film()
        gpu.render(ng.make_view(cam, 2, 4), seed=1)
        first = gpu.aov(ng.AOV_ALBEDO).copy()
        gpu.render(ng.make_view(cam, 3, 4, sample_offset=2), seed=1)
        both = gpu.aov(ng.AOV_ALBEDO)
        assert np.all(first[:, 3] == 2) and np.all(both[:, 3] == 5)
        assert np.all(both[:, :3] >= first[:, :3] - 1e-6)
    finally:
        gpu.set_option("aov", 0)


def test_graph_follows_a_film_layout_change(aov_gpu, cornell):
    # The captured bounce loop bakes in the feature films' offsets (film +
    # pixels of all views).  Render a large layout, then a smaller one without a
    # new scene upload: the second render's films must equal an eager render's
    # (graph off) of that layout.  Radiance arrives through float atomics (two
    # next-event records of a vertex land in one launch), so sums agree to
    # rounding; a stale offset would put features in the wrong place entirely.
    gpu = aov_gpu
    gpu.upload(cornell)
    cam = cornell.camera
    big = ng.lookat_camera_like(cam, 96, 80)
    small = ng.lookat_camera_like(cam, 40, 24)
    _say("case: aov graph, layout 96x80 then 40x24")
    gpu.reset_film()
    gpu.render(ng.make_view(big, 4, 8), seed=5)
    out = {}
    for graph in (1, 0):
        gpu.set_option("graph", graph)
        gpu.reset_film()
        img = gpu.render([ng.make_view(small, 2, 8), ng.make_view(small, 2, 8, sample_offset=2)], seed=9)
        out[graph] = (img, gpu.aov(ng.AOV_ALBEDO), gpu.aov(ng.AOV_NORMAL))
    gpu.set_option("graph", 1)
    for a, b in zip(out[1], out[0]):
        assert np.allclose(a, b, rtol=1e-5, atol=1e-7)
    assert np.all(out[1][0][:, 3] == 2) and out[1][1][:, 3].max() > 0
