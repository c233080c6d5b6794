"""GPU textures and alpha test against the oracle: trilinear lookups bit for
bit (same decode tables, same FP32 filter arithmetic without contraction),
textured and alpha-tested renders on the same samples, and the alpha pass
rate 1 - a^2."""
import math
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from texture_scenes import (alpha_camera, alpha_quad_scene, assemble, checker, floor_camera, noise_texture, quad,
                            textured_floor_scene)

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_texture_lookups_bit_exact(gpu):
    texs = [(noise_texture(37, 11, 1), True), (noise_texture(64, 64, 2), False), (noise_texture(5, 128, 3), True),
            (checker(16, 16, (255, 0, 0, 255), (0, 0, 255, 128)), False)]
    p, t, uv = quad(0.0, 1.0)
    sc = assemble([(p, t, uv)], ng.default_materials(1), texs)
    gpu.upload(sc)
    o = orc.OracleScene(sc)
    rng = np.random.default_rng(7)
    n = 20000
    ti = rng.integers(0, len(texs), n).astype(np.uint32)
    uvs = rng.uniform(-3, 3, (n, 2)).astype(np.float32)
    lod = rng.uniform(-1, 9, n).astype(np.float32)
    lod[::3] = np.floor(lod[::3])
    _say(f"case: {n} texture lookups over {len(texs)} textures")
    g, c = gpu.texture_sample(ti, uvs, lod), o.texture_sample(ti, uvs, lod)
    bad = np.nonzero(np.any(g.view(np.uint32) != c.view(np.uint32), axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:5], g[bad[:2]], c[bad[:2]])


@pytest.mark.parametrize("a_byte", [0, 128, 200, 255])
def test_alpha_quad_same_samples_and_rate(gpu, a_byte):
    sc = alpha_quad_scene(a_byte)
    gpu.upload(sc)
    gpu.reset_film()
    _say(f"case: alpha quad a = {a_byte}/255, 256 spp")
    g = gpu.render(ng.make_view(alpha_camera(), 256, 1), seed=4)
    o, _ = orc.OracleScene(sc).render(alpha_camera(), 256, 1, seed=4)
    assert np.array_equal(g[:, 3], o[:, 3])
    assert np.allclose(g[:, :3], o[:, :3], rtol=1e-5, atol=1e-5)
    L = g[:, :3].sum() / (3 * g[:, 3].sum())
    a = a_byte / 255.0
    expect = 1.0 if a < 0.01 else 1.0 - a * a
    assert abs(L - expect) < 5 * math.sqrt(max(expect * (1 - expect), 1e-6) / g[:, 3].sum()) + 1e-6


@pytest.mark.parametrize("normal", [None, (255, 128, 128), (100, 170, 230)])
def test_textured_floor_same_samples(gpu, normal):
    albedo = noise_texture(32, 32, 11)
    albedo[..., 3] = 255
    sc = textured_floor_scene(normal_rgb=normal, albedo=albedo, light_x=3.0)
    cam = floor_camera(32, 32)
    gpu.upload(sc)
    gpu.reset_film()
    _say(f"case: textured floor normal {normal}, depth 4, 64 spp")
    g = gpu.render(ng.make_view(cam, 64, 4), seed=8)
    o, _ = orc.OracleScene(sc).render(cam, 64, 4, seed=8)
    e = orc.rel_mse(g[:, :3] / 64, o[:, :3] / 64)
    _say(f"  relMSE {e:.3e}")
    assert e < 1e-5
