"""The oracle's denoiser features (shaders/Shaders.hlsl:99-106, 195-200,
320-324, 475-479, 630-638) on scenes where they are known exactly
(aov_scenes.py): a diffuse quad gives its albedo and its normal, an emitter
white and its normal, a miss the environment radiance and a zero normal, and a
mirror is looked through to the surface it shows."""
import numpy as np

import orc
from aov_scenes import quad_camera, quad_scene, sky_camera

ALB = np.array([0.2, 0.4, 0.6], np.float32)


def _aov(scene, cam, spp=4):
    img, _, alb, nrm = orc.OracleScene(scene).render(cam, spp, 4, seed=1, aov=True)
    assert np.all(alb[:, 3] == spp) and np.all(nrm[:, 3] == spp)
    return alb[:, :3] / spp, nrm[:, :3] / spp


def test_diffuse_quad():
    a, n = _aov(quad_scene("diffuse"), quad_camera())
    assert np.allclose(a, ALB, atol=1e-6)
    assert np.allclose(n, [0, 0, 1], atol=1e-6)


def test_emitter_quad():
    a, n = _aov(quad_scene("emitter"), quad_camera())
    assert np.allclose(a, 1.0, atol=1e-6)
    assert np.allclose(n, [0, 0, 1], atol=1e-6)


def test_miss():
    a, n = _aov(quad_scene("diffuse"), sky_camera())
    assert np.allclose(a, 0.7, atol=1e-6)
    assert np.allclose(n, 0.0)


def test_mirror_is_looked_through():
    a, n = _aov(quad_scene("mirror"), quad_camera())
    assert np.allclose(a, ALB, atol=1e-6)
    assert np.allclose(n, [0, 1, 0], atol=1e-6)
