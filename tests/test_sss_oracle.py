"""The oracle's random-walk subsurface material against closed forms
(scenes and formulas in sss_scenes.py)."""
import numpy as np
import pytest

import orc
from sss_scenes import slab_camera, slab_expected, slab_scene, sphere_scene


@pytest.mark.parametrize("thickness,mfp,ior", [(1.0, 1.0 / 0.7, 1.5), (2.0, 1.5, 1.33), (0.5, 4.0, 1.8)])
def test_slab_transmission_closed_form(thickness, mfp, ior):
    sc = slab_scene(thickness, mfp, ior)
    img, _ = orc.OracleScene(sc).render(slab_camera(), 4096, 8, seed=3)
    v = img[:, :3] / img[:, 3:]
    expect = slab_expected(thickness, mfp, ior)
    # 64 pixels x 4096 samples of a {0, 1}-valued estimate: sd < 0.001
    assert abs(v.mean() - expect) < 0.005, (v.mean(), expect)


def test_lossless_sphere_vanishes_in_white_furnace():
    sc, cam = sphere_scene(2.0, 0.5)
    img, _ = orc.OracleScene(sc).render(cam, 256, 16, seed=5)
    v = (img[:, :3] / img[:, 3:]).reshape(32, 32, 3)
    centre = v[10:22, 10:22]
    assert abs(centre.mean() - 1.0) < 0.01, centre.mean()


def test_absorbing_sphere_is_darker_and_tinted():
    # scatter colour (0.9, 0.5, 0.2): albedo per channel, so red survives the walk best
    sc, cam = sphere_scene(np.array([0.9, 0.5, 0.2], np.float32), 0.2)
    img, _ = orc.OracleScene(sc).render(cam, 128, 16, seed=7)
    c = (img[:, :3] / img[:, 3:]).reshape(32, 32, 3)[12:20, 12:20].reshape(-1, 3).mean(0)
    assert 0.0 < c[2] < c[1] < c[0] < 1.0, c
