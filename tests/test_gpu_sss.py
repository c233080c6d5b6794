"""The random-walk subsurface material on the GPU (shaders/Subsurface.hlsli,
shaders/Shaders.hlsl:513-612): the closed forms of sss_scenes.py, and the
oracle's image of an absorbing, tinted sphere in the mean (walks of many
steps diverge from the oracle's after FP32 rounding differences, as paths in
dense media do)."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from sss_scenes import slab_camera, slab_expected, slab_scene, sphere_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("thickness,mfp,ior", [(1.0, 1.0 / 0.7, 1.5), (0.5, 4.0, 1.8)])
def test_slab_transmission_closed_form(gpu, thickness, mfp, ior):
    gpu.upload(slab_scene(thickness, mfp, ior))
    gpu.reset_film()
    _say(f"case: slab L={thickness} d={mfp:.3f} ior={ior}, 8x8 x 16384 spp")
    img = gpu.render(ng.make_view(slab_camera(), 16384, 8), seed=3)
    v = img[:, :3] / img[:, 3:]
    expect = slab_expected(thickness, mfp, ior)
    _say(f"  {v.mean():.5f} vs {expect:.5f}")
    assert abs(v.mean() - expect) < 0.003


def test_lossless_sphere_vanishes_in_white_furnace(gpu):
    sc, cam = sphere_scene(2.0, 0.5)
    gpu.upload(sc)
    gpu.reset_film()
    _say("case: lossless subsurface sphere, white furnace, 32x32 x 1024 spp")
    img = gpu.render(ng.make_view(cam, 1024, 16), seed=5)
    c = (img[:, :3] / img[:, 3:]).reshape(32, 32, 3)[10:22, 10:22]
    _say(f"  centre mean {c.mean():.5f}")
    assert abs(c.mean() - 1.0) < 0.005


@pytest.mark.parametrize("rough,g", [(0.0, 0.0), (0.3, 0.6)])
def test_absorbing_sphere_matches_oracle_in_the_mean(gpu, rough, g):
    sc, cam = sphere_scene(np.array([0.9, 0.5, 0.2], np.float32), 0.2, rough=rough, g=g)
    gpu.upload(sc)
    gpu.reset_film()
    _say(f"case: tinted subsurface sphere rough={rough} g={g}, 32x32 x 512 spp vs oracle")
    gi = gpu.render(ng.make_view(cam, 512, 16), seed=7)
    oi, _ = orc.OracleScene(sc).render(cam, 512, 16, seed=7)
    g_c = (gi[:, :3] / gi[:, 3:]).reshape(32, 32, 3)[8:24, 8:24].reshape(-1, 3).mean(0)
    o_c = (oi[:, :3] / oi[:, 3:]).reshape(32, 32, 3)[8:24, 8:24].reshape(-1, 3).mean(0)
    _say(f"  gpu {g_c} oracle {o_c}")
    assert np.all(np.abs(g_c / o_c - 1.0) < 0.02)
