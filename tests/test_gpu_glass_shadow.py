"""GPU shadow rays through mirror and dielectric meshes against the oracle on
the same samples, and against the closed-form Fresnel factor of one pane."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from glass_scenes import floor_camera, pane_cosine, pane_scene
from test_glass_shadow_oracle import _fresnel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_render(gpu, sc, depth, spp=256, seed=3):
    gpu.upload(sc)
    gpu.reset_film()
    img = gpu.render(ng.make_view(floor_camera(), spp, depth), seed=seed)
    return img


@pytest.mark.parametrize("panes", [[(1.5, "glass")], [(1.0, "glass"), (2.0, "mirror")], [(1.5, "opaque")]])
def test_pane_scene_same_samples(gpu, panes):
    sc = pane_scene(panes)
    _say(f"case: panes {panes}, depth 4, 256 spp, same samples")
    g = _gpu_render(gpu, sc, 4)
    o, _ = orc.OracleScene(sc).render(floor_camera(), 256, 4, seed=3)
    gi, oi = g[:, :3] / 256, o[:, :3] / 256
    e = orc.rel_mse(gi, oi)
    _say(f"  relMSE {e:.3e}")
    assert e < 1e-5


def test_one_pane_fresnel_factor_gpu(gpu):
    _say("case: one glass pane vs none, depth 1")
    bare = _gpu_render(gpu, pane_scene([]), 1)
    pane = _gpu_render(gpu, pane_scene([(1.5, "glass")]), 1)
    ratio = pane[:, :3] / bare[:, :3]
    expect = 1 - _fresnel(pane_cosine(), 1.0, 1.5)
    _say(f"  ratio {ratio.min():.5f}..{ratio.max():.5f}, expected {expect:.5f}")
    assert np.allclose(ratio, expect, atol=5e-4)
