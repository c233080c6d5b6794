"""GPU checks of participating media: the wavefront path (free flight in the
shade kernel, ratio-tracking transmittance in the shadow kernel) against the
oracle on the same samples, and against the closed forms the oracle is pinned
to (Beer-Lambert, the volume white furnace)."""
import math
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from test_volume_oracle import _box_scene

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _render(gpu, scene, cam, spp, depth, seed):
    gpu.upload(scene)
    gpu.reset_film()
    return gpu.render(ng.make_view(cam, spp, depth), seed=seed)


@pytest.mark.parametrize("variant", [0, 2])
def test_medium_scene_same_samples(gpu, variant):
    # config 4 at a 32^3 grid: same pcg32 streams on both sides
    sc = ng.generate_scene(4, grid=32, env_width=128, env_height=64, width=48, height=48, variant=variant)
    _say(f"case: medium scene g={sc.media[0].g:.1f} 48x48 16spp depth 64")
    g = _render(gpu, sc, sc.camera, 16, 64, 3)
    o, _ = orc.OracleScene(sc).render(sc.camera, 16, 64, seed=3)
    gi, oi = g[:, :3] / 16, o[:, :3] / 16
    e = orc.rel_mse(gi, oi)
    close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-4, axis=1)
    _say(f"  relMSE {e:.3e}, close {close.mean():.4f}, means {gi.mean(0)} {oi.mean(0)}")
    # A path diverges from its oracle twin once an FP32 rounding difference in a
    # free-flight distance (logf, the trilinear lookup) flips one of the tens of
    # collision decisions per segment in this dense medium, so per-pixel
    # agreement is partial; the estimates agree in the mean.
    assert np.isfinite(gi).all()
    assert np.allclose(gi.mean(0), oi.mean(0), rtol=5e-3)
    assert e < 2e-2


def test_medium_scene_converges_to_oracle(gpu):
    sc = ng.generate_scene(4, grid=32, env_width=128, env_height=64, width=24, height=24, variant=1)
    _say("case: medium scene independent seeds 24x24 1024spp")
    g = _render(gpu, sc, sc.camera, 1024, 64, 21)
    o, _ = orc.OracleScene(sc).render(sc.camera, 1024, 64, seed=22)
    gi, oi = g[:, :3] / 1024, o[:, :3] / 1024
    e = orc.rel_mse(gi, oi)
    _say(f"  relMSE {e:.3e}, means {gi.mean(0)} {oi.mean(0)}")
    assert np.allclose(gi.mean(0), oi.mean(0), rtol=0.03)
    assert e < 2e-2


def test_beer_lambert_gpu(gpu):
    m = ng.MediumSpec([-5, -5, -0.5], [5, 5, 0.5], sigma_a=2.0, sigma_s=0.0)
    sc = _box_scene([m])
    cam = ng.lookat_camera([0, 0, -3], [0, 0, 3], [0, 1, 0], 10, 16, 16)
    _say("case: Beer-Lambert slab 16x16 4096spp")
    img = _render(gpu, sc, cam, 4096, 8, 1)
    L = img[:, :3] / img[:, 3:4]
    ys, xs = np.mgrid[0:16, 0:16]
    th = math.tan(math.radians(5))
    dx, dy = (1 - 2 * (xs + 0.5) / 16) * th, (1 - 2 * (ys + 0.5) / 16) * th
    expect = np.exp(-2.0 * np.sqrt(1 + dx ** 2 + dy ** 2)).reshape(-1)
    _say(f"  mean {L.mean():.5f} expected {expect.mean():.5f}")
    assert abs(L.mean() - expect.mean()) < 3e-3


def test_volume_white_furnace_gpu(gpu):
    rng = np.random.default_rng(5)
    dens = rng.random((8, 8, 8)).astype(np.float32)
    m = ng.MediumSpec([-1, -1, -1], [1, 1, 1], sigma_a=0.0, sigma_s=[2.0, 2.0, 2.0], g=0.3, density=dens, brick=4)
    sc = _box_scene([m])
    cam = ng.lookat_camera([0, 0, -4], [0, 0, 0], [0, 1, 0], 35, 12, 12)
    _say("case: volume furnace 12x12 8192spp")
    img = _render(gpu, sc, cam, 8192, 64, 2)
    L = img[:, :3] / img[:, 3:4]
    _say(f"  mean {L.mean():.5f} max dev {np.abs(L - 1).max():.4f}")
    assert abs(L.mean() - 1.0) < 5e-3
    assert np.abs(L - 1.0).max() < 0.08
