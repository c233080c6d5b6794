"""The batched textured scene (config 5, reduced) on the GPU: same-sample
parity with the oracle (thin lens, texture maps, alpha-tested cards, glass),
and sixteen views of uneven sizes and spp in one persistent render equal to
the views rendered one at a time."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batched():
    return ng.generate_scene(5, grid=64, subdiv=3, meshes=30, strands=2000, tex_size=64, env_width=128,
                             env_height=64, width=48, height=32)


def sixteen_views(scene, scale=0.1):
    # 256^2 .. 1920x1080 and 16 .. 256 spp, scaled down by `scale` in pixels and spp
    sizes = [(256, 256), (320, 240), (400, 300), (480, 270), (512, 512), (640, 360), (640, 480), (800, 450),
             (800, 600), (960, 540), (1024, 576), (1024, 768), (1280, 720), (1366, 768), (1600, 900), (1920, 1080)]
    spps = [16, 32, 48, 64, 80, 96, 112, 128, 144, 160, 176, 192, 208, 224, 240, 256]
    views = []
    for k, ((w, h), spp) in enumerate(zip(sizes, spps)):
        cam = ng.lookat_camera_like(scene.camera, max(8, int(w * scale)), max(8, int(h * scale)))
        views.append(ng.make_view(cam, max(1, int(spp * scale)), 8))
    return views


def test_batched_same_samples(gpu, batched):
    sc = batched
    gpu.upload(sc)
    gpu.reset_film()
    _say("case: batched scene 48x32, 8 spp, depth 8, same samples")
    g = gpu.render(ng.make_view(sc.camera, 8, 8), seed=2)
    o, _ = orc.OracleScene(sc).render(sc.camera, 8, 8, seed=2)
    gi, oi = g[:, :3] / 8, o[:, :3] / 8
    e = orc.rel_mse(gi, oi)
    close = np.all(np.abs(gi - oi) <= 1e-3 * np.abs(oi) + 1e-4, axis=1).mean()
    _say(f"  relMSE {e:.3e}, close {close:.4f}")
    assert np.isfinite(gi).all()
    assert e < 1e-3 and close > 0.97


def test_sixteen_views_in_one_render(gpu, batched):
    sc = batched
    gpu.upload(sc)
    views = sixteen_views(sc)
    _say(f"case: {len(views)} views, {sum(v.camera.width * v.camera.height * v.spp for v in views)} samples")
    gpu.reset_film()
    together = gpu.render(views, seed=9)
    off = 0
    for v in views:
        n = v.camera.width * v.camera.height
        gpu.reset_film()
        alone = gpu.render(v, seed=9)
        assert np.array_equal(together[off:off + n, 3], alone[:, 3])
        tol = 1e-5 * np.maximum(np.abs(together[off:off + n, :3]), np.abs(alone[:, :3])) + 1e-5
        assert (np.abs(together[off:off + n, :3] - alone[:, :3]) > tol).mean() < 1e-3
        off += n
