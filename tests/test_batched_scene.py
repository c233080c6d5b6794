"""The batched textured scene (config 5) at reduced size: geometry and texture
inventory of the generator, and an oracle render (thin lens, textures, alpha
cards, glass) that stays finite."""
import numpy as np
import pytest

import nori_gpu as ng
import orc


@pytest.fixture(scope="module")
def small_batched():
    return ng.generate_scene(5, grid=32, subdiv=2, meshes=12, strands=300, tex_size=32, env_width=64, env_height=32,
                             width=32, height=24)


def test_inventory(small_batched):
    sc = small_batched
    # the first ceil(36 * 12 / 199) = 3 icospheres one subdivision level finer
    assert len(sc.indices) == 2 * 32 * 32 + 3 * 20 * 4 ** 3 + 9 * 20 * 4 ** 2 + 2 * 300
    assert len(sc.materials) == 1 + 12 + 1
    assert len(sc.textures) == 129
    assert sum(srgb for _, srgb in sc.textures) == 64
    # normal maps decode to unit vectors (to 8-bit precision)
    for t, _ in sc.textures[64:128]:
        n = t[..., :3].astype(np.float64) / 255 * 2 - 1
        assert np.allclose(np.linalg.norm(n, axis=-1), 1.0, atol=0.02)
        assert (n[..., 2] > 0).all()
    a = sc.textures[128][0][..., 0]
    assert a.min() == 0 and a.max() >= 200
    assert sc.materials["alphaTexIndex"][-1] == 128
    assert sc.camera.lens_radius == pytest.approx(0.05) and sc.camera.focal_distance == pytest.approx(5.0)


def test_oracle_render_is_finite(small_batched):
    sc = small_batched
    img, stats = orc.OracleScene(sc).render(sc.camera, 2, 6, seed=1)
    assert np.isfinite(img).all() and (img[:, :3] >= 0).all()
    assert img[:, :3].mean() > 0
