"""Bit-reproducible renders.  The film accumulates every contribution in
64-bit fixed point (2^-28 steps), so the sum of a pixel's samples does not
depend on the order in which the shading and shadow kernels' atomics land:
the same render repeated, or rendered with other pool sizes, chunk lengths,
traversal schedules, shading sorts or without the graph, gives the same bits,
as the reference's renders do run to run.  Covered: surfaces with envmap and
area lights, textures with the alpha test and a thin lens, a medium, the
denoiser feature films (signed normals), the per-ray Li entry point."""
import sys

import numpy as np
import pytest

import nori_gpu as ng

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _assert_same(a, b, label):
    diff = _bits(a) != _bits(b)
    assert not diff.any(), f"{label}: {diff.sum()} of {diff.size} values differ (largest by " \
                           f"{np.abs(np.asarray(a) - np.asarray(b)).max():.3g})"


SETTINGS = [
    # batch, chunk_iters, traversal, sort, graph
    (1 << 21, 8, 1, 1, 1),
    (1 << 10, 2, 0, 1, 1),
    (3 << 12, 4, 1, 0, 1),
    (1 << 16, 16, 0, 0, 0),
]


@pytest.fixture
def restore(gpu):
    yield gpu
    gpu.set_option("batch", 1 << 21)
    gpu.set_option("chunk_iters", 8)
    gpu.set_option("traversal", 1)
    gpu.set_option("sort", 1)
    gpu.set_option("graph", 1)
    gpu.set_option("aov", 0)
    gpu.set_option("guiding", 0)


def _render_all(gpu, view, seed, aov=False):
    out = []
    for batch, chunk, trav, sort, graph in SETTINGS:
        gpu.set_option("batch", batch)
        gpu.set_option("chunk_iters", chunk)
        gpu.set_option("traversal", trav)
        gpu.set_option("sort", sort)
        gpu.set_option("graph", graph)
        gpu.reset_film()
        img = gpu.render(view, seed=seed)
        if aov:
            img = np.concatenate([img, gpu.aov(ng.AOV_ALBEDO), gpu.aov(ng.AOV_NORMAL)])
        out.append(img)
    return out


@pytest.mark.parametrize("name", ["disney", "batched", "medium"])
def test_render_bits_do_not_depend_on_schedule(restore, name):
    gpu = restore
    if name == "disney":
        sc = ng.generate_scene(2, subdiv=3, grid=4, env_width=256, env_height=128, width=64, height=64)
        view = ng.make_view(sc.camera, 16, 12)
    elif name == "batched":
        sc = ng.generate_scene(5, grid=128, subdiv=1, meshes=12, strands=400, tex_size=64, env_width=128, env_height=64,
                               width=64, height=48)
        view = ng.make_view(sc.camera, 8, 12)
    else:
        sc = ng.generate_scene(4, grid=32, env_width=128, env_height=64, width=48, height=48)
        view = ng.make_view(sc.camera, 16, 32)
    _say(f"case: {name}, {len(sc.indices)} triangles, {view.camera.width}x{view.camera.height}x{view.spp}, "
         f"{len(SETTINGS)} schedules")
    gpu.upload(sc)
    imgs = _render_all(gpu, view, seed=5)
    assert np.all(imgs[0][:, 3] == view.spp)
    assert np.isfinite(imgs[0]).all() and imgs[0][:, :3].max() > 0
    for k in range(1, len(imgs)):
        _assert_same(imgs[k], imgs[0], f"{name}: schedule {SETTINGS[k]} vs {SETTINGS[0]}")
    # the same schedule again
    _assert_same(_render_all(gpu, view, seed=5)[0], imgs[0], f"{name}: repeated render")


def test_feature_films_bits_do_not_depend_on_schedule(restore, disney_small):
    gpu = restore
    gpu.upload(disney_small)
    gpu.set_option("aov", 1)
    view = ng.make_view(ng.lookat_camera_like(disney_small.camera, 48, 40), 8, 8)
    _say("case: albedo and normal feature films over 4 schedules")
    imgs = _render_all(gpu, view, seed=9, aov=True)
    n = 48 * 40
    assert (imgs[0][2 * n:, :3] < 0).any(), "normals with negative components expected"
    for k in range(1, len(imgs)):
        _assert_same(imgs[k], imgs[0], f"features: schedule {SETTINGS[k]}")


def test_frozen_guide_render_is_reproducible(restore, disney_small):
    gpu = restore
    gpu.upload(disney_small)
    gpu.set_option("guiding", 1)
    gpu.set_option("guide_reset", 1)
    gpu.set_option("guide_train_steps", 0)
    view = ng.make_view(ng.lookat_camera_like(disney_small.camera, 48, 48), 8, 8)
    _say("case: guided render with the network frozen at its initialisation, 4 schedules")
    try:
        imgs = _render_all(gpu, view, seed=13)
    finally:
        gpu.set_option("guide_train_steps", 4)
    for k in range(1, len(imgs)):
        _assert_same(imgs[k], imgs[0], f"guided: schedule {SETTINGS[k]}")


def test_li_is_reproducible(restore, cornell):
    gpu = restore
    gpu.upload(cornell)
    rng = np.random.default_rng(3)
    n = 20000
    o = np.tile(np.array([0.0, 1.0, 3.5], np.float32), (n, 1))
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:, 2] = -np.abs(d[:, 2]) - 0.5
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.c_[o, np.full(n, 1e-4, np.float32), d, np.full(n, 1e4, np.float32)].astype(np.float32)
    seeds = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    _say(f"case: li on {n} rays, pools 2^21 and 2^11")
    a = gpu.li(rays, seeds, 8)
    gpu.set_option("batch", 1 << 11)
    b = gpu.li(rays, seeds, 8)
    assert np.abs(a).sum() > 0
    _assert_same(a, b, "li")
