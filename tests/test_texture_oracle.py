"""Pins of the oracle's texture path (src/gpu/DXRApp_Textures.cpp mips and
formats, the MIN_MAG_MIP_LINEAR / WRAP sampler, shaders/Shaders.hlsl:232-315
maps, shaders/Shaders.hlsl:674-724 alpha test): the box-filter mip chain by
hand, bilinear closed forms (texel centres, midpoints, wrap, constants),
trilinear blending, the sRGB decode formula, the albedo map through a white
furnace, the tangent direction of a normal map, and the alpha test's pass
rate 1 - a^2."""
import math

import numpy as np
import pytest

import nori_gpu as ng
import orc
from texture_scenes import (alpha_camera, alpha_quad_scene, assemble, checker, floor_camera, noise_texture, quad,
                            textured_floor_scene)


def _srgb(c):
    x = c / 255.0
    return np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)


def _scene_with(textures):
    p, t, uv = quad(0.0, 1.0)
    return orc.OracleScene(assemble([(p, t, uv)], ng.default_materials(1), textures))


def _box_mips(img):
    lv = [img.astype(np.int64)]
    while lv[-1].shape[0] > 1 or lv[-1].shape[1] > 1:
        a = lv[-1]
        h, w = a.shape[:2]
        mh, mw = max(h // 2, 1), max(w // 2, 1)
        out = np.zeros((mh, mw, 4), np.int64)
        for y in range(mh):
            for x in range(mw):
                x0, y0 = min(2 * x, w - 1), min(2 * y, h - 1)
                x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
                out[y, x] = (a[y0, x0] + a[y0, x1] + a[y1, x0] + a[y1, x1]) // 4
        lv.append(out)
    return lv


@pytest.mark.parametrize("shape", [(4, 4), (2, 8), (3, 5), (1, 7)])
def test_mip_chain_box_filter(shape):
    img = noise_texture(shape[1], shape[0], seed=sum(shape))
    o = _scene_with([(img, False)])
    mips = _box_mips(img)
    for l, m in enumerate(mips):
        h, w = m.shape[:2]
        # texel centres of level l, level of detail exactly l
        ys, xs = np.mgrid[0:h, 0:w]
        uv = np.c_[((xs + 0.5) / w).reshape(-1), ((ys + 0.5) / h).reshape(-1)]
        got = o.texture_sample(np.zeros(len(uv), np.uint32), uv, float(l))
        assert np.allclose(got, (m.reshape(-1, 4) / 255.0), atol=1e-6), l


def test_bilinear_midpoint_and_wrap():
    img = noise_texture(8, 4, seed=3)
    o = _scene_with([(img, False)])
    c = img.astype(np.float64) / 255.0
    # halfway between texels (2, 1) and (3, 1)
    got = o.texture_sample([0], [[3.0 / 8, 1.5 / 4]], 0.0)[0]
    assert np.allclose(got, 0.5 * (c[1, 2] + c[1, 3]), atol=1e-6)
    # between the last and the first column: wrap
    got = o.texture_sample([0], [[0.0, 1.5 / 4]], 0.0)[0]
    assert np.allclose(got, 0.5 * (c[1, 7] + c[1, 0]), atol=1e-6)
    # u and u + 3 are the same point
    uv = np.random.default_rng(1).random((64, 2)).astype(np.float32)
    a = o.texture_sample(np.zeros(64, np.uint32), uv, 0.3)
    b = o.texture_sample(np.zeros(64, np.uint32), uv + np.float32(3.0), 0.3)
    assert np.allclose(a, b, atol=2e-6)


def test_trilinear_blends_levels():
    img = noise_texture(16, 16, seed=9)
    o = _scene_with([(img, False)])
    uv = np.random.default_rng(2).random((32, 2)).astype(np.float32)
    z = np.zeros(32, np.uint32)
    l0, l1, mid = o.texture_sample(z, uv, 1.0), o.texture_sample(z, uv, 2.0), o.texture_sample(z, uv, 1.25)
    assert np.allclose(mid, 0.75 * l0 + 0.25 * l1, atol=1e-6)
    # beyond the chain: the 1x1 level
    assert np.allclose(o.texture_sample(z, uv, 40.0), _box_mips(img)[-1][0, 0] / 255.0, atol=1e-6)


def test_srgb_decode_and_constant_texture():
    vals = np.array([0, 1, 10, 11, 64, 128, 200, 255], np.uint8)
    for v in vals:
        img = np.full((4, 4, 4), v, np.uint8)
        o = _scene_with([(img, True)])
        got = o.texture_sample([0, 0], [[0.13, 0.77], [5.5, -2.25]], [0.0, 1.7])
        assert np.allclose(got[:, :3], _srgb(float(v)), rtol=1e-6, atol=1e-9), v
        assert np.allclose(got[:, 3], v / 255.0, atol=1e-7)  # alpha stays linear


def test_albedo_map_in_white_furnace():
    # a closed Lambertian box (all faces the same sRGB albedo texture) seen from
    # inside under no light but a unit envmap is not closed-form; instead a
    # single floor quad with an albedo texture, lit by a unit envmap from above
    # only through its front: the pixel value is rho * (1 - occlusion); with no
    # occluders, rho * 1 for a diffuse surface under a uniform unit sky.
    p, t, uv = quad(0.0, 50.0)
    mats = ng.default_materials(1)
    mats[0]["albedoTexIndex"] = 0
    tex = np.full((4, 4, 4), 128, np.uint8)
    env = np.full((8, 16, 3), 1.0, np.float32)
    sc = assemble([(p, t, uv)], mats, [(tex, True)], env=env)
    cam = ng.lookat_camera([0, 1, 0.001], [0, 0, 0], [0, 0, 1], 20.0, 8, 8)
    img, _ = orc.OracleScene(sc).render(cam, 1024, 2, seed=1)
    L = img[:, :3] / img[:, 3:4]
    # upper hemisphere of a unit sky onto a Lambertian floor: radiance rho
    assert abs(L.mean() - _srgb(128.0)) < 6e-3, (L.mean(), _srgb(128.0))


def test_normal_map_tilts_towards_u():
    # tangent-space normal (1, 0, 0) -> shading normal along +T = +x (u grows
    # with x): a light on +x lights the floor, one on -x does not (next event
    # only, depth 1)
    tilt = (255, 128, 128)
    plus = textured_floor_scene(normal_rgb=tilt, light_x=3.0)
    minus = textured_floor_scene(normal_rgb=tilt, light_x=-3.0)
    flat_minus = textured_floor_scene(normal_rgb=None, light_x=-3.0)
    cam = floor_camera()
    a, _ = orc.OracleScene(plus).render(cam, 64, 1, seed=2)
    b, _ = orc.OracleScene(minus).render(cam, 64, 1, seed=2)
    c, _ = orc.OracleScene(flat_minus).render(cam, 64, 1, seed=2)
    assert a[:, :3].mean() > 0.05
    assert b[:, :3].max() == 0.0
    assert c[:, :3].mean() > 0.0


@pytest.mark.parametrize("a_byte", [0, 2, 128, 200, 255])
def test_alpha_pass_rate(a_byte):
    sc = alpha_quad_scene(a_byte)
    img, _ = orc.OracleScene(sc).render(alpha_camera(), 256, 1, seed=4)
    L = img[:, :3].sum() / (3 * img[:, 3].sum())
    a = a_byte / 255.0
    expect = 1.0 if a < 0.01 else 1.0 - a * a
    n = img[:, 3].sum()
    assert abs(L - expect) < 5 * math.sqrt(max(expect * (1 - expect), 1e-6) / n) + 1e-6, (L, expect)
