"""Small textured scenes for the texture and alpha-test checks."""
import numpy as np

import nori_gpu as ng


def checker(w, h, a, b, cells=4):
    y, x = np.mgrid[0:h, 0:w]
    m = ((x * cells // w) + (y * cells // h)) % 2 == 0
    t = np.where(m[..., None], np.array(a, np.uint8), np.array(b, np.uint8))
    return np.ascontiguousarray(t.astype(np.uint8))


def noise_texture(w, h, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, (h, w, 4), dtype=np.uint8)


def quad(y, half, u_tiles=1.0):
    # [0, 1, 2] winding faces -y; u along +x, v along +z
    p = np.array([[-half, y, -half], [half, y, -half], [half, y, half], [-half, y, half]], np.float32)
    uv = np.array([[0, 0], [u_tiles, 0], [u_tiles, u_tiles], [0, u_tiles]], np.float32)
    return p, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), uv


def assemble(parts, mats, textures, env=None):
    pos, idx, uvs, off, cnt = [], [], [], [], []
    nv = nt = 0
    for p, t, uv in parts:
        pos.append(p)
        idx.append(t + nv)
        uvs.append(uv)
        off.append(nt)
        cnt.append(len(t))
        nv += len(p)
        nt += len(t)
    return ng.HostScene(np.concatenate(pos), np.concatenate(idx), np.array(off, np.uint32), np.array(cnt, np.uint32),
                        mats, uvs=np.concatenate(uvs), env_rgb=env, textures=textures)


def alpha_quad_scene(alpha_byte, env_value=1.0):
    """A black alpha-tested quad in front of a uniform envmap (camera looks at it along +z)."""
    p = np.array([[-2, -2, 0], [2, -2, 0], [2, 2, 0], [-2, 2, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    mats = ng.default_materials(1)
    mats["albedo"] = 0.0
    tex = np.zeros((8, 8, 4), np.uint8)
    tex[..., 0] = alpha_byte
    tex[..., 3] = 255
    mats[0]["alphaTexIndex"] = 0
    env = np.full((8, 16, 3), env_value, np.float32)
    return assemble([(p, t, uv)], mats, [(tex, False)], env=env)


def alpha_camera(w=16, h=16):
    return ng.lookat_camera([0, 0, -3], [0, 0, 0], [0, 1, 0], 20.0, w, h)


def textured_floor_scene(normal_rgb=None, albedo=None, light_x=3.0):
    """A diffuse floor (textures optional) lit by a small light at height 1 on the x axis side."""
    fp, ft, fuv = quad(0.0, 5.0)
    lp = np.array([[light_x, 0.9, -0.1], [light_x, 1.1, -0.1], [light_x, 1.1, 0.1], [light_x, 0.9, 0.1]], np.float32)
    # the light faces back towards the origin
    lt = np.array([[0, 2, 1], [0, 3, 2]] if light_x > 0 else [[0, 1, 2], [0, 2, 3]], np.uint32)
    luv = np.zeros((4, 2), np.float32)
    mats = ng.default_materials(2)
    textures = []
    if albedo is not None:
        mats[0]["albedoTexIndex"] = len(textures)
        textures.append((albedo, True))
    if normal_rgb is not None:
        nt = np.zeros((4, 4, 4), np.uint8)
        nt[..., :3] = normal_rgb
        nt[..., 3] = 255
        mats[0]["normalTexIndex"] = len(textures)
        textures.append((nt, False))
    mats[1]["isEmitter"] = 1
    mats[1]["radiance"] = 50.0
    return assemble([(fp, ft, fuv), (lp, lt, luv)], mats, textures)


def floor_camera(w=16, h=16):
    return ng.lookat_camera([0.0, 0.6, 0.3], [0.0, 0.0, 0.0], [0.0, 0.0, -1.0], 20.0, w, h)
