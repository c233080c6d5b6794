"""Scenes for the denoiser-feature checks: a camera at the origin looking
down -z at a quad (z = -2, facing the camera, 4 x 4) that fills the view, with
a uniform environment of radiance 0.7 behind it."""
import numpy as np

import nori_gpu as ng


def quad_scene(kind, albedo=(0.2, 0.4, 0.6)):
    """kind: 'diffuse' (the quad), 'emitter' (an area light), 'mirror' (a mirror
    quad tilted 45 degrees about x that shows a diffuse floor quad below it)."""
    env = np.full((8, 16, 3), 0.7, np.float32)
    if kind in ("diffuse", "emitter"):
        p = np.array([[-2, -2, -2], [2, -2, -2], [2, 2, -2], [-2, 2, -2]], np.float32)
        t = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
        m = ng.default_materials(1)
        m["albedo"] = albedo
        if kind == "emitter":
            m["isEmitter"] = 1
            m["radiance"] = 3.0
        return ng.HostScene(p, t, [0], [2], m, env_rgb=env)
    # mirror plane z - y = -2 (normal (0, -1, 1)/sqrt2): camera rays reflect
    # straight down onto a diffuse floor at y = -3
    mp = np.array([[-3, 3, 1], [3, 3, 1], [3, -3, -5], [-3, -3, -5]], np.float32)
    fp = np.array([[-20, -3, -30], [20, -3, -30], [20, -3, 10], [-20, -3, 10]], np.float32)
    p = np.concatenate([mp, fp])
    t = np.array([[0, 1, 2], [0, 2, 3], [4, 6, 5], [4, 7, 6]], np.uint32)
    m = ng.default_materials(2)
    m[0]["type"] = ng.MAT_MIRROR
    m[1]["albedo"] = albedo
    return ng.HostScene(p, t, [0, 2], [2, 2], m, env_rgb=env)


def quad_camera(w=8, h=8):
    return ng.lookat_camera([0.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.0, 1.0, 0.0], 30.0, w, h)


def sky_camera(w=8, h=8):
    # looking up +y past everything: every camera ray misses
    return ng.lookat_camera([0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], 30.0, w, h)
