"""Scenes for the shadow-ray pass-through checks: a diffuse floor lit by a small
area light standing at x = 3 and facing -x, seen from a camera above the origin;
vertical panes (planes x = const) stand between the floor and the light.  The
light faces sideways because a light parallel to the floor occludes its own
next-event rays (origin offset 0.002 along the normal, tmax = distance - 0.001:
shaders/Emitter.hlsli:122-131, shaders/GeometryUtils.hlsli:178-183)."""
import numpy as np

import nori_gpu as ng


def _floor(half):
    p = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    return p, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _wall(x, y0, y1, half_z, facing_minus_x):
    p = np.array([[x, y0, -half_z], [x, y1, -half_z], [x, y1, half_z], [x, y0, half_z]], np.float32)
    # [0, 1, 2] winding faces +x
    t = np.array([[0, 2, 1], [0, 3, 2]] if facing_minus_x else [[0, 1, 2], [0, 2, 3]], np.uint32)
    return p, t


def pane_scene(panes):
    """panes: list of (x, type) with type 'glass', 'mirror' or 'opaque'."""
    parts = [(_floor(5.0), "floor"), (_wall(3.0, 1.0, 1.2, 0.1, True), "light")]
    for x, kind in panes:
        parts.append((_wall(x, -1.0, 3.0, 2.0, False), kind))
    pos, idx, off, cnt = [], [], [], []
    nv = nt = 0
    for (p, t), _ in parts:
        pos.append(p)
        idx.append(t + nv)
        off.append(nt)
        cnt.append(len(t))
        nv += len(p)
        nt += len(t)
    mats = ng.default_materials(len(parts))
    for k, (_, kind) in enumerate(parts):
        if kind == "floor":
            mats[k]["albedo"] = 0.5
        elif kind == "light":
            mats[k]["isEmitter"] = 1
            mats[k]["radiance"] = 100.0
        elif kind == "glass":
            mats[k]["type"] = ng.MAT_DIELECTRIC
            mats[k]["intIOR"], mats[k]["extIOR"] = 1.5, 1.0
        elif kind == "mirror":
            mats[k]["type"] = ng.MAT_MIRROR
        else:
            mats[k]["albedo"] = 0.3
    return ng.HostScene(np.concatenate(pos), np.concatenate(idx), np.array(off, np.uint32), np.array(cnt, np.uint32),
                        mats)


def floor_camera(w=16, h=16):
    # looking down at the floor around the origin; no pane in between
    return ng.lookat_camera([0.0, 0.6, 0.3], [0.0, 0.0, 0.0], [0.0, 0.0, -1.0], 20.0, w, h)


def pane_cosine():
    # |cos| between the pane normal (x) and the floor-to-light direction at
    # the centre of the view
    d = np.array([3.0, 1.1, 0.0])
    return d[0] / np.linalg.norm(d)
