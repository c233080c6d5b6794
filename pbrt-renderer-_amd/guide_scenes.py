"""A scene where the light that reaches most surfaces arrives indirectly from a
small bright region, the case directional guiding is meant for: a closed room
(4 x 2.5 x 4, diffuse walls) with a square opening in its roof, lit by a
sky-and-sun envmap.  The sun (40 degrees above the horizon) shines through the
opening onto a patch of the left wall; the floor, the other walls and a 3 x 3
grid of rough Disney spheres (parameters drawn from the seed as in the
procedural Disney grid, roughness in [0.6, 1], no clearcoat) see the sun only
through that patch.  The camera is inside the
room.

BSDF sampling at a diffuse point finds the sunlit patch with probability about
its solid angle, and next-event estimation toward the sun is occluded there, so
plain path tracing is noisy; a guide that learns where the patch is from each
point can sample it directly."""
import numpy as np

import nori_gpu as ng

_EL, _AZ = np.radians(40.0), np.radians(10.0)
SUN = np.array([np.cos(_EL) * np.cos(_AZ), np.sin(_EL), np.cos(_EL) * np.sin(_AZ)])


def _quad(o, e1, e2):
    """Quad o, o+e1, o+e1+e2, o+e2; the face normal is e1 x e2."""
    o, e1, e2 = (np.asarray(x, np.float32) for x in (o, e1, e2))
    return np.stack([o, o + e1, o + e1 + e2, o + e2]), np.array([[0, 1, 2], [0, 2, 3]], np.uint32), None


def _sphere(c, r, n_lat=24, n_lon=48):
    """UV sphere with outward vertex normals."""
    th = np.linspace(0.0, np.pi, n_lat + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, n_lon + 1)[:-1]
    t, p = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(t) * np.cos(p), np.cos(t), np.sin(t) * np.sin(p)], -1).reshape(-1, 3)
    pos = np.asarray(c, np.float64) + r * n
    tris = []
    for i in range(n_lat):
        for j in range(n_lon):
            a, b = i * n_lon + j, i * n_lon + (j + 1) % n_lon
            c0, d0 = a + n_lon, b + n_lon
            if i > 0:
                tris.append([a, b, c0])
            if i < n_lat - 1:
                tris.append([b, d0, c0])
    return pos.astype(np.float32), np.array(tris, np.uint32), n.astype(np.float32)


def _env(w, h, sun_deg=3.0, sun_radiance=2000.0):
    """Sky gradient above the horizon, dim ground below, a sun disc of the given
    angular radius around SUN (directions as in the envmap lookup: theta from +y,
    phi from +x toward +z)."""
    y, x = np.mgrid[0:h, 0:w]
    th = np.pi * (y + 0.5) / h
    ph = 2.0 * np.pi * (x + 0.5) / w
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1)
    up = np.clip(d[..., 1], 0.0, 1.0)[..., None]
    sky = (1.0 - np.sqrt(up)) * np.array([0.85, 0.9, 1.0]) + np.sqrt(up) * np.array([0.25, 0.45, 0.95])
    sky = np.where(d[..., 1:2] > 0.0, sky, np.array([0.2, 0.18, 0.15]))
    disc = (d @ SUN) >= np.cos(np.radians(sun_deg))
    sky[disc] = sun_radiance
    return sky.astype(np.float32)


def sun_room(seed=0, width=96, height=96, env_w=512, env_h=256):
    rng = np.random.default_rng(seed)
    H, R, hole = 2.5, 2.0, 0.45
    parts = [
        (_quad([-R, 0, R], [2 * R, 0, 0], [0, 0, -2 * R]), "wall"),          # floor, normal +y
        (_quad([-R, 0, -R], [2 * R, 0, 0], [0, H, 0]), "wall"),              # back wall, normal +z
        (_quad([-R, 0, R], [0, 0, -2 * R], [0, H, 0]), "wall"),              # left wall, normal +x
        (_quad([R, 0, -R], [0, 0, 2 * R], [0, H, 0]), "wall"),               # right wall, normal -x
        (_quad([R, 0, R], [-2 * R, 0, 0], [0, H, 0]), "wall"),               # front wall, normal -z
        # roof around a square opening, normals -y
        (_quad([-R, H, -R], [R - hole, 0, 0], [0, 0, 2 * R]), "wall"),
        (_quad([hole, H, -R], [R - hole, 0, 0], [0, 0, 2 * R]), "wall"),
        (_quad([-hole, H, -R], [2 * hole, 0, 0], [0, 0, R - hole]), "wall"),
        (_quad([-hole, H, hole], [2 * hole, 0, 0], [0, 0, R - hole]), "wall"),
    ]
    for i in range(3):
        for j in range(3):
            parts.append((_sphere([(i - 1) * 1.1, 0.35, (j - 1) * 1.1 - 0.3], 0.35), "sphere"))
    pos, idx, nrm, off, cnt = [], [], [], [], []
    nv = nt = 0
    for (p, t, n), _ in parts:
        pos.append(p)
        idx.append(t + nv)
        if n is None:
            e = np.cross(p[1] - p[0], p[3] - p[0])
            n = np.tile(e / np.linalg.norm(e), (len(p), 1)).astype(np.float32)
        nrm.append(n)
        off.append(nt)
        cnt.append(len(t))
        nv += len(p)
        nt += len(t)
    mats = ng.default_materials(len(parts))
    for k, (_, kind) in enumerate(parts):
        if kind == "wall":
            mats[k]["albedo"] = 0.7
        else:
            mats[k]["type"] = ng.MAT_DISNEY
            mats[k]["albedo"] = rng.random(3)
            for f in ("metallic", "specular", "sheen"):
                mats[k][f] = rng.random()
            # rough lobes only: a glossy sphere would focus the sun into
            # caustics, whose fireflies no directional guide at a diffuse
            # vertex removes and which would swamp what guiding does gain
            mats[k]["clearcoat"] = 0.0
            mats[k]["roughness"] = 0.6 + 0.4 * rng.random()
            mats[k]["anisotropic"] = 0.8 * rng.random()
    cam = ng.lookat_camera([0.9, 1.3, 1.9], [-0.5, 0.5, -0.5], [0.0, 1.0, 0.0], 60.0, width, height)
    sc = ng.HostScene(np.concatenate(pos), np.concatenate(idx), np.array(off, np.uint32), np.array(cnt, np.uint32),
                      mats, normals=np.concatenate(nrm), env_rgb=_env(env_w, env_h), camera=cam)
    return sc
