"""Scenes for the random-walk subsurface checks (shaders/Subsurface.hlsli,
shaders/Shaders.hlsl:513-612) under a uniform unit environment:

* a slab (closed box, 10 x 10 x L) seen at normal incidence through a 2-degree
  view: with scatter colour 0 the interior albedo is 1 - exp(0) = 0, so only
  unscattered light leaves, and with a smooth boundary (reflectance R) the
  pixel value is R + (1-R)^2 e^{-L/d} / (1 - R e^{-L/d}) (d = mean free path:
  each crossing survives e^{-L/d}, each inside reflection has probability R);
* a smooth sphere whose scatter colour 2 makes the albedo 1 - exp(-34) = 1 in
  float: nothing is absorbed, every walk leaves, and the sphere vanishes in
  the white furnace (value 1)."""
import numpy as np

import nori_gpu as ng


def sss_material(albedo, mfp, ior=1.5, rough=0.0, g=0.0):
    m = ng.default_materials(1)
    m["type"] = ng.MAT_SUBSURFACE
    m["albedo"] = albedo
    m["subsurface"] = mfp
    m["intIOR"], m["extIOR"] = ior, 1.0
    m["roughness"] = rough
    m["specular"] = 0.0
    m["anisotropic"] = g
    m["sheen"] = m["sheenTint"] = m["clearcoat"] = 1.0  # tint (1, 1, 1)
    return m


def box(half_x, half_y, half_z):
    p = np.array([[x, y, z] for x in (-half_x, half_x) for y in (-half_y, half_y) for z in (-half_z, half_z)],
                 np.float32)
    # vertex k = 4 * (x > 0) + 2 * (y > 0) + (z > 0); two triangles per face
    t = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],
                  [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]], np.uint32)
    return p, t


def slab_scene(thickness, mfp, ior=1.5):
    p, t = box(5.0, 5.0, thickness / 2)
    env = np.ones((8, 16, 3), np.float32)
    return ng.HostScene(p, t, [0], [len(t)], sss_material(0.0, mfp, ior), env_rgb=env)


def slab_camera(w=8, h=8):
    return ng.lookat_camera([0.0, 0.0, 10.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], 2.0, w, h)


def slab_expected(thickness, mfp, ior=1.5):
    R = ((ior - 1.0) / (ior + 1.0)) ** 2
    e = np.exp(-thickness / mfp)
    return R + (1.0 - R) ** 2 * e / (1.0 - R * e)


def sphere_scene(albedo, mfp, rough=0.0, g=0.0):
    sc1 = ng.generate_scene(1, subdiv=3)
    k = 5  # first icosphere of the Cornell generator
    o, c = int(sc1.mesh_tri_offset[k]), int(sc1.mesh_tri_count[k])
    tri = sc1.indices[o:o + c]
    used = np.unique(tri)
    remap = np.zeros(tri.max() + 1, np.uint32)
    remap[used] = np.arange(len(used))
    pos = sc1.positions[used]
    env = np.ones((8, 16, 3), np.float32)
    scene = ng.HostScene(pos, remap[tri], [0], [c], sss_material(albedo, mfp, rough=rough, g=g),
                         normals=sc1.normals[used], env_rgb=env)
    ctr = pos.mean(0)
    cam = ng.lookat_camera(ctr + np.array([0, 0, 2.5]), ctr, [0, 1, 0], 20.0, 32, 32)
    return scene, cam
