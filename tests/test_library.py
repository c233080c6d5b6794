"""CPU checks of the product library: it loads, exports the C ABI, its host-side
scene generation and validation work, and compute entry points fail cleanly
without a GPU."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import nori_gpu as ng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions(header="nori_gpu.h", prefixes="gpu|scene|bvh8|guide"):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(rf"\b(nori_(?:{prefixes})_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ng.load_library()
    names = _declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(ng.EXPORTS) == names


def test_host_library_exports_every_declared_symbol():
    lib = ng.load_host_library()
    names = _declared_functions("nori_host.h", "xml|image|plugin|furnace")
    assert len(names) >= 11
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(ng.HOST_EXPORTS) == names


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = ng.load_library()
    h = ctypes.c_void_p()
    assert lib.nori_gpu_create(ctypes.byref(h), 0) == -5  # NORI_E_NODEVICE
    assert not h


@pytest.mark.parametrize("level", [1, 2, 4])
def test_icosphere_counts(level):
    s = ng.generate_scene(1, subdiv=level)
    counts = s.mesh_tri_count
    assert counts[5] == counts[6] == 20 * 4 ** level
    o, c = int(s.mesh_tri_offset[5]), int(counts[5])
    assert len(np.unique(s.indices[o:o + c])) == 10 * 4 ** level + 2
    # unit-sphere normals, vertices on the sphere of the generator's radius
    nrm = s.normals[np.unique(s.indices[o:o + c])]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-6)


def test_cornell_config():
    s = ng.generate_scene(1)
    assert len(s.indices) == 10 + 2 * 5120 + 2  # five quads, two level-4 icospheres, the light quad
    m = s.materials
    assert m["isEmitter"].sum() == 1
    light = m[m["isEmitter"] == 1][0]
    assert list(light["radiance"]) == [17, 12, 4]
    albedos = {tuple(float(x) for x in np.round(a.astype(np.float64), 2)) for a in m["albedo"]}
    assert {(0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15)} <= albedos
    assert s.camera.fov_deg == 40 and s.camera.width == s.camera.height == 256
    # emitter light faces down so next-event estimation from the room sees its front
    lo, lc = int(s.mesh_tri_offset[-1]), int(s.mesh_tri_count[-1])
    p = s.positions[s.indices[lo]]
    assert np.cross(p[1] - p[0], p[2] - p[0])[1] < 0


def test_disney_grid_config_small():
    s = ng.generate_scene(2, subdiv=2, grid=3, env_width=64, env_height=32)
    assert len(s.materials) == 1 + 9
    d = s.materials[1:]
    assert np.all(d["type"] == ng.MAT_DISNEY)
    assert np.all((d["roughness"] >= 0.05) & (d["roughness"] <= 1.0))
    assert np.all((d["anisotropic"] >= 0) & (d["anisotropic"] <= 0.8))
    for k in ("metallic", "specular", "sheen", "clearcoat"):
        assert np.all((d[k] >= 0) & (d[k] <= 1))
    assert s.env_rgb.shape == (32, 64, 3)
    full = ng.generate_scene(2, subdiv=1, grid=2)  # default 2048x1024 envmap
    sun = full.env_rgb.max(axis=2) == 5e4  # texels whose centre lies inside the 0.5 degree disc
    texel_sr = (2 * np.pi / 2048) * (np.pi / 1024) * np.sin(np.pi * (np.nonzero(sun)[0] + 0.5) / 1024)
    disc_sr = 2 * np.pi * (1 - np.cos(np.radians(0.5)))
    assert abs(texel_sr.sum() / disc_sr - 1) < 0.15


def test_full_size_disney_config_counts():
    s = ng.generate_scene(2, env_width=64, env_height=32)  # full geometry, small envmap
    assert len(s.indices) == 64 * 81920 + 2


def test_scene_generation_is_seeded():
    a = ng.generate_scene(2, subdiv=1, grid=2, env_width=32, env_height=16)
    b = ng.generate_scene(2, subdiv=1, grid=2, env_width=32, env_height=16)
    c = ng.generate_scene(2, seed=99, subdiv=1, grid=2, env_width=32, env_height=16)
    assert a.materials.tobytes() == b.materials.tobytes()
    assert a.materials.tobytes() != c.materials.tobytes()


def test_medium_config():
    sc = ng.generate_scene(4, grid=32, env_width=64, env_height=32, variant=2)
    assert len(sc.indices) == 8 and len(sc.media) == 1
    m = sc.media[0]
    assert m.density.shape == (32, 32, 32)
    assert m.density.min() == 0.0 and m.density.max() == 1.0
    assert 0.05 < (m.density > 0).mean() < 0.95
    assert np.allclose(np.array(m.sigma_s) + np.array(m.sigma_a), 20.0)
    assert np.allclose(np.array(m.sigma_s) / 20.0, 0.9)
    assert math.isclose(m.g, 0.8, rel_tol=1e-6) and m.brick == 16
    assert (sc.camera.width, sc.camera.height) == (512, 512)
    # seeded: the same options give the same grid
    assert np.array_equal(ng.generate_scene(4, grid=32, env_width=64, env_height=32).media[0].density, m.density)
