"""GPU checks of the drop-in boundary: an XML scene loaded through the Nori
plugin surface (libnori_host.so) renders on the GPU like the oracle renders the
same packed scene, and Integrator::Li on caller rays matches the oracle's."""
import sys

import numpy as np
import pytest

import nori_gpu as ng
import orc
from test_xml_scene import CORNELL, LIGHT, QUAD, uv_sphere

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture()
def xml_scene(tmp_path):
    (tmp_path / "floor.obj").write_text(QUAD)
    (tmp_path / "light.obj").write_text(LIGHT)
    (tmp_path / "sphere.obj").write_text(uv_sphere())
    (tmp_path / "cornell.xml").write_text(CORNELL)
    s = ng.XmlScene(tmp_path / "cornell.xml")
    yield s
    s.close()


def test_xml_render_matches_oracle(xml_scene):
    # path_mis, seed 0, 16 spp, depth 8 through the plugin; the oracle renders the
    # packed scene (flat grey envmap included) with the same pcg32 streams
    _say("case: xml render 32x24 16spp")
    img = xml_scene.render()
    hs = xml_scene.host_scene()
    o, _ = orc.OracleScene(hs).render(hs.camera, 16, 8, seed=0)
    oi = o[:, :3] / 16
    e = orc.rel_mse(img, oi)
    _say(f"  relMSE {e:.3e}, means {img.mean(0)} {oi.mean(0)}")
    assert e < 1e-4
    assert np.isfinite(img).all() and img.mean() > 0.05


def test_li_matches_oracle(gpu, disney_small):
    # Integrator::Li on caller rays, same seeds on both sides
    gpu.upload(disney_small)
    o = orc.OracleScene(disney_small)
    cam = ng.lookat_camera_like(disney_small.camera, 64, 64)
    ys, xs = np.mgrid[0:64, 0:64]
    rng = np.random.default_rng(12)
    smp = np.c_[xs.reshape(-1) + rng.random(4096), ys.reshape(-1) + rng.random(4096), np.full(4096, 0.5),
                np.full(4096, 0.5)].astype(np.float32)
    rays = np.repeat(orc.camera_rays(cam, smp), 8, axis=0)
    seeds = rng.integers(0, 2 ** 63, len(rays), dtype=np.uint64)
    _say(f"case: li n={len(rays)}")
    g = gpu.li(rays, seeds, 12)
    c = o.li(rays, seeds, 12)
    e = orc.rel_mse(g, c)
    close = np.all(np.abs(g - c) <= 1e-3 * np.abs(c) + 1e-5, axis=1)
    _say(f"  relMSE {e:.3e}, close {close.mean():.4f}")
    assert close.mean() > 0.95
    assert e < 1e-3


def test_xml_li_is_unbiased(xml_scene):
    # Li through the plugin (seeds from the Independent sampler) against the
    # oracle's Li on the same rays with other seeds: equal in expectation
    hs = xml_scene.host_scene()
    cam = hs.camera
    ys, xs = np.mgrid[0:cam.height, 0:cam.width]
    n = cam.width * cam.height
    smp = np.c_[xs.reshape(-1) + 0.5, ys.reshape(-1) + 0.5, np.full(n, 0.5), np.full(n, 0.5)].astype(np.float32)
    rays = np.tile(orc.camera_rays(cam, smp), (64, 1))
    _say(f"case: xml li n={len(rays)}")
    g = xml_scene.li(rays, seed=5)
    c = orc.OracleScene(hs).li(rays, np.arange(len(rays), dtype=np.uint64) * 7919 + 3, 8)
    gm, cm = g.mean(0), c.mean(0)
    _say(f"  means {gm} {cm}")
    assert np.isfinite(g).all() and (g >= 0).all()
    assert np.allclose(gm, cm, rtol=0.05)


def test_hair_file_scene_matches_oracle(tmp_path):
    # a .hair strand file through the "hair" mesh plugin (tubes with fiber
    # tangents), Chiang hair BSDF, flat grey environment: same samples as the oracle
    from test_xml_scene import _hair_file
    rng = np.random.default_rng(3)
    strands = []
    for k in range(40):
        root = np.array([rng.uniform(-0.5, 0.5), 0.0, rng.uniform(-0.5, 0.5)])
        pts = [root + np.array([0.05 * np.sin(j + k), 0.1 * j, 0.03 * j]) for j in range(12)]
        strands.append(np.array(pts, np.float32))
    _hair_file(tmp_path / "h.hair", strands, default_thickness=0.02)
    xml = ("<scene><integrator type='path_mis'><integer name='maxDepth' value='8'/></integrator>"
           "<camera type='perspective'><float name='fov' value='40'/><integer name='width' value='32'/>"
           "<integer name='height' value='32'/><transform name='toWorld'><lookat origin='0, 0.6, 3' "
           "target='0, 0.6, 0' up='0, 1, 0'/></transform></camera>"
           "<sampler type='independent'><integer name='sampleCount' value='16'/></sampler>"
           "<mesh type='hair'><string name='filename' value='h.hair'/><integer name='sides' value='5'/>"
           "<bsdf type='hair_bsdf'/></mesh></scene>")
    s = ng.XmlScene(text=xml, directory=str(tmp_path))
    try:
        _say("case: hair file 40 strands x 11 segments, 32x32 16spp")
        img = s.render()
        hs = s.host_scene()
    finally:
        s.close()
    o, _ = orc.OracleScene(hs).render(hs.camera, 16, 8, seed=0)
    e = orc.rel_mse(img, o[:, :3] / 16)
    _say(f"  relMSE {e:.3e}")
    assert e < 1e-3 and img.mean() > 0.01


@pytest.mark.parametrize("fmt", ["png", "jpg"])
def test_textured_xml_scene_matches_oracle(tmp_path, fmt):
    # albedo (sRGB) and normal maps from PNG or JPEG files on a floor quad under
    # the flat grey environment and a small area light: same samples as the oracle
    rng = np.random.default_rng(5)
    alb = rng.integers(0, 256, (32, 32, 3))
    nrm = np.stack([rng.integers(96, 160, (16, 16)), rng.integers(96, 160, (16, 16)),
                    np.full((16, 16), 230)], -1)
    if fmt == "png":
        from png_writer import encode
        (tmp_path / "alb.png").write_bytes(encode(alb, 2, 8))
        (tmp_path / "nrm.png").write_bytes(encode(nrm, 2, 8))
    else:
        Image = pytest.importorskip("PIL.Image")
        Image.fromarray(alb.astype(np.uint8)).save(tmp_path / "alb.jpg", quality=90)
        Image.fromarray(nrm.astype(np.uint8)).save(tmp_path / "nrm.jpg", quality=90)
    (tmp_path / "floor.obj").write_text(QUAD)
    (tmp_path / "light.obj").write_text(LIGHT)
    xml = ("<scene><integrator type='path_mis'><integer name='maxDepth' value='6'/></integrator>"
           "<camera type='perspective'><float name='fov' value='50'/><integer name='width' value='32'/>"
           "<integer name='height' value='32'/><transform name='toWorld'><lookat origin='0, 2, 2.5' "
           "target='0, 0, 0' up='0, 1, 0'/></transform></camera>"
           "<sampler type='independent'><integer name='sampleCount' value='16'/></sampler>"
           "<mesh type='obj'><string name='filename' value='floor.obj'/><bsdf type='diffuse'>"
           f"<string name='albedoTexture' value='alb.{fmt}'/><string name='normalTexture' value='nrm.{fmt}'/></bsdf></mesh>"
           "<mesh type='obj'><string name='filename' value='light.obj'/><emitter type='area'>"
           "<color name='radiance' value='8 8 8'/></emitter></mesh></scene>")
    s = ng.XmlScene(text=xml, directory=str(tmp_path))
    try:
        _say(f"case: textured xml scene ({fmt}) 32x32 16spp")
        img = s.render()
        hs = s.host_scene()
    finally:
        s.close()
    assert len(hs.textures) == 2 and hs.materials[0]["albedoTexIndex"] == 0
    o, _ = orc.OracleScene(hs).render(hs.camera, 16, 6, seed=0)
    e = orc.rel_mse(img, o[:, :3] / 16)
    _say(f"  relMSE {e:.3e}")
    assert e < 1e-4 and img.mean() > 0.01
