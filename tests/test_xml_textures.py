"""Texture files of the XML scene format (albedoTexture and the other map
properties; src/gpu/DXRApp_Textures.cpp:34-45, 452-486): PNG of every colour
type and bit depth and every scanline filter, binary PNM, one texture per
resolved path, sRGB for albedo maps only, and files that do not decode leave
the map unset."""
import numpy as np
import pytest

import nori_gpu as ng
from png_writer import encode

QUAD = "v -1 0 1\nv 1 0 1\nv 1 0 -1\nv -1 0 -1\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/1 2/2 3/3 4/4\n"


def _scene(tmp_path, bsdf_props, nmesh=1):
    (tmp_path / "q.obj").write_text(QUAD)
    meshes = "".join(f"<mesh type='obj'><string name='filename' value='q.obj'/><bsdf type='diffuse'>{p}</bsdf></mesh>"
                     for p in (bsdf_props if isinstance(bsdf_props, list) else [bsdf_props] * nmesh))
    return ng.XmlScene(text=f"<scene><camera type='perspective'/>{meshes}</scene>", directory=str(tmp_path)).host_scene()


def _rgba(samples, ctype, depth, palette=None, trns=None):
    """The RGBA8 a decoder must produce (16-bit: high byte; <8-bit grey: scaled to 255)."""
    s = samples.astype(np.int64)
    if ctype == 3:
        pal = np.asarray(palette)
        out = np.concatenate([pal[s[..., 0]], np.full(s.shape[:2] + (1,), 255)], -1)
        if trns is not None:
            t = np.full(len(pal), 255)
            t[:len(trns)] = trns
            out[..., 3] = t[s[..., 0]]
        return out.astype(np.uint8)
    if depth == 16:
        s = s >> 8
    elif depth < 8:
        s = s * 255 // ((1 << depth) - 1)
    ch = s.shape[2]
    if ch <= 2:
        out = np.concatenate([s[..., :1]] * 3 + [s[..., 1:2] if ch == 2 else np.full(s.shape[:2] + (1,), 255)], -1)
    else:
        out = np.concatenate([s[..., :3], s[..., 3:4] if ch == 4 else np.full(s.shape[:2] + (1,), 255)], -1)
    return out.astype(np.uint8)


CASES = [(0, 1), (0, 2), (0, 4), (0, 8), (0, 16), (2, 8), (2, 16), (3, 1), (3, 4), (3, 8), (4, 8), (4, 16),
         (6, 8), (6, 16)]


@pytest.mark.parametrize("ctype,depth", CASES)
def test_png_decoding(tmp_path, ctype, depth):
    rng = np.random.default_rng(ctype * 100 + depth)
    h, w = 5, 7
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    top = (1 << depth) if ctype != 3 else min(1 << depth, 12)
    samples = rng.integers(0, top, (h, w, ch))
    palette = rng.integers(0, 256, (12, 3)) if ctype == 3 else None
    trns = rng.integers(0, 256, 5) if ctype == 3 else None
    (tmp_path / "t.png").write_bytes(encode(samples, ctype, depth, palette, trns))
    hs = _scene(tmp_path, "<string name='albedoTexture' value='t.png'/>")
    assert hs.materials[0]["albedoTexIndex"] == 0 and len(hs.textures) == 1
    img, srgb = hs.textures[0]
    assert srgb and img.shape == (h, w, 4)
    assert np.array_equal(img, _rgba(samples, ctype, depth, palette, trns))


def test_pnm_and_map_kinds_and_dedup(tmp_path):
    px = np.arange(3 * 4 * 3, dtype=np.uint8).reshape(3, 4, 3) * 7
    (tmp_path / "c.ppm").write_bytes(b"P6\n# comment\n4 3\n255\n" + px.tobytes())
    (tmp_path / "g.pgm").write_bytes(b"P5 4 3 255\n" + px[..., 0].tobytes())
    hs = _scene(tmp_path, ["<string name='albedoTexture' value='c.ppm'/><string name='normalTexture' value='g.pgm'/>",
                           "<string name='roughnessTexture' value='c.ppm'/>"])
    m = hs.materials
    assert len(hs.textures) == 2
    assert m[0]["albedoTexIndex"] == 0 and m[0]["normalTexIndex"] == 1 and m[1]["roughnessTexIndex"] == 0
    assert hs.textures[0][1] and not hs.textures[1][1]
    assert np.array_equal(hs.textures[0][0][..., :3], px) and np.all(hs.textures[0][0][..., 3] == 255)
    assert np.array_equal(hs.textures[1][0][..., 0], px[..., 0])


def test_files_that_do_not_load_leave_the_map_unset(tmp_path):
    (tmp_path / "x.jpg").write_bytes(b"\xff\xd8\xff\xe0 not decoded here")
    hs = _scene(tmp_path, "<string name='albedoTexture' value='x.jpg'/><string name='alphaTexture' value='none.png'/>")
    assert hs.materials[0]["albedoTexIndex"] == ng.NO_TEXTURE and hs.materials[0]["alphaTexIndex"] == ng.NO_TEXTURE
    assert len(hs.textures) == 0
