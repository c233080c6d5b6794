"""JPEG texture files of the XML scene format (src/gpu/DXRApp_Textures.cpp:34-45
loads them with stb_image, which the reference does not vendor).  The decoder
follows the IJG decoder (islow integer IDCT, fancy chroma upsampling, 16-bit
YCbCr conversion), so its RGBA8 is compared byte for byte with libjpeg-turbo's
through Pillow, on files Pillow writes: 4:4:4, 4:2:2 and 4:2:0 chroma,
greyscale, restart intervals, odd sizes down to 1x1.  The sizes cover the
upsampler's edge cases: chroma one and two samples wide (replicated rather
than filtered) and odd luma sizes whose chroma row or column is partial."""
import itertools

import numpy as np
import pytest

import nori_gpu as ng

Image = pytest.importorskip("PIL.Image")

QUAD = "v -1 0 1\nv 1 0 1\nv 1 0 -1\nv -1 0 -1\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/1 2/2 3/3 4/4\n"


def _decode(tmp_path, name):
    (tmp_path / "q.obj").write_text(QUAD)
    hs = ng.XmlScene(text="<scene><camera type='perspective'/><mesh type='obj'><string name='filename' value='q.obj'/>"
                          f"<bsdf type='diffuse'><string name='albedoTexture' value='{name}'/></bsdf></mesh></scene>",
                     directory=str(tmp_path)).host_scene()
    return hs.textures[0][0] if hs.textures else None


def _image(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    ramp = np.stack([xx * 7 % 256, yy * 11 % 256, (xx + yy) * 5 % 256], -1)
    return (rng.integers(0, 256, (h, w, 3)) // 3 + ramp // 2).astype(np.uint8)


SIZES = [(1, 1), (2, 3), (3, 4), (4, 2), (5, 6), (6, 1), (13, 9), (33, 17), (64, 48)]


@pytest.mark.parametrize("size,sub,quality", list(itertools.product(SIZES, [0, 1, 2], [20, 90])))
def test_colour_jpeg_matches_libjpeg(tmp_path, size, sub, quality):
    w, h = size
    Image.fromarray(_image(w, h, w * 31 + h)).save(tmp_path / "t.jpg", quality=quality, subsampling=sub)
    expected = np.asarray(Image.open(tmp_path / "t.jpg").convert("RGB"))
    got = _decode(tmp_path, "t.jpg")
    assert got is not None and got.shape == (h, w, 4)
    assert np.array_equal(got[..., :3], expected) and np.all(got[..., 3] == 255)


@pytest.mark.parametrize("size", [(1, 1), (7, 5), (40, 31)])
def test_grey_jpeg_matches_libjpeg(tmp_path, size):
    w, h = size
    Image.fromarray(_image(w, h, 5)).convert("L").save(tmp_path / "g.jpg", quality=75)
    expected = np.asarray(Image.open(tmp_path / "g.jpg"))
    got = _decode(tmp_path, "g.jpg")
    assert got is not None
    for c in range(3):
        assert np.array_equal(got[..., c], expected)


@pytest.mark.parametrize("kw", [dict(restart_marker_blocks=3), dict(restart_marker_rows=1),
                                dict(restart_marker_blocks=1, subsampling=2)])
def test_restart_intervals(tmp_path, kw):
    Image.fromarray(_image(45, 37, 9)).save(tmp_path / "r.jpg", quality=80, **kw)
    assert b"\xff\xdd" in (tmp_path / "r.jpg").read_bytes()
    expected = np.asarray(Image.open(tmp_path / "r.jpg").convert("RGB"))
    got = _decode(tmp_path, "r.jpg")
    assert got is not None and np.array_equal(got[..., :3], expected)


def test_flat_colour_is_exact(tmp_path):
    # one DC coefficient per block: every decoder returns the same flat value
    Image.fromarray(np.full((16, 24, 3), (200, 120, 40), np.uint8)).save(tmp_path / "f.jpg", quality=100,
                                                                         subsampling=0)
    got = _decode(tmp_path, "f.jpg")
    expected = np.asarray(Image.open(tmp_path / "f.jpg").convert("RGB"))
    assert np.array_equal(got[..., :3], expected)
    assert np.all(np.abs(got[..., :3].astype(int) - (200, 120, 40)) <= 1)


def test_progressive_and_corrupt_files_leave_the_map_unset(tmp_path):
    Image.fromarray(_image(20, 20, 1)).save(tmp_path / "p.jpg", progressive=True)
    assert _decode(tmp_path, "p.jpg") is None
    data = (tmp_path / "p.jpg").read_bytes()
    Image.fromarray(_image(20, 20, 1)).save(tmp_path / "b.jpg")
    good = (tmp_path / "b.jpg").read_bytes()
    (tmp_path / "c.jpg").write_bytes(good[:len(good) // 3])  # cut inside the frame
    _decode(tmp_path, "c.jpg")  # must not crash; a truncated scan decodes as zeros or fails
    assert data[:2] == b"\xff\xd8"


def _dc_drift_jpeg(blocks):
    """Greyscale baseline JPEG, 8 x (8 blocks) pixels, written by hand: a 16-bit
    quantisation table of 65535s, a DC Huffman table whose one code '0' means
    category 11, an AC table whose one code '0' is end-of-block, and every
    block's DC difference +2047, so the DC predictor climbs by 2047 a block."""
    def seg(marker, body):
        return bytes([0xFF, marker]) + (len(body) + 2).to_bytes(2, "big") + body
    dqt = seg(0xDB, bytes([0x10]) + b"\xff\xff" * 64)
    sof = seg(0xC0, bytes([8]) + (8).to_bytes(2, "big") + (8 * blocks).to_bytes(2, "big") + bytes([1, 1, 0x11, 0]))
    counts = bytes([1] + [0] * 15)
    dht = seg(0xC4, bytes([0x00]) + counts + bytes([11])) + seg(0xC4, bytes([0x10]) + counts + bytes([0]))
    sos = seg(0xDA, bytes([1, 1, 0x00, 0, 63, 0]))
    bits = ("0" + "1" * 11 + "0") * blocks
    bits += "1" * (-len(bits) % 8)
    data = bytearray()
    for i in range(0, len(bits), 8):
        b = int(bits[i:i + 8], 2)
        data += bytes([b, 0]) if b == 0xFF else bytes([b])
    return b"\xff\xd8" + dqt + sof + dht + sos + bytes(data) + b"\xff\xd9"


def test_dc_predictor_drift_is_rejected(tmp_path):
    # 8 blocks: predictor 16376, coefficient 16376 * 65535 (beyond 32 bits
    # before the IDCT): decodes without overflow
    (tmp_path / "ok.jpg").write_bytes(_dc_drift_jpeg(8))
    got = _decode(tmp_path, "ok.jpg")
    assert got is not None and got.shape == (8, 64, 4)
    # 40 blocks: the predictor passes 32767 at block 17, which no 8-bit image
    # produces: the file is refused and the map left unset
    (tmp_path / "drift.jpg").write_bytes(_dc_drift_jpeg(40))
    assert _decode(tmp_path, "drift.jpg") is None
