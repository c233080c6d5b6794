"""CPU checks of the Nori plugin surface in libnori_host.so: the XML scene
format (tags, properties, transforms, errors with file and line), the OBJ
loader, the packing into the GPU path's flat scene, and image I/O.

Semantics follow src/cpu/parser.cpp, src/cpu/obj.cpp, src/cpu/proplist.cpp and
src/gpu/DXRApp_Scene.cpp of the reference; expected values here come from the
file contents and closed-form transforms, not from the library's own code."""
import math
import struct

import numpy as np
import pytest

import nori_gpu as ng

QUAD = "v -1 0 1\nv 1 0 1\nv 1 0 -1\nv -1 0 -1\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/1 2/2 3/3 4/4\n"
LIGHT = "v -0.25 1.99 -0.25\nv 0.25 1.99 -0.25\nv 0.25 1.99 0.25\nv -0.25 1.99 0.25\nf 1 2 3 4\n"


def uv_sphere(n_lat=8, n_lon=12):
    lines = []
    for i in range(n_lat + 1):
        th = math.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * math.pi * j / n_lon
            lines.append(f"v {math.sin(th) * math.cos(ph):.7f} {math.cos(th):.7f} {math.sin(th) * math.sin(ph):.7f}")
            lines.append(f"vn {math.sin(th) * math.cos(ph):.7f} {math.cos(th):.7f} {math.sin(th) * math.sin(ph):.7f}")
    for i in range(n_lat):
        for j in range(n_lon):
            a, b = i * n_lon + j + 1, i * n_lon + (j + 1) % n_lon + 1
            c, d = a + n_lon, b + n_lon
            lines.append(f"f {a}//{a} {c}//{c} {d}//{d} {b}//{b}")
    return "\n".join(lines) + "\n"


CORNELL = """<?xml version="1.0" encoding="utf-8"?>
<!-- a small test scene -->
<scene>
    <integrator type="path_mis">
        <integer name="maxDepth" value="8"/>
    </integrator>
    <camera type="perspective">
        <float name="fov" value="40"/>
        <transform name="toWorld">
            <lookat origin="0, 1, 3.7" target="0, 1, 0" up="0, 1, 0"/>
        </transform>
        <integer name="width" value="32"/>
        <integer name="height" value="24"/>
    </camera>
    <sampler type="independent">
        <integer name="sampleCount" value="16"/>
    </sampler>
    <mesh type="obj">
        <string name="filename" value="floor.obj"/>
        <bsdf type="diffuse">
            <color name="albedo" value="0.73 0.73 0.73"/>
        </bsdf>
    </mesh>
    <mesh type="obj">
        <string name="filename" value="light.obj"/>
        <emitter type="area">
            <color name="radiance" value="17 12 4"/>
        </emitter>
    </mesh>
    <mesh type="obj">
        <string name="filename" value="sphere.obj"/>
        <transform name="toWorld">
            <scale value="0.3 0.3 0.3"/>
            <translate value="0 0.3 0"/>
        </transform>
        <bsdf type="disney">
            <color name="baseColor" value="0.8 0.2 0.1"/>
            <float name="roughness" value="0.3"/>
            <float name="metallic" value="0.5"/>
        </bsdf>
    </mesh>
</scene>
"""


@pytest.fixture()
def scene_dir(tmp_path):
    (tmp_path / "floor.obj").write_text(QUAD)
    (tmp_path / "light.obj").write_text(LIGHT)
    (tmp_path / "sphere.obj").write_text(uv_sphere())
    (tmp_path / "cornell.xml").write_text(CORNELL)
    return tmp_path


def test_plugin_registry():
    names = set(ng.plugin_names())
    # every class the reference registers with NORI_REGISTER_CLASS, bar the
    # .hair strand loader, plus the GPU integrators
    expected = {"scene", "perspective", "homogeneous", "diffuse", "independent", "dielectric", "disney",
                "subsurface", "area", "gaussian", "hair_bsdf", "microfacet", "heterogeneous", "mirror", "obj",
                "path_mis", "path_guided", "direct"}
    assert expected <= names, expected - names


def test_load_scene_and_pack(scene_dir):
    s = ng.XmlScene(scene_dir / "cornell.xml")
    i = s.info()
    assert (i.width, i.height, i.spp, i.meshes, i.emitters) == (32, 24, 16, 3, 1)
    assert i.integrator.decode().startswith("path_mis[maxDepth=8")
    assert i.envmap_loaded == 0  # no envmap file next to the scene: flat grey fallback
    hs = s.host_scene()
    # quad: 4 vertices, 2 triangles; light: 4 / 2; sphere: 9*12 vertices, 8*12*2 triangles
    assert list(hs.mesh_tri_count) == [2, 2, 8 * 12 * 2]
    assert len(hs.positions) == 4 + 4 + 9 * 12
    # sphere positions: scale 0.3 then translate (0, 0.3, 0)
    sph = hs.positions[8:]
    assert np.allclose(np.linalg.norm(sph - [0, 0.3, 0], axis=1), 0.3, atol=1e-6)
    m = hs.materials
    assert list(m["type"]) == [ng.MAT_DIFFUSE, ng.MAT_DIFFUSE, ng.MAT_DISNEY]
    assert np.allclose(m["albedo"][0], 0.73) and np.allclose(m["albedo"][2], [0.8, 0.2, 0.1])
    assert m["isEmitter"][1] == 1 and np.allclose(m["radiance"][1], [17, 12, 4])
    assert math.isclose(m["roughness"][2], 0.3, rel_tol=1e-6) and math.isclose(m["metallic"][2], 0.5, rel_tol=1e-6)
    # OBJ V is flipped for the GPU path (src/gpu/DXRApp_Scene.cpp:402-408)
    assert np.allclose(hs.uvs[:4], [[0, 1], [1, 1], [1, 0], [0, 0]])
    # flat grey envmap
    assert hs.env_rgb.shape == (1, 1, 3) and np.allclose(hs.env_rgb, 0.5)
    # camera: Nori lookat
    ref = ng.lookat_camera([0, 1, 3.7], [0, 1, 0], [0, 1, 0], 40, 32, 24)
    assert np.allclose(np.array(hs.camera.to_world[:]), np.array(ref.to_world[:]), atol=1e-6)
    assert hs.camera.fov_deg == 40 and (hs.camera.width, hs.camera.height) == (32, 24)
    s.close()


def _scene(body, tmp_path, name="t.xml"):
    (tmp_path / name).write_text(body)
    return tmp_path / name


def _one_mesh(transform, tmp_path, obj=QUAD):
    (tmp_path / "m.obj").write_text(obj)
    xml = f"""<scene><camera type="perspective"/><mesh type="obj"><string name="filename" value="m.obj"/>
    <transform name="toWorld">{transform}</transform></mesh></scene>"""
    return ng.XmlScene(text=xml, directory=str(tmp_path)).host_scene()


def test_transform_composition(tmp_path):
    src = np.array([[-1, 0, 1], [1, 0, 1], [1, 0, -1], [-1, 0, -1]], np.float64)
    # operations compose as op * current: translate first, then scale
    hs = _one_mesh('<translate value="1 0 0"/><scale value="2 3 4"/>', tmp_path)
    assert np.allclose(hs.positions, (src + [1, 0, 0]) * [2, 3, 4], atol=1e-6)
    # rotation by +90 degrees about y maps x to -z
    hs = _one_mesh('<rotate axis="0 1 0" angle="90"/>', tmp_path)
    exp = np.stack([src[:, 2], src[:, 1], -src[:, 0]], 1)
    assert np.allclose(hs.positions, exp, atol=1e-6)
    # a matrix is given row by row
    hs = _one_mesh('<matrix value="1 0 0 5  0 1 0 6  0 0 1 7  0 0 0 1"/>', tmp_path)
    assert np.allclose(hs.positions, src + [5, 6, 7], atol=1e-6)


def test_obj_normals(tmp_path):
    # no normals in the file: angle-weighted face normals (+y for the quad)
    hs = _one_mesh("", tmp_path)
    assert np.allclose(hs.normals, [[0, 1, 0]] * 4, atol=1e-6)
    # normals from the file go through the inverse transpose of toWorld
    obj = "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 1 1 0\nf 1//1 2//1 3//1\n"
    hs = _one_mesh('<scale value="2 1 1"/>', tmp_path, obj)
    n = np.array([0.5, 1.0, 0.0]) / np.linalg.norm([0.5, 1.0, 0.0])
    assert np.allclose(hs.normals, [n] * 3, atol=1e-6)


def test_obj_quads_and_dedup(tmp_path):
    # two quads sharing an edge: 6 unique (position, uv, normal) triples, 4 triangles
    obj = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 2 0 0\nv 2 1 0\nf 1 2 3 4\nf 2 5 6 3\n"
    hs = _one_mesh("", tmp_path, obj)
    tri = hs.indices.reshape(-1, 3)
    assert len(hs.positions) == 6 and len(tri) == 4
    assert [list(hs.positions[t].sum(0)) for t in tri[:2]] == [[2, 1, 0], [1, 2, 0]]


def test_obj_vertex_and_triangle_order(tmp_path):
    # vertices in order of the first use of each (p, t, n) triple, triangles in
    # face order, polygons fanned from their first corner; negative indices
    # count back from the last position read
    obj = ("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 1.5 0\n"
           "f 3 4 1\nf -5 -4 -3 -1 -2\n")
    hs = _one_mesh("", tmp_path, obj)
    assert np.allclose(hs.positions, [[1, 1, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0], [0.5, 1.5, 0]])
    assert hs.indices.reshape(-1, 3).tolist() == [[0, 1, 2], [2, 3, 0], [2, 0, 4], [2, 4, 1]]


@pytest.mark.parametrize("xml,needle", [
    ("<scene><camera type='perspective'/>", "line 1"),
    ("<scene><camera type='nosuch'/></scene>", 'A constructor for class "nosuch" could not be found'),
    ("<scene>\n  <camera type='perspective' colour='red'/></scene>", 'unexpected attribute "colour"'),
    ("<scene><mesh type='obj'><string name='filename'/></mesh></scene>", 'missing attribute "value"'),
    ("<scene><camera type='perspective'><float name='fov' value='1'><integer name='x' value='1'/></float>"
     "</camera></scene>", 'requires a Nori object as parent'),
    ("<float name='x' value='1'/>", "must be a Nori object"),
    ("<scene><camera type='perspective'><integer name='fov' value='3'/></camera></scene>",
     "Property 'fov' has the wrong type"),
    ("<scene><camera type='perspective'><float name='fov' value='3x'/></camera></scene>",
     'Could not parse floating point value "3x"'),
    ("<scene><camera type='perspective'><transform name='toWorld'><float name='a' value='1'/></transform>"
     "</camera></scene>", "transform nodes can only contain transform operations"),
    ("<scene><bsdf type='diffuse'/></scene>", "Scene::addChild(<bsdf>) is not supported"),
    ("<scene><camera type='perspective'/><camera type='perspective'/></scene>", "only be one camera"),
    ("<scene><camera type='diffuse'/></scene>", "Unexpectedly constructed an object of type <bsdf>"),
    ("<scene><sampler type='independent'/></scene>", "No camera was specified"),
])
def test_parse_errors(xml, needle, tmp_path):
    with pytest.raises(ng.NoriError) as e:
        ng.XmlScene(_scene(xml, tmp_path))
    msg = str(e.value)
    assert needle in msg, msg
    assert "t.xml" in msg


def test_error_names_line_and_column(tmp_path):
    xml = "<scene>\n  <camera type='perspective'/>\n  <mesh type='obj'>\n    <string name='filename' value='none.obj'/>\n  </mesh>\n</scene>\n"
    with pytest.raises(ng.NoriError) as e:
        ng.XmlScene(_scene(xml, tmp_path))
    msg = str(e.value)
    assert "Unable to open OBJ file" in msg and "line 3, col 3" in msg, msg


def test_area_light_needs_radiance(tmp_path):
    (tmp_path / "m.obj").write_text(LIGHT)
    xml = "<scene><camera type='perspective'/><mesh type='obj'><string name='filename' value='m.obj'/>" \
          "<emitter type='area'/></mesh></scene>"
    with pytest.raises(ng.NoriError, match="Property 'radiance' is missing"):
        ng.XmlScene(_scene(xml, tmp_path))


def test_envmap_from_file(tmp_path):
    rng = np.random.default_rng(0)
    env = rng.random((4, 8, 3)).astype(np.float32) * 3
    ng.write_image(tmp_path / "sky.pfm", env, 8, 4)
    xml = "<scene><string name='envmap' value='sky.pfm'/><float name='envmapScale' value='2.5'/>" \
          "<camera type='perspective'/></scene>"
    s = ng.XmlScene(_scene(xml, tmp_path))
    assert s.info().envmap_loaded == 1
    hs = s.host_scene()
    assert np.array_equal(hs.env_rgb, env) and hs.env_scale == 2.5


def test_pfm_round_trip(tmp_path):
    img = np.random.default_rng(1).random((5, 7, 3)).astype(np.float32)
    ng.write_image(tmp_path / "a.pfm", img, 7, 5)
    assert np.array_equal(ng.read_image(tmp_path / "a.pfm"), img)
    raw = (tmp_path / "a.pfm").read_bytes()
    assert raw.startswith(b"PF\n7 5\n-1\n")
    # rows are stored bottom to top, little endian
    assert np.array_equal(np.frombuffer(raw[len(b"PF\n7 5\n-1\n"):], "<f4").reshape(5, 7, 3)[::-1], img)


def test_exr_layout(tmp_path):
    img = np.random.default_rng(2).random((3, 4, 3)).astype(np.float32)
    ng.write_image(tmp_path / "a.exr", img, 4, 3)
    raw = (tmp_path / "a.exr").read_bytes()
    assert struct.unpack("<II", raw[:8]) == (20000630, 2)
    end = raw.index(b"screenWindowWidth\x00float\x00") + len(b"screenWindowWidth\x00float\x00") + 4 + 4 + 1
    offsets = struct.unpack("<3Q", raw[end:end + 24])
    for y, off in enumerate(offsets):
        yy, size = struct.unpack("<ii", raw[off:off + 8])
        assert yy == y and size == 4 * 12
        planes = np.frombuffer(raw[off + 8:off + 8 + size], "<f4").reshape(3, 4)  # B, G, R
        assert np.array_equal(planes[::-1].T, img[y])


def _rgbe(c):
    m = max(c)
    if m < 1e-32:
        return bytes(4)
    e = math.frexp(m)[1]
    s = math.ldexp(1.0, -e) * 256.0
    return bytes([int(c[0] * s), int(c[1] * s), int(c[2] * s), e + 128])


def test_hdr_reader_flat_and_rle(tmp_path):
    rng = np.random.default_rng(3)
    w, h = 16, 3
    px = rng.random((h, w, 3)) * 4
    enc = [[_rgbe(px[y, x]) for x in range(w)] for y in range(h)]
    head = f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode()
    flat = head + b"".join(b"".join(r) for r in enc)
    # run-length scanlines: per channel, literal runs of <= 128 bytes
    rle = head
    for r in enc:
        rle += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            vals = bytes(p[c] for p in r)
            rle += bytes([len(vals)]) + vals
    expect = np.array([[[b[k] * math.ldexp(1.0, b[3] - 136) if b[3] else 0.0 for k in range(3)] for b in r]
                       for r in enc], np.float32)
    for name, data in (("flat.hdr", flat), ("rle.hdr", rle)):
        (tmp_path / name).write_bytes(data)
        assert np.array_equal(ng.read_image(tmp_path / name), expect), name


def test_media_packing(tmp_path):
    (tmp_path / "m.obj").write_text(QUAD)
    xml = """<scene><camera type='perspective'/>
      <mesh type='obj'><string name='filename' value='m.obj'/></mesh>
      <medium type='homogeneous'><color name='sigmaA' value='0.1 0.2 0.3'/><float name='sigmaS' value='0.5'/>
        <float name='g' value='0.4'/></medium></scene>"""
    hs = ng.XmlScene(text=xml, directory=str(tmp_path)).host_scene()
    m = hs.media[0]
    # no explicit bounds: the scene's bounding box (the quad, flat in y)
    assert m.vmin == [-1, 0, -1] and m.vmax == [1, 0, 1]
    assert np.allclose(m.sigma_a, [0.1, 0.2, 0.3]) and np.allclose(m.sigma_s, 0.5) and math.isclose(m.g, 0.4, rel_tol=1e-6)
    assert m.density is None
    # heterogeneous with a VOL1 grid: "VOL1", W, H, D, bbox (6 floats), W*H*D floats
    grid = np.random.default_rng(0).random((3, 4, 5)).astype(np.float32)
    with open(tmp_path / "d.vol", "wb") as f:
        f.write(b"VOL1" + struct.pack("<3I", 5, 4, 3) + struct.pack("<6f", 0, 0, 0, 1, 1, 1) + grid.tobytes())
    xml = """<scene><camera type='perspective'/>
      <medium type='heterogeneous'><string name='volume' value='d.vol'/>
        <point name='boundsMin' value='-1, 0, -2'/><point name='boundsMax' value='1, 3, 2'/></medium></scene>"""
    hs = ng.XmlScene(text=xml, directory=str(tmp_path)).host_scene()
    m = hs.media[0]
    assert m.vmin == [-1, 0, -2] and m.vmax == [1, 3, 2] and m.brick == 4
    assert np.allclose(m.sigma_s, 5.0) and math.isclose(m.g, 0.85, rel_tol=1e-6)
    assert np.array_equal(m.density, grid)


def test_subsurface_bsdf_record(tmp_path):
    # src/cpu/subsurface.cpp:1-92: albedo -> A, radius -> subsurface, g ->
    # anisotropic, tint -> sheen / sheenTint / clearcoat, type 6
    xml = ("<scene><camera type='perspective'/><mesh type='obj'><string name='filename' value='m.obj'/>"
           "<bsdf type='subsurface'><color name='albedo' value='0.9 0.6 0.5'/><float name='radius' value='0.25'/>"
           "<float name='g' value='0.3'/><float name='intIOR' value='1.4'/><color name='tint' value='1 0.8 0.6'/>"
           "<float name='roughness' value='0.2'/></bsdf></mesh></scene>")
    (tmp_path / "m.obj").write_text(QUAD)
    hs = ng.XmlScene(text=xml, directory=str(tmp_path)).host_scene()
    m = hs.materials[0]
    assert m["type"] == ng.MAT_SUBSURFACE
    assert np.allclose(m["albedo"], [0.9, 0.6, 0.5])
    assert m["subsurface"] == pytest.approx(0.25) and m["anisotropic"] == pytest.approx(0.3)
    assert m["intIOR"] == pytest.approx(1.4) and m["roughness"] == pytest.approx(0.2)
    assert (m["sheen"], m["sheenTint"], m["clearcoat"]) == pytest.approx((1.0, 0.8, 0.6))


def _hair_file(path, strands, thickness=None, default_thickness=0.1, flags_extra=0, signature=b"HAIR",
               npoints_override=None):
    """Writes a .hair file (Yuksel's format: 128-byte header, segments, points, thickness)."""
    pts = np.concatenate([np.asarray(s, np.float32) for s in strands])
    flags = 1 | 2 | (4 if thickness is not None else 0) | flags_extra
    npts = len(pts) if npoints_override is None else npoints_override
    hdr = signature + struct.pack("<IIII f f 3f", len(strands), npts, flags, 0, default_thickness, 1.0, 1, 1, 1)
    hdr += b"\0" * (128 - len(hdr))
    body = np.array([len(s) - 1 for s in strands], np.uint16).tobytes() + pts.tobytes()
    if thickness is not None:
        body += np.asarray(thickness, np.float32).tobytes()
    path.write_bytes(hdr + body)


HAIR_XML = ("<scene><camera type='perspective'/><mesh type='hair'><string name='filename' value='h.hair'/>"
            "<integer name='sides' value='{sides}'/>{extra}<bsdf type='hair_bsdf'/></mesh></scene>")


def test_hair_file_tubes(tmp_path):
    # src/cpu/hair.cpp:150-322: (points) x sides vertices, 2 x sides triangles per
    # segment, rings normal to the fiber tangent at radius thickness / 2,
    # u along the strand, v = (1 - cos angle) / 2
    s0 = [[0, 0, 0], [0, 1, 0], [0, 2, 0.5], [0, 3, 1.5]]
    s1 = [[1, 0, 0], [1, 0, 1], [1, 0, 2]]
    thick = [0.2, 0.2, 0.1, 0.1, 0.4, 0.4, 0.4]
    _hair_file(tmp_path / "h.hair", [s0, s1], thick)
    hs = ng.XmlScene(text=HAIR_XML.format(sides=4, extra=""), directory=str(tmp_path)).host_scene()
    assert len(hs.positions) == (4 + 3) * 4 and len(hs.indices) == (3 + 2) * 2 * 4
    assert hs.tangents is not None and hs.normals is not None and hs.uvs is not None
    # ring j of strand 0 around control point j: |p - c| = thickness / 2, normal
    # outward and orthogonal to the tangent, tangent = central difference
    p, n, t, uv = hs.positions, hs.normals, hs.tangents, hs.uvs
    for j in range(4):
        ring = slice(4 * j, 4 * j + 4)
        c = np.asarray(s0[j], np.float32)
        assert np.allclose(np.linalg.norm(p[ring] - c, axis=1), thick[j] / 2, atol=1e-5)
        assert np.allclose(np.linalg.norm(n[ring], axis=1), 1.0, atol=1e-5)
        assert np.allclose((n[ring] * t[ring]).sum(1), 0.0, atol=1e-5)
        assert np.allclose(np.cross(p[ring] - c, n[ring]), 0.0, atol=1e-5)
        d = np.asarray(s0[min(j + 1, 3)], np.float32) - np.asarray(s0[max(j - 1, 0)], np.float32)
        assert np.allclose(t[ring], d / np.linalg.norm(d), atol=1e-5)
        assert np.allclose(uv[ring, 0], j / 3)
        assert np.allclose(uv[ring, 1], [0.0, 0.5, 1.0, 0.5], atol=1e-6)  # angles 0, 90, 180, 270 degrees
    # the second strand starts at vertex 16 and its triangles index its own rings
    assert hs.indices[24:].min() >= 16
    assert hs.materials[0]["type"] == ng.MAT_HAIR


def test_hair_file_options_and_errors(tmp_path):
    s0 = [[0, 0, 0], [0, 1, 0]]
    s1 = [[1, 0, 0], [1, 1, 0], [1, 2, 0]]
    _hair_file(tmp_path / "h.hair", [s0, s1], default_thickness=0.5)
    hs = ng.XmlScene(text=HAIR_XML.format(sides=6, extra="<integer name='maxStrands' value='1'/>"
                                          "<float name='radiusScale' value='2'/>"),
                     directory=str(tmp_path)).host_scene()
    assert len(hs.positions) == 2 * 6 and len(hs.indices) == 2 * 6
    assert np.allclose(np.linalg.norm(hs.positions[:6] - np.array([0, 0, 0]), axis=1), 0.5)  # 0.5 / 2 * 2
    for kw, needle in [(dict(signature=b"HAIX"), "signature"), (dict(npoints_override=6), "points")]:
        _hair_file(tmp_path / "h.hair", [s0, s1], **kw)
        with pytest.raises(ng.NoriError) as e:
            ng.XmlScene(text=HAIR_XML.format(sides=6, extra=""), directory=str(tmp_path))
        assert needle in str(e.value)


def test_utf8_byte_order_mark_is_skipped(tmp_path):
    # scene files saved with a UTF-8 BOM and an XML declaration load as pugixml
    # loads them (the reference's own scenes/final_scene_real/scene2.xml starts so)
    (tmp_path / "q.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "s.xml").write_bytes(b"\xef\xbb\xbf<?xml version='1.0' encoding='utf-8'?>\n<scene>"
                                     b"<camera type='perspective'/><mesh type='obj'>"
                                     b"<string name='filename' value='q.obj'/></mesh></scene>")
    s = ng.XmlScene(str(tmp_path / "s.xml"))
    try:
        assert len(s.host_scene().materials) == 1
    finally:
        s.close()
