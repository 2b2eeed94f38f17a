"""The .ymtmp scene parser (ymrt_ymtmp_load, ymrt_amd/csrc/ymtmp.cpp) against
the oracle's restatement (oracle/ymtmp.py) and against what pins that:

  * transforms: glm itself, built from the reference tree into oracle/_ref
    (bit for bit), and closed forms (translate * scale, quarter turns);
  * the fixture tests/golden/CornellBox_Test.ymtmp (YMRT's own test scene,
    src/model/CornellBox_Test.ymtmp): block counts and values read off the
    file;
  * the statement rules of src/core/YMParser.cpp on small texts, each case
    naming the rule it exercises.
No GPU is needed: the parser is host code in the library.
"""
import ctypes as C
import math
import os
import struct
import zlib

import numpy as np
import pytest

import oracle
import ymrt_amd
from oracle import ymtmp as oymtmp
from ymrt_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FIXTURE = os.path.join(HERE, "golden", "CornellBox_Test.ymtmp")
REF_GLM = os.path.join(ROOT, "oracle", "_ref", "libref_glm.so")


def _mat(cm):
    return np.asarray(cm, np.float32).reshape(4, 4).T


def _product_text(text, scene_path=None):
    return ymrt_amd.YmtmpFile(text=text, scene_path=scene_path)


def _xform_block(name, t, s, r, parent=None):
    v = lambda a: "<%s>" % ",".join(repr(float(x)) for x in a)
    p = '\tparent_transform_index = "%s";\n' % parent if parent else ""
    return ('Begin(Transform);\n\tname = "%s";\n\ttranslate = %s;\n\tscale = %s;\n\trotate = %s;\n%sEnd(Transform);\n'
            % (name, v(t), v(s), v(r), p))


# ------------------------------------------------------------- transforms
def _random_transforms(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = np.round(rng.uniform(-20, 20, 3), 4)
        s = np.round(rng.uniform(0.05, 8, 3), 4)
        r = np.round(rng.uniform(0, 360, 3), 3)
        parent = int(rng.integers(0, i)) if i > 0 and rng.random() < 0.5 else None
        out.append((t, s, r, parent))
    return out


@pytest.mark.skipif(not os.path.exists(REF_GLM), reason="oracle/_ref not built (needs the reference tree's glm)")
def test_transforms_equal_reference_glm():
    """InitTransformMatrix's composition through the reference tree's glm
    (oracle/_ref) equals the oracle's restatement and the library's parser
    bit for bit, parents included."""
    ref = C.CDLL(REF_GLM)
    ref.ref_glm_transform.argtypes = [C.c_void_p] * 6
    xs = _random_transforms(40, 1)
    text = ""
    ref_out = []
    for i, (t, s, r, parent) in enumerate(xs):
        t32, s32, r32 = (np.asarray(a, np.float32) for a in (t, s, r))
        par = None if parent is None else ref_out[parent][0]
        o, inv = np.zeros(16, np.float32), np.zeros(16, np.float32)
        ref.ref_glm_transform(t32.ctypes.data, s32.ctypes.data, r32.ctypes.data,
                              None if par is None else par.ctypes.data, o.ctypes.data, inv.ctypes.data)
        ref_out.append((o, inv))
        oo, oi = oymtmp.glm_transform(t32, s32, r32, par)
        assert np.array_equal(oo, o) and np.array_equal(oi, inv), i
        text += _xform_block("x%d" % i, t32, s32, r32, None if parent is None else "x%d" % parent)
    f = _product_text(text)
    for i, (o, inv) in enumerate(ref_out):
        k, idx = f.find("x%d" % i)
        po, pi = f.transform(idx)
        assert k == "transform" and np.array_equal(po, _mat(o)) and np.array_equal(pi, _mat(inv)), i


def test_transform_closed_forms():
    """T * S is exact; a quarter turn about z maps x to y; M * M^-1 = I."""
    o, i = oymtmp.glm_transform((1.5, -2.0, 3.25), (2.0, 4.0, 0.5), (0.0, 0.0, 0.0))
    M, I = _mat(o), _mat(i)
    assert np.array_equal(M, np.array([[2, 0, 0, 1.5], [0, 4, 0, -2], [0, 0, 0.5, 3.25], [0, 0, 0, 1]], np.float32))
    assert np.array_equal(I[:3, :3], np.diag([0.5, 0.25, 2.0]).astype(np.float32))
    assert np.allclose(I[:3, 3], (-0.75, 0.5, -6.5))
    o, _ = oymtmp.glm_transform((0, 0, 0), (1, 1, 1), (0.0, 0.0, 90.0))
    assert np.allclose(_mat(o) @ (1, 0, 0, 0), (0, 1, 0, 0), atol=1e-7)
    # Rz applied after Ry after Rx (T Rz Ry Rx S): x-rotation first
    o, _ = oymtmp.glm_transform((0, 0, 0), (1, 1, 1), (90.0, 0.0, 90.0))
    assert np.allclose(_mat(o) @ (0, 1, 0, 0), (0, 0, 1, 0), atol=1e-6)   # Rx: y -> z, then Rz keeps z
    assert np.allclose(_mat(o) @ (1, 0, 0, 0), (0, 1, 0, 0), atol=1e-6)   # Rx keeps x, Rz: x -> y
    for t, s, r, _ in _random_transforms(20, 2):
        o, i = oymtmp.glm_transform(t, s, r)
        assert np.allclose(_mat(o).astype(np.float64) @ _mat(i), np.eye(4), atol=2e-5)


# ------------------------------------------------------------- the fixture
@pytest.fixture(scope="module")
def cornell():
    return oymtmp.load(FIXTURE)


def test_fixture_known_answers(cornell):
    """Values read off CornellBox_Test.ymtmp."""
    L, sa = cornell, cornell.arrays
    assert L.warnings == []
    assert (len(sa.meshes), len(sa.spheres), len(sa.instances), len(sa.materials), len(sa.textures),
            len(L.transforms), len(sa.distant_lights)) == (1, 1, 9, 9, 9, 11, 1)
    cam = sa.camera
    assert cam["position"] == (0.0, 0.0, 17.0) and cam["w"] == (0.0, 0.0, 1.0)
    assert cam["fov"] == float(np.float32(np.float32(34.999989) * np.float32(0.017453292519943295)))
    m = sa.meshes[0]
    assert m["pidx"].shape == (12, 3) and m["pos"].shape == (36, 3) and m["uv"].shape == (36, 2)
    assert np.array_equal(np.sort(np.abs(m["pos"]).ravel()), np.full(108, 0.5, np.float32))
    k, gi = L.names["ground"]
    inst = sa.instances[gi]
    assert np.array_equal(inst["o2w"], np.array([[135, 0, 0, 0], [0, 1, 0, -5.55], [0, 0, 135, 0], [0, 0, 0, 1]],
                                                  np.float32))
    mat = sa.materials[inst["material"]]
    checker = sa.textures[mat.diffuse_albedo_tex]
    assert checker["type"] == abi.TEXTURE_CHECKERBOARD and checker["params"]["frequency"] == 63.0
    black, white = (sa.textures[c]["value"] for c in checker["child"])
    assert np.allclose(black, 0.05) and np.allclose(white, 0.95)
    bump = sa.textures[mat.bump_mapping_tex]
    assert [sa.textures[c]["value"][0] for c in bump["child"]] == [float(np.float32(0.15)), float(np.float32(0.35))]
    assert math.isclose(mat.ior_n, 1.65, rel_tol=1e-7)
    # children of cornellbox_transform (the identity) keep their own matrix
    _, bw = L.names["back_wall"]
    assert np.array_equal(sa.instances[bw]["o2w"][:3, 3], np.array([0, 0, -10.025], np.float32))
    _, li = L.names["box_light"]
    lm = sa.materials[sa.instances[li]["material"]]
    assert lm.type == abi.MATERIAL_LIGHT and lm.intensity == 10.0
    _, si = L.names["righe_sphere"]
    assert sa.instances[si]["geometry_type"] == abi.GEOMETRY_SPHERE and sa.spheres[0] == 1.0
    d = sa.distant_lights[0]
    assert d["intensity"] == 2.0 and d["theta"] == 5.0


def _same_arrays(a, b):
    assert len(a.meshes) == len(b.meshes)
    for x, y in zip(a.meshes, b.meshes):
        assert x.keys() == y.keys()
        for k in x:
            assert np.array_equal(x[k], y[k]), k
    assert a.spheres == b.spheres
    assert len(a.materials) == len(b.materials)
    for x, y in zip(a.materials, b.materials):
        if y.type == abi.MATERIAL_LIGHT:   # LightMtl carries colour and intensity only
            assert x.type == y.type and list(x.light_color) == list(y.light_color) and x.intensity == y.intensity
        else:
            assert bytes(x) == bytes(y)
    assert len(a.textures) == len(b.textures)
    for x, y in zip(a.textures, b.textures):
        assert x["type"] == y["type"] and x["child"] == y["child"], (x, y)
        assert np.array_equal(np.float32(x["value"]), np.float32(y["value"]))
        for k, v in y["params"].items():
            assert np.float32(x["params"].get(k, v)) == np.float32(v), k
    assert len(a.instances) == len(b.instances)
    for x, y in zip(a.instances, b.instances):
        for k in ("mesh", "geometry_type", "material", "inner_volume", "boundary"):
            assert x[k] == y[k], k
        assert np.array_equal(x["o2w"], y["o2w"]) and np.array_equal(x["w2o"], y["w2o"])
    assert len(a.distant_lights) == len(b.distant_lights)
    for x, y in zip(a.distant_lights, b.distant_lights):
        assert x["color"] == y["color"] and x["intensity"] == y["intensity"] and x["theta"] == y["theta"]
        assert np.array_equal(x["o2w"], y["o2w"]) and np.array_equal(x["w2o"], y["w2o"])
    for k in ("position", "w"):
        assert np.array_equal(np.float32(a.camera[k]), np.float32(b.camera[k])), k
    assert np.float32(a.camera["fov"]) == np.float32(b.camera["fov"])


def test_loader_matches_oracle_on_fixture(cornell):
    f = ymrt_amd.YmtmpFile(FIXTURE)
    assert f.warnings == []
    _same_arrays(f.arrays(), cornell.arrays)
    for name, (kind, index) in cornell.names.items():
        assert f.find(name) == (kind, index), name


# --------------------------------------------------------- statement rules
BASE = ('Begin(Transform); name = "t"; End(Transform);\n'
        'Begin(ConstantTextureRgb); name = "c0"; color = <0.1,0.2,0.3>; End(ConstantTextureRgb);\n'
        'Begin(ConstantTextureFloat); name = "c1"; End(ConstantTextureFloat);\n')


def _both(text):
    """(product YmtmpFile or exception, oracle Loaded or exception)"""
    try:
        p = _product_text(text)
    except ymrt_amd.YmrtError as e:
        p = e
    try:
        o = oymtmp.load_text(text)
    except oymtmp.ParseError as e:
        o = e
    return p, o


@pytest.mark.parametrize("text,why", [
    ("Begin(Camera);\nfov = 40;\nBegin(Sphere);", "nested Begin (check_statements_valid)"),
    ("Begin(Sphere);\nname = \"s\";\nEnd(Camera);", "End of another type"),
    ("Begin(Sphere);\nname = \"s\";", "unclosed block"),
    ("name = \"s\";", "attribute outside a block"),
    ("Begin(Sphere);\nfov = 1;\nEnd(Sphere);", "attribute of another type"),
    ("Begin(Sphere);\nradius = 1..5;\nEnd(Sphere);", "value of no accepted shape"),
    ("Begin(Sphere);\nradius = 1e5;\nEnd(Sphere);", "exponents are not part of a number (is_numeric_float)"),
    ("Begin(Sphere)\nname = \"s\";\nEnd(Sphere);", "missing ';' joins two statements"),
    ("Begin(Bogus);\nEnd(Bogus);", "unknown block type"),
    ('Begin(NoiseTextureWave);\nname = "w";\nfreq_0 = 1.0;\nEnd(NoiseTextureWave);',
     "attribute names hold letters and '_' only (is_attribute_assign), so the wave's freq_0 is illegal"),
    ("Begin(Sphere);\nradius = <1,2>;\nEnd(Sphere);", "vector of two numbers"),
])
def test_statement_errors_fail_the_load(text, why):
    p, o = _both(text)
    assert isinstance(p, ymrt_amd.YmrtError), why
    assert isinstance(o, oymtmp.ParseError), why


@pytest.mark.parametrize("text,nwarn,why", [
    ('Begin(Sphere);\nradius = 2.0;\nEnd(Sphere);', 1, "a block without a name is skipped"),
    ('Begin(Sphere);\nname = "t";\nEnd(Sphere);', 1, "a name declared twice"),
    ('Begin(CheckerBoardTexture);\nname = "k";\ntexture_black_idx = "c0";\nEnd(CheckerBoardTexture);', 1,
     "a two-input texture with one input"),
    ('Begin(CheckerBoardTexture);\nname = "k";\ntexture_black_idx = "nope";\ntexture_white_idx = "c0";\n'
     'End(CheckerBoardTexture);', 1, "an undeclared name"),
    ('Begin(PrimitiveInstance);\nname = "i";\ntransform_index = "t";\nEnd(PrimitiveInstance);', 1,
     "an instance without geometry"),
    ('Begin(DistantLight);\nname = "d";\nEnd(DistantLight);', 1, "a distant light without transform"),
    ('Begin(Mesh);\nname = "m";\nmesh_triangles = [0,1,2];\nmesh_position_indices = [0,1,2];\n'
     'mesh_positions = [0,0,0,1,0,0,0,1,0];\nEnd(Mesh);', 1, "a mesh without normals"),
])
def test_reader_errors_skip_the_block(text, nwarn, why):
    p, o = _both(BASE + text)
    assert len(p.warnings) == nwarn and len(o.warnings) == nwarn, why
    _same_arrays(p.arrays(), o.arrays)


def test_values_defaults_and_clamps():
    """Whitespace and foreign characters vanish, '#' lines are comments,
    defaults and clamps of the readers, lacunarity of fBm ignored by its
    reader, noise normalized by default, Cube built in."""
    text = BASE + """
# a comment line ; with = odd ( characters
Begin(Transform);
  na me = "s c";   translate = <1,2,3>;  scale = <0,-1,2>; rotate = <-5,400,30>;
  parent_transform_index = "t";
End(Transform);
Begin(NoiseTextureFBM); name = "f"; texture_black_idx = "c0"; texture_white_idx = "c1";
  frequency = -3; lacunarity = 9.5; gain = 0.25f; layer_count = 3; End(NoiseTextureFBM);
Begin(NoiseTexture); name = "n"; texture_black_idx = "c0"; texture_white_idx = "c1"; End(NoiseTexture);
Begin(NoiseTextureMarble); name = "mb"; texture_black_idx = "c0"; texture_white_idx = "c1"; y_turb = false;
  frequency = -2; End(NoiseTextureMarble);
Begin(DefaultMaterial); name = "m"; roughness_x = 1.5; ior_n = -1; metalness = -2; coat_ior = 0;
  ior_priority = 3; diffuse_albedo_tex = "f"; alpha_x_tex = "n"; coat_normal_tex_type = "bump_map";
End(DefaultMaterial);
Begin(Cube); name = "box"; End(Cube);
Begin(Sphere); name = "sp"; radius = -1; End(Sphere);
Begin(PrimitiveInstance); name = "i0"; geometry_index = "box"; transform_index = "sc"; material_index = "m";
End(PrimitiveInstance);
Begin(PrimitiveInstance); name = "i1"; geometry_index = "sp"; transform_index = "t"; material_index = "m";
  treat_as_boundary = true; End(PrimitiveInstance);
Begin(Camera); from = <1,2,3>; look = <1,2,2>; fov = 0; End(Camera);
"""
    p, o = _both(text)
    assert p.warnings == [] and o.warnings == []
    _same_arrays(p.arrays(), o.arrays)
    sa = o.arrays
    assert o.names["sc"] == ("transform", 1)
    _, tf = o.names["f"]
    fb = sa.textures[tf]["params"]
    assert fb["frequency"] == np.float32(0.001) and fb["lacunarity"] == 2.0 and fb["gain"] == 0.25
    assert fb["layer_count"] == 3
    assert sa.textures[o.names["n"][1]]["params"]["normalized"] == 1
    mb = sa.textures[o.names["mb"][1]]["params"]
    assert mb["turbulence_axes"] == 5 and mb["frequency"] == -2.0          # no clamp in the marble reader
    m = sa.materials[0]
    assert (m.roughness_x, m.metalness, m.ior_priority) == (1.0, 0.0, 3)
    assert m.ior_n == np.float32(0.0001) and m.coat_ior == np.float32(0.0001)
    assert m.roughness_x_tex == o.names["n"][1] and m.diffuse_albedo_tex == tf
    assert sa.spheres == [float(np.float32(0.0001))]
    assert sa.instances[1]["boundary"] == 1
    # scale clamps to 0.0001, rotate to [0, 360]
    k, xi = p.find("sc")
    ref = oymtmp.glm_transform((1, 2, 3), (0.0001, 0.0001, 2), (0, 360, 30), o.transforms[0][0])
    assert np.array_equal(p.transform(xi)[0], _mat(ref[0]))
    assert sa.camera["fov"] == float(np.float32(np.float32(0.01) * np.float32(0.017453292519943295)))


def test_density_texture_volume():
    """A Volume's density_texture_index names the texture whose x channel is
    its density (src/core/YMParser.cpp:2210, Volume::density_texture_idx);
    without it the volume is homogeneous.  Both readers agree."""
    text = (BASE + 'Begin(Volume); name = "v"; transform_index = "t"; sigma_t = <0.5,0.5,0.5>; '
            'density_texture_index = "c1"; End(Volume);\n'
            'Begin(Volume); name = "w"; transform_index = "t"; End(Volume);')
    f = _product_text(text)
    sd = f.desc()
    assert sd.volume_count == 2
    c1 = f.find("c1")[1]
    assert sd.volumes[0].density_texture == c1 and sd.volumes[0].density_grid == -1
    assert sd.volumes[1].density_texture == -1
    o = oracle.ymtmp.load_text(text)
    assert [v["density_texture"] for v in o.arrays.volumes] == [c1, -1]


def test_file_type_and_missing_file(tmp_path):
    with pytest.raises(ymrt_amd.YmrtError, match="file type"):
        ymrt_amd.YmtmpFile(str(tmp_path / "scene.txt"))
    with pytest.raises(ymrt_amd.YmrtError, match="open"):
        ymrt_amd.YmtmpFile(str(tmp_path / "missing.ymtmp"))


# ----------------------------------------------------------- image textures
def _png(path, img, ctype, palette=None, trns=None):
    """Minimal PNG writer (filter 'up' on odd rows, 'sub' on even) so the
    decoder's filters are exercised."""
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    raw = b""
    prev = np.zeros(w * ch, np.int32)
    for y in range(h):
        row = img[y].reshape(-1).astype(np.int32)
        if y % 2:
            raw += b"\x02" + bytes(((row - prev) % 256).astype(np.uint8))
        else:
            left = np.concatenate([np.zeros(ch, np.int32), row[:-ch]])
            raw += b"\x01" + bytes(((row - left) % 256).astype(np.uint8))
        prev = row

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)
    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if palette is not None:
        out += chunk(b"PLTE", bytes(palette.reshape(-1)))
    if trns is not None:
        out += chunk(b"tRNS", bytes(trns))
    out += chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b"")
    path.write_bytes(out)


def test_png_image_textures(tmp_path):
    """ImageTexture reads <dir>/texture_<stem>_<ext>/<file> (get_texture_folder_from_scene_file)
    with rows flipped (stbi_set_flip_vertically_on_load) and stb's channel
    counts; a file that does not load leaves the name meaning "no texture"."""
    folder = tmp_path / "texture_room_ymtmp"
    folder.mkdir()
    rng = np.random.default_rng(3)
    rgb = rng.integers(0, 256, (5, 7, 3), dtype=np.uint8)
    rgba = rng.integers(0, 256, (4, 3, 4), dtype=np.uint8)
    grey = rng.integers(0, 256, (6, 2), dtype=np.uint8)
    pal = rng.integers(0, 256, (4, 3), dtype=np.uint8)
    pidx = rng.integers(0, 4, (3, 5), dtype=np.uint8)
    _png(folder / "rgb.png", rgb, 2)
    _png(folder / "rgba.png", rgba, 6)
    _png(folder / "grey.png", grey, 0)
    _png(folder / "pal.png", pidx, 3, palette=pal, trns=[10, 20])
    text = "".join('Begin(ImageTexture); name = "%s"; texture_file_name = "%s.png"; %s End(ImageTexture);\n'
                   % (n, n, cs) for n, cs in (("rgb", ""), ("rgba", 'color_space = "linear";'), ("grey", ""),
                                              ("pal", ""), ("missing", "")))
    text += 'Begin(DefaultMaterial); name = "m"; diffuse_albedo_tex = "missing"; End(DefaultMaterial);'
    scene = tmp_path / "room.ymtmp"
    scene.write_text(text)
    f = ymrt_amd.YmtmpFile(str(scene))
    assert len(f.warnings) == 1 and "missing" in f.warnings[0]
    assert f.find("missing") == ("texture", abi.INVALID)
    sa = f.arrays()
    assert sa.materials[0].diffuse_albedo_tex == abi.INVALID
    t = sa.textures
    assert len(t) == 4
    assert np.array_equal(t[0]["texels"], rgb[::-1]) and t[0]["cs"] == abi.COLOR_SRGB
    assert np.array_equal(t[1]["texels"], rgba[::-1]) and t[1]["cs"] == abi.COLOR_LINEAR
    assert np.array_equal(t[2]["texels"][..., 0], grey[::-1])
    expect = np.concatenate([pal[pidx], np.where(pidx < 2, np.array([10, 20, 255, 255])[pidx], 255)[..., None]], -1)
    assert np.array_equal(t[3]["texels"], expect[::-1].astype(np.uint8))
    assert all(x["wrap"] == abi.WRAP_REPEAT for x in t)
