"""Test infrastructure (the CPU oracle): a Python restatement of YMRT's .ymtmp
scene parser, used only to check the library's parser (ymrt_ymtmp_load).

SceneParser::parsing_scene (src/core/YMParser.cpp:2330-2433): each line keeps
only the statement alphabet (is_valid_character, 52-55), lines starting with
'#' are dropped, the rest is concatenated and split at ';'.  Statements are
Begin(Type) / End(Type) (is_type_bracket, 630-669) or attribute=value
(is_attribute_assign, 685-721; the value shapes of 420-628), blocks do not
nest and take only their type's attributes (check_statements_valid, 765-814).
load_scene_asset (2280-2328) reads the blocks in order with the readers of
971-2231: defaults, clamps, one name space, references to earlier names; a
reader that fails skips its block.  Transforms come from or_xform.c (glm's
operation order).
"""
import ctypes as C
import re

import numpy as np

from ymrt_amd import abi

from . import lib

_libc = C.CDLL(None)
_libc_strtof = _libc.strtof
_libc_strtof.restype = C.c_float
_libc_strtof.argtypes = [C.c_char_p, C.c_void_p]

KEEP = re.compile(r"[A-Za-z0-9=<>,.\"()#_;\[\]+\-\\/:]")

ATTRS = {
    "Camera": ["from", "look", "fov"],
    "Transform": ["name", "translate", "scale", "rotate", "parent_transform_index"],
    "Mesh": ["name", "mesh_triangles", "mesh_position_indices", "mesh_positions", "mesh_normal_indices",
             "mesh_normals", "mesh_texcoord_indices", "mesh_texcoords"],
    "Cube": ["name"],
    "Sphere": ["name", "radius"],
    "PrimitiveInstance": ["name", "geometry_index", "transform_index", "material_index", "inner_volume_index",
                          "treat_as_boundary"],
    "DefaultMaterial": ["name", "diffuse_albedo", "specular_albedo", "roughness_x", "roughness_y", "ior_n",
                        "metalness", "specular_weight", "transmission_weight", "ior_priority", "coat_albedo",
                        "coat_weight", "coat_thickness", "coat_ior", "coat_roughness_x", "coat_roughness_y",
                        "diffuse_albedo_tex", "alpha_x_tex", "specular_albedo_tex", "alpha_y_tex",
                        "specular_weight_tex", "metalness_tex", "transmission_weight_tex", "normal_mapping_tex",
                        "bump_mapping_tex", "coat_albedo_tex", "coat_weight_tex", "coat_thickness_tex",
                        "coat_roughness_x_tex", "coat_roughness_y_tex", "coat_normal_tex", "coat_normal_tex_type"],
    "LightMaterial": ["name", "light_color", "intensity"],
    "ImageTexture": ["name", "texture_file_name", "color_space"],
    "ConstantTextureFloat": ["name", "value"],
    "ConstantTextureRgb": ["name", "color"],
    "CheckerBoardTexture": ["name", "texture_black_idx", "texture_white_idx", "frequency"],
    "NoiseTexture": ["name", "texture_black_idx", "texture_white_idx", "frequency", "normalized"],
    "NoiseTextureFBM": ["name", "texture_black_idx", "texture_white_idx", "frequency", "lacunarity", "gain",
                        "layer_count", "amplitude", "offset"],
    "NoiseTextureTurbulence": ["name", "texture_black_idx", "texture_white_idx", "frequency", "lacunarity", "gain",
                               "layer_count", "amplitude", "offset"],
    "NoiseTextureMarble": ["name", "texture_black_idx", "texture_white_idx", "frequency", "lacunarity", "gain",
                           "layer_count", "variation", "x_turb", "y_turb", "z_turb"],
    "NoiseTextureWood": ["name", "texture_black_idx", "texture_white_idx", "frequency", "variation"],
    "NoiseTexturePolkaDot": ["name", "texture_black_idx", "texture_white_idx", "frequency", "radius"],
    "NoiseTextureWave": ["name", "texture_black_idx", "texture_white_idx", "freq_0", "lacunarity_0", "gain_0",
                         "layer_count_0", "freq_1", "lacunarity_1", "gain_1", "layer_count_1"],
    "DistantLight": ["name", "light_color", "transform_index", "intensity", "theta_max"],
    "Volume": ["name", "transform_index", "sigma_t", "albedo", "g", "density_texture_index"],
}


class ParseError(Exception):
    pass


class Skip(Exception):
    pass


def _f32(x):
    return float(np.float32(x))


def _signed_digits(s):
    n = 0
    for i, ch in enumerate(s):
        d = ch.isdigit()
        n += d
        if i == 0 and not d and ch not in "+-":
            return None
        if i > 0 and not d:
            return None
    return n


def is_int(v):
    return bool(v) and not (len(v) > 1 and v[0] == "0") and all(c.isdigit() for c in v)


def is_float(v):
    if not v:
        return False
    if "." not in v:
        n = _signed_digits(v)
        return n is not None and n > 0
    left, right = v.split(".", 1)
    if not left and not right:
        return False
    nl = 0
    if left:
        nl = _signed_digits(left)
        if nl is None:
            return False
    nr = 0
    for i, ch in enumerate(right):
        d = ch.isdigit()
        nr += d
        if i == len(right) - 1 and not d and ch != "f":
            return False
        if i < len(right) - 1 and not d:
            return False
    return nl + nr > 0


def _items(v, close):
    out, start = [], 1
    for end in range(1, len(v)):
        if v[end] in ("," + close):
            out.append(v[start:end])
            start = end + 1
    return out


def is_vector(v):
    if len(v) < 2 or v[0] != "<" or v[-1] != ">":
        return False
    it = _items(v, ">")
    return all(is_float(x) for x in it) and len(it) == 3


def _quoted(v, path):
    if len(v) < 2 or v[0] != '"' or v[-1] != '"':
        return False
    extra = "_.-+" + ("\\/:" if path else "")
    return all(c.isalnum() or c in extra for c in v[1:-1])


def is_array(v):
    if len(v) < 2 or v[0] != "[" or v[-1] != "]":
        return False
    return all(is_float(x) or is_int(x) for x in _items(v, "]"))


def is_value(v):
    return (is_float(v) or is_vector(v) or _quoted(v, False) or _quoted(v, True) or v in ("true", "false")
            or is_int(v) or is_array(v))


def statements(text):
    content = ""
    for line in text.split("\n"):
        kept = "".join(KEEP.findall(line))
        if not kept or kept[0] == "#":
            continue
        content += kept
    parts = content.split(";")
    out = []
    for i, s in enumerate([p for p in parts if p]):
        if s[0] == "#":
            continue
        lb, rb = s.find("("), s.rfind(")")
        if lb != -1 and rb != -1 and lb < rb and rb == len(s) - 1:
            ty, be = s[lb + 1:rb], s[:lb]
            if ty in ATTRS and be in ("Begin", "End"):
                out.append((True, be, ty, i))
                continue
        eq = s.find("=")
        if eq == -1:
            raise ParseError("illegal statement %d: %s" % (i, s))
        attr, val = s[:eq], s[eq + 1:]
        if not all(c.isalpha() or c == "_" for c in attr) or not is_value(val):
            raise ParseError("illegal statement %d: %s" % (i, s))
        out.append((False, attr, val, i))
    open_ = None
    for br, left, right, line in out:
        if br:
            if left == "Begin":
                if open_ is not None:
                    raise ParseError("nested block at %d" % line)
                open_ = right
            else:
                if open_ != right:
                    raise ParseError("End does not match at %d" % line)
                open_ = None
        else:
            if open_ is None or left not in ATTRS[open_]:
                raise ParseError("unknown attribute %s at %d" % (left, line))
    if open_ is not None:
        raise ParseError("block not closed")
    return out


def glm_transform(t, s, r, parent=None):
    """(object_to_world, world_to_object) column-major float32[16] each."""
    out, inv = np.zeros(16, np.float32), np.zeros(16, np.float32)
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    par = None if parent is None else (C.c_float * 16)(*[float(x) for x in parent])
    lib().or_glm_transform(f3(t), f3(s), f3(r), par, out.ctypes.data_as(C.c_void_p), inv.ctypes.data_as(C.c_void_p))
    return out, inv


class Loaded:
    def __init__(self):
        self.arrays = abi.SceneArrays()
        self.names = {}          # name -> (kind, index)
        self.transforms = []     # (o2w, w2o) column-major
        self.warnings = []
        self.meshes = 0
        c = self.arrays
        c.camera = dict(position=(0.0, 0.0, 10.0), w=(0.0, 0.0, 6.0), fov=_f32(45.0 * _f32(0.017453292519943295)),
                        aspect=1.0)


def _cube():
    """GeometryData::InitCube (src/GeometryDefines.h:274-519): 36 corners."""
    faces = [
        ([(-1, -1, -1), (1, 1, -1), (1, -1, -1), (1, 1, -1), (-1, -1, -1), (-1, 1, -1)], (0, 0, -1),
         [(0, 1), (1, 0), (1, 1), (1, 0), (0, 1), (0, 0)]),
        ([(-1, -1, 1), (1, -1, 1), (1, 1, 1), (1, 1, 1), (-1, 1, 1), (-1, -1, 1)], (0, 0, 1),
         [(0, 0), (1, 0), (1, 1), (1, 1), (0, 1), (0, 0)]),
        ([(-1, 1, 1), (-1, 1, -1), (-1, -1, -1), (-1, -1, -1), (-1, -1, 1), (-1, 1, 1)], (-1, 0, 0),
         [(1, 0), (1, 1), (0, 1), (0, 1), (0, 0), (1, 0)]),
        ([(1, 1, 1), (1, -1, -1), (1, 1, -1), (1, -1, -1), (1, 1, 1), (1, -1, 1)], (1, 0, 0),
         [(1, 1), (0, 0), (1, 0), (0, 0), (1, 1), (0, 1)]),
        ([(-1, -1, -1), (1, -1, -1), (1, -1, 1), (1, -1, 1), (-1, -1, 1), (-1, -1, -1)], (0, -1, 0),
         [(0, 0), (1, 0), (1, 1), (1, 1), (0, 1), (0, 0)]),
        ([(-1, 1, -1), (1, 1, 1), (1, 1, -1), (1, 1, 1), (-1, 1, -1), (-1, 1, 1)], (0, 1, 0),
         [(0, 1), (1, 0), (1, 1), (1, 0), (0, 1), (0, 0)]),
    ]
    pos = np.array([p for f in faces for p in f[0]], np.float32) * 0.5
    nrm = np.array([f[1] for f in faces for _ in range(6)], np.float32)
    uv = np.array([u for f in faces for u in f[2]], np.float32)
    idx = np.arange(36, dtype=np.uint32)
    return pos, idx, nrm, uv


class _Block:
    def __init__(self, L, stmts, ty):
        self.L, self.stmts, self.ty = L, stmts, ty

    def skip(self, msg):
        raise Skip("%s: %s" % (self.ty, msg))

    @staticmethod
    def _stof(v):
        # std::stof is strtof: one correct rounding of the decimal to float
        return float(_libc_strtof(v.encode(), None))

    def num(self, v):
        if not (is_float(v) or is_int(v)):
            self.skip("expected a number")
        return self._stof(v)

    def integer(self, v):
        if not (is_float(v) or is_int(v)):
            self.skip("expected an integer")
        return int(re.match(r"[+-]?\d+", v).group(0))

    def vec(self, v):
        if not is_vector(v):
            self.skip("expected a vector")
        return [self._stof(x) for x in _items(v, ">")]

    def string(self, v):
        if len(v) < 2 or v[0] != '"' or v[-1] != '"':
            self.skip("expected a string")
        return v[1:-1]

    def new_name(self, v):
        n = self.string(v)
        if n in self.L.names:
            self.skip("name already declared")
        return n

    def ref(self, v, kinds):
        n = self.string(v)
        if n not in self.L.names:
            self.skip("undeclared name")
        k, i = self.L.names[n]
        if k not in kinds:
            self.skip("wrong kind")
        return k, i

    def ints(self, v):
        if not is_array(v):
            self.skip("expected an array")
        return [int(re.match(r"[+-]?\d+", x).group(0)) for x in _items(v, "]")]

    def floats(self, v):
        if not is_array(v):
            self.skip("expected an array")
        return [self._stof(x) for x in _items(v, "]")]


def _max(a, b):
    return b if a < b else a


def _clamp(v, lo, hi):
    return lo if v < lo else (hi if v > hi else v)


def _read(L, ty, stmts):
    b = _Block(L, stmts, ty)
    sa = L.arrays
    name = None
    if ty == "Camera":
        frm, look, fov = [0.0, 0.0, 10.0], [0.0, 0.0, 4.0], 45.0
        for a, v in stmts:
            if a == "from":
                frm = b.vec(v)
            elif a == "look":
                look = b.vec(v)
            elif a == "fov":
                fov = _max(b.num(v), _f32(0.01))
        w = tuple(_f32(np.float32(f) - np.float32(l)) for f, l in zip(frm, look))
        sa.camera = dict(position=tuple(frm), w=w, fov=_f32(np.float32(fov) * np.float32(0.017453292519943295)),
                         aspect=1.0)
        return
    if ty == "Transform":
        t, s, r, parent = [0.0] * 3, [1.0] * 3, [0.0] * 3, None
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a == "translate":
                t = b.vec(v)
            elif a == "scale":
                s = [_max(_f32(0.0001), x) for x in b.vec(v)]
            elif a == "rotate":
                r = [_clamp(x, 0.0, 360.0) for x in b.vec(v)]
            elif a == "parent_transform_index":
                parent = b.ref(v, ("transform",))[1]
        if not name:
            b.skip("requires a name")
        o, i = glm_transform(t, s, r, None if parent is None else L.transforms[parent][0])
        L.transforms.append((o, i))
        L.names[name] = ("transform", len(L.transforms) - 1)
        return
    if ty in ("Mesh", "Cube"):
        arr = {}
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a in ("mesh_positions", "mesh_normals", "mesh_texcoords"):
                arr[a] = b.floats(v)
            elif a != "name":
                arr[a] = b.ints(v)
        if not name:
            b.skip("requires a name")
        if ty == "Cube":
            pos, idx, nrm, uv = _cube()
            sa.add_mesh(pos, idx, nrm=nrm, nidx=idx, uv=uv, tidx=idx)
        else:
            tris = arr.get("mesh_triangles", [])
            tris = tris[:len(tris) // 3 * 3]
            nt = len(tris) // 3
            pidx, nidx = arr.get("mesh_position_indices", []), arr.get("mesh_normal_indices", [])
            tidx = arr.get("mesh_texcoord_indices", [])
            pos, nrm, uv = arr.get("mesh_positions", []), arr.get("mesh_normals", []), arr.get("mesh_texcoords", [])
            if len(pidx) != 3 * nt or len(nidx) != 3 * nt or (tidx and len(tidx) != 3 * nt) or not pos or not nrm \
                    or nt == 0:
                b.skip("bad mesh")
            np_, nn, nu = len(pos) // 3, len(nrm) // 3, len(uv) // 2
            cp = [pidx[k] for k in tris]
            cn = [nidx[k] for k in tris]
            if max(tris) >= len(pidx) or max(cp) >= np_ or max(cn) >= nn:
                b.skip("index out of range")
            kw = dict(nrm=np.array(nrm[:nn * 3], np.float32), nidx=np.array(cn, np.uint32))
            if tidx:
                ct = [tidx[k] for k in tris]
                if max(ct) >= nu:
                    b.skip("index out of range")
                kw.update(uv=np.array(uv[:nu * 2], np.float32), tidx=np.array(ct, np.uint32))
            sa.add_mesh(np.array(pos[:np_ * 3], np.float32), np.array(cp, np.uint32), **kw)
        L.names[name] = ("mesh", len(sa.meshes) - 1)
        return
    if ty == "Sphere":
        radius = 1.0
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a == "radius":
                radius = _max(_f32(0.0001), b.num(v))
        if not name:
            b.skip("requires a name")
        sa.add_sphere(radius)
        L.names[name] = ("sphere", len(sa.spheres) - 1)
        return
    if ty == "PrimitiveInstance":
        geom = xf = None
        mat, vol, boundary = abi.INVALID, -1, False
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a == "geometry_index":
                geom = b.ref(v, ("mesh", "sphere"))
            elif a == "transform_index":
                xf = b.ref(v, ("transform",))[1]
            elif a == "material_index":
                mat = b.ref(v, ("material",))[1]
            elif a == "inner_volume_index":
                vol = b.ref(v, ("volume",))[1]
            elif a == "treat_as_boundary":
                boundary = v == "true"
        if not name or geom is None or xf is None:
            b.skip("requires name, geometry and transform")
        g = abi.SphereRef(geom[1]) if geom[0] == "sphere" else geom[1]
        o, i = L.transforms[xf]
        sa.add_instance(g, mat, o.reshape(4, 4).T, i.reshape(4, 4).T, inner_volume=vol, boundary=boundary)
        L.names[name] = ("instance", len(sa.instances) - 1)
        return
    if ty == "DefaultMaterial":
        kw = dict(diffuse_albedo=(0.5, 0.5, 0.5), specular_albedo=(1.0, 1.0, 1.0), roughness_x=0.2,
                  roughness_y=0.2, ior_n=1.3, metalness=0.0, specular_weight=1.0, transmission_weight=0.0,
                  ior_priority=0, coat_albedo=(1.0, 1.0, 1.0), coat_weight=0.0, coat_thickness=1.0, coat_ior=1.6,
                  coat_roughness_x=0.2, coat_roughness_y=0.2)
        texmap = {"alpha_x_tex": "roughness_x_tex", "alpha_y_tex": "roughness_y_tex"}
        unit = ("roughness_x", "roughness_y", "metalness", "specular_weight", "transmission_weight", "coat_weight",
                "coat_roughness_x", "coat_roughness_y")
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a in ("diffuse_albedo", "specular_albedo", "coat_albedo"):
                kw[a] = tuple(_max(x, 0.0) for x in b.vec(v))
            elif a in unit:
                kw[a] = _clamp(b.num(v), 0.0, 1.0)
            elif a == "ior_n":
                kw[a] = _max(_f32(0.0001), b.num(v))
            elif a == "coat_ior":
                kw[a] = _max(b.num(v), _f32(0.0001))
            elif a == "coat_thickness":
                kw[a] = _max(b.num(v), 0.0)
            elif a == "ior_priority":
                kw[a] = max(0, b.integer(v))
            elif a == "coat_normal_tex":
                b.ref(v, ("texture",))
            elif a == "coat_normal_tex_type":
                if b.string(v) not in ("normal_map", "bump_map"):
                    b.skip("invalid normal tex type")
            elif a.endswith("_tex"):
                kw[texmap.get(a, a)] = b.ref(v, ("texture",))[1]
        if not name:
            b.skip("requires a name")
        sa.add_material(abi.default_material(**kw))
        L.names[name] = ("material", len(sa.materials) - 1)
        return
    if ty == "LightMaterial":
        color, intensity = (1.0, 1.0, 1.0), 1.0
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a == "light_color":
                color = tuple(_max(x, 0.0) for x in b.vec(v))
            elif a == "intensity":
                intensity = _max(b.num(v), 0.0)
        if not name:
            b.skip("requires a name")
        sa.add_material(abi.light_material(color, intensity))
        L.names[name] = ("material", len(sa.materials) - 1)
        return
    if ty == "DistantLight":
        color, intensity, theta, xf = (1.0, 1.0, 1.0), 1.0, 5.0, None
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a == "light_color":
                color = tuple(_max(x, 0.0) for x in b.vec(v))
            elif a == "transform_index":
                xf = b.ref(v, ("transform",))[1]
            elif a == "intensity":
                intensity = _max(0.0, b.num(v))
            elif a == "theta_max":
                theta = _max(0.0, b.num(v))
        if not name or xf is None:
            b.skip("requires name and transform")
        o, i = L.transforms[xf]
        sa.add_distant_light(color, intensity, theta, o.reshape(4, 4).T, i.reshape(4, 4).T)
        L.names[name] = ("distant_light", len(sa.distant_lights) - 1)
        return
    if ty == "Volume":
        sig, alb, g, xf, dens = (0.05,) * 3, (1.0,) * 3, 0.0, None, None
        for a, v in stmts:
            if a == "name":
                name = b.new_name(v)
            elif a == "transform_index":
                xf = b.ref(v, ("transform",))[1]
            elif a == "sigma_t":
                sig = tuple(_max(0.0, x) for x in b.vec(v))
            elif a == "albedo":
                alb = tuple(_max(0.0, x) for x in b.vec(v))
            elif a == "g":
                g = _clamp(b.num(v), 0.0, 1.0)
            elif a == "density_texture_index":
                dens = b.ref(v, ("texture",))
        if not name or xf is None:
            b.skip("requires name and transform")
        tex = -1 if dens is None or dens[1] is None or dens[1] == abi.INVALID else int(dens[1])
        sa.add_volume(sig, alb, g, grid=-1, world_to_object=L.transforms[xf][1].reshape(4, 4).T,
                      density_texture=tex)
        L.names[name] = ("volume", len(sa.volumes) - 1)
        return
    if ty == "ImageTexture":
        raise ParseError("ImageTexture is not restated in the oracle")
    # constant and procedural textures
    T = abi
    spec = {
        "ConstantTextureFloat": (T.TEXTURE_CONSTANT_FLOAT, {}),
        "ConstantTextureRgb": (T.TEXTURE_CONSTANT_RGB, {}),
        "CheckerBoardTexture": (T.TEXTURE_CHECKERBOARD, dict(frequency=4.0)),
        "NoiseTexture": (T.TEXTURE_NOISE, dict(frequency=1.0, normalized=1)),
        "NoiseTextureFBM": (T.TEXTURE_FBM, dict(frequency=16.0, lacunarity=2.0, gain=0.5, layer_count=4,
                                                amplitude=1.0, offset=0.0)),
        "NoiseTextureTurbulence": (T.TEXTURE_TURBULENCE, dict(frequency=16.0, lacunarity=2.0, gain=0.5,
                                                              layer_count=4, amplitude=1.0, offset=0.0)),
        "NoiseTextureMarble": (T.TEXTURE_MARBLE, dict(frequency=16.0, lacunarity=2.0, gain=0.5, layer_count=4,
                                                      variation=10.0, turbulence_axes=7)),
        "NoiseTextureWood": (T.TEXTURE_WOOD, dict(frequency=2.0, variation=10.0)),
        "NoiseTexturePolkaDot": (T.TEXTURE_POLKA_DOT, dict(frequency=8.0, radius=0.35)),
        "NoiseTextureWave": (T.TEXTURE_WAVE, dict(frequency=0.1, lacunarity=2.0, gain=0.5, layer_count=2,
                                                  frequency1=2.0, lacunarity1=2.0, gain1=0.5, layer_count1=4)),
    }[ty]
    ttype, p = spec[0], dict(spec[1])
    value = [1.0, 0.0, 0.0] if ttype == T.TEXTURE_CONSTANT_FLOAT else [0.5, 0.5, 0.5]
    if ttype not in (T.TEXTURE_CONSTANT_FLOAT, T.TEXTURE_CONSTANT_RGB):
        value = [0.0, 0.0, 0.0]
    child = [None, None]
    clampf = ttype in (T.TEXTURE_CHECKERBOARD, T.TEXTURE_NOISE, T.TEXTURE_FBM, T.TEXTURE_TURBULENCE,
                       T.TEXTURE_WOOD)
    fields = {"gain": "gain", "gain_0": "gain", "amplitude": "amplitude", "offset": "offset",
              "variation": "variation", "radius": "radius", "lacunarity_0": "lacunarity", "freq_1": "frequency1",
              "lacunarity_1": "lacunarity1", "gain_1": "gain1"}
    for a, v in stmts:
        if a == "name":
            name = b.new_name(v)
        elif a == "value":
            value[0] = b.num(v)
        elif a == "color":
            value = b.vec(v)
        elif a == "texture_black_idx":
            child[0] = b.ref(v, ("texture",))[1]
        elif a == "texture_white_idx":
            child[1] = b.ref(v, ("texture",))[1]
        elif a in ("frequency", "freq_0"):
            f = b.num(v)
            p["frequency"] = _max(_f32(0.001), f) if clampf else f
        elif a == "normalized":
            p["normalized"] = 1 if v == "true" else 0
        elif a in ("layer_count", "layer_count_0"):
            p["layer_count"] = b.integer(v)
        elif a == "layer_count_1":
            p["layer_count1"] = b.integer(v)
        elif a in ("x_turb", "y_turb", "z_turb"):
            bit = {"x": 1, "y": 2, "z": 4}[a[0]]
            p["turbulence_axes"] = (p["turbulence_axes"] | bit) if v == "true" else (p["turbulence_axes"] & ~bit)
        elif a in fields:
            p[fields[a]] = b.num(v)
    if not name:
        b.skip("requires a name")
    if ttype >= T.TEXTURE_CHECKERBOARD and (child[0] is None or child[1] is None):
        b.skip("requires both inputs")
    sa.add_texture_node(ttype, tuple(child), tuple(value), **p)
    L.names[name] = ("texture", len(sa.textures) - 1)


def load_text(text):
    """Loaded scene of a .ymtmp text: .arrays, .names, .transforms, .warnings."""
    st = statements(text)
    L = Loaded()
    i = 0
    while i < len(st):
        br, left, right, _ = st[i]
        if not (br and left == "Begin"):
            i += 1
            continue
        e = i + 1
        while e < len(st) and not (st[e][0] and st[e][1] == "End"):
            e += 1
        try:
            _read(L, right, [(s[1], s[2]) for s in st[i + 1:e]])
        except Skip as ex:
            L.warnings.append(str(ex))
        i = e + 1
    return L


def load(path):
    with open(path) as f:
        return load_text(f.read())
