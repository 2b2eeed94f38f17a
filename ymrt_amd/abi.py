"""ctypes mirror of include/ymrt.h (types only) and a holder that turns numpy
scene arrays into a ``ymrt_scene_desc`` while keeping the arrays alive.

The descriptor carries the data of YMRT's host arrays (SceneModule,
src/core/SceneModule.h:109-230): meshes as per-corner index lists
(src/GeometryDefines.h:197-273), instances with their transform pair
(src/PrimtiveInstance.h:9-25), DefaultMtl/LightMtl materials
(src/Material.h:21-359), image textures, volumes (src/Volume.h:9-18) and the
camera (src/Camera.h:15-323).
"""
import ctypes as C
import numpy as np

INVALID = 0xFFFFFFFF
MATERIAL_DEFAULT = 0
MATERIAL_LIGHT = 1
WRAP_REPEAT, WRAP_CLAMP = 0, 1
COLOR_SRGB, COLOR_LINEAR = 0, 1
SAMPLER_PCG, SAMPLER_HALTON = 0, 1

u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)


class MeshDesc(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_uint32), ("position_count", C.c_uint32),
        ("position_indices", u32p), ("positions", f32p),
        ("normal_count", C.c_uint32), ("normal_indices", u32p), ("normals", f32p),
        ("texcoord_count", C.c_uint32), ("texcoord_indices", u32p), ("texcoords", f32p),
    ]


class SphereDesc(C.Structure):
    _fields_ = [("radius", C.c_float)]


GEOMETRY_MESH, GEOMETRY_SPHERE = 0, 1


class InstanceDesc(C.Structure):
    _fields_ = [
        ("geometry", C.c_uint32), ("material", C.c_uint32),
        ("inner_volume", C.c_int32), ("treat_as_boundary", C.c_int32),
        ("geometry_type", C.c_uint32), ("reserved", C.c_uint32),
        ("object_to_world", C.c_float * 16), ("world_to_object", C.c_float * 16),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("diffuse_albedo", C.c_float * 3), ("diffuse_albedo_tex", C.c_uint32),
        ("specular_albedo", C.c_float * 3), ("specular_albedo_tex", C.c_uint32),
        ("roughness_x", C.c_float), ("roughness_x_tex", C.c_uint32),
        ("roughness_y", C.c_float), ("roughness_y_tex", C.c_uint32),
        ("specular_weight", C.c_float), ("specular_weight_tex", C.c_uint32),
        ("metalness", C.c_float), ("metalness_tex", C.c_uint32),
        ("transmission_weight", C.c_float), ("transmission_weight_tex", C.c_uint32),
        ("ior_n", C.c_float), ("ior_priority", C.c_uint32),
        ("coat_albedo", C.c_float * 3), ("coat_albedo_tex", C.c_uint32),
        ("coat_weight", C.c_float), ("coat_weight_tex", C.c_uint32),
        ("coat_thickness", C.c_float), ("coat_thickness_tex", C.c_uint32),
        ("coat_roughness_x", C.c_float), ("coat_roughness_x_tex", C.c_uint32),
        ("coat_roughness_y", C.c_float), ("coat_roughness_y_tex", C.c_uint32),
        ("coat_ior", C.c_float),
        ("normal_mapping_tex", C.c_uint32), ("bump_mapping_tex", C.c_uint32),
        ("light_color", C.c_float * 3), ("intensity", C.c_float),
    ]


TEXTURE_IMAGE, TEXTURE_CONSTANT_FLOAT, TEXTURE_CONSTANT_RGB, TEXTURE_CHECKERBOARD = 0, 1, 2, 3
TEXTURE_NOISE, TEXTURE_FBM, TEXTURE_TURBULENCE, TEXTURE_MARBLE = 4, 5, 6, 7
TEXTURE_WOOD, TEXTURE_POLKA_DOT, TEXTURE_WAVE = 8, 9, 10


class TextureDesc(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("channels", C.c_uint32),
        ("texels", C.POINTER(C.c_uint8)), ("wrap_mode", C.c_int32), ("color_space", C.c_int32),
        ("scale_u", C.c_float), ("scale_v", C.c_float), ("offset_u", C.c_float), ("offset_v", C.c_float),
        ("type", C.c_uint32), ("value", C.c_float * 3), ("child", C.c_uint32 * 2),
        ("frequency", C.c_float), ("lacunarity", C.c_float), ("gain", C.c_float), ("amplitude", C.c_float),
        ("offset", C.c_float), ("variation", C.c_float), ("radius", C.c_float),
        ("layer_count", C.c_int32), ("normalized", C.c_int32), ("turbulence_axes", C.c_int32),
        ("frequency1", C.c_float), ("lacunarity1", C.c_float), ("gain1", C.c_float), ("layer_count1", C.c_int32),
        ("texels_f32", C.POINTER(C.c_float)),
    ]


class GridDesc(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("density", f32p)]


class VolumeDesc(C.Structure):
    _fields_ = [
        ("sigma_t", C.c_float * 3), ("albedo", C.c_float * 3), ("g", C.c_float),
        ("density_grid", C.c_int32), ("world_to_object", C.c_float * 16),
        ("density_texture", C.c_int32),
    ]


class CameraDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("w", C.c_float * 3), ("fov", C.c_float),
                ("aspect_ratio", C.c_float)]


class DistantLightDesc(C.Structure):
    _fields_ = [
        ("light_color", C.c_float * 3), ("intensity", C.c_float), ("theta_max", C.c_float),
        ("object_to_world", C.c_float * 16), ("world_to_object", C.c_float * 16),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("mesh_count", C.c_uint32), ("meshes", C.POINTER(MeshDesc)),
        ("sphere_count", C.c_uint32), ("spheres", C.POINTER(SphereDesc)),
        ("instance_count", C.c_uint32), ("instances", C.POINTER(InstanceDesc)),
        ("material_count", C.c_uint32), ("materials", C.POINTER(MaterialDesc)),
        ("texture_count", C.c_uint32), ("textures", C.POINTER(TextureDesc)),
        ("grid_count", C.c_uint32), ("grids", C.POINTER(GridDesc)),
        ("volume_count", C.c_uint32), ("volumes", C.POINTER(VolumeDesc)),
        ("distant_light_count", C.c_uint32), ("distant_lights", C.POINTER(DistantLightDesc)),
        ("camera", CameraDesc),
        ("halton_seed", C.c_uint32),
    ]


class Ray(C.Structure):
    _fields_ = [("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("tmin", C.c_float),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("triangle", C.c_uint32), ("instance", C.c_uint32),
                ("back", C.c_uint32)]


class RenderSettings(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_int32),
        ("ray_depth", C.c_int32), ("sampler_type", C.c_int32),
        ("enable_shape_light", C.c_int32), ("enable_distant_light", C.c_int32),
        ("enable_volume_scattering", C.c_int32),
        ("enable_env_light", C.c_int32), ("enable_russian_roulette", C.c_int32),
        ("enable_indirect_clamp", C.c_int32), ("indirect_limit", C.c_float),
        ("gamma", C.c_float), ("exposure", C.c_float), ("path_capacity", C.c_uint32),
        ("persistent_waves", C.c_uint32),
    ]


class FrameStats(C.Structure):
    _fields_ = [
        ("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64), ("transmittance_rays", C.c_uint64),
        ("paths", C.c_uint64), ("passes", C.c_uint32), ("chunks", C.c_uint32),
        ("ms_raygen", C.c_float), ("ms_extend", C.c_float), ("ms_shade", C.c_float),
        ("ms_shadow", C.c_float), ("ms_resolve", C.c_float),
        ("launches_extend", C.c_uint32), ("launches_shade", C.c_uint32),
        ("launches_shadow", C.c_uint32), ("ms_total", C.c_float),
        ("ms_medium", C.c_float), ("launches_medium", C.c_uint32),
        ("extend_nodes", C.c_uint64), ("extend_tris", C.c_uint64),
        ("shadow_nodes", C.c_uint64), ("shadow_tris", C.c_uint64),
        ("ms_transmit", C.c_float), ("launches_transmit", C.c_uint32),
        ("transmit_nodes", C.c_uint64), ("transmit_tris", C.c_uint64),
        ("ms_adjoint", C.c_float), ("reserved", C.c_uint32),
    ]


class BatchItem(C.Structure):
    """ymrt_batch_item: one frame of one scene for ymrt_render_batch."""
    _fields_ = [
        ("scene", C.c_void_p), ("settings", C.POINTER(RenderSettings)), ("frame_index", C.c_int),
        ("timing", C.c_int), ("noise_rgba", C.c_void_p), ("stats", C.POINTER(FrameStats)),
    ]


class AdjointTargets(C.Structure):
    """ymrt_adjoint_targets: the parameters ymrt_render_adjoint differentiates
    (host material lists, texture and volume indices) and their device
    gradient arrays."""
    _fields_ = [
        ("albedo_materials", C.POINTER(C.c_uint32)), ("albedo_count", C.c_uint32),
        ("roughness_materials", C.POINTER(C.c_uint32)), ("roughness_count", C.c_uint32),
        ("texture", C.c_uint32), ("volume", C.c_uint32),
        ("grad_albedo", C.c_void_p), ("grad_roughness", C.c_void_p),
        ("grad_texture", C.c_void_p), ("grad_density", C.c_void_p),
    ]


def lambert_only(m):
    """A DefaultMtl whose one lobe is its untextured diffuse albedo (the
    albedo targets ymrt_render_adjoint accepts)."""
    return (m.type == MATERIAL_DEFAULT and m.specular_weight == 0.0 and m.metalness == 0.0
            and m.transmission_weight == 0.0 and m.coat_weight == 0.0 and m.diffuse_albedo_tex == INVALID
            and m.specular_weight_tex == INVALID and m.metalness_tex == INVALID
            and m.transmission_weight_tex == INVALID and m.coat_weight_tex == INVALID)


def metal_only(m):
    """Metalness 1 without coat, isotropic untextured roughness (the roughness
    targets ymrt_render_adjoint accepts)."""
    return (m.type == MATERIAL_DEFAULT and m.metalness == 1.0 and m.metalness_tex == INVALID
            and m.coat_weight == 0.0 and m.coat_weight_tex == INVALID and m.roughness_x == m.roughness_y
            and m.roughness_x_tex == INVALID and m.roughness_y_tex == INVALID)


class SceneInfo(C.Structure):
    _fields_ = [
        ("triangle_count", C.c_uint64), ("node_count", C.c_uint64),
        ("device_bytes", C.c_uint64), ("max_stack", C.c_uint32),
        ("shape_light_count", C.c_uint32), ("build_ms", C.c_float),
        ("camera_medium_records", C.c_uint32),
    ]


RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("triangle", "<u4"), ("instance", "<u4"), ("back", "<u4")])
assert RAY_DTYPE.itemsize == C.sizeof(Ray) == 32
assert HIT_DTYPE.itemsize == C.sizeof(Hit) == 16


def _ptr(arr, ctype):
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


def default_material(**kw):
    """DefaultMtl with the defaults of Material::InitDefaultMtl (src/Material.h:278-312)."""
    m = MaterialDesc()
    m.type = MATERIAL_DEFAULT
    vals = dict(diffuse_albedo=(0.5, 0.5, 0.5), specular_albedo=(1.0, 1.0, 1.0),
                roughness_x=0.2, roughness_y=0.2, ior_n=1.3, metalness=0.0,
                specular_weight=1.0, transmission_weight=0.0, ior_priority=0,
                coat_albedo=(1.0, 1.0, 1.0), coat_weight=0.0, coat_thickness=1.0,
                coat_ior=1.6, coat_roughness_x=0.2, coat_roughness_y=0.2,
                light_color=(0.0, 0.0, 0.0), intensity=0.0)
    vals.update(kw)
    for name, _ in MaterialDesc._fields_:
        if name.endswith("_tex"):
            setattr(m, name, vals.get(name, INVALID))
    for k, v in vals.items():
        if k.endswith("_tex"):
            continue
        if isinstance(v, (tuple, list)):
            getattr(m, k)[:] = [float(x) for x in v]
        else:
            setattr(m, k, v)
    return m


def light_material(color=(1.0, 1.0, 1.0), intensity=1.0):
    """LightMtl (Material::InitLightMtl, src/Material.h:353-358)."""
    m = default_material()
    m.type = MATERIAL_LIGHT
    m.light_color[:] = [float(c) for c in color]
    m.intensity = float(intensity)
    return m


class SphereRef(int):
    """Index of a sphere in SceneArrays.spheres (distinguishes it from a mesh index)."""


class SceneArrays:
    """Numpy-backed scene; ``desc()`` returns a SceneDesc pointing into it.

    meshes: list of dicts with int32/uint32 'pidx' (T,3), float32 'pos' (P,3),
    optional 'nidx'/'nrm', 'tidx'/'uv'.  instances: list of dicts with mesh,
    material, object_to_world (4x4 row-major numpy, converted to column-major),
    world_to_object, inner_volume, boundary.
    """

    def __init__(self):
        self.meshes = []
        self.spheres = []
        self.instances = []
        self.materials = []
        self.textures = []
        self.grids = []
        self.volumes = []
        self.distant_lights = []
        self.camera = dict(position=(0, 0, 0), w=(0, 0, 1), fov=np.radians(45.0), aspect=1.0)
        self.halton_seed = 0
        self._keep = []

    # -- builders -------------------------------------------------------
    def add_mesh(self, pos, pidx, nrm=None, nidx=None, uv=None, tidx=None):
        m = dict(pos=np.ascontiguousarray(pos, np.float32).reshape(-1, 3),
                 pidx=np.ascontiguousarray(pidx, np.uint32).reshape(-1, 3))
        if nrm is not None:
            m["nrm"] = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
            m["nidx"] = np.ascontiguousarray(nidx if nidx is not None else pidx, np.uint32).reshape(-1, 3)
        if uv is not None:
            m["uv"] = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
            m["tidx"] = np.ascontiguousarray(tidx if tidx is not None else pidx, np.uint32).reshape(-1, 3)
        self.meshes.append(m)
        return len(self.meshes) - 1

    def add_material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def add_sphere(self, radius):
        """Analytic sphere of the given radius about the object-space origin;
        pass the returned handle to add_instance like a mesh index."""
        self.spheres.append(float(radius))
        return SphereRef(len(self.spheres) - 1)

    def add_instance(self, mesh, material, object_to_world=None, world_to_object=None,
                     inner_volume=-1, boundary=False):
        o2w = np.eye(4, dtype=np.float32) if object_to_world is None else np.asarray(object_to_world, np.float32)
        if world_to_object is None:
            world_to_object = np.linalg.inv(o2w.astype(np.float64)).astype(np.float32)
        o2w = o2w.copy()
        world_to_object = np.array(world_to_object, np.float32)
        for m in (o2w, world_to_object):  # exact affine bottom row
            m[3, :] = (0.0, 0.0, 0.0, 1.0)
        gtype = GEOMETRY_SPHERE if isinstance(mesh, SphereRef) else GEOMETRY_MESH
        self.instances.append(dict(mesh=int(mesh), geometry_type=gtype, material=material, o2w=o2w,
                                   w2o=np.asarray(world_to_object, np.float32),
                                   inner_volume=inner_volume, boundary=int(boundary)))
        return len(self.instances) - 1

    def add_texture(self, texels, wrap=WRAP_REPEAT, color_space=COLOR_SRGB,
                    scale=(1.0, 1.0), offset=(0.0, 0.0)):
        """Image texture: uint8 texels (decoded by color_space), or float32
        texels taken as linear values (a float texture, color_space unused)."""
        if np.asarray(texels).dtype == np.float32:
            t = np.ascontiguousarray(texels, np.float32)
            if t.ndim == 2:
                t = t[:, :, None]
            self.textures.append(dict(texels=t, wrap=wrap, cs=COLOR_LINEAR, scale=scale, offset=offset))
            return len(self.textures) - 1
        t = np.ascontiguousarray(texels, np.uint8)
        if t.ndim == 2:
            t = t[:, :, None]
        self.textures.append(dict(texels=t, wrap=wrap, cs=color_space, scale=scale, offset=offset))
        return len(self.textures) - 1

    def add_texture_node(self, type, child=(None, None), value=(0.0, 0.0, 0.0), **params):
        """Constant or procedural texture node; child holds the indices of the
        black and white inputs (None: black).  params: frequency, lacunarity,
        gain, amplitude, offset, variation, radius, layer_count, normalized,
        turbulence_axes, frequency1, lacunarity1, gain1, layer_count1."""
        self.textures.append(dict(type=type, child=tuple(INVALID if c is None else int(c) for c in child),
                                  value=tuple(float(v) for v in value), params=params))
        return len(self.textures) - 1

    def add_grid(self, density):
        d = np.ascontiguousarray(density, np.float32)
        assert d.ndim == 3  # (nz, ny, nx)
        self.grids.append(d)
        return len(self.grids) - 1

    def add_volume(self, sigma_t, albedo, g, grid=-1, world_to_object=None, density_texture=-1):
        w2o = np.eye(4, dtype=np.float32) if world_to_object is None else np.asarray(world_to_object, np.float32)
        self.volumes.append(dict(sigma_t=sigma_t, albedo=albedo, g=g, grid=grid, w2o=w2o,
                                 density_texture=int(density_texture)))
        return len(self.volumes) - 1

    def add_distant_light(self, color, intensity=1.0, theta_max_deg=5.0, object_to_world=None,
                          world_to_object=None):
        """Distant light shining along -y of its frame (object_to_world)."""
        o2w = np.eye(4, dtype=np.float32) if object_to_world is None else np.asarray(object_to_world, np.float32)
        o2w = o2w.copy()
        o2w[3, :] = (0.0, 0.0, 0.0, 1.0)
        if world_to_object is None:
            w2o = np.linalg.inv(o2w.astype(np.float64)).astype(np.float32)
            w2o[3, :] = (0.0, 0.0, 0.0, 1.0)
        else:
            w2o = np.array(world_to_object, np.float32)
        self.distant_lights.append(dict(color=tuple(float(c) for c in color), intensity=float(intensity),
                                        theta=float(theta_max_deg), o2w=o2w, w2o=w2o))
        return len(self.distant_lights) - 1

    def set_camera(self, position, look_at, fov_deg, aspect):
        p = np.asarray(position, np.float32)
        w = p - np.asarray(look_at, np.float32)
        self.camera = dict(position=tuple(p.tolist()), w=tuple(w.tolist()),
                           fov=float(np.float32(np.radians(fov_deg))), aspect=float(aspect))

    @property
    def triangle_count(self):
        return int(sum(self.meshes[i["mesh"]]["pidx"].shape[0] for i in self.instances
                       if i["geometry_type"] == GEOMETRY_MESH))

    # -- descriptor -----------------------------------------------------
    def desc(self):
        keep = []
        md = (MeshDesc * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            d = md[i]
            d.triangle_count = m["pidx"].shape[0]
            d.position_count = m["pos"].shape[0]
            d.position_indices = _ptr(m["pidx"], C.c_uint32)
            d.positions = _ptr(m["pos"], C.c_float)
            if "nrm" in m:
                d.normal_count = m["nrm"].shape[0]
                d.normal_indices = _ptr(m["nidx"], C.c_uint32)
                d.normals = _ptr(m["nrm"], C.c_float)
            if "uv" in m:
                d.texcoord_count = m["uv"].shape[0]
                d.texcoord_indices = _ptr(m["tidx"], C.c_uint32)
                d.texcoords = _ptr(m["uv"], C.c_float)
        inst = (InstanceDesc * max(1, len(self.instances)))()
        for i, it in enumerate(self.instances):
            d = inst[i]
            d.geometry, d.material = it["mesh"], it["material"]
            d.geometry_type = it.get("geometry_type", GEOMETRY_MESH)
            d.inner_volume, d.treat_as_boundary = it["inner_volume"], it["boundary"]
            # numpy row-major M[r, c] -> column-major m[c*4 + r]
            d.object_to_world[:] = it["o2w"].T.reshape(-1).tolist()
            d.world_to_object[:] = it["w2o"].T.reshape(-1).tolist()
        mats = (MaterialDesc * max(1, len(self.materials)))(*self.materials)
        texs = (TextureDesc * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            d = texs[i]
            d.child[:] = [INVALID, INVALID]
            d.amplitude, d.gain, d.lacunarity = 1.0, 0.5, 2.0
            if "texels" in t:
                h, w, c = t["texels"].shape
                d.type = TEXTURE_IMAGE
                d.width, d.height, d.channels = w, h, c
                if t["texels"].dtype == np.float32:
                    d.texels_f32 = _ptr(t["texels"], C.c_float)
                else:
                    d.texels = _ptr(t["texels"], C.c_uint8)
                d.wrap_mode, d.color_space = t["wrap"], t["cs"]
                d.scale_u, d.scale_v = t["scale"]
                d.offset_u, d.offset_v = t["offset"]
                continue
            d.type = t["type"]
            d.value[:] = list(t["value"])
            d.child[:] = list(t["child"])
            for k, v in t["params"].items():
                setattr(d, k, v)
        grids = (GridDesc * max(1, len(self.grids)))()
        for i, g in enumerate(self.grids):
            grids[i].nz, grids[i].ny, grids[i].nx = g.shape
            grids[i].density = _ptr(g, C.c_float)
        vols = (VolumeDesc * max(1, len(self.volumes)))()
        for i, v in enumerate(self.volumes):
            d = vols[i]
            d.sigma_t[:] = [float(x) for x in v["sigma_t"]]
            d.albedo[:] = [float(x) for x in v["albedo"]]
            d.g = v["g"]
            d.density_grid = v["grid"]
            d.density_texture = v.get("density_texture", -1)
            d.world_to_object[:] = v["w2o"].T.reshape(-1).tolist()
        sd = SceneDesc()
        sd.mesh_count, sd.meshes = len(self.meshes), md
        sph = (SphereDesc * max(1, len(self.spheres)))()
        for i, r in enumerate(self.spheres):
            sph[i].radius = float(r)
        sd.sphere_count, sd.spheres = len(self.spheres), sph
        sd.instance_count, sd.instances = len(self.instances), inst
        sd.material_count, sd.materials = len(self.materials), mats
        sd.texture_count, sd.textures = len(self.textures), texs
        sd.grid_count, sd.grids = len(self.grids), grids
        sd.volume_count, sd.volumes = len(self.volumes), vols
        dls = (DistantLightDesc * max(1, len(self.distant_lights)))()
        for i, dl in enumerate(self.distant_lights):
            d = dls[i]
            d.light_color[:] = list(dl["color"])
            d.intensity, d.theta_max = dl["intensity"], dl["theta"]
            d.object_to_world[:] = dl["o2w"].T.reshape(-1).tolist()
            d.world_to_object[:] = dl["w2o"].T.reshape(-1).tolist()
        sd.distant_light_count, sd.distant_lights = len(self.distant_lights), dls
        cam = self.camera
        sd.camera.position[:] = [float(x) for x in cam["position"]]
        sd.camera.w[:] = [float(x) for x in cam["w"]]
        sd.camera.fov = cam["fov"]
        sd.camera.aspect_ratio = cam["aspect"]
        sd.halton_seed = self.halton_seed
        keep.extend([md, sph, inst, mats, texs, grids, vols, dls])
        self._keep = keep
        return sd


def render_settings(width, height, spp, depth, sampler=SAMPLER_HALTON, shape_light=1,
                    volume=0, env_light=0, rr=0, path_capacity=0, distant_light=0, persistent_waves=0):
    rs = RenderSettings()
    rs.width, rs.height, rs.spp, rs.ray_depth = width, height, spp, depth
    rs.sampler_type = sampler
    rs.enable_shape_light = shape_light
    rs.enable_distant_light = distant_light
    rs.enable_volume_scattering = volume
    rs.enable_env_light = env_light
    rs.enable_russian_roulette = rr
    rs.enable_indirect_clamp = 0
    rs.indirect_limit = 10.0
    rs.gamma = 2.2
    rs.exposure = 1.0
    rs.path_capacity = path_capacity
    rs.persistent_waves = persistent_waves
    return rs


class TuneResult(C.Structure):
    """ymrt_tune_result."""
    _fields_ = [("path_capacity", C.c_uint32), ("persistent_waves", C.c_uint32), ("ms", C.c_float),
                ("ms_default", C.c_float), ("trials", C.c_uint32)]


def _np(ptr, n, dtype):
    if n == 0 or not ptr:
        return np.zeros(0, dtype)
    ct = {np.float32: C.c_float, np.uint32: C.c_uint32, np.uint8: C.c_uint8}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).astype(dtype, copy=True)


def _mat(cm):
    """column-major 16 floats -> row-major 4x4"""
    return np.array(list(cm), np.float32).reshape(4, 4).T.copy()


def arrays_from_desc(sd):
    """A SceneArrays holding copies of everything a SceneDesc points to."""
    sa = SceneArrays()
    for i in range(sd.mesh_count):
        m = sd.meshes[i]
        nt = m.triangle_count
        pos = _np(m.positions, 3 * m.position_count, np.float32)
        pidx = _np(m.position_indices, 3 * nt, np.uint32)
        kw = {}
        if m.normals and m.normal_indices:
            kw.update(nrm=_np(m.normals, 3 * m.normal_count, np.float32),
                      nidx=_np(m.normal_indices, 3 * nt, np.uint32))
        if m.texcoords and m.texcoord_indices:
            kw.update(uv=_np(m.texcoords, 2 * m.texcoord_count, np.float32),
                      tidx=_np(m.texcoord_indices, 3 * nt, np.uint32))
        sa.add_mesh(pos, pidx, **kw)
    for i in range(sd.sphere_count):
        sa.add_sphere(sd.spheres[i].radius)
    for i in range(sd.material_count):
        m = MaterialDesc()
        C.memmove(C.byref(m), C.byref(sd.materials[i]), C.sizeof(MaterialDesc))
        sa.add_material(m)
    for i in range(sd.texture_count):
        t = sd.textures[i]
        if t.type == TEXTURE_IMAGE and t.texels_f32:
            tex = _np(t.texels_f32, t.width * t.height * t.channels, np.float32).reshape(t.height, t.width,
                                                                                           t.channels)
            sa.add_texture(tex, wrap=t.wrap_mode, scale=(t.scale_u, t.scale_v), offset=(t.offset_u, t.offset_v))
        elif t.type == TEXTURE_IMAGE:
            tex = _np(t.texels, t.width * t.height * t.channels, np.uint8).reshape(t.height, t.width, t.channels)
            sa.add_texture(tex, wrap=t.wrap_mode, color_space=t.color_space, scale=(t.scale_u, t.scale_v),
                           offset=(t.offset_u, t.offset_v))
        else:
            params = {k: getattr(t, k) for k in ("frequency", "lacunarity", "gain", "amplitude", "offset", "variation",
                                                 "radius", "layer_count", "normalized", "turbulence_axes",
                                                 "frequency1", "lacunarity1", "gain1", "layer_count1")}
            sa.add_texture_node(t.type, tuple(None if c == INVALID else c for c in t.child), tuple(t.value),
                                **params)
    for i in range(sd.volume_count):
        v = sd.volumes[i]
        sa.add_volume(tuple(v.sigma_t), tuple(v.albedo), v.g, grid=v.density_grid,
                      world_to_object=_mat(v.world_to_object), density_texture=v.density_texture)
    for i in range(sd.instance_count):
        d = sd.instances[i]
        geom = SphereRef(d.geometry) if d.geometry_type == GEOMETRY_SPHERE else d.geometry
        sa.add_instance(geom, d.material, _mat(d.object_to_world), _mat(d.world_to_object),
                        inner_volume=d.inner_volume, boundary=bool(d.treat_as_boundary))
    for i in range(sd.distant_light_count):
        d = sd.distant_lights[i]
        sa.add_distant_light(tuple(d.light_color), d.intensity, d.theta_max, _mat(d.object_to_world),
                             _mat(d.world_to_object))
    c = sd.camera
    sa.camera = dict(position=tuple(c.position), w=tuple(c.w), fov=float(c.fov), aspect=float(c.aspect_ratio))
    sa.halton_seed = sd.halton_seed
    return sa
