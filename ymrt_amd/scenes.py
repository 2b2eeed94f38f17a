"""Procedural scenes and ray sets for the workloads the tracer is measured on.

Every generator is deterministic (numpy PCG64 / RandomState seeded as stated)
and returns an ``abi.SceneArrays`` that both the HIP path and the CPU oracle
consume, so the two see bit-identical float inputs.

  cornell_box   -- 5 wall quads, two boxes, a 32x16 UV sphere, one 0.3 m
                   area light of radiance 15; Lambert / GGX conductor /
                   dielectric materials.
  interior      -- about 260k triangles in 40 instances of 8 procedural meshes
                   with random affine transforms, four 2048^2 RGBA8
                   value-noise textures, two quad lights, homogeneous fog.
  heightfield   -- 1001x1001-vertex fBm height field (2M triangles) and the
                   incoherent ray sets of the traversal stress test.
"""
import numpy as np

from . import abi


# ----------------------------------------------------------------- primitives
def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3) with the face normal of the first."""
    pos = np.array([p0, p1, p2, p3], np.float32)
    n = np.cross(pos[1] - pos[0], pos[2] - pos[0])
    n = (n / np.linalg.norm(n)).astype(np.float32)
    nrm = np.repeat(n[None], 4, axis=0)
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return pos, idx, nrm, uv


def box(lo, hi):
    """Axis-aligned box, 12 outward-facing triangles, per-face normals."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    c = [np.array([hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]], np.float32)
         for i in range(8)]
    faces = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    pos, idx, nrm, uv = [], [], [], []
    for f in faces:
        p, i, n, t = quad(*[c[k] for k in f])
        idx.append(i + len(pos) * 4)
        pos.append(p)
        nrm.append(n)
        uv.append(t)
    return (np.concatenate(pos), np.concatenate(idx).astype(np.uint32), np.concatenate(nrm), np.concatenate(uv))


def uv_sphere(center, radius, nu=32, nv=16):
    """Latitude-longitude sphere with nu*nv*2 - 2*nu triangles (960 for 32x16)."""
    center = np.asarray(center, np.float64)
    verts, nrms, uvs = [], [], []
    for j in range(nv + 1):
        th = np.pi * j / nv
        for i in range(nu + 1):
            ph = 2 * np.pi * (i % nu) / nu  # the seam column repeats column 0 exactly
            n = np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)])
            if j == 0 or j == nv:  # poles: one exact point
                n = np.array([0.0, 1.0 if j == 0 else -1.0, 0.0])
            verts.append(center + radius * n)
            nrms.append(n)
            uvs.append([i / nu, j / nv])
    idx = []
    for j in range(nv):
        for i in range(nu):
            a = j * (nu + 1) + i
            b = a + nu + 1
            if j != 0:
                idx.append([a, a + 1, b])
            if j != nv - 1:
                idx.append([a + 1, b + 1, b])
    return (np.array(verts, np.float32), np.array(idx, np.uint32), np.array(nrms, np.float32),
            np.array(uvs, np.float32))


def _add(sa, geom):
    pos, idx, nrm, uv = geom
    return sa.add_mesh(pos, idx, nrm=nrm, nidx=idx, uv=uv, tidx=idx)


# -------------------------------------------------------------- Cornell box
def cornell_box(width=256, height=256, halton_seed=0, analytic_sphere=False, sphere_light=False):
    """Unit Cornell box [0,1]^3 open at z=1 (the camera side).  analytic_sphere
    puts an analytic sphere where the glass UV sphere is; sphere_light replaces
    the ceiling quad light with a sphere light of radius 0.08."""
    sa = abi.SceneArrays()
    sa.halton_seed = halton_seed
    white = sa.add_material(abi.default_material(diffuse_albedo=(0.7, 0.7, 0.7), specular_weight=0.0))
    red = sa.add_material(abi.default_material(diffuse_albedo=(0.7, 0.1, 0.1), specular_weight=0.0))
    green = sa.add_material(abi.default_material(diffuse_albedo=(0.1, 0.7, 0.1), specular_weight=0.0))
    metal = sa.add_material(abi.default_material(diffuse_albedo=(0.9, 0.9, 0.9), specular_albedo=(1.0, 1.0, 1.0),
                                                 roughness_x=0.3, roughness_y=0.3, metalness=1.0))
    glass = sa.add_material(abi.default_material(diffuse_albedo=(1.0, 1.0, 1.0), specular_albedo=(1.0, 1.0, 1.0),
                                                 roughness_x=0.05, roughness_y=0.05, transmission_weight=1.0,
                                                 ior_n=1.5, ior_priority=1))
    light = sa.add_material(abi.light_material((1.0, 1.0, 1.0), 15.0))
    walls = [
        (quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), white),      # floor, normal +y
        (quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), white),      # ceiling, normal -y
        (quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), white),      # back wall, normal +z
        (quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), red),        # left wall, normal +x
        (quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), green),      # right wall, normal -x
    ]
    for geom, mat in walls:
        sa.add_instance(_add(sa, geom), mat)
    sa.add_instance(_add(sa, box((0.55, 0.0, 0.45), (0.85, 0.3, 0.75))), white)
    sa.add_instance(_add(sa, box((0.15, 0.0, 0.15), (0.45, 0.6, 0.45))), metal)
    if analytic_sphere:
        sa.add_instance(sa.add_sphere(0.15), glass, _translate(0.65, 0.45, 0.55))
    else:
        sa.add_instance(_add(sa, uv_sphere((0.65, 0.45, 0.55), 0.15)), glass)
    if sphere_light:
        m = _translate(0.5, 0.85, 0.5)
        sa.add_instance(sa.add_sphere(0.08), sa.add_material(abi.light_material((1.0, 1.0, 1.0), 30.0)), m)
    else:
        lq = quad((0.35, 0.999, 0.35), (0.65, 0.999, 0.35), (0.65, 0.999, 0.65), (0.35, 0.999, 0.65))
        sa.add_instance(_add(sa, lq), light)
    sa.set_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), 40.0, width / height)
    return sa


def normal_map_texels(n=64, waves=4, slope=0.35):
    """Tangent-space normal map of n x n texels (linear RGB8): a bump field of
    `waves` sine periods per side, n = normalize(-slope cos, -slope cos, 1)."""
    u = (np.arange(n) + 0.5) / n
    gx = -slope * np.cos(2 * np.pi * waves * u)[None, :] * np.ones((n, 1))
    gy = -slope * np.cos(2 * np.pi * waves * u)[:, None] * np.ones((1, n))
    v = np.stack([gx, gy, np.ones((n, n))], -1)
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return np.round((v * 0.5 + 0.5) * 255.0).astype(np.uint8)


def texture_node_set():
    """A Cornell box plus one texture graph of every node type, for evaluating
    textures directly; returns (arrays, {name: texture index})."""
    sa = cornell_box(8, 8)
    T = abi
    c = lambda *rgb: sa.add_texture_node(T.TEXTURE_CONSTANT_RGB, value=rgb)
    n = {}
    n["black"], n["white"] = c(0.0, 0.0, 0.0), c(1.0, 1.0, 1.0)
    bw = (n["black"], n["white"])
    n["float"] = sa.add_texture_node(T.TEXTURE_CONSTANT_FLOAT, value=(0.375, 0.0, 0.0))
    n["rgb"] = c(0.25, 0.5, 0.75)
    n["checker"] = sa.add_texture_node(T.TEXTURE_CHECKERBOARD, (n["rgb"], n["float"]), frequency=np.pi)
    n["noise"] = sa.add_texture_node(T.TEXTURE_NOISE, bw, frequency=3.0)
    n["noise_n"] = sa.add_texture_node(T.TEXTURE_NOISE, bw, frequency=3.0, normalized=1)
    fb = dict(frequency=2.5, lacunarity=2.1, gain=0.45, layer_count=5, amplitude=0.9, offset=0.05)
    n["fbm"] = sa.add_texture_node(T.TEXTURE_FBM, bw, **fb)
    n["turb"] = sa.add_texture_node(T.TEXTURE_TURBULENCE, bw, **fb)
    n["marble"] = sa.add_texture_node(T.TEXTURE_MARBLE, bw, frequency=2.0, lacunarity=2.0, gain=0.5,
                                      layer_count=4, variation=5.0, turbulence_axes=5)
    n["wood"] = sa.add_texture_node(T.TEXTURE_WOOD, bw, frequency=1.5, variation=7.0)
    n["polka"] = sa.add_texture_node(T.TEXTURE_POLKA_DOT, bw, frequency=5.0, radius=0.3)
    n["wave"] = sa.add_texture_node(T.TEXTURE_WAVE, bw, frequency=4.0, lacunarity=2.0, gain=0.5, layer_count=3,
                                    frequency1=1.5, lacunarity1=1.8, gain1=0.7, layer_count1=2)
    # nested: a checkerboard whose black input is empty and whose white input
    # is an fBm node over a noise node and a constant
    n["inner"] = sa.add_texture_node(T.TEXTURE_FBM, (n["noise_n"], n["rgb"]), **fb)
    n["nested"] = sa.add_texture_node(T.TEXTURE_CHECKERBOARD, (None, n["inner"]), frequency=2.0)
    tex = np.zeros((4, 4, 3), np.uint8)
    tex[..., 0] = np.arange(16).reshape(4, 4) * 16
    n["image"] = sa.add_texture(tex, color_space=T.COLOR_LINEAR, scale=(2.0, 3.0), offset=(0.1, -0.2))
    n["image_in_graph"] = sa.add_texture_node(T.TEXTURE_MARBLE, (n["image"], n["fbm"]), frequency=1.0,
                                              lacunarity=2.0, gain=0.5, layer_count=2, variation=2.0,
                                              turbulence_axes=2)
    return sa, n


def textured_box(width=64, height=64, halton_seed=0):
    """Cornell box whose materials are driven by texture graphs: a checkerboard
    floor whose white input is an fBm node, a marble wall, a wood-textured
    roughness, polka dots on the ceiling (surface uv), turbulence with a clear
    coat, a wave-textured analytic sphere, a normal-mapped box (8-bit linear
    image) and a bump-mapped back wall (height from a noise node)."""
    T = abi
    sa = abi.SceneArrays()
    sa.halton_seed = halton_seed
    c = lambda *rgb: sa.add_texture_node(T.TEXTURE_CONSTANT_RGB, value=rgb)
    f = lambda v: sa.add_texture_node(T.TEXTURE_CONSTANT_FLOAT, value=(v, 0.0, 0.0))
    fbm = sa.add_texture_node(T.TEXTURE_FBM, (c(0.5, 0.45, 0.3), c(0.9, 0.85, 0.7)), frequency=9.0,
                              lacunarity=2.0, gain=0.5, layer_count=4, amplitude=0.8, offset=0.1)
    checker = sa.add_texture_node(T.TEXTURE_CHECKERBOARD, (c(0.15, 0.15, 0.2), fbm), frequency=18.0)
    marble = sa.add_texture_node(T.TEXTURE_MARBLE, (c(0.6, 0.1, 0.1), c(0.9, 0.85, 0.8)), frequency=3.0,
                                 lacunarity=2.0, gain=0.5, layer_count=4, variation=6.0, turbulence_axes=3)
    wood = sa.add_texture_node(T.TEXTURE_WOOD, (f(0.1), f(0.7)), frequency=2.5, variation=6.0)
    polka = sa.add_texture_node(T.TEXTURE_POLKA_DOT, (c(0.7, 0.7, 0.7), c(0.2, 0.3, 0.8)), frequency=6.0,
                                radius=0.3)
    turb = sa.add_texture_node(T.TEXTURE_TURBULENCE, (c(0.1, 0.1, 0.5), c(0.9, 0.8, 0.2)), frequency=7.0,
                               lacunarity=2.1, gain=0.55, layer_count=3, amplitude=1.0, offset=0.0)
    wave = sa.add_texture_node(T.TEXTURE_WAVE, (c(0.2, 0.6, 0.3), c(0.9, 0.9, 0.9)), frequency=12.0,
                               lacunarity=2.0, gain=0.5, layer_count=3, frequency1=5.0, lacunarity1=1.9,
                               gain1=0.6, layer_count1=2)
    bump = sa.add_texture_node(T.TEXTURE_NOISE, (None, f(0.02)), frequency=25.0, normalized=1)
    nmap = sa.add_texture(normal_map_texels(), color_space=T.COLOR_LINEAR)
    floor = sa.add_material(abi.default_material(diffuse_albedo_tex=checker, specular_weight=0.0))
    ceil = sa.add_material(abi.default_material(diffuse_albedo_tex=polka, specular_weight=0.0))
    back = sa.add_material(abi.default_material(diffuse_albedo=(0.7, 0.7, 0.7), specular_weight=0.0,
                                                bump_mapping_tex=bump))
    left = sa.add_material(abi.default_material(diffuse_albedo_tex=marble, roughness_x=0.4, roughness_y=0.4))
    right = sa.add_material(abi.default_material(diffuse_albedo=(0.1, 0.6, 0.15), roughness_x_tex=wood,
                                                 roughness_y_tex=wood))
    coated = sa.add_material(abi.default_material(diffuse_albedo_tex=turb, specular_weight=0.0, coat_weight=1.0,
                                                  coat_roughness_x=0.1, coat_roughness_y=0.1))
    metal = sa.add_material(abi.default_material(diffuse_albedo=(0.9, 0.9, 0.9), specular_albedo=(1.0, 1.0, 1.0),
                                                 roughness_x=0.25, roughness_y=0.25, metalness=1.0,
                                                 normal_mapping_tex=nmap))
    ball = sa.add_material(abi.default_material(diffuse_albedo_tex=wave, specular_weight=0.5))
    light = sa.add_material(abi.light_material((1.0, 1.0, 1.0), 15.0))
    # the floor's object space is lifted by 0.05 so that p_object.y is not a
    # zero of the checkerboard's sin(y * frequency)
    sa.add_instance(_add(sa, quad((0, 0.05, 0), (0, 0.05, 1), (1, 0.05, 1), (1, 0.05, 0))), floor,
                    _translate(0.0, -0.05, 0.0))
    walls = [
        (quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), ceil),
        (quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), back),
        (quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), left),
        (quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), right),
    ]
    for geom, mat in walls:
        sa.add_instance(_add(sa, geom), mat)
    sa.add_instance(_add(sa, box((0.55, 0.0, 0.45), (0.85, 0.3, 0.75))), coated)
    sa.add_instance(_add(sa, box((0.15, 0.0, 0.15), (0.45, 0.6, 0.45))), metal)
    sa.add_instance(sa.add_sphere(0.15), ball, _translate(0.65, 0.45, 0.55))
    lq = quad((0.35, 0.999, 0.35), (0.65, 0.999, 0.35), (0.65, 0.999, 0.65), (0.35, 0.999, 0.65))
    sa.add_instance(_add(sa, lq), light)
    sa.set_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), 40.0, width / height)
    return sa


def cornell_media(width=64, height=64, grid=True, halton_seed=0, sigma=2.0, density_texture=False):
    """Cornell box with a participating medium in the box [0.2,0.8] x [0.05,0.65]
    x [0.2,0.8] (a boundary instance): an 8^3 voxel density grid in [0, 1]
    scaled by sigma when grid is True, else homogeneous sigma * 0.5.
    density_texture: the density is instead an fBm texture graph (3 layers,
    frequency 4 over the unit cube of the box) under its upper bound, as
    YMRT's density_texture_idx volumes."""
    sa = cornell_box(width, height, halton_seed)
    lo, hi = np.array((0.2, 0.05, 0.2)), np.array((0.8, 0.65, 0.8))
    w2o = np.eye(4, dtype=np.float32)
    w2o[:3, :3] = np.diag(1.0 / (hi - lo))
    w2o[:3, 3] = -lo / (hi - lo)
    if density_texture:
        black = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(0.0, 0.0, 0.0))
        white = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(1.0, 0.0, 0.0))
        fbm = sa.add_texture_node(abi.TEXTURE_FBM, (black, white), frequency=4.0, lacunarity=2.0, gain=0.5,
                                  amplitude=0.6, offset=0.3, layer_count=3)
        vol = sa.add_volume((sigma, sigma, sigma), (0.9, 0.7, 0.5), 0.2, world_to_object=w2o,
                            density_texture=fbm)
    elif grid:
        rng = np.random.RandomState(5)
        g = sa.add_grid(rng.rand(8, 8, 8).astype(np.float32))
        vol = sa.add_volume((sigma, sigma, sigma), (0.9, 0.7, 0.5), 0.2, grid=g, world_to_object=w2o)
    else:
        vol = sa.add_volume((0.5 * sigma,) * 3, (0.9, 0.7, 0.5), 0.2, world_to_object=w2o)
    mat = sa.add_material(abi.default_material(ior_n=1.0))
    sa.add_instance(_add(sa, box(tuple(lo), tuple(hi))), mat, inner_volume=vol, boundary=True)
    return sa


def _translate(x, y, z, scale=1.0):
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] *= scale
    m[:3, 3] = (x, y, z)
    return m


def outdoor(width=64, height=48, fog=False, halton_seed=0):
    """Ground plane with boxes and analytic spheres under a distant light (cone
    half-angle 5 degrees, from above at an angle) and one small quad light;
    fog adds a homogeneous medium box that shadow rays must cross."""
    sa = abi.SceneArrays()
    sa.halton_seed = halton_seed
    ground = sa.add_material(abi.default_material(diffuse_albedo=(0.6, 0.6, 0.55), specular_weight=0.0))
    red = sa.add_material(abi.default_material(diffuse_albedo=(0.7, 0.2, 0.15), roughness_x=0.3, roughness_y=0.3))
    metal = sa.add_material(abi.default_material(diffuse_albedo=(0.9, 0.8, 0.5), specular_albedo=(1.0, 1.0, 1.0),
                                                 roughness_x=0.2, roughness_y=0.2, metalness=1.0))
    glass = sa.add_material(abi.default_material(diffuse_albedo=(1.0, 1.0, 1.0), roughness_x=0.05,
                                                 roughness_y=0.05, transmission_weight=1.0, ior_n=1.5,
                                                 ior_priority=1))
    sa.add_instance(_add(sa, quad((-20, 0, -20), (-20, 0, 20), (20, 0, 20), (20, 0, -20))), ground)
    sa.add_instance(_add(sa, box((-1.5, 0.0, -1.0), (-0.5, 1.2, 0.0))), red)
    sa.add_instance(sa.add_sphere(0.6), metal, _translate(0.8, 0.6, -0.3))
    sa.add_instance(sa.add_sphere(0.4), glass, _translate(0.0, 0.4, 1.0))
    lq = quad((2.0, 2.5, -1.0), (2.5, 2.5, -1.0), (2.5, 2.5, -0.5), (2.0, 2.5, -0.5))
    sa.add_instance(_add(sa, lq), sa.add_material(abi.light_material((1.0, 0.9, 0.8), 20.0)))
    # light frame: -y of the frame points to the light (up and to the side)
    ang = np.radians(35.0)
    rot = np.array([[1, 0, 0, 0], [0, -np.cos(ang), np.sin(ang), 0], [0, -np.sin(ang), -np.cos(ang), 0],
                    [0, 0, 0, 1]], np.float32)
    sa.add_distant_light((1.0, 0.95, 0.9), 3.0, 5.0, rot)
    if fog:
        vol = sa.add_volume((0.3, 0.3, 0.3), (0.9, 0.9, 0.9), 0.2)
        sa.add_instance(_add(sa, box((-2.0, 0.01, -1.5), (2.0, 1.8, 2.0))),
                        sa.add_material(abi.default_material(ior_n=1.0)), inner_volume=vol, boundary=True)
    sa.set_camera((0.0, 1.6, 5.0), (0.0, 0.5, 0.0), 45.0, width / height)
    return sa


# --------------------------------------------------------------- value noise
def value_noise_2d(n, octaves, seed, lattice=16):
    """fBm of bilinearly interpolated value noise on an n x n grid, in [0, 1]."""
    rng = np.random.RandomState(seed)
    out = np.zeros((n, n), np.float64)
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = lattice * (2 ** o)
        lat = rng.rand(cells + 1, cells + 1)
        x = np.linspace(0, cells, n, endpoint=False)
        i0 = np.floor(x).astype(int)
        f = x - i0
        f = f * f * (3 - 2 * f)
        a = lat[i0][:, i0]
        b = lat[i0][:, i0 + 1]
        c = lat[i0 + 1][:, i0]
        d = lat[i0 + 1][:, i0 + 1]
        fy = f[:, None]
        fx = f[None, :]
        out += amp * ((a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy)
        total += amp
        amp *= 0.5
    return out / total


# -------------------------------------------------------------- height field
def heightfield(verts_per_side=1001, octaves=6, amplitude=0.1, seed=1):
    """(V-1)^2*2 triangles on [0,1]^2 (x,z), height y from a 512^2 fBm lattice."""
    sa = abi.SceneArrays()
    n = verts_per_side
    lat = value_noise_2d(512, octaves, seed)
    g = np.linspace(0.0, 1.0, n)
    # bilinear sample of the 512^2 lattice at the vertex positions
    s = g * 511.0
    i0 = np.minimum(np.floor(s).astype(int), 510)
    f = s - i0
    a = lat[i0][:, i0]
    b = lat[i0][:, i0 + 1]
    c = lat[i0 + 1][:, i0]
    d = lat[i0 + 1][:, i0 + 1]
    fz = f[:, None]
    fx = f[None, :]
    h = amplitude * ((a * (1 - fx) + b * fx) * (1 - fz) + (c * (1 - fx) + d * fx) * fz)
    X, Z = np.meshgrid(g, g)
    pos = np.stack([X, h, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    k = np.arange(n - 1)
    I, J = np.meshgrid(k, k)
    v00 = (J * n + I).reshape(-1)
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    tri = np.empty((2 * len(v00), 3), np.uint32)
    tri[0::2] = np.stack([v00, v01, v11], axis=-1)
    tri[1::2] = np.stack([v00, v11, v10], axis=-1)
    mesh = sa.add_mesh(pos, tri)
    mat = sa.add_material(abi.default_material())
    sa.add_instance(mesh, mat)
    sa.set_camera((0.5, 0.5, 2.0), (0.5, 0.0, 0.5), 45.0, 1.0)
    return sa


def bounds(sa):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for it in sa.instances:
        if it["geometry_type"] == abi.GEOMETRY_SPHERE:
            r = sa.spheres[it["mesh"]]
            p = np.array([[x, y, z] for x in (-r, r) for y in (-r, r) for z in (-r, r)], np.float64)
        else:
            p = sa.meshes[it["mesh"]]["pos"].astype(np.float64)
        ph = np.c_[p, np.ones(len(p))] @ it["o2w"].astype(np.float64).T
        lo = np.minimum(lo, ph[:, :3].min(0))
        hi = np.maximum(hi, ph[:, :3].max(0))
    return lo, hi


def stress_rays(sa, n, seed_origin=2, seed_dir=3, seed_target=4):
    """Closest-hit rays (origins uniform in the scene AABB, directions uniform
    on S^2) and any-hit rays (segments to uniform points in the AABB)."""
    lo, hi = bounds(sa)
    o = lo + np.random.RandomState(seed_origin).rand(n, 3) * (hi - lo)
    u = np.random.RandomState(seed_dir).rand(n, 2)
    z = 1.0 - 2.0 * u[:, 0]
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    ph = 2.0 * np.pi * u[:, 1]
    d = np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=-1)
    closest = np.zeros(n, abi.RAY_DTYPE)
    closest["o"] = o.astype(np.float32)
    closest["d"] = d.astype(np.float32)
    closest["tmin"] = np.float32(1e-7)
    closest["tmax"] = np.float32(1e36)
    tgt = lo + np.random.RandomState(seed_target).rand(n, 3) * (hi - lo)
    seg = tgt - o
    length = np.linalg.norm(seg, axis=1)
    shadow = np.zeros(n, abi.RAY_DTYPE)
    shadow["o"] = o.astype(np.float32)
    shadow["d"] = (seg / length[:, None]).astype(np.float32)
    shadow["tmin"] = np.float32(1e-7)
    shadow["tmax"] = length.astype(np.float32)
    return closest, shadow


# ------------------------------------------------------------- interior (C2)
def _grid_mesh(nu, nv, fn):
    """Parametric surface p = fn(u, v) on an (nu+1) x (nv+1) vertex grid."""
    u = np.linspace(0.0, 1.0, nu + 1)
    v = np.linspace(0.0, 1.0, nv + 1)
    U, V = np.meshgrid(u, v)
    P = fn(U, V)
    pos = np.stack(P, axis=-1).reshape(-1, 3)
    uv = np.stack([U, V], axis=-1).reshape(-1, 2)
    i = np.arange(nu)
    j = np.arange(nv)
    I, J = np.meshgrid(i, j)
    a = (J * (nu + 1) + I).reshape(-1)
    b = a + 1
    c = a + nu + 1
    d = c + 1
    idx = np.empty((2 * len(a), 3), np.int64)
    idx[0::2] = np.stack([a, c, d], -1)
    idx[1::2] = np.stack([a, d, b], -1)
    return pos, idx, uv


def _smooth_normals(pos, idx):
    n = np.zeros_like(pos)
    fn = np.cross(pos[idx[:, 1]] - pos[idx[:, 0]], pos[idx[:, 2]] - pos[idx[:, 0]])
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    ln = np.linalg.norm(n, axis=1)
    ln[ln == 0] = 1.0
    return n / ln[:, None]


def _finish(pos, idx, uv):
    pos = pos - 0.5 * (pos.min(0) + pos.max(0))
    pos = pos / np.abs(pos).max()
    nrm = _smooth_normals(pos, idx)
    return (pos.astype(np.float32), idx.astype(np.uint32), nrm.astype(np.float32), uv.astype(np.float32))


def interior_meshes():
    """Eight procedural meshes of about 6.5k triangles each, unit extent."""
    tau = 2 * np.pi
    out = []
    # 1 sphere
    out.append(_finish(*_grid_mesh(64, 51, lambda u, v: (np.sin(np.pi * v) * np.cos(tau * u), np.cos(np.pi * v),
                                                          np.sin(np.pi * v) * np.sin(tau * u)))))
    # 2 torus
    out.append(_finish(*_grid_mesh(64, 51, lambda u, v: ((1 + 0.35 * np.cos(tau * v)) * np.cos(tau * u),
                                                          0.35 * np.sin(tau * v),
                                                          (1 + 0.35 * np.cos(tau * v)) * np.sin(tau * u)))))
    # 3 displaced grid
    hf = value_noise_2d(65, 3, 101, lattice=4)
    out.append(_finish(*_grid_mesh(57, 57, lambda u, v: (u, 0.25 * hf[(v * 57).astype(int), (u * 57).astype(int)], v))))
    # 4 cylinder (open)
    out.append(_finish(*_grid_mesh(64, 51, lambda u, v: (np.cos(tau * u), 2 * v - 1, np.sin(tau * u)))))
    # 5 bumpy sphere
    bump = value_noise_2d(129, 3, 102, lattice=8)
    out.append(_finish(*_grid_mesh(64, 51, lambda u, v: tuple(
        (1 + 0.15 * bump[(v * 128).astype(int), (u * 128).astype(int)]) * c
        for c in (np.sin(np.pi * v) * np.cos(tau * u), np.cos(np.pi * v), np.sin(np.pi * v) * np.sin(tau * u))))))
    # 6 cone
    out.append(_finish(*_grid_mesh(64, 51, lambda u, v: ((1 - v) * np.cos(tau * u), 2 * v - 1, (1 - v) * np.sin(tau * u)))))
    # 7 twisted ribbon
    out.append(_finish(*_grid_mesh(64, 51, lambda u, v: ((v - 0.5) * np.cos(tau * u), 2 * u - 1,
                                                          (v - 0.5) * np.sin(tau * u)))))
    # 8 spring (tube along a helix)
    def spring(u, v):
        t = 3 * tau * u
        c = np.stack([np.cos(t), 0.15 * t - 1.4, np.sin(t)])
        r = 0.2
        return (c[0] + r * np.cos(tau * v) * np.cos(t), c[1] + r * np.sin(tau * v), c[2] + r * np.cos(tau * v) * np.sin(t))
    out.append(_finish(*_grid_mesh(127, 25, spring)))
    return out


def value_noise_texture(n, seed, octaves=4):
    """RGBA8 value-noise texture (n x n): three fBm channels, alpha 255."""
    chans = [value_noise_2d(n, octaves, seed * 3 + c) for c in range(3)]
    rgb = np.stack(chans, axis=-1)
    lo, hi = rgb.min(), rgb.max()
    rgb = (rgb - lo) / max(hi - lo, 1e-9)
    tex = np.empty((n, n, 4), np.uint8)
    tex[..., :3] = np.clip(rgb * 255.0 + 0.5, 0, 255).astype(np.uint8)
    tex[..., 3] = 255
    return tex


def _rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def interior(width=1920, height=1080, tex_size=2048, instances=40, halton_seed=0, fog=True):
    """Procedural room: 40 random affine instances of 8 meshes (~260k
    triangles), four value-noise RGBA8 textures, two ceiling quad lights and
    homogeneous fog (sigma_t 0.05, albedo 0.8, g 0.3) filling a boundary box
    that contains the camera."""
    sa = abi.SceneArrays()
    sa.halton_seed = halton_seed
    texs = [sa.add_texture(value_noise_texture(tex_size, s)) for s in (7, 8, 9, 10)]
    meshes = [_add(sa, m) for m in interior_meshes()]
    rng = np.random.RandomState(42)
    mrng = np.random.RandomState(43)
    for i in range(instances):
        if mrng.rand() < 0.5:
            mat = abi.default_material(diffuse_albedo=(1.0, 1.0, 1.0), diffuse_albedo_tex=texs[i % 4],
                                       specular_weight=0.0, ior_n=1.0)
        else:
            r = float(mrng.uniform(0.05, 0.8))
            mat = abi.default_material(diffuse_albedo=tuple(mrng.uniform(0.4, 0.95, 3)),
                                       specular_albedo=(1.0, 1.0, 1.0), roughness_x=r, roughness_y=r,
                                       metalness=1.0, ior_n=1.0)
            if i % 3 == 0:
                mat.diffuse_albedo_tex = texs[(i + 1) % 4]
        m = np.eye(4)
        m[:3, :3] = _rotation(rng) * rng.uniform(0.5, 2.0) * 0.45
        m[:3, 3] = (rng.uniform(-4.0, 4.0), rng.uniform(0.6, 3.0), rng.uniform(-4.0, 2.5))
        sa.add_instance(meshes[i % len(meshes)], sa.add_material(mat), m.astype(np.float32))
    # room: inward-facing box [-5,5] x [0,4] x [-5,5], textured walls
    pos, idx, nrm, uv = box((-5.0, 0.0, -5.0), (5.0, 4.0, 5.0))
    idx = idx[:, ::-1].copy()
    room = sa.add_mesh(pos, idx, nrm=-nrm, nidx=idx, uv=uv * 4.0, tidx=idx)
    wall = sa.add_material(abi.default_material(diffuse_albedo=(0.8, 0.8, 0.8), diffuse_albedo_tex=texs[0],
                                                specular_weight=0.0, ior_n=1.0))
    sa.add_instance(room, wall)
    light = sa.add_material(abi.light_material((1.0, 0.95, 0.85), 12.0))
    for x0, z0 in ((-2.5, -1.5), (1.5, 0.5)):
        _ = sa.add_instance(_add(sa, quad((x0, 3.99, z0), (x0 + 1.0, 3.99, z0), (x0 + 1.0, 3.99, z0 + 1.0),
                                          (x0, 3.99, z0 + 1.0))), light)
    if fog:
        vol = sa.add_volume((0.05, 0.05, 0.05), (0.8, 0.8, 0.8), 0.3)
        fb = box((-5.2, -0.2, -5.2), (5.2, 4.2, 5.2))
        fogmat = sa.add_material(abi.default_material(ior_n=1.0))
        sa.add_instance(_add(sa, fb), fogmat, inner_volume=vol, boundary=True)
    sa.set_camera((0.0, 1.7, 4.6), (0.0, 1.4, 0.0), 60.0, width / height)
    return sa


# --------------------------------------------------------- multi-scene batch
def batch_scene(triangles, width, height, seed):
    """One scene of the multi-scene workload: the Cornell box (walls, two
    boxes, glass sphere, quad light; 1k triangles) whose floor carries a
    displaced grid of about `triangles` triangles in all (log-uniform over
    the batch), glossy or diffuse by seed."""
    sa = cornell_box(width, height, halton_seed=seed)
    extra = max(0, int(triangles) - sa.triangle_count)
    if extra >= 2:
        n = max(2, int(round(np.sqrt(extra / 2.0))) + 1)          # n x n vertices, 2 (n-1)^2 triangles
        rng = np.random.RandomState(seed)
        g = np.linspace(0.02, 0.98, n)
        X, Z = np.meshgrid(g, g)
        ph = rng.uniform(0.0, 6.28, 4)
        Y = 0.004 + 0.02 * (0.5 + 0.25 * np.sin(9.0 * X + ph[0]) * np.sin(7.0 * Z + ph[1])
                            + 0.25 * np.sin(23.0 * X + ph[2]) * np.cos(19.0 * Z + ph[3]))
        pos = np.stack([X, Y, Z], axis=-1).reshape(-1, 3).astype(np.float32)
        k = np.arange(n - 1)
        I, J = np.meshgrid(k, k)
        v00 = (J * n + I).reshape(-1)
        v10, v01 = v00 + 1, v00 + n
        v11 = v01 + 1
        idx = np.concatenate([np.stack([v00, v01, v11], 1), np.stack([v00, v11, v10], 1)]).astype(np.uint32)
        if rng.rand() < 0.5:
            mat = abi.default_material(diffuse_albedo=tuple(rng.uniform(0.3, 0.9, 3)), specular_weight=0.0)
        else:
            r = float(rng.uniform(0.1, 0.6))
            mat = abi.default_material(diffuse_albedo=tuple(rng.uniform(0.5, 0.95, 3)), roughness_x=r,
                                       roughness_y=r, metalness=1.0)
        sa.add_instance(sa.add_mesh(pos, idx), sa.add_material(mat))
    return sa


def batch_workload(count=64, seed=11):
    """The multi-scene workload: `count` scenes with triangle counts
    log-uniform in [1k, 500k], resolution uniform in {256^2, 512^2, 1024^2},
    spp uniform in {4, 16, 64}, depth 6.  Returns [(tris, res, spp, depth)]."""
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(count):
        tris = int(round(np.exp(rng.uniform(np.log(1e3), np.log(5e5)))))
        res = int(rng.choice([256, 512, 1024]))
        spp = int(rng.choice([4, 16, 64]))
        out.append((tris, res, spp, 6))
    return out


# ------------------------------------------------------- inverse rendering
def inverse_params(tex_size=256, grid_n=64, n_rough=8, seed=3):
    """Parameters of the differentiable Cornell workload: an RGB float albedo
    texture (tex_size^2, values in [0.15, 0.85]), n_rough GGX roughnesses in
    [0.15, 0.6] and a grid_n^3 density grid in [0, 1] (a few smooth blobs)."""
    rng = np.random.RandomState(seed)
    lattice = max(2, min(16, tex_size // 4))
    tex = np.stack([value_noise_2d(tex_size, 3, seed * 7 + c, lattice=lattice) for c in range(3)], -1)
    tex = (0.15 + 0.7 * tex).astype(np.float32)
    rough = rng.uniform(0.15, 0.6, n_rough).astype(np.float32)
    c = (np.arange(grid_n) + 0.5) / grid_n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    dens = np.zeros((grid_n, grid_n, grid_n), np.float64)
    for _ in range(6):
        p = rng.uniform(0.2, 0.8, 3)
        r = rng.uniform(0.12, 0.3)
        dens += np.exp(-((x - p[0]) ** 2 + (y - p[1]) ** 2 + (z - p[2]) ** 2) / (2 * r * r))
    dens = (dens / dens.max()).astype(np.float32)
    return dict(texture=tex, roughness=rough, density=dens)


def perturb_params(params, frac=0.3, seed=5):
    """Each parameter times (1 + u), u uniform in [-frac, frac]; albedo texels
    clipped to [0, 1], roughness to [0.01, 0.99], density to >= 0."""
    rng = np.random.RandomState(seed)
    out = {}
    for k, v in params.items():
        out[k] = (v * (1.0 + rng.uniform(-frac, frac, v.shape))).astype(np.float32)
    out["texture"] = np.clip(out["texture"], 0.0, 1.0).astype(np.float32)
    out["roughness"] = np.clip(out["roughness"], 0.01, 0.99).astype(np.float32)
    out["density"] = np.maximum(out["density"], 0.0).astype(np.float32)
    return out


def cornell_inverse(width=512, height=512, params=None, sigma_max=2.0, halton_seed=0):
    """The Cornell box of the differentiable workload: the floor's albedo is a
    float RGB texture, len(roughness) metal spheres of their own GGX roughness
    stand on it in two rows, and a voxel-grid medium (gray sigma_t =
    sigma_max times the density) fills the box's middle.  Returns the scene
    and its targets: dict(albedo=[Lambert materials], roughness=[metal
    materials], texture=index, volume=index, grid=index)."""
    params = inverse_params() if params is None else params
    sa = abi.SceneArrays()
    sa.halton_seed = halton_seed
    white = sa.add_material(abi.default_material(diffuse_albedo=(0.7, 0.7, 0.7), specular_weight=0.0))
    red = sa.add_material(abi.default_material(diffuse_albedo=(0.7, 0.1, 0.1), specular_weight=0.0))
    green = sa.add_material(abi.default_material(diffuse_albedo=(0.1, 0.7, 0.1), specular_weight=0.0))
    tex = sa.add_texture(np.ascontiguousarray(params["texture"], np.float32))
    floor = sa.add_material(abi.default_material(diffuse_albedo=(1.0, 1.0, 1.0), diffuse_albedo_tex=tex,
                                                 specular_weight=0.0))
    light = sa.add_material(abi.light_material((1.0, 1.0, 1.0), 15.0))
    walls = [
        (quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), floor),
        (quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), white),
        (quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), white),
        (quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), red),
        (quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), green),
    ]
    for geom, mat in walls:
        sa.add_instance(_add(sa, geom), mat)
    rough = [float(r) for r in params["roughness"]]
    metals = []
    per_row = (len(rough) + 1) // 2
    sphere = sa.add_sphere(0.07)
    for i, r in enumerate(rough):
        m = sa.add_material(abi.default_material(diffuse_albedo=(0.95, 0.75, 0.45), specular_albedo=(1.0, 1.0, 1.0),
                                                 roughness_x=r, roughness_y=r, metalness=1.0))
        metals.append(m)
        row, col = divmod(i, per_row)
        x = 0.15 + 0.7 * (col + 0.5) / per_row
        z = 0.35 + 0.3 * row
        sa.add_instance(sphere, m, _translate(x, 0.07, z))
    lq = quad((0.35, 0.999, 0.35), (0.65, 0.999, 0.35), (0.65, 0.999, 0.65), (0.35, 0.999, 0.65))
    sa.add_instance(_add(sa, lq), light)
    lo, hi = np.array((0.1, 0.3, 0.1)), np.array((0.9, 0.8, 0.9))
    w2o = np.eye(4, dtype=np.float32)
    w2o[:3, :3] = np.diag(1.0 / (hi - lo))
    w2o[:3, 3] = -lo / (hi - lo)
    grid = sa.add_grid(np.ascontiguousarray(params["density"], np.float32))
    vol = sa.add_volume((sigma_max,) * 3, (0.85, 0.85, 0.85), 0.2, grid=grid, world_to_object=w2o)
    bmat = sa.add_material(abi.default_material(ior_n=1.0))
    sa.add_instance(_add(sa, box(tuple(lo), tuple(hi))), bmat, inner_volume=vol, boundary=True)
    sa.set_camera((0.5, 0.5, 2.2), (0.5, 0.5, 0.0), 40.0, width / height)
    return sa, dict(albedo=[white, red, green], roughness=metals, texture=tex, volume=vol, grid=grid)


def absorbing_slab(width=48, height=16, d0=0.5, dmax=1.0, sigma=2.0, le=1.0):
    """A camera looking along -z through a purely absorbing slab (z in
    [-0.2, 0.2], albedo 0, gray sigma_t = sigma times a 5 x 1 x 1 voxel
    grid along x) at an emissive quad (z = -1): the first four voxels hold d0
    and cover every camera ray, the fifth holds dmax (the majorant) outside
    the view.  A ray's radiance is le * exp(-sigma d0 l) for its chord l in
    the slab.  Returns (scene, grid x bounds of the slab)."""
    sa = abi.SceneArrays()
    light = sa.add_material(abi.light_material((1.0, 1.0, 1.0), le))
    sa.add_instance(_add(sa, quad((-2, -2, -1), (2, -2, -1), (2, 2, -1), (-2, 2, -1))), light)
    lo, hi = np.array((-1.6, -1.0, -0.2)), np.array((2.4, 1.0, 0.2))
    w2o = np.eye(4, dtype=np.float32)
    w2o[:3, :3] = np.diag(1.0 / (hi - lo))
    w2o[:3, 3] = -lo / (hi - lo)
    dens = np.full((1, 1, 5), d0, np.float32)
    dens[0, 0, 4] = dmax
    g = sa.add_grid(dens)
    vol = sa.add_volume((sigma,) * 3, (0.0, 0.0, 0.0), 0.0, grid=g, world_to_object=w2o)
    sa.add_instance(_add(sa, box(tuple(lo), tuple(hi))), sa.add_material(abi.default_material(ior_n=1.0)),
                    inner_volume=vol, boundary=True)
    sa.set_camera((0.0, 0.0, 5.0), (0.0, 0.0, 0.0), 10.0, width / height)
    return sa, (float(lo[0]), float(hi[0]))


def lit_floor(width=32, height=32, material=None, slab=None, light_half=0.4, fov=70.0):
    """A floor at z = -1 (facing +z) under a quad light at z = +1 facing it,
    seen by a camera between them (z = -0.5, looking down): with ray depth 1
    every image term is a next-event ray from the floor to the light (of
    half-width light_half).  material: the floor's (default Lambert 0.6);
    slab: a voxel-grid density
    (nz, ny, nx) filling z in [-0.2, 0.2] that the shadow rays cross (gray
    sigma_t 2, albedo 0.8)."""
    sa = abi.SceneArrays()
    m = material if material is not None else abi.default_material(diffuse_albedo=(0.6, 0.6, 0.6),
                                                                   specular_weight=0.0)
    floor = sa.add_material(m)
    sa.add_instance(_add(sa, quad((-3, -3, -1), (3, -3, -1), (3, 3, -1), (-3, 3, -1))), floor)
    light = sa.add_material(abi.light_material((1.0, 0.9, 0.8), 8.0))
    h = float(light_half)
    sa.add_instance(_add(sa, quad((-h, -h, 1), (-h, h, 1), (h, h, 1), (h, -h, 1))), light)
    if slab is not None:
        lo, hi = np.array((-2.0, -2.0, -0.2)), np.array((2.0, 2.0, 0.2))
        w2o = np.eye(4, dtype=np.float32)
        w2o[:3, :3] = np.diag(1.0 / (hi - lo))
        w2o[:3, 3] = -lo / (hi - lo)
        g = sa.add_grid(np.ascontiguousarray(slab, np.float32))
        vol = sa.add_volume((2.0,) * 3, (0.8, 0.8, 0.8), 0.0, grid=g, world_to_object=w2o)
        sa.add_instance(_add(sa, box(tuple(lo), tuple(hi))), sa.add_material(abi.default_material(ior_n=1.0)),
                        inner_volume=vol, boundary=True)
    sa.set_camera((0.0, 0.0, -0.5), (0.0, 0.0, -1.0), fov, width / height)
    return sa


def open_medium(width=48, height=32, sigma=0.3):
    """A boundary quad (a medium's only face, facing the camera) with a
    homogeneous medium behind it and nothing else: camera rays pass through
    the quad into the medium and then leave the scene with no surface hit
    while inside it, lit by the environment (enable env_light) and by a
    distant light."""
    sa = abi.SceneArrays()
    vol = sa.add_volume((sigma,) * 3, (0.8, 0.7, 0.6), 0.3)
    bmat = sa.add_material(abi.default_material(ior_n=1.0))
    sa.add_instance(_add(sa, quad((-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0))), bmat, inner_volume=vol,
                    boundary=True)
    rot = np.eye(4, dtype=np.float32)
    sa.add_distant_light((1.0, 0.95, 0.9), 2.0, 5.0, rot)
    sa.set_camera((0.3, 0.2, 3.0), (0.0, 0.0, 0.0), 30.0, width / height)
    return sa
