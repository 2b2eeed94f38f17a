"""Python access to the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg, always as the checker.  The product
(ymrt_amd) never imports this package.
"""
import ctypes as C
import os
import subprocess
import threading

import numpy as np

from ymrt_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None
_lock = threading.Lock()


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(_LIB_PATH):
                build()
            L = C.CDLL(_LIB_PATH)
            vp = C.c_void_p
            L.or_scene_create.restype = vp
            L.or_scene_create.argtypes = [C.POINTER(abi.SceneDesc)]
            L.or_scene_destroy.argtypes = [vp]
            L.or_trace_closest_batch.argtypes = [vp, vp, C.c_uint32, vp]
            L.or_trace_any_batch.argtypes = [vp, vp, C.c_uint32, vp]
            L.or_brute_closest_batch.argtypes = [vp, vp, C.c_uint32, vp]
            L.or_render_rows.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, C.c_uint32,
                                         C.c_uint32, vp, vp]
            L.or_halton_table_size.restype = C.c_uint32
            L.or_halton_table.argtypes = [C.c_uint32, vp]
            L.or_radical_inverse.restype = C.c_float
            L.or_radical_inverse.argtypes = [C.c_uint64, C.c_uint32, vp]
            L.or_halton_value.restype = C.c_float
            L.or_halton_value.argtypes = [C.c_uint64, C.c_uint32, vp]
            L.or_radical_inverse2.restype = C.c_float
            L.or_radical_inverse2.argtypes = [C.c_uint64]
            L.or_radical_inverse3.restype = C.c_float
            L.or_radical_inverse3.argtypes = [C.c_uint64]
            L.or_prime.restype = C.c_uint32
            L.or_prime.argtypes = [C.c_uint32]
            L.or_prime_sum.restype = C.c_uint32
            L.or_prime_sum.argtypes = [C.c_uint32]
            L.or_mt_seed.argtypes = [vp, C.c_uint32]
            L.or_mt_next.restype = C.c_uint32
            L.or_mt_next.argtypes = [vp]
            L.or_enum_init.argtypes = [vp, C.c_int, C.c_int]
            L.or_enum_index.restype = C.c_uint64
            L.or_enum_index.argtypes = [vp, C.c_int, C.c_int, C.c_int]
            L.or_pcg4d.argtypes = [vp]
            L.or_bsdf_eval.restype = C.c_int
            L.or_bsdf_eval.argtypes = [C.POINTER(abi.MaterialDesc), V3, V3, C.c_float, C.c_float,
                                       C.c_int, vp, vp]
            L.or_bsdf_sample.restype = C.c_int
            L.or_bsdf_sample.argtypes = [C.POINTER(abi.MaterialDesc), V3, C.c_float, C.c_float, C.c_int,
                                         C.c_float, C.c_float, C.c_float, vp, vp, vp]
            L.or_hg_eval.restype = C.c_float
            L.or_hg_eval.argtypes = [C.c_float, V3, V3, vp]
            L.or_hg_sample.restype = C.c_int
            L.or_hg_sample.argtypes = [C.c_float, C.c_float, C.c_float, V3, vp, vp]
            L.or_srgb_to_linear.restype = C.c_float
            L.or_srgb_to_linear.argtypes = [C.c_float]
            L.or_texture_eval.restype = V3
            L.or_texture_eval.argtypes = [vp, C.c_uint32, C.c_float, C.c_float]
            L.or_texture_graph.restype = V3
            L.or_texture_graph.argtypes = [vp, C.c_uint32, TexCoord]
            L.or_probe_uv_differentials.argtypes = [vp, vp, vp, vp, vp, vp, C.c_uint32, C.c_uint32,
                                                    C.c_float, C.c_float, vp]
            L.or_glm_transform.argtypes = [vp, vp, vp, vp, vp, vp]
            L.or_first_hits.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, vp]
            L.or_primary_rays.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, C.c_uint32, vp]
            L.or_render_adjoint.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, C.c_uint32, C.c_uint32,
                                            vp, C.POINTER(AdjTargets), vp, vp]
            L.or_adjoint_prepare.restype = C.c_int
            L.or_adjoint_prepare.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.POINTER(AdjTargets)]
            L.or_adjoint_free.argtypes = [C.POINTER(AdjTargets)]
            L.or_lambert_only.restype = C.c_int
            L.or_lambert_only.argtypes = [C.POINTER(abi.MaterialDesc)]
            L.or_perlin.restype = C.c_float
            L.or_perlin.argtypes = [V3]
            L.or_accumulate.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_int]
            L.or_tonemap.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float]
            L.or_probe_ggx_d.restype = C.c_float
            L.or_probe_ggx_d.argtypes = [C.c_float] * 5
            L.or_probe_fresnel_dd.restype = C.c_float
            L.or_probe_fresnel_dd.argtypes = [C.c_float, C.c_float]
            L.or_probe_fresnel_dc.argtypes = [vp, vp, C.c_float, vp]
            L.or_probe_edge_tint.argtypes = [vp, vp, vp, vp]
            L.or_probe_light_clamp.argtypes = [C.POINTER(abi.RenderSettings), vp, C.c_int, vp]
            L.or_probe_free_flight.restype = C.c_int
            L.or_probe_free_flight.argtypes = [vp, vp, C.c_float, C.c_float, vp, vp]
            _lib = L
    return _lib


class AdjTargets(C.Structure):
    """or_adj_targets (oracle.h)."""
    _fields_ = [("alb_slot", C.c_void_p), ("rough_slot", C.c_void_p), ("n_alb", C.c_uint32),
                ("n_rough", C.c_uint32), ("tex", C.c_uint32), ("vol", C.c_uint32), ("off_rough", C.c_size_t),
                ("off_tex", C.c_size_t), ("off_vox", C.c_size_t), ("n_params", C.c_size_t)]


class V3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class TexCoord(C.Structure):
    _fields_ = [("s", C.c_float), ("t", C.c_float), ("p_world", V3), ("p_object", V3)]


def v3(a):
    return V3(float(a[0]), float(a[1]), float(a[2]))


class MT(C.Structure):
    _fields_ = [("s", C.c_uint32 * 624), ("i", C.c_int)]


class Enumerator(C.Structure):
    _fields_ = [("x_scale", C.c_float), ("y_scale", C.c_float), ("m_x", C.c_int), ("m_y", C.c_int),
                ("exp2", C.c_int), ("exp3", C.c_int), ("increment", C.c_int)]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Scene:
    """Oracle scene built from the same SceneArrays the product consumes."""

    def __init__(self, arrays):
        self.arrays = arrays
        self._desc = arrays.desc()
        self.h = lib().or_scene_create(C.byref(self._desc))

    def close(self):
        if self.h:
            lib().or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_closest(self, rays, brute=False):
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        fn = lib().or_brute_closest_batch if brute else lib().or_trace_closest_batch
        fn(self.h, _p(rays), len(rays), _p(hits))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays)
        occ = np.zeros(len(rays), np.uint32)
        lib().or_trace_any_batch(self.h, _p(rays), len(rays), _p(occ))
        return occ

    def render(self, settings, frame=1, rows=None, threads=1):
        """Noise image (H, W, 4) float32 of one frame; ``rows`` limits the rows
        rendered (others stay zero).  ``threads`` > 1 splits rows over Python
        threads (ctypes releases the GIL)."""
        h, w = settings.height, settings.width
        out = np.zeros((h, w, 4), np.float32)
        r0, r1 = (0, h) if rows is None else rows
        counts = np.zeros((max(1, threads), 8), np.uint64)

        def work(k, a, b):
            lib().or_render_rows(self.h, C.byref(settings), frame, a, b, _p(out), _p(counts[k]))

        if threads <= 1:
            work(0, r0, r1)
        else:
            bounds = np.linspace(r0, r1, threads + 1).astype(int)
            ts = [threading.Thread(target=work, args=(k, bounds[k], bounds[k + 1])) for k in range(threads)]
            for t in ts:
                t.start()
            for t in ts:
                t.join()
        self.last_counts = counts.sum(axis=0)
        return out

    def render_adjoint(self, settings, adj, frame=1, threads=1, albedo=None, roughness=(), texture=None,
                       volume=None):
        """(noise image, gradient) of the path-replay adjoint: adj is
        dLoss/dImage (H, W, 3).  Targets: albedo, a list of Lambert-only
        materials (default: every Lambert-only material); roughness, a list of
        metal-only materials; texture, an image texture used as the diffuse
        albedo of Lambert materials; volume, a volume with a gray grid
        density.  The gradient is a dict of float64 arrays: "albedo"
        (len(albedo), 3), "roughness" (len(roughness),), "texture" (H, W, C),
        "density" (nz, ny, nx).  Raises ValueError naming a target the adjoint
        cannot differentiate."""
        h, w = settings.height, settings.width
        adj = np.ascontiguousarray(adj, np.float32).reshape(h, w, 3)
        out = np.zeros((h, w, 4), np.float32)
        if albedo is None:
            albedo = [i for i, m in enumerate(self.arrays.materials) if lambert_only(m)]
        alb = np.ascontiguousarray(albedo, np.uint32)
        rough = np.ascontiguousarray(roughness, np.uint32)
        INV = 0xFFFFFFFF
        tg = AdjTargets()
        bad = lib().or_adjoint_prepare(self.h, _p(alb) if len(alb) else None, len(alb),
                                       _p(rough) if len(rough) else None, len(rough),
                                       INV if texture is None else texture, INV if volume is None else volume,
                                       C.byref(tg))
        try:
            if bad:
                names = (["albedo material %d" % m for m in alb] + ["roughness material %d" % m for m in rough]
                         + ["texture %s" % texture, "volume %s" % volume])
                raise ValueError("the adjoint cannot differentiate " + names[bad - 1])
            grads = np.zeros((max(1, threads), max(tg.n_params, 1)), np.float64)

            def work(k, a, b):
                lib().or_render_adjoint(self.h, C.byref(settings), frame, a, b, _p(adj), C.byref(tg),
                                        _p(grads[k]), _p(out))

            bounds = np.linspace(0, h, max(1, threads) + 1).astype(int)
            ts = [threading.Thread(target=work, args=(k, bounds[k], bounds[k + 1]))
                  for k in range(max(1, threads))]
            for t in ts:
                t.start()
            for t in ts:
                t.join()
            g = grads.sum(axis=0)
            res = {"albedo": g[:tg.off_rough].reshape(-1, 3), "roughness": g[tg.off_rough:tg.off_tex]}
            if texture is not None:
                td = self.arrays.textures[texture]["texels"]
                res["texture"] = g[tg.off_tex:tg.off_vox].reshape(td.shape[0], td.shape[1], -1)
            if volume is not None:
                gd = self.arrays.grids[self.arrays.volumes[volume]["grid"]]
                res["density"] = g[tg.off_vox:tg.n_params].reshape(gd.shape)
            return out, res
        finally:
            lib().or_adjoint_free(C.byref(tg))

    def primary_rays(self, settings, frame=1, sample=0):
        """Camera rays (H*W records of abi.RAY_DTYPE) of one sample per pixel."""
        rays = np.zeros(settings.height * settings.width, abi.RAY_DTYPE)
        lib().or_primary_rays(self.h, C.byref(settings), frame, sample, _p(rays))
        return rays

    def first_hits(self, settings, frame=1):
        """(H, W) int32 instance ids of sample 0's camera rays, -1 on a miss."""
        ids = np.zeros((settings.height, settings.width), np.uint32)
        lib().or_first_hits(self.h, C.byref(settings), frame, _p(ids))
        return ids.view(np.int32)

    def texture_eval(self, tex, u, v):
        r = lib().or_texture_eval(self.h, tex, u, v)
        return np.array([r.x, r.y, r.z], np.float32)

    def uv_differentials(self, hp, sn, ng, dpdu, dpdv, width, height, sx=0.25, sy=0.25):
        """(du/dx, dv/dx, du/dy, dv/dy) of the scene camera at hit point hp."""
        a = [np.ascontiguousarray(x, np.float32) for x in (hp, sn, ng, dpdu, dpdv)]
        out = np.zeros(4, np.float32)
        lib().or_probe_uv_differentials(self.h, *[_p(x) for x in a], width, height, sx, sy, _p(out))
        return out

    def texture_graph(self, tex, coords):
        """Texture graph rooted at ``tex`` at each row of coords (n, 8):
        (s, t, p_world xyz, p_object xyz); returns (n, 3) float32."""
        coords = np.asarray(coords, np.float32).reshape(-1, 8)
        out = np.zeros((len(coords), 3), np.float32)
        for i, c in enumerate(coords):
            tc = TexCoord(float(c[0]), float(c[1]), v3(c[2:5]), v3(c[5:8]))
            r = lib().or_texture_graph(self.h, tex, tc)
            out[i] = (r.x, r.y, r.z)
        return out


def halton_table(seed):
    n = lib().or_halton_table_size()
    t = np.zeros(n, np.uint32)
    lib().or_halton_table(seed, _p(t))
    return t


def halton_value(index, dim, table):
    return lib().or_halton_value(index, dim, _p(table))


def enum_index(width, height, px, py, sample):
    e = Enumerator()
    lib().or_enum_init(C.byref(e), width, height)
    return lib().or_enum_index(C.byref(e), px, py, sample), e


def mt19937(seed, n):
    m = MT()
    lib().or_mt_seed(C.byref(m), seed)
    return np.array([lib().or_mt_next(C.byref(m)) for _ in range(n)], np.uint64)


def lambert_only(material):
    return bool(lib().or_lambert_only(C.byref(material)))


def perlin(p):
    """Perlin noise of each point of p (n, 3)."""
    p = np.asarray(p, np.float32).reshape(-1, 3)
    return np.array([lib().or_perlin(v3(x)) for x in p], np.float32)


def pcg4d(v):
    a = np.ascontiguousarray(v, np.uint32).copy()
    lib().or_pcg4d(_p(a))
    return a


def bsdf_eval(mat, wo, wi, ray_ior=1.0, ex_ior=1.0, back=0):
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().or_bsdf_eval(C.byref(mat), v3(wo), v3(wi), ray_ior, ex_ior, back, _p(f), _p(pdf))
    return f, float(pdf[0])


def bsdf_sample(mat, wo, u, ray_ior=1.0, ex_ior=1.0, back=0):
    w = np.zeros(3, np.float32)
    wi = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    ok = lib().or_bsdf_sample(C.byref(mat), v3(wo), ray_ior, ex_ior, back, float(u[0]), float(u[1]),
                              float(u[2]), _p(w), _p(wi), _p(pdf))
    return bool(ok), w, wi, float(pdf[0])


def hg_eval(g, wo, wi):
    pdf = np.zeros(1, np.float32)
    return lib().or_hg_eval(g, v3(wo), v3(wi), _p(pdf))


def hg_sample(g, u0, u1, wo):
    wi = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    lib().or_hg_sample(g, u0, u1, v3(wo), _p(wi), _p(pdf))
    return wi, float(pdf[0])


def srgb_to_linear(s):
    return lib().or_srgb_to_linear(s)


def accumulate(noise, accum, frame):
    h, w, _ = noise.shape
    lib().or_accumulate(_p(noise), _p(accum), w, h, frame)


def tonemap(img, gamma, exposure):
    h, w, _ = img.shape
    lib().or_tonemap(_p(img), w, h, gamma, exposure)


def ggx_d(h, ax, ay):
    return lib().or_probe_ggx_d(float(h[0]), float(h[1]), float(h[2]), ax, ay)


def fresnel_dielectric(eta, cos_i):
    return lib().or_probe_fresnel_dd(eta, cos_i)


def fresnel_conductor(eta, k, cos_i):
    e = np.ascontiguousarray(eta, np.float32)
    kk = np.ascontiguousarray(k, np.float32)
    out = np.zeros(3, np.float32)
    lib().or_probe_fresnel_dc(_p(e), _p(kk), cos_i, _p(out))
    return out


def edge_tint(reflectivity, tint):
    r = np.ascontiguousarray(reflectivity, np.float32)
    t = np.ascontiguousarray(tint, np.float32)
    n = np.zeros(3, np.float32)
    k = np.zeros(3, np.float32)
    lib().or_probe_edge_tint(_p(r), _p(t), _p(n), _p(k))
    return n, k


def free_flight(sigma_t, albedo, tmax, u):
    """SampleTrHomogeneous for one medium with the sampler value u:
    (scattered, distance, weight rgb)."""
    s = np.ascontiguousarray(sigma_t, np.float32)
    a = np.ascontiguousarray(albedo, np.float32)
    dist = np.zeros(1, np.float32)
    w = np.zeros(3, np.float32)
    r = lib().or_probe_free_flight(_p(s), _p(a), tmax, u, _p(dist), _p(w))
    return bool(r), float(dist[0]), w


def light_clamp(settings, c, depth):
    """indirect_light_clamp (src/Shading.cu:61-81) of one rgb contribution."""
    cc = np.ascontiguousarray(c, np.float32)
    out = np.zeros(3, np.float32)
    lib().or_probe_light_clamp(C.byref(settings), _p(cc), depth, _p(out))
    return out
