"""ymrt_amd -- MI355X (gfx950) wavefront path tracer behind YMRT's scene API.

Python binding over the C ABI of ``lib/libymrt_hip.so`` (declared in
include/ymrt.h).  Device buffers are torch tensors on ``cuda`` (plumbing only:
every kernel is the library's own HIP code).  There is no CPU fallback: when
the library is missing, ``lib()`` raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libymrt_hip.so")
CSRC = os.path.join(_HERE, "csrc")

# Every symbol include/ymrt.h declares.
EXPORTS = [
    "ymrt_last_error", "ymrt_abi_version", "ymrt_context_create", "ymrt_context_destroy",
    "ymrt_scene_create", "ymrt_scene_destroy", "ymrt_scene_get_info", "ymrt_trace_closest",
    "ymrt_trace_any", "ymrt_trace_closest_counted", "ymrt_render_frame", "ymrt_accumulate",
    "ymrt_tonemap", "ymrt_halton_values", "ymrt_halton_enumerate", "ymrt_texture_eval",
    "ymrt_ymtmp_load", "ymrt_ymtmp_parse", "ymrt_ymtmp_scene", "ymrt_ymtmp_warning_count",
    "ymrt_ymtmp_warning", "ymrt_ymtmp_find", "ymrt_ymtmp_transform", "ymrt_ymtmp_free",
    "ymrt_render_frame_aov", "ymrt_save_png", "ymrt_render_adjoint", "ymrt_scene_update_materials",
    "ymrt_render_batch", "ymrt_autotune", "ymrt_scene_release_buffers", "ymrt_scene_update_texture", "ymrt_render_frame_noalloc",
    "ymrt_scene_update_grid",
]

_lib = None


class YmrtError(RuntimeError):
    pass


def build(jobs=8):
    import subprocess
    subprocess.check_call(["make", "-s", "-j%d" % jobs, "-C", CSRC])


def lib():
    """The loaded C-ABI library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise YmrtError("libymrt_hip.so is not built (run __graft_entry__.build() or make -C %s)" % CSRC)
        # torch ships its own HIP runtime under the same soname: loaded first,
        # it is the one process-wide runtime the library binds to; loaded
        # after the library's, torch would find no GPU
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.ymrt_last_error.restype = C.c_char_p
        L.ymrt_abi_version.restype = C.c_int
        L.ymrt_context_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.ymrt_context_destroy.argtypes = [vp]
        L.ymrt_scene_create.argtypes = [vp, C.POINTER(abi.SceneDesc), C.POINTER(vp)]
        L.ymrt_scene_destroy.argtypes = [vp]
        L.ymrt_scene_get_info.argtypes = [vp, C.POINTER(abi.SceneInfo)]
        L.ymrt_scene_release_buffers.argtypes = [vp]
        L.ymrt_scene_update_texture.argtypes = [vp, C.c_uint32, vp, vp]
        L.ymrt_scene_update_grid.argtypes = [vp, C.c_uint32, vp, vp]
        L.ymrt_trace_closest.argtypes = [vp, vp, C.c_uint32, vp, vp]
        L.ymrt_trace_any.argtypes = [vp, vp, C.c_uint32, vp, vp]
        L.ymrt_trace_closest_counted.argtypes = [vp, vp, C.c_uint32, vp, vp, vp]
        L.ymrt_render_frame.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, vp, vp, C.c_int,
                                        C.POINTER(abi.FrameStats)]
        L.ymrt_render_frame_noalloc.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, vp, vp,
                                                C.POINTER(abi.FrameStats)]
        L.ymrt_accumulate.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_int, C.c_int, vp]
        L.ymrt_tonemap.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float, vp]
        L.ymrt_halton_values.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp]
        L.ymrt_halton_enumerate.argtypes = [C.c_uint32, C.c_uint32, vp, vp, vp, C.c_uint32, vp, vp]
        L.ymrt_texture_eval.argtypes = [vp, C.c_uint32, vp, C.c_uint32, vp, vp]
        L.ymrt_scene_update_materials.argtypes = [vp, C.c_uint32, C.c_uint32, C.POINTER(abi.MaterialDesc), vp]
        L.ymrt_render_batch.argtypes = [vp, C.POINTER(abi.BatchItem), C.c_uint32, C.c_uint32, vp]
        L.ymrt_autotune.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, vp, C.POINTER(abi.TuneResult)]
        L.ymrt_ymtmp_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.ymrt_ymtmp_parse.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.POINTER(vp)]
        L.ymrt_ymtmp_scene.argtypes = [vp]
        L.ymrt_ymtmp_warning_count.argtypes = [vp]
        L.ymrt_ymtmp_warning.argtypes = [vp, C.c_uint32]
        L.ymrt_ymtmp_find.argtypes = [vp, C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.ymrt_ymtmp_transform.argtypes = [vp, C.c_uint32, vp, vp]
        L.ymrt_ymtmp_free.argtypes = [vp]
        L.ymrt_render_frame_aov.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, vp, vp, vp, C.c_int,
                                            C.POINTER(abi.FrameStats)]
        L.ymrt_save_png.argtypes = [C.c_char_p, vp, C.c_uint32, C.c_uint32]
        L.ymrt_render_adjoint.argtypes = [vp, C.POINTER(abi.RenderSettings), C.c_int, vp, vp,
                                          C.POINTER(abi.AdjointTargets), vp, C.c_int, C.POINTER(abi.FrameStats)]
        for name in EXPORTS:
            if name not in ("ymrt_last_error", "ymrt_abi_version"):
                getattr(L, name).restype = C.c_int
        L.ymrt_ymtmp_scene.restype = C.POINTER(abi.SceneDesc)
        L.ymrt_ymtmp_warning_count.restype = C.c_uint32
        L.ymrt_ymtmp_warning.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise YmrtError("ymrt error %d: %s" % (rc, lib().ymrt_last_error().decode()))


def _dev_ptr(t):
    import torch
    assert t.is_cuda and t.is_contiguous(), "expected a contiguous cuda tensor"
    return C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().ymrt_context_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().ymrt_context_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


    def render_batch(self, jobs, max_slots=0):
        """ymrt_render_batch over jobs = [(scene, settings, frame), ...]:
        returns one (H, W, 4) cuda noise image and FrameStats per job;
        max_slots bounds the frame chunks per shared launch (0: 64)."""
        import torch
        n = len(jobs)
        items = (abi.BatchItem * max(1, n))()
        keep = []
        for i, (scene, settings, frame) in enumerate(jobs):
            img = torch.empty((settings.height, settings.width, 4), dtype=torch.float32, device="cuda")
            st = abi.FrameStats()
            keep.append((img, st, settings))
            items[i].scene = scene.h.value
            items[i].settings = C.pointer(settings)
            items[i].frame_index = frame
            items[i].timing = 0
            items[i].noise_rgba = img.data_ptr()
            items[i].stats = C.pointer(st)
        _check(lib().ymrt_render_batch(self.h, items, n, max_slots, _stream()))
        return [(img, st) for img, st, _ in keep]


class Scene:
    """Device-resident scene (ymrt_scene_create) built from abi.SceneArrays."""

    def __init__(self, ctx, arrays):
        self.ctx = ctx
        self.arrays = arrays
        self._desc = arrays.desc()
        self.h = C.c_void_p()
        _check(lib().ymrt_scene_create(ctx.h, C.byref(self._desc), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().ymrt_scene_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = abi.SceneInfo()
        _check(lib().ymrt_scene_get_info(self.h, C.byref(i)))
        return i

    def update_texture(self, texture, texels):
        """ymrt_scene_update_texture: new linear values of a float image texture."""
        t = np.ascontiguousarray(texels, np.float32)
        assert t.size == self.arrays.textures[texture]["texels"].size
        _check(lib().ymrt_scene_update_texture(self.h, texture, t.ctypes.data_as(C.c_void_p), _stream()))
        self.arrays.textures[texture]["texels"] = t.reshape(self.arrays.textures[texture]["texels"].shape)

    def release_buffers(self):
        """ymrt_scene_release_buffers: free the path state of this scene's frames."""
        _check(lib().ymrt_scene_release_buffers(self.h))

    def update_materials(self, first, materials):
        """ymrt_scene_update_materials: materials (abi.MaterialDesc records)
        replace [first, first + len(materials)); self.arrays follows."""
        n = len(materials)
        buf = (abi.MaterialDesc * max(1, n))(*materials)
        _check(lib().ymrt_scene_update_materials(self.h, first, n, buf, _stream()))
        for i, m in enumerate(materials):
            self.arrays.materials[first + i] = m

    # ----------------------------------------------------------- ray queries
    @staticmethod
    def rays_to_device(rays):
        import torch
        a = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
        return torch.from_numpy(a.copy()).cuda()

    def trace_closest(self, rays_dev, n=None, counted=False):
        """rays_dev: (n, 8) float32 cuda tensor; returns (n, 4) int32 cuda tensor
        of (t bits, triangle, instance, back) and, if counted, (n, 2) visits."""
        import torch
        n = rays_dev.shape[0] if n is None else n
        hits = torch.empty((n, 4), dtype=torch.int32, device=rays_dev.device)
        if counted:
            visits = torch.empty((n, 2), dtype=torch.int32, device=rays_dev.device)
            _check(lib().ymrt_trace_closest_counted(self.h, _dev_ptr(rays_dev), n, _dev_ptr(hits),
                                                    _dev_ptr(visits), _stream()))
            return hits, visits
        _check(lib().ymrt_trace_closest(self.h, _dev_ptr(rays_dev), n, _dev_ptr(hits), _stream()))
        return hits

    def trace_any(self, rays_dev, n=None):
        import torch
        n = rays_dev.shape[0] if n is None else n
        occ = torch.empty((n,), dtype=torch.int32, device=rays_dev.device)
        _check(lib().ymrt_trace_any(self.h, _dev_ptr(rays_dev), n, _dev_ptr(occ), _stream()))
        return occ

    @staticmethod
    def hits_to_numpy(hits):
        h = hits.cpu().numpy().astype(np.int32)
        out = np.zeros(len(h), abi.HIT_DTYPE)
        out["t"] = h[:, 0].view(np.float32)
        out["triangle"] = h[:, 1].view(np.uint32)
        out["instance"] = h[:, 2].view(np.uint32)
        out["back"] = h[:, 3].view(np.uint32)
        return out

    # -------------------------------------------------------------- sampler
    def halton_values(self, indices, dim0, dims):
        import torch
        idx = torch.from_numpy(np.ascontiguousarray(indices, np.int64)).cuda()
        out = torch.empty((len(indices), dims), dtype=torch.float32, device="cuda")
        _check(lib().ymrt_halton_values(self.h, _dev_ptr(idx), len(indices), dim0, dims, _dev_ptr(out), _stream()))
        return out.cpu().numpy()

    # ------------------------------------------------------------ textures
    def texture_eval(self, tex, coords):
        """coords: (n, 8) float32 (s, t, p_world xyz, p_object xyz); returns
        (n, 3) float32 rgb of the texture graph rooted at ``tex``."""
        import torch
        c = torch.from_numpy(np.ascontiguousarray(coords, np.float32).reshape(-1, 8).copy()).cuda()
        n = c.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device="cuda")
        _check(lib().ymrt_texture_eval(self.h, tex, _dev_ptr(c), n, _dev_ptr(out), _stream()))
        return out.cpu().numpy()

    # ------------------------------------------------------------ rendering
    def render_frame(self, settings, frame=1, noise=None, timing=False):
        import torch
        if noise is None:
            noise = torch.empty((settings.height, settings.width, 4), dtype=torch.float32, device="cuda")
        stats = abi.FrameStats()
        _check(lib().ymrt_render_frame(self.h, C.byref(settings), frame, _dev_ptr(noise), _stream(),
                                       1 if timing else 0, C.byref(stats)))
        return noise, stats

    def render_frame_noalloc(self, settings, frame=1, noise=None):
        """ymrt_render_frame_noalloc: the same image as render_frame, with the
        path state in the library's fixed device memory (no allocation)."""
        import torch
        if noise is None:
            noise = torch.empty((settings.height, settings.width, 4), dtype=torch.float32, device="cuda")
        stats = abi.FrameStats()
        _check(lib().ymrt_render_frame_noalloc(self.h, C.byref(settings), frame, _dev_ptr(noise), _stream(),
                                               C.byref(stats)))
        return noise, stats

    def render_adjoint(self, settings, adjoint, frame=1, albedo=None, roughness=(), texture=None, volume=None,
                       grads=None, timing=False):
        """(noise image, gradients, stats) of ymrt_render_adjoint; adjoint is
        dLoss/dImage, a (H, W, 3) cuda tensor.  Targets: albedo, a list of
        Lambert-only materials (default: every Lambert-only material of the
        scene); roughness, a list of metal-only materials; texture, an image
        texture used as the albedo of Lambert materials; volume, a volume
        with a gray sigma_t and a voxel-grid density.  The gradients are a
        dict of float32 cuda tensors: "albedo" (len(albedo), 3), "roughness"
        (len(roughness),), "texture" (H, W, C), "density" (nz, ny, nx);
        grads (the same dict) is added to instead of fresh zeros."""
        import torch
        h, w = settings.height, settings.width
        adjoint = adjoint.contiguous()
        assert adjoint.shape == (h, w, 3) and adjoint.dtype == torch.float32
        noise = torch.empty((h, w, 4), dtype=torch.float32, device="cuda")
        if albedo is None:
            albedo = [i for i, m in enumerate(self.arrays.materials) if abi.lambert_only(m)]
        albedo, roughness = [int(m) for m in albedo], [int(m) for m in roughness]
        if grads is None:
            z = lambda *shape: torch.zeros(shape, dtype=torch.float32, device="cuda")
            grads = {"albedo": z(len(albedo), 3), "roughness": z(len(roughness))}
            if texture is not None:
                grads["texture"] = z(*self.arrays.textures[texture]["texels"].shape)
            if volume is not None:
                g = self.arrays.volumes[volume]["grid"]
                grads["density"] = z(*self.arrays.grids[g].shape)
        alb = (C.c_uint32 * max(1, len(albedo)))(*albedo)
        rgh = (C.c_uint32 * max(1, len(roughness)))(*roughness)
        tg = abi.AdjointTargets()
        tg.albedo_materials, tg.albedo_count = alb, len(albedo)
        tg.roughness_materials, tg.roughness_count = rgh, len(roughness)
        tg.texture = abi.INVALID if texture is None else int(texture)
        tg.volume = abi.INVALID if volume is None else int(volume)
        tg.grad_albedo = _dev_ptr(grads["albedo"]) if albedo else None
        tg.grad_roughness = _dev_ptr(grads["roughness"]) if roughness else None
        tg.grad_texture = _dev_ptr(grads["texture"]) if texture is not None else None
        tg.grad_density = _dev_ptr(grads["density"]) if volume is not None else None
        stats = abi.FrameStats()
        _check(lib().ymrt_render_adjoint(self.h, C.byref(settings), frame, _dev_ptr(noise), _dev_ptr(adjoint),
                                         C.byref(tg), _stream(), 1 if timing else 0, C.byref(stats)))
        return noise, grads, stats

    def update_grid(self, grid, density):
        """ymrt_scene_update_grid: new densities (nz, ny, nx) of a voxel grid."""
        d = np.ascontiguousarray(density, np.float32)
        assert d.shape == self.arrays.grids[grid].shape
        _check(lib().ymrt_scene_update_grid(self.h, grid, d.ctypes.data_as(C.c_void_p), _stream()))
        self.arrays.grids[grid] = d

    def autotune(self, settings, frame=1):
        """abi.TuneResult of ymrt_autotune: the fastest path_capacity and
        persistent_waves for these settings (0: the library default)."""
        res = abi.TuneResult()
        _check(lib().ymrt_autotune(self.h, C.byref(settings), frame, _stream(), C.byref(res)))
        return res

    def render_frame_aov(self, settings, frame=1):
        """(noise image, int32 instance-id image, stats); ids are -1 on a miss."""
        import torch
        noise = torch.empty((settings.height, settings.width, 4), dtype=torch.float32, device="cuda")
        ids = torch.empty((settings.height, settings.width), dtype=torch.int32, device="cuda")
        stats = abi.FrameStats()
        _check(lib().ymrt_render_frame_aov(self.h, C.byref(settings), frame, _dev_ptr(noise), _dev_ptr(ids),
                                           _stream(), 0, C.byref(stats)))
        return noise, ids, stats


class YmtmpFile:
    """A .ymtmp scene file loaded by the library's parser (ymrt_ymtmp_load);
    needs no GPU.  ``arrays()`` copies the scene into an abi.SceneArrays."""

    KINDS = ("transform", "mesh", "sphere", "material", "texture", "instance", "distant_light", "volume")

    def __init__(self, path=None, text=None, scene_path=None):
        self.h = C.c_void_p()
        if text is not None:
            b = text.encode() if isinstance(text, str) else bytes(text)
            sp = scene_path.encode() if scene_path else None
            _check(lib().ymrt_ymtmp_parse(b, len(b), sp, C.byref(self.h)))
        else:
            _check(lib().ymrt_ymtmp_load(os.fspath(path).encode(), C.byref(self.h)))

    def close(self):
        if self.h:
            lib().ymrt_ymtmp_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def warnings(self):
        n = lib().ymrt_ymtmp_warning_count(self.h)
        return [lib().ymrt_ymtmp_warning(self.h, i).decode() for i in range(n)]

    def find(self, name):
        """(kind, index) of a named asset; kind is one of KINDS."""
        k, i = C.c_uint32(), C.c_uint32()
        _check(lib().ymrt_ymtmp_find(self.h, name.encode(), C.byref(k), C.byref(i)))
        return self.KINDS[k.value], i.value

    def transform(self, index):
        """(object_to_world, world_to_object) as row-major 4x4 float32."""
        a, b = np.zeros(16, np.float32), np.zeros(16, np.float32)
        _check(lib().ymrt_ymtmp_transform(self.h, index, a.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p)))
        return a.reshape(4, 4).T.copy(), b.reshape(4, 4).T.copy()

    def desc(self):
        return lib().ymrt_ymtmp_scene(self.h).contents

    def arrays(self):
        return abi.arrays_from_desc(self.desc())


def halton_enumerate(width, height, px, py, sample):
    import torch
    n = len(px)
    t = [torch.from_numpy(np.ascontiguousarray(a, np.int32)).cuda() for a in (px, py, sample)]
    out = torch.empty((n,), dtype=torch.int64, device="cuda")
    _check(lib().ymrt_halton_enumerate(width, height, _dev_ptr(t[0]), _dev_ptr(t[1]), _dev_ptr(t[2]), n,
                                       _dev_ptr(out), _stream()))
    return out.cpu().numpy().view(np.uint64)


def save_png(path, rgba):
    """SavePngImage of a host (H, W, 4) float32 image (row 0 at the bottom)."""
    a = np.ascontiguousarray(rgba, np.float32)
    h, w = a.shape[:2]
    _check(lib().ymrt_save_png(os.fspath(path).encode(), a.ctypes.data_as(C.c_void_p), w, h))


def accumulate(noise, accum, frame, max_frames=-1):
    h, w, _ = noise.shape
    _check(lib().ymrt_accumulate(_dev_ptr(noise), _dev_ptr(accum), w, h, frame, max_frames, _stream()))


def tonemap(img, gamma=2.2, exposure=1.0):
    h, w, _ = img.shape
    _check(lib().ymrt_tonemap(_dev_ptr(img), w, h, gamma, exposure, _stream()))
