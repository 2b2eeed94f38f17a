"""torch.autograd bridge for the path-replay adjoint (ymrt_render_adjoint).

``render(scene, settings, targets, albedo=..., roughness=..., texture=...,
density=...)`` writes the given parameters into the scene (diffuse albedos
and GGX roughnesses of the listed materials through
ymrt_scene_update_materials, the texels of a float image texture through
ymrt_scene_update_texture, the densities of a voxel grid through
ymrt_scene_update_grid), renders one frame and returns the (H, W, 3) image
as a tensor that ``backward()`` differentiates with respect to every
parameter tensor passed: the backward pass renders the same frame (same
frame index, so the same samples) with the incoming image gradient as the
adjoint and replays its paths.  ``targets`` names what the tensors drive:
dict(albedo=[Lambert-only materials], roughness=[metal-only materials],
texture=image texture index, volume=volume index, grid=its grid index), as
scenes.cornell_inverse returns it; a target the adjoint cannot differentiate
raises ymrt_amd.YmrtError from the backward pass.
"""
import ctypes as C

import torch

from . import abi


def _write_materials(scene, ids, albedo, roughness):
    """diffuse_albedo of ids[0] <- albedo[k]; roughness_x/y of ids[1] <- roughness[k]."""
    alb_ids, rough_ids = ids
    for k, m in enumerate(alb_ids if albedo is not None else ()):
        rec = abi.MaterialDesc()
        C.memmove(C.byref(rec), C.byref(scene.arrays.materials[m]), C.sizeof(abi.MaterialDesc))
        rec.diffuse_albedo[:] = [float(x) for x in albedo[k]]
        scene.update_materials(m, [rec])
    for k, m in enumerate(rough_ids if roughness is not None else ()):
        rec = abi.MaterialDesc()
        C.memmove(C.byref(rec), C.byref(scene.arrays.materials[m]), C.sizeof(abi.MaterialDesc))
        rec.roughness_x = rec.roughness_y = float(roughness[k])
        scene.update_materials(m, [rec])


def _write_params(scene, targets, vals):
    albedo, roughness, texture, density = vals
    _write_materials(scene, (targets.get("albedo", ()), targets.get("roughness", ())), albedo, roughness)
    if texture is not None:
        scene.update_texture(targets["texture"], texture)
    if density is not None:
        scene.update_grid(targets["grid"], density)


def _host(t):
    return None if t is None else t.detach().to("cpu", torch.float32).numpy().copy()


class SceneFrame(torch.autograd.Function):
    """image = render(albedo, roughness, texture, density); the gradients of
    a loss of the image by path replay."""

    @staticmethod
    def forward(ctx, albedo, roughness, texture, density, scene, settings, frame, targets):
        vals = tuple(_host(t) for t in (albedo, roughness, texture, density))
        _write_params(scene, targets, vals)
        img, _ = scene.render_frame(settings, frame)
        ctx.scene, ctx.settings, ctx.frame, ctx.targets, ctx.vals = scene, settings, frame, targets, vals
        ctx.meta = [None if t is None else (t.device, t.dtype, t.shape) for t in (albedo, roughness, texture, density)]
        return img[..., :3].contiguous()

    @staticmethod
    def backward(ctx, grad_img):
        scene, tg, vals = ctx.scene, ctx.targets, ctx.vals
        # the scene may have been changed since the forward pass
        _write_params(scene, tg, vals)
        adj = grad_img.detach().to("cuda", torch.float32).contiguous()
        _, g, _ = scene.render_adjoint(
            ctx.settings, adj, ctx.frame,
            albedo=tg.get("albedo", []) if vals[0] is not None else [],
            roughness=tg.get("roughness", []) if vals[1] is not None else [],
            texture=tg.get("texture") if vals[2] is not None else None,
            volume=tg.get("volume") if vals[3] is not None else None)
        out = []
        for key, meta in zip(("albedo", "roughness", "texture", "density"), ctx.meta):
            if meta is None:
                out.append(None)
            else:
                device, dtype, shape = meta
                out.append(g[key].reshape(shape).to(device, dtype))
        return (*out, None, None, None, None)


def render(scene, settings, targets, albedo=None, roughness=None, texture=None, density=None, frame=1):
    """(H, W, 3) float32 cuda image of one frame, differentiable in each
    parameter tensor given: albedo (len(targets["albedo"]), 3), roughness
    (len(targets["roughness"]),), texture (H, W, C) of targets["texture"],
    density (nz, ny, nx) of targets["grid"]."""
    return SceneFrame.apply(albedo, roughness, texture, density, scene, settings, frame, targets)


def render_albedo(scene, settings, albedo, material_ids, frame=1):
    """render() with only the diffuse albedos of material_ids as parameters."""
    material_ids = list(material_ids)
    assert albedo.shape == (len(material_ids), 3)
    return render(scene, settings, dict(albedo=material_ids), albedo=albedo, frame=frame)
