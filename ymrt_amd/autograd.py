"""torch.autograd bridge for the path-replay adjoint (ymrt_render_adjoint).

``render_albedo(scene, settings, albedo, material_ids)`` renders one frame
with the rows of ``albedo`` written into the diffuse albedo of the listed
materials (ymrt_scene_update_materials) and returns the (H, W, 3) image as a
tensor that ``backward()`` differentiates with respect to ``albedo``: the
backward pass replays the same frame (same frame index, so the same samples)
with the incoming image gradient as the adjoint.  Materials that are not
Lambert-only DefaultMtl get a zero gradient (their albedo also shapes
sampling, which the adjoint does not follow).
"""
import ctypes as C

import torch

from . import abi


def _with_albedo(scene, material_ids, values):
    """Writes values[k] into the diffuse albedo of material material_ids[k]."""
    for k, m in enumerate(material_ids):
        rec = abi.MaterialDesc()
        C.memmove(C.byref(rec), C.byref(scene.arrays.materials[m]), C.sizeof(abi.MaterialDesc))
        rec.diffuse_albedo[:] = [float(x) for x in values[k]]
        scene.update_materials(m, [rec])


class AlbedoFrame(torch.autograd.Function):
    """image = render(albedo); d(loss)/d(albedo) by path replay."""

    @staticmethod
    def forward(ctx, albedo, scene, settings, frame, material_ids):
        vals = albedo.detach().to("cpu", torch.float32).numpy()
        _with_albedo(scene, material_ids, vals)
        img, _ = scene.render_frame(settings, frame)
        ctx.scene, ctx.settings, ctx.frame = scene, settings, frame
        ctx.material_ids = list(material_ids)
        ctx.vals = vals.copy()
        ctx.albedo_meta = (albedo.device, albedo.dtype)
        return img[..., :3].contiguous()

    @staticmethod
    def backward(ctx, grad_img):
        scene = ctx.scene
        # the scene may have been changed since the forward pass
        _with_albedo(scene, ctx.material_ids, ctx.vals)
        adj = grad_img.detach().to("cuda", torch.float32).contiguous()
        _, grad, _ = scene.render_adjoint(ctx.settings, adj, ctx.frame)
        ids = torch.tensor(ctx.material_ids, dtype=torch.long, device=grad.device)
        device, dtype = ctx.albedo_meta
        return grad.index_select(0, ids).to(device, dtype), None, None, None, None


def render_albedo(scene, settings, albedo, material_ids, frame=1):
    """(H, W, 3) float32 cuda image of one frame, differentiable in albedo
    (a (len(material_ids), 3) tensor)."""
    material_ids = list(material_ids)
    assert albedo.shape == (len(material_ids), 3)
    return AlbedoFrame.apply(albedo, scene, settings, frame, material_ids)
