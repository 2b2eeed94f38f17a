"""The path-replay adjoint on the GPU (ymrt_render_adjoint), through the C
ABI.

* Against the oracle's adjoint, which runs the same estimator per path and
  is itself pinned by finite differences and closed forms
  (tests/test_oracle_adjoint.py): the differentiable Cornell workload with
  every target at once (Lambert albedos, a float albedo texture, eight
  metal roughnesses, a voxel-grid density), the absorbing slab, shadow rays
  through a grid slab and a metal floor.  Both sides replay the same paths
  with the same floats; the GPU sums its events with float atomics in no
  fixed order, so gradients agree to about 1e-4 relative, not bit for bit.
* The differentiable workload as configured (512 x 512, 16 spp, depth 6,
  256^2 texture, 8 roughnesses, 64^3 grid) against central finite
  differences of the GPU forward with common random numbers (eps 1e-3,
  FP64 accumulation of the loss): texels exactly (an albedo moves no path),
  roughness and density within their Monte Carlo noise.
* Targets the adjoint cannot differentiate are rejected by name.
* torch.autograd through ymrt_amd.autograd.render.
"""
import copy

import numpy as np
import pytest
import torch

import oracle
import ymrt_amd
from ymrt_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _adj(h, w, seed):
    return np.random.default_rng(seed).uniform(0.5, 1.5, (h, w, 3)).astype(np.float32)


def _gpu_vs_oracle(gpu_ctx, sa, rs, adj, **targets):
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    img, g, stats = sc.render_adjoint(rs, torch.from_numpy(adj).cuda(), 1, timing=True, **targets)
    ref = sc.render_frame(rs, 1)[0]
    sc.close()
    oimg, og = oracle.Scene(sa).render_adjoint(rs, adj, 1, threads=16, **targets)
    assert torch.equal(img, ref)
    same = float(np.all(np.abs(img.cpu().numpy() - oimg) <= 1e-4 * np.abs(oimg) + 1e-6, axis=-1).mean())
    assert same >= 0.999
    out = {}
    for k, ov in og.items():
        if ov.size == 0:
            continue
        gv = g[k].cpu().numpy().astype(np.float64).reshape(ov.shape)
        scale = max(float(np.abs(ov).max()), 1e-12)
        err = float(np.linalg.norm(gv - ov) / max(np.linalg.norm(ov), 1e-12))
        worst = float(np.abs(gv - ov).max() / scale)
        print("%s: |g| %.4g  rel l2 err %.2e  max err / max|g| %.2e  replay %.2f ms" %
              (k, np.linalg.norm(ov), err, worst, stats.ms_adjoint))
        out[k] = (gv, ov, err, worst)
    return out


def test_adjoint_all_targets_match_oracle(gpu_ctx):
    params = scenes.inverse_params(32, 8, 8)
    sa, tg = scenes.cornell_inverse(48, 48, params)
    rs = abi.render_settings(48, 48, 8, 6, volume=1)
    res = _gpu_vs_oracle(gpu_ctx, sa, rs, _adj(48, 48, 7), albedo=tg["albedo"], roughness=tg["roughness"],
                         texture=tg["texture"], volume=tg["volume"])
    assert set(res) == {"albedo", "roughness", "texture", "density"}
    for k, (gv, ov, err, worst) in res.items():
        assert np.abs(ov).max() > 0, k
        assert err <= 1e-3, (k, err)
        assert worst <= 1e-3, (k, worst)


@pytest.mark.parametrize("case", ["slab", "shadow", "metal", "cornell"])
def test_adjoint_pinned_scenes_match_oracle(gpu_ctx, case):
    if case == "slab":
        sa, _ = scenes.absorbing_slab(48, 16)
        rs = abi.render_settings(48, 16, 64, 4, volume=1)
        kw = dict(albedo=[], volume=0)
    elif case == "shadow":
        slab = np.random.default_rng(4).uniform(0.2, 0.8, (1, 4, 4)).astype(np.float32)
        slab[0, 0, 0] = 1.0
        sa = scenes.lit_floor(24, 24, slab=slab)
        rs = abi.render_settings(24, 24, 64, 1, volume=1)
        kw = dict(albedo=[], volume=0)
    elif case == "metal":
        m = abi.default_material(diffuse_albedo=(0.9, 0.8, 0.6), specular_albedo=(1.0, 1.0, 1.0), roughness_x=0.3,
                                 roughness_y=0.3, metalness=1.0)
        sa = scenes.lit_floor(24, 24, material=m, light_half=2.5, fov=40.0)
        rs = abi.render_settings(24, 24, 256, 1)
        kw = dict(albedo=[], roughness=[0])
    else:
        sa = scenes.cornell_box(48, 48)
        rs = abi.render_settings(48, 48, 8, 6)
        kw = {}
    h, w = rs.height, rs.width
    res = _gpu_vs_oracle(gpu_ctx, sa, rs, _adj(h, w, 9), **kw)
    for k, (gv, ov, err, worst) in res.items():
        assert err <= 1e-3 and worst <= 1e-3, (case, k, err, worst)


def test_adjoint_rejects_targets_it_cannot_differentiate(gpu_ctx):
    sa, tg = scenes.cornell_inverse(16, 16, scenes.inverse_params(8, 4, 2))
    rs = abi.render_settings(16, 16, 1, 2, volume=1)
    adj = torch.ones((16, 16, 3), dtype=torch.float32, device="cuda")
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    with pytest.raises(ymrt_amd.YmrtError, match="albedo of material %d" % tg["roughness"][0]):
        sc.render_adjoint(rs, adj, albedo=[tg["roughness"][0]])
    with pytest.raises(ymrt_amd.YmrtError, match="roughness of material 0"):
        sc.render_adjoint(rs, adj, albedo=[], roughness=[0])
    with pytest.raises(ymrt_amd.YmrtError, match="Russian roulette"):
        bad = abi.render_settings(16, 16, 1, 2, volume=1, rr=1)
        sc.render_adjoint(bad, adj, albedo=[0])
    sc.close()
    sa.materials[tg["roughness"][0]].roughness_x_tex = tg["texture"]
    sa.materials[tg["roughness"][0]].roughness_y_tex = tg["texture"]
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    with pytest.raises(ymrt_amd.YmrtError, match="texture %d" % tg["texture"]):
        sc.render_adjoint(rs, adj, albedo=[], texture=tg["texture"])
    sc.close()
    sa2, tg2 = scenes.cornell_inverse(16, 16, scenes.inverse_params(8, 4, 2))
    sa2.volumes[tg2["volume"]]["sigma_t"] = (2.0, 1.0, 2.0)
    sc = ymrt_amd.Scene(gpu_ctx, sa2)
    with pytest.raises(ymrt_amd.YmrtError, match="not gray"):
        sc.render_adjoint(rs, adj, albedo=[], volume=tg2["volume"])
    sc.close()


def _c5_loss_fn(gpu_ctx, rs, adj_t):
    def loss(p):
        sa, _ = scenes.cornell_inverse(rs.width, rs.height, p)
        sc = ymrt_amd.Scene(gpu_ctx, sa)
        img = sc.render_frame(rs, 1)[0][..., :3].double()
        sc.close()
        return float((img * adj_t).sum())

    return loss


def test_adjoint_workload_against_finite_differences(gpu_ctx):
    """The differentiable workload as configured (512 x 512, depth 6, 256^2
    texture, 8 roughnesses, 64^3 grid) against central finite differences of
    the GPU forward with the same samples, the loss summed in FP64.

    Texels (16 spp, eps 1e-3): an albedo moves no path, so the difference
    quotient is the derivative of the same sample set; rel 1e-2 (measured
    2e-5).  Roughness and density move paths: a GGX-sampled direction turns
    and the delta tracker's accept decisions flip, so a difference quotient
    at eps 1e-3 is dominated by the few paths that jump.  They are compared
    at 256 spp with eps 3e-2 (relative to roughness 0.15-0.6 and to the
    densities): the sum of the 8 roughness gradients within 5 %, each of
    the four largest within 10 %, and the derivative along all densities
    (the majorant voxel held) within 8 % (measured 1.4 %, 1-2 % and 4 %)."""
    params = scenes.inverse_params()
    sa, tg = scenes.cornell_inverse(512, 512, params)
    adj = _adj(512, 512, 11)
    adj_t = torch.from_numpy(adj).cuda()
    targets = dict(albedo=tg["albedo"], roughness=tg["roughness"], texture=tg["texture"], volume=tg["volume"])

    def adjoint(spp):
        rs = abi.render_settings(512, 512, spp, 6, volume=1)
        sc = ymrt_amd.Scene(gpu_ctx, sa)
        _, g, _ = sc.render_adjoint(rs, adj_t, 1, **targets)
        sc.close()
        return rs, {k: v.cpu().numpy().astype(np.float64) for k, v in g.items()}

    rs, g = adjoint(16)
    loss = _c5_loss_fn(gpu_ctx, rs, adj_t.double())
    gt = g["texture"]
    for flat in np.argsort(-np.abs(gt).ravel())[:3]:
        ii = np.unravel_index(flat, gt.shape)
        up, dn = copy.deepcopy(params), copy.deepcopy(params)
        up["texture"][ii] += 1e-3
        dn["texture"][ii] -= 1e-3
        f = (loss(up) - loss(dn)) / 2e-3
        print("texel", ii, "adjoint %.6g fd %.6g" % (gt[ii], f))
        assert abs(f - gt[ii]) <= 1e-2 * abs(f)

    rs, g = adjoint(256)
    loss = _c5_loss_fn(gpu_ctx, rs, adj_t.double())
    eps = 3e-2
    fd = np.zeros(len(tg["roughness"]))
    for k in range(len(fd)):
        up, dn = copy.deepcopy(params), copy.deepcopy(params)
        up["roughness"][k] += eps
        dn["roughness"][k] -= eps
        fd[k] = (loss(up) - loss(dn)) / (2 * eps)
    gr = g["roughness"]
    print("roughness adjoint", np.round(gr, 1).tolist(), "fd", np.round(fd, 1).tolist())
    assert abs(gr.sum() - fd.sum()) <= 0.05 * abs(fd.sum())
    for k in np.argsort(-np.abs(fd))[:4]:
        assert abs(gr[k] - fd[k]) <= 0.1 * abs(fd[k]), (k, gr[k], fd[k])
    dens = params["density"]
    mask = np.ones_like(dens)
    mask[np.unravel_index(np.argmax(dens), dens.shape)] = 0.0
    up, dn = copy.deepcopy(params), copy.deepcopy(params)
    up["density"] = (dens * (1 + eps * mask)).astype(np.float32)
    dn["density"] = (dens * (1 - eps * mask)).astype(np.float32)
    fdd = (loss(up) - loss(dn)) / (2 * eps)
    gdd = float((g["density"] * dens * mask).sum())
    print("density along all voxels: adjoint %.2f fd %.2f" % (gdd, fdd))
    assert abs(gdd - fdd) <= 0.08 * abs(fdd)


def test_autograd_matches_adjoint_and_finite_differences(gpu_ctx):
    """loss.backward() through ymrt_amd.autograd.render gives the adjoint's
    gradients for every parameter tensor, and central finite differences of
    the loss in the Lambert albedos (no path moves)."""
    from ymrt_amd.autograd import render
    params = scenes.inverse_params(16, 8, 4)
    sa, tg = scenes.cornell_inverse(40, 40, params)
    rs = abi.render_settings(40, 40, 8, 6, volume=1)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    w = torch.from_numpy(_adj(40, 40, 3)).cuda()
    alb0 = torch.tensor([[0.7, 0.7, 0.7], [0.7, 0.1, 0.1], [0.1, 0.7, 0.1]], dtype=torch.float64)
    t = {"albedo": alb0.clone().requires_grad_(True),
         "roughness": torch.from_numpy(params["roughness"]).double().requires_grad_(True),
         "texture": torch.from_numpy(params["texture"]).double().requires_grad_(True),
         "density": torch.from_numpy(params["density"]).double().requires_grad_(True)}

    def loss_of(**kw):
        return (render(sc, rs, tg, frame=2, **kw).double() * w.double()).sum()

    loss = loss_of(**t)
    loss.backward()
    _, g, _ = sc.render_adjoint(rs, w, 2, albedo=tg["albedo"], roughness=tg["roughness"], texture=tg["texture"],
                                volume=tg["volume"])
    for k in t:
        a, b = t[k].grad.numpy(), g[k].cpu().numpy().astype(np.float64).reshape(t[k].shape)
        assert np.linalg.norm(a - b) <= 1e-3 * max(np.linalg.norm(b), 1e-9), k
    h = 1e-2
    with torch.no_grad():
        for i in range(3):
            for c in range(3):
                up, dn = alb0.clone(), alb0.clone()
                up[i, c] += h
                dn[i, c] -= h
                f = (loss_of(albedo=up).item() - loss_of(albedo=dn).item()) / (2 * h)
                assert abs(f - t["albedo"].grad[i, c].item()) <= 5e-3 * abs(f), (i, c, f)
    sc.close()
