"""The oracle's path-replay adjoint (or_render_adjoint) against what fixes
each of its gradients:

* albedo and texels: central finite differences of the oracle's own renders
  with common random numbers.  A Lambert albedo (constant or a texel of its
  texture) changes no sampling decision, so the frame is a polynomial in it
  and the difference quotient with eps = 1e-3 matches the derivative to float
  noise (rel 2e-3).
* density, free flight: a purely absorbing slab seen against an emitter,
  where a ray's radiance is exp(-sigma d l), so each voxel's expected
  gradient is -sigma l_voxel L summed over the camera rays (their chords in
  each voxel computed from the oracle's own primary rays); delta tracking's
  null-collision events must reproduce it (rel 2 %, Monte Carlo).
* density, shadow rays: ratio tracking's transmittance is a smooth function
  of a voxel's density under common random numbers (for voxels below the
  majorant), so finite differences match per sample (rel 1e-3).
* GGX roughness: finite differences at 1024 spp under a light that fills the
  metal's reflection lobe (so a sampled direction's hit or miss of the light
  does not jump with the roughness); both estimators are unbiased and agree
  to their noise (rel 4 %).
* non-differentiable targets are rejected by name.
Known limit, shared by the GPU adjoint: the voxel that sets the majorant
never has a null collision, so its gradient lacks the null-collision term.
"""
import copy

import numpy as np
import pytest

import oracle
from ymrt_amd import abi, scenes


def _loss(sa, rs, adj):
    img = oracle.Scene(sa).render(rs, 1, threads=8)
    return float((img[..., :3].astype(np.float64) * adj).sum())


def test_adjoint_matches_finite_differences():
    rs = abi.render_settings(24, 24, 8, 5)
    rng = np.random.default_rng(0)
    adj = rng.uniform(0.5, 1.5, (24, 24, 3)).astype(np.float32)
    sa = scenes.cornell_box(24, 24)
    lam = [oracle.lambert_only(m) for m in sa.materials]
    assert lam == [True, True, True, False, False, False]
    img, gd = oracle.Scene(sa).render_adjoint(rs, adj, 1, threads=8)
    g = gd["albedo"]          # the Lambert-only materials 0-2, in order
    assert g.shape == (3, 3)
    # the adjoint's frame is the plain frame
    assert np.array_equal(img, oracle.Scene(sa).render(rs, 1, threads=8))
    eps = 1e-3
    for mi in range(3):
        for c in range(3):
            up, dn = scenes.cornell_box(24, 24), scenes.cornell_box(24, 24)
            up.materials[mi].diffuse_albedo[c] += eps
            dn.materials[mi].diffuse_albedo[c] -= eps
            fd = (_loss(up, rs, adj) - _loss(dn, rs, adj)) / (2 * eps)
            assert abs(fd - g[mi, c]) <= 1e-3 * abs(fd), (mi, c, fd, g[mi, c])


def test_adjoint_zero_albedo_channel():
    """A channel with albedo 0 gets gradient 0 (the estimator divides by the
    albedo); the other channels are unaffected."""
    sa = scenes.cornell_box(16, 16)
    sa.materials[1].diffuse_albedo[2] = 0.0
    rs = abi.render_settings(16, 16, 4, 4)
    adj = np.ones((16, 16, 3), np.float32)
    g = oracle.Scene(sa).render_adjoint(rs, adj, 1, threads=4)[1]["albedo"]
    assert g[1, 2] == 0.0 and g[1, 0] > 0.0


def _adj(h, w, seed):
    return np.random.default_rng(seed).uniform(0.5, 1.5, (h, w, 3)).astype(np.float32)


def test_texture_gradient_matches_finite_differences():
    params = scenes.inverse_params(16, 8, 8)
    rs = abi.render_settings(32, 32, 16, 6, volume=1)
    adj = _adj(32, 32, 2)
    sa, tg = scenes.cornell_inverse(32, 32, params)
    g = oracle.Scene(sa).render_adjoint(rs, adj, 1, threads=8, albedo=[], texture=tg["texture"])[1]["texture"]
    assert g.shape == params["texture"].shape

    def loss(p):
        img = oracle.Scene(scenes.cornell_inverse(32, 32, p)[0]).render(rs, 1, threads=8)
        return float((img[..., :3].astype(np.float64) * adj).sum())

    for flat in np.argsort(-np.abs(g).ravel())[:4]:
        ii = np.unravel_index(flat, g.shape)
        up, dn = copy.deepcopy(params), copy.deepcopy(params)
        up["texture"][ii] += 1e-3
        dn["texture"][ii] -= 1e-3
        fd = (loss(up) - loss(dn)) / 2e-3
        assert abs(fd - g[ii]) <= 2e-3 * abs(fd), (ii, fd, g[ii])


def test_density_gradient_absorbing_slab_closed_form():
    w, h, spp, sigma, d0 = 48, 16, 256, 2.0, 0.5
    sa, (x0, x1) = scenes.absorbing_slab(w, h, d0=d0, dmax=1.0, sigma=sigma)
    rs = abi.render_settings(w, h, spp, 4, volume=1)
    adj = _adj(h, w, 3)
    osc = oracle.Scene(sa)
    img, gd = osc.render_adjoint(rs, adj, 1, threads=8, albedo=[], volume=0)
    g = gd["density"].ravel()
    nx = g.size
    width = (x1 - x0) / nx
    expect = np.zeros(nx)
    mean_l = 0.0
    asum = adj.sum(-1).ravel().astype(np.float64)
    for s in range(spp):
        r = osc.primary_rays(rs, 1, s)
        o, d = r["o"].astype(np.float64), r["d"].astype(np.float64)
        ta, tb = (0.2 - o[:, 2]) / d[:, 2], (-0.2 - o[:, 2]) / d[:, 2]
        xa, xb = o[:, 0] + ta * d[:, 0], o[:, 0] + tb * d[:, 0]
        chord = np.abs(tb - ta) * np.linalg.norm(d, axis=1)
        L = np.exp(-sigma * d0 * chord)
        mean_l += L.mean() / spp
        lo, hi = np.minimum(xa, xb), np.maximum(xa, xb)
        for i in range(nx):
            v0, v1 = x0 + width * i, x0 + width * (i + 1)
            ov = np.clip(np.minimum(hi, v1) - np.maximum(lo, v0), 0.0, None)
            frac = np.where(hi > lo, ov / np.maximum(hi - lo, 1e-30), ((lo >= v0) & (lo < v1)).astype(float))
            expect[i] += np.sum(asum / spp * (-sigma * frac * chord * L))
    assert abs(float(img[..., 0].mean()) - mean_l) <= 0.01 * mean_l
    # the voxel holding the majorant is outside every ray
    assert g[nx - 1] == 0.0 and expect[nx - 1] == 0.0
    assert np.all(np.abs(g[:-1] - expect[:-1]) <= 0.02 * np.abs(expect[:-1])), (g, expect)


def test_density_gradient_shadow_rays_matches_finite_differences():
    rng = np.random.default_rng(4)
    slab = rng.uniform(0.2, 0.8, (1, 4, 4)).astype(np.float32)
    slab[0, 0, 0] = 1.0                      # the majorant, not perturbed
    w = h = 24
    adj = _adj(h, w, 5)
    rs = abi.render_settings(w, h, 64, 1, volume=1)

    def loss(sl):
        img = oracle.Scene(scenes.lit_floor(w, h, slab=sl)).render(rs, 1, threads=8)
        return float((img[..., :3].astype(np.float64) * adj).sum())

    g = oracle.Scene(scenes.lit_floor(w, h, slab=slab)).render_adjoint(rs, adj, 1, threads=8, albedo=[],
                                                                       volume=0)[1]["density"]
    checked = 0
    for ii in [(0, 1, 1), (0, 2, 2), (0, 1, 2), (0, 2, 1)]:
        up, dn = slab.copy(), slab.copy()
        up[ii] += 1e-3
        dn[ii] -= 1e-3
        fd = (loss(up) - loss(dn)) / 2e-3
        assert abs(fd - g[ii]) <= 1e-3 * abs(fd) + 1e-6, (ii, fd, g[ii])
        checked += abs(fd) > 1.0
    assert checked >= 3


@pytest.mark.parametrize("r0", [0.3, 0.6])
def test_roughness_gradient_matches_finite_differences(r0):
    w = h = 24
    adj = _adj(h, w, 1)
    rs = abi.render_settings(w, h, 1024, 1)

    def metal(r):
        return abi.default_material(diffuse_albedo=(0.9, 0.8, 0.6), specular_albedo=(1.0, 1.0, 1.0), roughness_x=r,
                                    roughness_y=r, metalness=1.0)

    def scene(r):
        return scenes.lit_floor(w, h, material=metal(r), light_half=2.5, fov=40.0)

    def loss(r):
        img = oracle.Scene(scene(r)).render(rs, 1, threads=8)
        return float((img[..., :3].astype(np.float64) * adj).sum())

    g = oracle.Scene(scene(r0)).render_adjoint(rs, adj, 1, threads=8, albedo=[], roughness=[0])[1]["roughness"]
    fd = (loss(r0 + 1e-2) - loss(r0 - 1e-2)) / 2e-2
    assert abs(fd - g[0]) <= 0.04 * abs(fd), (fd, g)


def test_adjoint_rejects_targets_it_cannot_differentiate():
    sa, tg = scenes.cornell_inverse(16, 16, scenes.inverse_params(8, 4, 2))
    rs = abi.render_settings(16, 16, 1, 2, volume=1)
    adj = np.ones((16, 16, 3), np.float32)
    osc = oracle.Scene(sa)
    with pytest.raises(ValueError, match="albedo material %d" % tg["roughness"][0]):
        osc.render_adjoint(rs, adj, albedo=[tg["roughness"][0]])
    with pytest.raises(ValueError, match="roughness material 0"):
        osc.render_adjoint(rs, adj, albedo=[], roughness=[0])
    # the texture also drives a metal's roughness
    sa.materials[tg["roughness"][0]].roughness_x_tex = tg["texture"]
    with pytest.raises(ValueError, match="texture"):
        oracle.Scene(sa).render_adjoint(rs, adj, albedo=[], texture=tg["texture"])
    # a coloured sigma_t
    sa2, tg2 = scenes.cornell_inverse(16, 16, scenes.inverse_params(8, 4, 2))
    sa2.volumes[tg2["volume"]]["sigma_t"] = (2.0, 1.0, 2.0)
    with pytest.raises(ValueError, match="volume"):
        oracle.Scene(sa2).render_adjoint(rs, adj, albedo=[], volume=tg2["volume"])
