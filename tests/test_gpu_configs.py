"""The configured workloads and the closed-form scenes on the GPU, through the
C ABI.

* The Cornell box as the CPU workload defines it (256 x 256, 16 spp,
  depth 8, Halton, permutation seed 0) against the oracle.
* The full interior workload scene (40 instances, 260 k triangles, four
  2048^2 textures, homogeneous fog around the camera) at 1920 x 1080 and one
  sample per pixel against the oracle: the scene of the bench line, every
  texture at full size.
* Russian roulette, the indirect clamp and the environment background, on
  together, against the oracle.
* The scenes whose image mean is a closed form or an FP64 quadrature
  (tests/test_oracle_pins.py): the GPU image against that value directly.

Per-path agreement with the oracle is measured as the share of pixels within
1e-4 relative.  Both sides evaluate transcendentals of float arguments in
double precision and round once, so they agree bit for bit but for the
indirect clamp's powers (float powf on each side).
"""
import numpy as np
import pytest
import torch

import oracle
import ymrt_amd
from ymrt_amd import abi, scenes

from test_oracle_pins import closed_form_cases

pytestmark = pytest.mark.gpu


def close_frac(a, b):
    d = np.abs(a[..., :3] - b[..., :3])
    return float(np.all(d <= 1e-4 * np.abs(b[..., :3]) + 1e-6, axis=-1).mean())


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a[..., :3] - b[..., :3]) ** 2)) / max(np.mean(np.abs(b[..., :3])), 1e-12))


def block_rel_rmse(a, b, k=16):
    """Relative RMSE of k x k block means (paths that diverged between the two
    renders average out with the rest of their block)."""
    h, w = (a.shape[0] // k) * k, (a.shape[1] // k) * k
    ab = a[:h, :w, :3].reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))
    bb = b[:h, :w, :3].reshape(h // k, k, w // k, k, 3).mean(axis=(1, 3))
    return float(np.sqrt(np.mean((ab - bb) ** 2)) / max(np.mean(np.abs(bb)), 1e-12))


def test_cornell_cpu_workload_as_configured(gpu_ctx):
    sa = scenes.cornell_box(256, 256, halton_seed=0)
    rs = abi.render_settings(256, 256, 16, 8, sampler=abi.SAMPLER_HALTON)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    g = sc.render_frame(rs, 1)[0].cpu().numpy()
    sc.close()
    o = oracle.Scene(sa).render(rs, 1, threads=16)
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("cornell 256x256x16 depth 8: close pixels %.5f  rel rmse %.2e" % (same, err))
    assert np.isfinite(g).all()
    assert same >= 0.999
    assert err <= 1e-4


def test_interior_full_scene_1080p(gpu_ctx):
    sa = scenes.interior()          # 1920 x 1080, 40 instances, 2048^2 textures, fog
    assert sa.triangle_count > 250000
    rs = abi.render_settings(1920, 1080, 1, 8, volume=1)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    img, stats = sc.render_frame(rs, 1)
    g = img.cpu().numpy()
    sc.close()
    o = oracle.Scene(sa).render(rs, 1, threads=16)
    same, err, berr = close_frac(g, o), rel_rmse(g, o), block_rel_rmse(g, o)
    mean_diff = abs(float(g[..., :3].mean()) - float(o[..., :3].mean())) / float(o[..., :3].mean())
    print("interior 1920x1080x1 fog: close pixels %.5f  rel rmse %.3e  16x16-block rel rmse %.3e  "
          "mean diff %.2e  ext %d shadow %d transmittance %d" %
          (same, err, berr, mean_diff, stats.extension_rays, stats.shadow_rays, stats.transmittance_rays))
    assert np.isfinite(g).all()
    assert stats.transmittance_rays > 0
    assert same >= 0.999
    assert berr <= 1e-4
    assert mean_diff <= 1e-5


def test_rr_clamp_env_parity(gpu_ctx):
    """Cornell box (open towards the camera, so rays escape to the
    environment) with Russian roulette past depth 3, the indirect clamp at
    limit 0.5 and the environment background on."""
    sa = scenes.cornell_box(64, 64)
    rs = abi.render_settings(64, 64, 16, 3, sampler=abi.SAMPLER_HALTON, env_light=1, rr=1)
    rs.enable_indirect_clamp = 1
    rs.indirect_limit = 0.5
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    g = sc.render_frame(rs, 1)[0].cpu().numpy()
    sc.close()
    o = oracle.Scene(sa).render(rs, 1, threads=16)
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("cornell rr + clamp + env: close pixels %.5f  rel rmse %.2e" % (same, err))
    assert np.isfinite(g).all()
    # the clamp's powers are the GPU's float powf, the oracle's glibc powf
    assert same >= 0.999
    assert err <= 0.002


@pytest.mark.parametrize("name", sorted(closed_form_cases()))
def test_closed_form_scene_on_gpu(gpu_ctx, name):
    sa, rs, frames, expect, rtol = closed_form_cases()[name]()
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    means = [sc.render_frame(rs, f)[0].cpu().numpy()[..., :3].astype(np.float64).mean(axis=(0, 1)) for f in frames]
    sc.close()
    got = np.mean(means, axis=0)
    print("%s: gpu %s expect %s" % (name, got, expect))
    assert np.allclose(got, expect, rtol=rtol), (got, expect)


@pytest.mark.parametrize("sigma", [2.0, 6.0])
def test_texture_density_medium_parity(gpu_ctx, sigma):
    """A medium whose density is an fBm texture graph (YMRT's
    density_texture_idx): delta tracking under the texture's upper bound for
    free flight and the shadow-ray transmittance, against the oracle."""
    sa = scenes.cornell_media(48, 48, density_texture=True, sigma=sigma)
    rs = abi.render_settings(48, 48, 4, 6, volume=1)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    g, stats = sc.render_frame(rs, 1)
    g = g.cpu().numpy()
    sc.close()
    o = oracle.Scene(sa).render(rs, 1, threads=16)
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("texture-density medium sigma %.1f: close pixels %.5f  rel rmse %.2e  transmittance %d" %
          (sigma, same, err, stats.transmittance_rays))
    assert np.isfinite(g).all()
    assert stats.transmittance_rays > 0
    assert same >= 0.999
    assert err <= 1e-4


def test_batch_workload_subset_equals_single_frames(gpu_ctx):
    """Twelve scenes of the multi-scene workload (triangle counts and the
    mix of resolutions and sample counts as configured, sizes divided by 8
    and 4) in shared launches: every image equals ymrt_render_frame's bit
    for bit, also when the round holds only five chunks."""
    jobs = []
    for i, (tris, res, spp, depth) in enumerate(scenes.batch_workload(12, seed=11)):
        r = res // 8
        sc = ymrt_amd.Scene(gpu_ctx, scenes.batch_scene(min(tris, 60000), r, r, i))
        jobs.append((sc, abi.render_settings(r, r, max(1, spp // 4), depth), 1 + i % 3))
    out = gpu_ctx.render_batch(jobs)
    out5 = gpu_ctx.render_batch(jobs, max_slots=5)
    for (sc, rs, frame), (img, st), (img5, _) in zip(jobs, out, out5):
        ref, ref_st = sc.render_frame(rs, frame)
        assert torch.equal(img, ref)
        assert torch.equal(img5, ref)
        assert st.extension_rays == ref_st.extension_rays and st.shadow_rays == ref_st.shadow_rays
    for sc, _, _ in jobs:
        sc.close()
