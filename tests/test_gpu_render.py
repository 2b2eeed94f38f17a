"""Path-traced frames on the GPU against the CPU oracle, through the C ABI.

The oracle restates the reference's path loop per path; the HIP path runs it
as a wavefront.  Both use the same Halton table, scene floats and operation
order, so paths follow the same vertices; their floats differ in the last bits
where GPU and glibc transcendentals (sin, cos, log, exp) round differently,
and a few paths diverge outright.  Bars:
  * >= 95% of pixels within 1e-4 relative, relative RMSE <= 1% (Cornell, 8 spp);
  * media (fogged textured interior, homogeneous and voxel-grid boxes):
    relative RMSE <= 1-2%, with a looser per-pixel bar since every segment
    in a medium runs expf/logf;
  * accumulate and tonemap equal to the oracle's;
  * two GPU frames bit-identical (no float atomics anywhere).
"""
import numpy as np
import pytest
import torch

import oracle
import ymrt_amd
from ymrt_amd import abi, scenes

pytestmark = pytest.mark.gpu


def render_both(gpu_ctx, sa, rs, frame=1):
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    img, stats = sc.render_frame(rs, frame)
    g = img.cpu().numpy()
    o = oracle.Scene(sa).render(rs, frame, threads=8)
    return g, o, stats


def close_frac(a, b):
    d = np.abs(a[..., :3] - b[..., :3])
    return float(np.all(d <= 1e-4 * np.abs(b[..., :3]) + 1e-6, axis=-1).mean())


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a[..., :3] - b[..., :3]) ** 2)) / max(np.mean(np.abs(b[..., :3])), 1e-12))


def test_cornell_halton_parity(gpu_ctx):
    sa = scenes.cornell_box(64, 64)
    rs = abi.render_settings(64, 64, 8, 8, sampler=abi.SAMPLER_HALTON)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    assert np.allclose(g[..., 3], 1.0)
    same = close_frac(g, o)
    err = rel_rmse(g, o)
    print("close pixels %.4f  rel rmse %.5f  stats ext %d shadow %d" %
          (same, err, stats.extension_rays, stats.shadow_rays))
    assert same >= 0.95
    assert err <= 0.01


def test_cornell_pcg_parity(gpu_ctx):
    sa = scenes.cornell_box(48, 32)
    rs = abi.render_settings(48, 32, 4, 6, sampler=abi.SAMPLER_PCG)
    g, o, _ = render_both(gpu_ctx, sa, rs, frame=3)
    assert close_frac(g, o) >= 0.90
    assert rel_rmse(g, o) <= 0.02


def test_deterministic_frames(gpu_ctx):
    sa = scenes.cornell_box(64, 64)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    rs = abi.render_settings(64, 64, 4, 8)
    a = sc.render_frame(rs, 2)[0].cpu().numpy()
    b = sc.render_frame(rs, 2)[0].cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # a small path capacity splits the frame into many chunks: same image
    rs2 = abi.render_settings(64, 64, 4, 8, path_capacity=4096)
    c = sc.render_frame(rs2, 2)[0].cpu().numpy()
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_accumulate_tonemap(gpu_ctx):
    rng = np.random.RandomState(0)
    h, w = 17, 23
    noise = [rng.rand(h, w, 4).astype(np.float32) * 3 for _ in range(3)]
    acc_o = np.zeros((h, w, 4), np.float32)
    acc_g = torch.zeros((h, w, 4), dtype=torch.float32, device="cuda")
    for f, n in enumerate(noise, start=1):
        oracle.accumulate(n, acc_o, f)
        ymrt_amd.accumulate(torch.from_numpy(n).cuda(), acc_g, f)
    assert np.array_equal(acc_g.cpu().numpy(), acc_o)
    tm_o = acc_o.copy()
    oracle.tonemap(tm_o, 2.2, 1.5)
    ymrt_amd.tonemap(acc_g, 2.2, 1.5)
    assert np.allclose(acc_g.cpu().numpy(), tm_o, rtol=2e-6, atol=1e-7)


def _media_report(name, g, o, stats):
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("%s: close pixels %.4f  rel rmse %.5f  ext %d shadow %d transmittance %d" %
          (name, same, err, stats.extension_rays, stats.shadow_rays, stats.transmittance_rays))
    return same, err


def test_interior_fog_parity(gpu_ctx):
    """Textured interior inside homogeneous fog (camera in the medium)."""
    sa = scenes.interior(64, 40, tex_size=64, instances=8)
    rs = abi.render_settings(64, 40, 4, 8, volume=1)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = _media_report("interior fog", g, o, stats)
    assert stats.transmittance_rays > 0
    # every fog segment runs expf/logf, whose last bits differ between the GPU
    # and glibc, so fewer pixels agree to 1e-4 while the image error stays low
    assert same >= 0.70
    assert err <= 0.01


@pytest.mark.parametrize("grid", [False, True], ids=["homogeneous", "grid"])
def test_cornell_medium_parity(gpu_ctx, grid):
    """A medium box seen from outside: boundary pass-through, free flight
    (homogeneous or voxel-grid delta tracking), transmittance of shadow rays."""
    sa = scenes.cornell_media(48, 48, grid=grid)
    rs = abi.render_settings(48, 48, 4, 6, volume=1)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = _media_report("cornell medium grid=%d" % grid, g, o, stats)
    assert same >= 0.50
    assert err <= 0.02


@pytest.mark.parametrize("kw", [dict(analytic_sphere=True), dict(sphere_light=True),
                                dict(analytic_sphere=True, sphere_light=True)],
                         ids=["glass-sphere", "sphere-light", "both"])
def test_cornell_sphere_parity(gpu_ctx, kw):
    """Analytic spheres: a glass sphere (shading data, nested dielectric) and a
    sphere light (cone sampling, EvalLi)."""
    sa = scenes.cornell_box(48, 48, **kw)
    rs = abi.render_settings(48, 48, 8, 6)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("cornell %s: close pixels %.4f  rel rmse %.5f" % (kw, same, err))
    assert same >= 0.90
    assert err <= 0.01
