"""Path-traced frames on the GPU against the CPU oracle, through the C ABI.

The oracle restates the reference's path loop per path; the HIP path runs it
as a wavefront.  Both use the same Halton table, scene floats, operation
order and transcendentals (float arguments evaluated in double precision and
rounded once), so paths follow the same vertices with the same floats.  Bars:
  * >= 99.9% of pixels within 1e-4 relative and relative RMSE <= 1e-4 for
    every scene (surfaces, media, spheres, distant lights, texture graphs);
    the slack covers a last-bit difference in a rare branch;
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
    assert same >= 0.999
    assert err <= 1e-4


def test_cornell_pcg_parity(gpu_ctx):
    sa = scenes.cornell_box(48, 32)
    rs = abi.render_settings(48, 32, 4, 6, sampler=abi.SAMPLER_PCG)
    g, o, _ = render_both(gpu_ctx, sa, rs, frame=3)
    print("pcg cornell: close pixels %.4f  rel rmse %.5f" % (close_frac(g, o), rel_rmse(g, o)))
    assert close_frac(g, o) >= 0.99
    assert rel_rmse(g, o) <= 0.01


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
    assert same >= 0.999
    assert err <= 1e-4


@pytest.mark.parametrize("grid", [False, True], ids=["homogeneous", "grid"])
def test_cornell_medium_parity(gpu_ctx, grid):
    """A medium box seen from outside: boundary pass-through, free flight
    (homogeneous or voxel-grid delta tracking), transmittance of shadow rays."""
    sa = scenes.cornell_media(48, 48, grid=grid)
    rs = abi.render_settings(48, 48, 4, 6, volume=1)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = _media_report("cornell medium grid=%d" % grid, g, o, stats)
    assert same >= 0.999
    assert err <= 1e-4


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
    assert same >= 0.999
    assert err <= 1e-4


def test_many_lights_parity(gpu_ctx):
    """More shape lights than the scene record's inline selection CDF holds
    (8): the pick takes the binary search over the CDF in HBM.  Nine ceiling
    tiles of their own intensities (18 triangles) besides the main light."""
    sa = scenes.cornell_box(48, 48)
    at = (0.05, 0.2, 0.75)  # 0.1-wide tiles clear of each other and of the main light
    for k in range(9):
        x0, z0 = at[k % 3], at[k // 3]
        tile = scenes.quad((x0, 0.998, z0), (x0 + 0.1, 0.998, z0), (x0 + 0.1, 0.998, z0 + 0.1), (x0, 0.998, z0 + 0.1))
        sa.add_instance(scenes._add(sa, tile), sa.add_material(abi.light_material((1.0, 0.9, 0.8), 4.0 + k)))
    rs = abi.render_settings(48, 48, 8, 6)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert ymrt_amd.Scene(gpu_ctx, sa).info().shape_light_count == 20
    assert np.isfinite(g).all()
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("cornell with 20 light triangles: close pixels %.4f  rel rmse %.5f" % (same, err))
    assert same >= 0.999
    assert err <= 1e-4


@pytest.mark.parametrize("fog", [False, True], ids=["clear", "fog"])
def test_outdoor_distant_light_parity(gpu_ctx, fog):
    """Distant light (EvalDistantLight / EvalVolumeDistantLight: cone sampling,
    unbounded shadow rays, transmittance to infinity through a medium box) with
    a quad light, boxes and analytic spheres on a ground plane."""
    sa = scenes.outdoor(64, 48, fog=fog)
    rs = abi.render_settings(64, 48, 8, 6, distant_light=1, volume=int(fog))
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("outdoor fog=%d: close pixels %.4f  rel rmse %.5f" % (fog, same, err))
    # the ground plane lies at y = 0, where the spawn offset is zero: shadow
    # rays that graze it hit or miss it on their last bits
    assert same >= 0.999
    assert err <= 1e-4


def test_texture_graph_parity(gpu_ctx):
    """Every texture node type on the GPU (ymrt_texture_eval: the post-order
    program on an LDS value stack) against the oracle's TextureEval (the
    reference's explicit stack of node copies) at 4096 random coordinates.
    Smooth nodes agree to 1e-4 everywhere; nodes with jumps (checkerboard
    sign, wood's fractional part, polka-dot edges) may flip at the few points
    where a contracted multiply-add rounds differently."""
    sa, nodes = scenes.texture_node_set()
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    osc = oracle.Scene(sa)
    rng = np.random.default_rng(11)
    c = np.zeros((4096, 8), np.float32)
    c[:, 0:2] = rng.uniform(-2, 3, (4096, 2))
    c[:, 2:8] = rng.uniform(-5, 5, (4096, 6))
    jumps = {"checker", "wood", "polka", "nested"}
    for name, tex in nodes.items():
        g = sc.texture_eval(tex, c)
        o = osc.texture_graph(tex, c)
        ok = np.all(np.abs(g - o) <= 1e-4 * np.maximum(1.0, np.abs(o)), axis=1)
        print("texture %-14s close %.5f  max err %.2e" % (name, ok.mean(), np.abs(g - o).max()))
        assert np.isfinite(g).all()
        assert ok.mean() >= (0.995 if name in jumps else 1.0), name
    osc.close()


def test_textured_box_parity(gpu_ctx):
    """Materials driven by texture graphs (every procedural node type), a
    normal-mapped box and a bump-mapped wall (height differences over the
    camera's texture-coordinate footprint)."""
    sa = scenes.textured_box(64, 64)
    rs = abi.render_settings(64, 64, 8, 6)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("textured box: close pixels %.4f  rel rmse %.5f" % (same, err))
    assert same >= 0.999
    assert err <= 1e-4


def test_ymtmp_cornell_fixture_parity(gpu_ctx):
    """YMRT's own test scene (tests/golden/CornellBox_Test.ymtmp) loaded by
    the library's parser and rendered on the GPU, against the oracle's
    rendering of the oracle parser's scene: checkerboard albedo, metalness
    and bump textures on the ground, an analytic sphere, a cube light and a
    distant light."""
    import os
    from oracle import ymtmp as oymtmp
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "CornellBox_Test.ymtmp")
    sa = ymrt_amd.YmtmpFile(path).arrays()
    osa = oymtmp.load(path).arrays
    rs = abi.render_settings(64, 64, 8, 6, distant_light=1)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    g = sc.render_frame(rs, 1)[0].cpu().numpy()
    o = oracle.Scene(osa).render(rs, 1, threads=8)
    assert np.isfinite(g).all()
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("ymtmp cornell: close pixels %.4f  rel rmse %.5f  mean %.4f" % (same, err, float(o[..., :3].mean())))
    assert float(o[..., :3].mean()) > 0.01
    assert same >= 0.999
    assert err <= 1e-4


def test_instance_id_aov(gpu_ctx):
    """ymrt_render_frame_aov: the id image is sample 0's camera-ray instance,
    as the oracle traces it, and the noise image equals ymrt_render_frame's."""
    sa = scenes.cornell_box(48, 40, analytic_sphere=True)
    rs = abi.render_settings(48, 40, 4, 4)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    img, ids, _ = sc.render_frame_aov(rs, 2)
    ref_img = sc.render_frame(rs, 2)[0]
    assert torch.equal(img, ref_img)
    o = oracle.Scene(sa).first_hits(rs, 2)
    g = ids.cpu().numpy()
    assert (g == o).mean() >= 0.999, (g != o).sum()
    assert set(np.unique(g).tolist()) <= set(range(len(sa.instances))) | {-1}
    assert len(np.unique(g)) >= 6


def test_update_materials(gpu_ctx):
    """ymrt_scene_update_materials: a scene whose wall albedo is updated
    renders the frame of a scene built with that albedo, bit for bit; a type
    change and an out-of-range update are rejected."""
    sa = scenes.cornell_box(32, 32)
    rs = abi.render_settings(32, 32, 4, 5)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    before = sc.render_frame(rs, 1)[0].clone()
    red = abi.default_material(diffuse_albedo=(0.3, 0.5, 0.2), specular_weight=0.0)
    sc.update_materials(1, [red])
    after = sc.render_frame(rs, 1)[0]
    sb = scenes.cornell_box(32, 32)
    sb.materials[1] = abi.default_material(diffuse_albedo=(0.3, 0.5, 0.2), specular_weight=0.0)
    built = ymrt_amd.Scene(gpu_ctx, sb).render_frame(rs, 1)[0]
    assert torch.equal(after, built)
    assert not torch.equal(after, before)
    with pytest.raises(ymrt_amd.YmrtError, match="type"):
        sc.update_materials(1, [abi.light_material((1.0, 1.0, 1.0), 2.0)])
    with pytest.raises(ymrt_amd.YmrtError, match="range"):
        sc.update_materials(len(sa.materials), [red])


@pytest.mark.parametrize("kw", [dict(), dict(sphere_light=True)], ids=["quad_light", "sphere_light"])
def test_update_light_material(gpu_ctx, kw):
    """A light material's colour updated in place: next-event estimation and
    emission on hit use the new radiance, so the frame equals that of a scene
    built with it.  The light-selection CDF keeps its build-time values, so
    the new colour keeps the light's power (sum of the channels times the
    intensity, 3.0 in both) and the two scenes' CDFs are the same floats."""
    sa = scenes.cornell_box(32, 32, **kw)
    rs = abi.render_settings(32, 32, 4, 5)
    used = {inst["material"] for inst in sa.instances}
    lights = [i for i, m in enumerate(sa.materials) if m.type == abi.MATERIAL_LIGHT and i in used]
    assert len(lights) == 1
    intensity = sa.materials[lights[0]].intensity
    assert list(sa.materials[lights[0]].light_color) == [1.0, 1.0, 1.0]
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    before = sc.render_frame(rs, 1)[0].clone()
    sc.update_materials(lights[0], [abi.light_material((1.5, 1.0, 0.5), intensity)])
    after = sc.render_frame(rs, 1)[0]
    sb = scenes.cornell_box(32, 32, **kw)
    sb.materials[lights[0]] = abi.light_material((1.5, 1.0, 0.5), intensity)
    built = ymrt_amd.Scene(gpu_ctx, sb).render_frame(rs, 1)[0]
    assert torch.equal(after, built)
    assert not torch.equal(after, before)


def test_render_batch_equals_single_frames(gpu_ctx):
    """ymrt_render_batch: frames of several scenes (one scene twice, with two
    frame indices) in shared launches equal ymrt_render_frame's bit for bit,
    packed three chunks to a round and all in one round; a bad item fails
    the call with its index in the message."""
    jobs = []
    cb = ymrt_amd.Scene(gpu_ctx, scenes.cornell_box(40, 32))
    od = ymrt_amd.Scene(gpu_ctx, scenes.outdoor(48, 32))
    tb = ymrt_amd.Scene(gpu_ctx, scenes.textured_box(32, 32))
    it = ymrt_amd.Scene(gpu_ctx, scenes.interior(64, 40, tex_size=64, instances=6))
    jobs = [(cb, abi.render_settings(40, 32, 4, 5), 1),
            (od, abi.render_settings(48, 32, 4, 5, distant_light=1), 1),
            (tb, abi.render_settings(32, 32, 2, 4), 3),
            (it, abi.render_settings(64, 40, 4, 6, volume=1), 1),
            (cb, abi.render_settings(40, 32, 4, 5), 2)]
    out = gpu_ctx.render_batch(jobs, max_slots=3)
    out_one = gpu_ctx.render_batch(jobs)
    for (sc, rs, frame), (img, st), (img1, st1) in zip(jobs, out, out_one):
        ref, ref_st = sc.render_frame(rs, frame)
        assert torch.equal(img, ref)
        assert torch.equal(img1, ref)
        assert st.extension_rays == ref_st.extension_rays > 0
        assert st1.shadow_rays == ref_st.shadow_rays
    assert not torch.equal(out[0][0], out[4][0])
    bad = list(jobs)
    bad[2] = (tb, abi.render_settings(32, 32, 2, 4), 0)
    with pytest.raises(ymrt_amd.YmrtError, match="batch item 2"):
        gpu_ctx.render_batch(bad)


def test_diffuse_only_shading_equals_general(gpu_ctx):
    """Hits on diffuse-only materials shade in a variant with the diffuse lobe
    alone; it must give the general variant's floats.  The same Cornell box
    with the walls' coat thickness read from a constant texture (the same
    value, which keeps the walls out of the diffuse-only class) renders the
    same image bit for bit."""
    from ymrt_amd import abi as T
    sa = scenes.cornell_box(48, 40)
    sb = scenes.cornell_box(48, 40)
    one = sb.add_texture_node(T.TEXTURE_CONSTANT_FLOAT, value=(1.0, 0.0, 0.0))
    walls = 0
    for m in sb.materials:
        if m.type == T.MATERIAL_DEFAULT and m.specular_weight == 0.0 and m.metalness == 0.0:
            assert m.coat_thickness == 1.0
            m.coat_thickness_tex = one
            walls += 1
    assert walls >= 3
    rs = abi.render_settings(48, 40, 8, 6)
    a = ymrt_amd.Scene(gpu_ctx, sa).render_frame(rs, 1)[0]
    b = ymrt_amd.Scene(gpu_ctx, sb).render_frame(rs, 1)[0]
    assert torch.equal(a, b)


@pytest.mark.gpu
def test_metal_only_shading_equals_general(gpu_ctx):
    """Hits on metal-only materials (metalness 1, no coat) shade in a variant
    with the metal and diffuse lobes fixed at compile time; it must give the
    general variant's floats.  The walls of a Cornell box made rough metal,
    once plain and once with the metalness read from a constant texture of
    the same value (which keeps them out of the metal class), render the same
    image bit for bit."""
    from ymrt_amd import abi as T
    sa = scenes.cornell_box(48, 40)
    sb = scenes.cornell_box(48, 40)
    one = sb.add_texture_node(T.TEXTURE_CONSTANT_FLOAT, value=(1.0, 0.0, 0.0))
    walls = 0
    for s in (sa, sb):
        for m in s.materials:
            if m.type == T.MATERIAL_DEFAULT and m.specular_weight == 0.0 and m.metalness == 0.0:
                m.metalness = 1.0
                m.roughness_x = m.roughness_y = 0.35
                m.specular_albedo = (0.9, 0.6, 0.3)
                m.coat_weight = 0.0
                if s is sb:
                    m.metalness_tex = one
                    walls += 1
    assert walls >= 3
    rs = abi.render_settings(48, 40, 8, 6)
    a = ymrt_amd.Scene(gpu_ctx, sa).render_frame(rs, 1)[0]
    b = ymrt_amd.Scene(gpu_ctx, sb).render_frame(rs, 1)[0]
    assert torch.isfinite(a).all() and a[..., :3].abs().sum() > 0
    assert torch.equal(a, b)


@pytest.mark.gpu
def test_autotune_and_launch_settings_keep_pixels(gpu_ctx):
    """path_capacity and persistent_waves change timing, never a pixel; the
    autotuner returns one of its candidates for each and a positive time."""
    sa = scenes.interior(64, 48, tex_size=32, instances=4)
    rs = abi.render_settings(64, 48, 8, 6)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    ref = sc.render_frame(rs, 2)[0].cpu().numpy()
    paths = 4 * 3 * 256 * 8
    for cap, waves in [(4096, 0), (0, 1024), (paths // 4, 64), (0, 1)]:
        rs2 = abi.render_settings(64, 48, 8, 6, path_capacity=cap, persistent_waves=waves)
        img = sc.render_frame(rs2, 2)[0].cpu().numpy()
        assert np.array_equal(img.view(np.uint32), ref.view(np.uint32)), (cap, waves)
    with pytest.raises(ymrt_amd.YmrtError):
        sc.render_frame(abi.render_settings(64, 48, 8, 6, persistent_waves=5121), 2)
    res = sc.autotune(rs, 2)
    print("autotune: capacity %d waves %d  %.3f ms (default %.3f ms), %d trials" %
          (res.path_capacity, res.persistent_waves, res.ms, res.ms_default, res.trials))
    assert res.path_capacity in (0, paths // 2, paths // 4)
    assert res.persistent_waves in (0, 4096, 2048, 1024)
    assert 0.0 < res.ms <= res.ms_default
    assert res.trials == 1 + 2 * 3 + 2 * 3
    rs3 = abi.render_settings(64, 48, 8, 6, path_capacity=res.path_capacity, persistent_waves=res.persistent_waves)
    img = sc.render_frame(rs3, 2)[0].cpu().numpy()
    assert np.array_equal(img.view(np.uint32), ref.view(np.uint32))


def test_float_texture_and_texel_update(gpu_ctx):
    """A linear float image texture drives the white walls' albedo; the GPU
    image matches the oracle's before and after ymrt_scene_update_texture
    replaces its texels."""
    sa = scenes.cornell_box(48, 48)
    rng = np.random.RandomState(5)
    t0 = rng.rand(16, 16, 3).astype(np.float32)
    ti = sa.add_texture(t0, scale=(3.0, 2.0))
    sa.materials[0].diffuse_albedo_tex = ti
    rs = abi.render_settings(48, 48, 4, 5)
    sc = ymrt_amd.Scene(gpu_ctx, sa)
    g0 = sc.render_frame(rs, 1)[0].cpu().numpy()
    o0 = oracle.Scene(sa).render(rs, 1, threads=8)
    t1 = (rng.rand(16, 16, 3) * 0.5).astype(np.float32)
    sc.update_texture(ti, t1)
    g1 = sc.render_frame(rs, 1)[0].cpu().numpy()
    sa.textures[ti]["texels"] = t1
    o1 = oracle.Scene(sa).render(rs, 1, threads=8)
    sc.close()
    for g, o in ((g0, o0), (g1, o1)):
        same, err = close_frac(g, o), rel_rmse(g, o)
        print("float texture: close pixels %.4f  rel rmse %.2e" % (same, err))
        assert np.isfinite(g).all()
        assert same >= 0.999
        assert err <= 1e-3
    assert float(np.abs(g1 - g0).mean()) > 1e-3


def test_open_medium_parity(gpu_ctx):
    """Rays that enter a medium through its only boundary face and then hit
    nothing while inside it (free flight to infinity, volume vertices lit
    by a distant light and the environment), and rays that miss it."""
    sa = scenes.open_medium(48, 32)
    rs = abi.render_settings(48, 32, 8, 4, volume=1, env_light=1, distant_light=1)
    g, o, stats = render_both(gpu_ctx, sa, rs)
    assert np.isfinite(g).all()
    same, err = close_frac(g, o), rel_rmse(g, o)
    print("open medium: close pixels %.4f  rel rmse %.2e" % (same, err))
    assert same >= 0.999
    assert err <= 1e-4
