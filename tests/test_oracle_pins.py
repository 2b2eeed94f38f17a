"""Pins of the oracle's path loop against closed forms and quadratures (no GPU).

Each scene is built so that one part of the restated path loop decides the
image and a closed form (or an FP64 quadrature) gives its expected value;
the oracle's image must match within Monte Carlo error.  The parts and the
reference code they follow:
  * camera medium records and the boundary pass-through (Camera::
    InitCameraRayTransfer, src/Camera.cpp:15-78; the pass-through of
    src/Shading.cu:2150-2170; BoundaryTransitionBunch,
    src/ShadingUtilities.cuh:954-990): absorbing media between the camera
    and an emitter attenuate it by exp(-sigma_t d) per channel;
  * the mixture of overlapping media (SampleTrHomogeneous,
    src/VolumeEval.cuh:18-78, sums sigma_t over the top-priority records);
  * next-event estimation at volume vertices (EvalVolumeShapeLight,
    src/Shading.cu:419-513) against an FP64 single-scattering quadrature;
  * the multi-bounce loop (throughput composition over depth,
    src/Shading.cu:2240-2302) in an integrating sphere, whose wall radiance
    after D bounces is Le (r/R)^2 rho (1 - rho^D) / (1 - rho);
  * Russian roulette (src/Shading.cu:2265-2276): unbiased, so with a short
    ray_depth it reaches the infinite-bounce sum rho / (1 - rho);
  * the environment background (Background, src/Shading.cu:35-38) seen by a
    Lambert floor: the hemisphere integral of the background;
  * nested dielectrics: a surface of lower priority inside a dielectric is
    skipped (src/Shading.cu:2150-2158), so an index-matched inner sphere of
    lower priority changes nothing;
  * voxel-grid media: free flight through a grid with distinct voxel
    densities along x and z, and shadow transmittance through layers along y,
    against exp(-sum sigma_i dt_i) (piecewise-constant optical depth).
"""
import math

import numpy as np
import pytest

import oracle
from ymrt_amd import abi, scenes


def _lambert(albedo):
    return abi.default_material(diffuse_albedo=albedo, specular_weight=0.0, metalness=0.0, ior_n=1.0)


def _boundary_material(ior=1.0, priority=0):
    # index-matched boundary: no refraction, only the medium change
    return abi.default_material(ior_n=ior, ior_priority=priority, specular_weight=0.0)


def _box_w2o(lo, hi):
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    w2o = np.eye(4, dtype=np.float32)
    w2o[:3, :3] = np.diag(1.0 / (hi - lo))
    w2o[:3, 3] = -lo / (hi - lo)
    return w2o


def _add_medium_box(sa, lo, hi, sigma, albedo=(0.0, 0.0, 0.0), g=0.0, grid=-1, priority=0):
    vol = sa.add_volume(tuple(float(s) for s in sigma), tuple(float(a) for a in albedo), g, grid=grid,
                        world_to_object=_box_w2o(lo, hi))
    sa.add_instance(sa.add_mesh(*scenes.box(tuple(lo), tuple(hi))[:2]), sa.add_material(_boundary_material(1.0, priority)),
                    inner_volume=vol, boundary=True)
    return vol


def _emitter_facing_minus_z(sa, z, half=3.0, le=1.0):
    # quad in the plane z, seen from z' > z (the camera looks along -z)
    q = scenes.quad((-half, -half, z), (half, -half, z), (half, half, z), (-half, half, z))
    sa.add_instance(sa.add_mesh(*q[:2]), sa.add_material(abi.light_material((1.0, 1.0, 1.0), le)))


def _render_mean(sa, rs, frames=(1,)):
    sc = oracle.Scene(sa)
    imgs = [sc.render(rs, f, threads=8)[..., :3].astype(np.float64) for f in frames]
    sc.close()
    return np.mean([im.mean(axis=(0, 1)) for im in imgs], axis=0), imgs


def _check_case(case):
    sa, rs, frames, expect, rtol = case()
    got, _ = _render_mean(sa, rs, frames)
    assert np.allclose(got, expect, rtol=rtol), (got, expect)
    return sa, rs, frames, expect, rtol


# Scenes whose expected image mean is a closed form or an FP64 quadrature;
# tests/test_gpu_configs.py renders the same cases on the GPU.
def closed_form_cases():
    return {
        "camera_inside_absorbing_medium": case_camera_inside_absorbing_medium,
        "boundary_pass_through": case_boundary_pass_through,
        "overlapping_media": case_overlapping_media,
        "volume_vertex_nee": case_volume_vertex_nee,
        "volume_vertex_distant_nee": case_volume_vertex_distant_nee,
        "integrating_sphere_depth5": lambda: case_integrating_sphere(5),
        "russian_roulette": case_russian_roulette,
        "env_background": case_env_background,
        "grid_free_flight": case_grid_free_flight,
        "grid_shadow_layers": case_grid_shadow_layers,
        "texture_density_constant": case_texture_density_constant,
        "texture_density_checker": case_texture_density_checker,
        "texture_density_shadow": case_texture_density_shadow,
    }


# ----------------------------------------------- media between camera and light
def case_camera_inside_absorbing_medium():
    """The camera sits inside an absorbing box (albedo 0); an emitter inside
    the box at distance d: every camera ray that reaches it carries
    exp(-sigma_c d), any scattering event kills the path (albedo 0).  The
    camera's starting medium comes only from the camera medium records."""
    sigma = np.array([0.2, 0.6, 1.1])
    d = 1.5
    sa = abi.SceneArrays()
    _add_medium_box(sa, (-5, -5, -5), (5, 5, 5), sigma)
    _emitter_facing_minus_z(sa, -d)
    sa.set_camera((0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 1.0, 1.0)
    rs = abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1)
    return sa, rs, (1,), np.exp(-sigma * d), 0.03


def test_camera_inside_absorbing_medium():
    _check_case(case_camera_inside_absorbing_medium)


def case_boundary_pass_through():
    """Camera outside an absorbing slab of thickness 0.8 with the emitter
    behind it: the camera ray passes the slab's two faces without consuming
    depth (ray_depth 1 still reaches the light) and picks up exp(-sigma 0.8)."""
    sigma = np.array([0.9, 0.5, 0.1])
    sa = abi.SceneArrays()
    _add_medium_box(sa, (-4, -4, -1.0), (4, 4, -0.2), sigma)
    _emitter_facing_minus_z(sa, -3.0)
    sa.set_camera((0.0, 0.0, 2.0), (0.0, 0.0, -1.0), 1.0, 1.0)
    rs = abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1)
    return sa, rs, (1,), np.exp(-sigma * 0.8), 0.03


def test_boundary_pass_through_attenuates():
    sa, _, _, _, _ = _check_case(case_boundary_pass_through)
    # without volume scattering the media are ignored: the emitter at full radiance
    got0, _ = _render_mean(sa, abi.render_settings(8, 8, 4, 1, sampler=abi.SAMPLER_PCG, volume=0))
    assert np.allclose(got0, 1.0, rtol=1e-6)


def case_overlapping_media():
    """Two absorbing boxes of equal priority overlap over 0.5 of the camera
    ray: the overlap has sigma1 + sigma2, so the emitter behind both is seen
    through exp(-(sigma1 * 1.0 + sigma2 * 0.9))."""
    s1 = np.array([0.4, 0.2, 0.7])
    s2 = np.array([0.3, 0.8, 0.1])
    sa = abi.SceneArrays()
    _add_medium_box(sa, (-4, -4, -1.0), (4, 4, 0.0), s1)      # z in [-1, 0]
    _add_medium_box(sa, (-4, -4, -1.4), (4, 4, -0.5), s2)     # z in [-1.4, -0.5], overlap 0.5
    _emitter_facing_minus_z(sa, -3.0)
    sa.set_camera((0.0, 0.0, 2.0), (0.0, 0.0, -1.0), 1.0, 1.0)
    rs = abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1)
    return sa, rs, (1,), np.exp(-(s1 * 1.0 + s2 * 0.9)), 0.03


def test_overlapping_media_mix():
    _check_case(case_overlapping_media)


# ------------------------------------------------------- volume-vertex NEE
def _single_scatter_f64(o, d, t_exit, sigma_t, sigma_s, c, radius, le, n_t=400, n_th=96):
    """FP64 single-scattering radiance along the ray o + t d, t in [0, t_exit],
    in a homogeneous isotropic medium lit by a sphere light (centre c, radius,
    radiance le), with transmittance to the light's surface: Gauss-Legendre in
    t and in the angle from the axis towards the light."""
    xt, wt = np.polynomial.legendre.leggauss(n_t)
    ts = 0.5 * t_exit * (xt + 1.0)
    wts = 0.5 * t_exit * wt
    xa, wa = np.polynomial.legendre.leggauss(n_th)
    total = np.zeros(3)
    for t, w in zip(ts, wts):
        x = o + d * t
        dist = np.linalg.norm(c - x)
        th_max = math.asin(min(radius / dist, 1.0))
        th = 0.5 * th_max * (xa + 1.0)
        wth = 0.5 * th_max * wa
        s = dist * np.cos(th) - np.sqrt(np.maximum(radius * radius - dist * dist * np.sin(th) ** 2, 0.0))
        inner = np.array([np.sum(wth * 2.0 * math.pi * np.sin(th) * np.exp(-sig * s)) for sig in sigma_t])
        total += w * np.exp(-sigma_t * t) * sigma_s * inner * le / (4.0 * math.pi)
    return total


def case_volume_vertex_nee():
    """Camera rays through a thin isotropic medium (g = 0) that miss a small
    sphere light inside it: at ray_depth 1 the image is single scattering
    only, light sampling and phase sampling combined with the power
    heuristic, transmittance to the light by the homogeneous evaluation."""
    sigma_t = np.array([0.15, 0.3, 0.45])
    albedo = np.array([0.9, 0.8, 0.7])
    sa = abi.SceneArrays()
    _add_medium_box(sa, (-3, -3, -3), (3, 3, 3), sigma_t, albedo=albedo, g=0.0)
    c = np.array([0.0, 0.6, -1.0])
    m = np.eye(4, dtype=np.float32)
    m[:3, 3] = c
    radius, le = 0.25, 20.0
    sa.add_instance(sa.add_sphere(radius), sa.add_material(abi.light_material((1.0, 1.0, 1.0), le)), m)
    o = np.array([0.0, 0.0, 2.0])
    sa.set_camera(tuple(o), (0.0, 0.0, -1.0), 0.05, 1.0)
    rs = abi.render_settings(4, 4, 2048, 1, sampler=abi.SAMPLER_PCG, volume=1)
    d = np.array([0.0, 0.0, -1.0])
    t_exit = 5.0   # camera at z = 2 inside the box, exit at z = -3
    expect = _single_scatter_f64(o, d, t_exit, sigma_t, sigma_t * albedo, c, radius, le)
    return sa, rs, (1, 2), expect, 0.03


def test_volume_vertex_nee_single_scattering():
    _check_case(case_volume_vertex_nee)


def case_volume_vertex_distant_nee():
    """The same thin isotropic medium lit by a distant light (cone of 5
    degrees about +y) instead: EvalVolumeDistantLight (src/Shading.cu:202-290)
    with the reference's DistantLight (SampleLi returns Le, EvalLi Le * pdf,
    src/Light.h:344-385), so each vertex adds sigma_s / (4 pi) * Le times the
    cone average of the transmittance to the top of the medium box; camera
    rays run along z at height 0, 3 below the top."""
    sigma_t = np.array([0.1, 0.2, 0.3])
    albedo = np.array([0.9, 0.8, 0.7])
    sa = abi.SceneArrays()
    _add_medium_box(sa, (-3, -3, -3), (3, 3, 3), sigma_t, albedo=albedo, g=0.0)
    flip = np.diag([1.0, -1.0, -1.0, 1.0]).astype(np.float32)   # light frame's -y points up: light from above
    theta = 5.0
    le = 4.0
    sa.add_distant_light((1.0, 1.0, 1.0), le, theta, flip)
    sa.set_camera((0.0, 0.0, 2.0), (0.0, 0.0, -1.0), 0.05, 1.0)
    rs = abi.render_settings(4, 4, 2048, 1, sampler=abi.SAMPLER_PCG, volume=1, shape_light=0, distant_light=1)
    th_max = math.radians(theta)
    x, w = np.polynomial.legendre.leggauss(64)
    th = 0.5 * th_max * (x + 1.0)
    wth = 0.5 * th_max * w
    omega = 2.0 * math.pi * (1.0 - math.cos(th_max))
    cone_tr = np.array([np.sum(wth * 2.0 * math.pi * np.sin(th) * np.exp(-sig * 3.0 / np.cos(th))) / omega
                        for sig in sigma_t])
    t_exit = 5.0
    along = (1.0 - np.exp(-sigma_t * t_exit)) / sigma_t      # integral of exp(-sigma_t t) over the ray
    expect = sigma_t * albedo / (4.0 * math.pi) * le * cone_tr * along
    return sa, rs, (1, 2), expect, 0.03


def test_volume_vertex_distant_nee():
    _check_case(case_volume_vertex_distant_nee)


# ----------------------------------------------------- multi-bounce loop
def _integrating_sphere_scene(rho, depth, rr=0, spp=256):
    """A Lambert sphere (radius 1, albedo rho) with a sphere light of radius
    0.01 and radiance 1e4 at its centre, so the first-bounce irradiance is
    uniform; the camera sits beside the light, looking away from it."""
    sa = abi.SceneArrays()
    sa.add_instance(sa.add_sphere(1.0), sa.add_material(_lambert((rho, rho, rho))))
    sa.add_instance(sa.add_sphere(0.01), sa.add_material(abi.light_material((1.0, 1.0, 1.0), 1.0e4)))
    sa.set_camera((0.0, 0.05, 0.0), (1.0, -0.25, 0.2), 60.0, 1.0)
    return sa, abi.render_settings(8, 8, spp, depth, sampler=abi.SAMPLER_PCG, rr=rr)


def _sphere_wall(rho, depth):
    # Le (r/R)^2 = 1: each bounce multiplies the uniform wall radiance by rho;
    # the light blocks (r/R)^2 / 4 of each bounce's hemisphere: exact to 1e-4
    return rho * (1.0 - rho ** depth) / (1.0 - rho)


def case_integrating_sphere(depth):
    sa, rs = _integrating_sphere_scene(0.5, depth)
    return sa, rs, (1,), np.full(3, _sphere_wall(0.5, depth)), 0.02


@pytest.mark.parametrize("depth", [1, 2, 5])
def test_integrating_sphere_depth(depth):
    _check_case(lambda: case_integrating_sphere(depth))


def case_russian_roulette():
    """With Russian roulette on, paths go on past ray_depth with probability
    min(max(throughput), 1) and are reweighted: ray_depth 2 reaches the
    infinite-bounce wall radiance rho / (1 - rho)."""
    rho = 0.6
    sa, rs = _integrating_sphere_scene(rho, 2, rr=1, spp=512)
    return sa, rs, (1, 2), np.full(3, rho / (1.0 - rho)), 0.03


def test_russian_roulette_unbiased():
    """Russian roulette reaches the infinite sum; with it off, ray_depth 2
    stops at rho (1 + rho)."""
    _check_case(case_russian_roulette)
    rho = 0.6
    sa, rs = _integrating_sphere_scene(rho, 2, rr=0)
    off, _ = _render_mean(sa, rs)
    assert np.allclose(off, rho * (1.0 + rho), rtol=0.02), off


# ------------------------------------------------------- environment light
def case_env_background():
    """An infinite Lambert floor (albedo 0.5) under the environment only:
    at ray_depth 2 the floor sends cosine-distributed rays into the sky, so
    L = rho / pi * integral of Background(w) cos over the hemisphere, with
    Background = mix(1, (0.05, 0.2, 0.85), (w.y + 1) / 2):
    L = rho * (1 + (c - 1) * 5 / 6)."""
    sa = abi.SceneArrays()
    sa.add_instance(sa.add_mesh(*scenes.quad((-1e3, 0, -1e3), (-1e3, 0, 1e3), (1e3, 0, 1e3), (1e3, 0, -1e3))[:2]),
                    sa.add_material(_lambert((0.5, 0.5, 0.5))))
    sa.set_camera((0.0, 1.0, 0.0), (0.0, 0.0, 0.001), 10.0, 1.0)
    rs = abi.render_settings(8, 8, 512, 2, sampler=abi.SAMPLER_PCG, shape_light=0, env_light=1)
    c = np.array([0.05, 0.2, 0.85])
    return sa, rs, (1,), 0.5 * (1.0 + (c - 1.0) * 5.0 / 6.0), 0.02


def test_env_background_on_lambert_floor():
    sa, _, _, _, _ = _check_case(case_env_background)
    # a camera ray into the sky sees the background itself
    sa.set_camera((0.0, 1.0, 0.0), (0.0, 2.0, 0.001), 0.01, 1.0)
    got_sky, _ = _render_mean(sa, abi.render_settings(2, 2, 1, 1, sampler=abi.SAMPLER_PCG, env_light=1))
    assert np.allclose(got_sky, [0.05, 0.2, 0.85], rtol=1e-4), got_sky


# ---------------------------------------------------------- nested dielectrics
def _glass_ball(inner):
    sa = abi.SceneArrays()
    sa.add_instance(sa.add_mesh(*scenes.quad((-5, -1, -5), (-5, -1, 5), (5, -1, 5), (5, -1, -5))[:2]),
                    sa.add_material(_lambert((0.6, 0.5, 0.4))))
    lq = scenes.quad((-1, 3, -1), (1, 3, -1), (1, 3, 1), (-1, 3, 1))
    sa.add_instance(sa.add_mesh(*lq[:2]), sa.add_material(abi.light_material((1.0, 1.0, 1.0), 8.0)))
    glass = dict(transmission_weight=1.0, specular_weight=1.0, roughness_x=0.0, roughness_y=0.0,
                 metalness=0.0)
    sa.add_instance(sa.add_sphere(0.8), sa.add_material(abi.default_material(ior_n=1.5, ior_priority=2, **glass)))
    if inner:
        # same IOR, lower priority: skipped whenever the outer glass is current
        sa.add_instance(sa.add_sphere(0.4), sa.add_material(abi.default_material(ior_n=1.5, ior_priority=1, **glass)))
    sa.set_camera((0.0, 0.3, 3.0), (0.0, 0.0, 0.0), 35.0, 1.0)
    return sa


def test_lower_priority_inner_sphere_is_skipped():
    rs = abi.render_settings(24, 24, 64, 6, sampler=abi.SAMPLER_PCG)
    sc0 = oracle.Scene(_glass_ball(False))
    a = sc0.render(rs, 1, threads=8)[..., :3].astype(np.float64)
    sc0.close()
    sc1 = oracle.Scene(_glass_ball(True))
    b = sc1.render(rs, 1, threads=8)[..., :3].astype(np.float64)
    sc1.close()
    # same samples along the same paths: the skip splits a segment at the inner
    # surface, so only the last bits of later hit points move
    close = np.all(np.abs(a - b) <= 1e-3 * np.abs(a) + 1e-5, axis=-1).mean()
    assert close >= 0.97, close
    assert abs(a.mean() - b.mean()) <= 0.005 * a.mean()
    # a higher-priority inner sphere is not skipped and changes the image
    sa = _glass_ball(True)
    sa.materials[-1].ior_priority = 3
    sa.materials[-1].ior_n = 2.2
    sc2 = oracle.Scene(sa)
    c = sc2.render(rs, 1, threads=8)[..., :3].astype(np.float64)
    sc2.close()
    assert np.all(np.abs(a - c) <= 1e-3 * np.abs(a) + 1e-5, axis=-1).mean() < 0.9


# ------------------------------------------------------------- voxel grids
def case_grid_free_flight():
    """Camera rays along +x through an absorbing grid box of 2 x 1 x 3
    voxels (x, y, z) with distinct densities, aimed at the middle z layer:
    the emitter behind it is seen through exp(-sigma (d[z=1,x=0] + d[z=1,x=1]) dx).
    A slip in the voxel index order (x <-> z) or an off-by-one reads other
    voxels and fails."""
    dens = np.array([[[0.3, 1.7]], [[0.5, 2.5]], [[1.1, 0.2]]], np.float32)   # (nz, ny, nx)
    sigma = 0.8
    lo, hi = (0.0, -1.0, -1.5), (2.0, 1.0, 1.5)    # voxels 1 x 2 x 1 world units
    sa = abi.SceneArrays()
    g = sa.add_grid(dens)
    _add_medium_box(sa, lo, hi, (sigma, sigma, sigma), grid=g)
    q = scenes.quad((4.0, -3, 3), (4.0, -3, -3), (4.0, 3, -3), (4.0, 3, 3))  # faces -x
    sa.add_instance(sa.add_mesh(*q[:2]), sa.add_material(abi.light_material((1.0, 1.0, 1.0), 1.0)))
    sa.set_camera((-2.0, 0.0, 0.0), (1.0, 0.0, 0.0), 1.0, 1.0)
    rs = abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1)
    expect = math.exp(-sigma * (dens[1, 0, 0] * 1.0 + dens[1, 0, 1] * 1.0))
    return sa, rs, (1,), np.full(3, expect), 0.03


def test_grid_free_flight_distinct_voxels():
    _check_case(case_grid_free_flight)


def case_grid_shadow_layers():
    """A Lambert floor lit by a square emitter through a wide absorbing slab of
    two density layers along y: every shadow ray crosses both layers, so the
    floor radiance is rho / pi * Le * integral over the emitter of
    h^2 / r^4 * exp(-tau r / h) with tau = sigma (d0 dy0 + d1 dy1)."""
    dens = np.array([[[0.4], [1.6]]], np.float32)       # (nz, ny, nx): two layers along y
    sigma = 0.9
    lo, hi = (-50.0, 1.0, -50.0), (50.0, 1.6, 50.0)        # layers 0.3 thick
    sa = abi.SceneArrays()
    g = sa.add_grid(dens)
    _add_medium_box(sa, lo, hi, (sigma, sigma, sigma), grid=g)
    sa.add_instance(sa.add_mesh(*scenes.quad((-60, 0, -60), (-60, 0, 60), (60, 0, 60), (60, 0, -60))[:2]),
                    sa.add_material(_lambert((0.5, 0.5, 0.5))))
    a, h = 1.0, 2.0
    lq = scenes.quad((-a, h, -a), (a, h, -a), (a, h, a), (-a, h, a))
    sa.add_instance(sa.add_mesh(*lq[:2]), sa.add_material(abi.light_material((1.0, 1.0, 1.0), 1.0)))
    sa.set_camera((0.0, 0.5, 0.0001), (0.0, 0.0, 0.0), 0.2, 1.0)
    rs = abi.render_settings(8, 8, 512, 1, sampler=abi.SAMPLER_PCG, volume=1)
    tau = sigma * (0.4 * 0.3 + 1.6 * 0.3)
    x, w = np.polynomial.legendre.leggauss(64)
    u = a * x
    uu, vv = np.meshgrid(u, u)
    r = np.sqrt(uu * uu + vv * vv + h * h)
    integral = np.sum(np.outer(w, w) * a * a * (h * h / r ** 4) * np.exp(-tau * r / h))
    return sa, rs, (1,), np.full(3, 0.5 / math.pi * integral), 0.03


def test_grid_shadow_transmittance_layers():
    _check_case(case_grid_shadow_layers)




# --------------------------------------------------- texture-density media
def _texture_medium_slab(sa, sigma, tex, lo=(-4, -4, -1.0), hi=(4, 4, 3.0)):
    # object space = world space: the texture is evaluated at world points
    vol = sa.add_volume(tuple(float(x) for x in sigma), (0.0, 0.0, 0.0), 0.0, density_texture=tex,
                        world_to_object=np.eye(4, dtype=np.float32))
    sa.add_instance(sa.add_mesh(*scenes.box(tuple(lo), tuple(hi))[:2]), sa.add_material(_boundary_material()),
                    inner_volume=vol, boundary=True)


def _reference_escape_weight(sigma_m, t):
    """Mean weight of a camera ray that crosses t of a medium whose density
    equals its majorant (every tentative collision real) under the
    reference's tracker, SampleTrHeterogeneous (src/VolumeEval.cuh:80-214):
    channel c is picked with probability w_c = sm_c / sum(sm), the ray
    escapes with probability exp(-sm_c t), and the escape weight is
    nf_k exp(-sm_k t) / sum_c w_c nf_c exp(-sm_c t) with the normalisation
    factor nf = 1 / (sm + 1) in both the transmittance and the pdf.  That
    factor does not cancel between channels, so for chromatic extinction the
    mean is not exp(-sm_k t); for gray media it is."""
    w = sigma_m / sigma_m.sum()
    e = np.exp(-sigma_m * t)
    nf = 1.0 / (sigma_m + 1.0)
    return np.dot(w, e) * nf * e / np.dot(w, nf * e)


def case_texture_density_constant():
    """A constant density texture c behaves as the medium of extinction
    sigma c under the heterogeneous tracker (its majorant GetUpperBound is
    c, so every tentative collision is real): the emitter behind a 4-deep
    slab is seen through the tracker's escape weight, which for this
    chromatic sigma is not exp(-sigma c 4) (see _reference_escape_weight)."""
    sigma = np.array([0.5, 0.25, 0.125])
    sa = abi.SceneArrays()
    tex = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(0.6, 0.0, 0.0))
    _texture_medium_slab(sa, sigma, tex)
    _emitter_facing_minus_z(sa, -3.0)
    sa.set_camera((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), 1.0, 1.0)
    rs = abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1)
    expect = _reference_escape_weight(sigma * 0.6, 4.0)
    assert not np.allclose(expect, np.exp(-sigma * 0.6 * 4.0), rtol=0.05)   # the quirk is visible here
    return sa, rs, (1,), expect, 0.03


def test_gray_constant_density_is_unbiased():
    """For gray extinction the tracker's escape weight is exp(-sigma c t)."""
    sm = np.array([0.3, 0.3, 0.3])
    assert np.allclose(_reference_escape_weight(sm, 4.0), np.exp(-sm * 4.0), rtol=1e-12)
    sigma = np.array([0.5, 0.5, 0.5])
    sa = abi.SceneArrays()
    tex = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(0.6, 0.0, 0.0))
    _texture_medium_slab(sa, sigma, tex)
    _emitter_facing_minus_z(sa, -3.0)
    sa.set_camera((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), 1.0, 1.0)
    got, _ = _render_mean(sa, abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1))
    assert np.allclose(got, np.exp(-sigma * 0.6 * 4.0), rtol=0.03), got


def case_texture_density_checker():
    """A checkerboard density (frequency pi: white where sin(pi x) sin(pi y)
    sin(pi z) < 0) of black 0.25 and white 1.0 under its majorant 1.  The
    camera rays run along -z through x = y = 0.5, so the density alternates
    per unit of z: over z in [-1, 3] two cells of each, optical depth
    sigma (2 * 0.25 + 2 * 1.0).  Delta tracking with null collisions decides
    this image; gray extinction, where the tracker is unbiased."""
    sigma = np.array([0.6, 0.6, 0.6])
    sa = abi.SceneArrays()
    black = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(0.25, 0.0, 0.0))
    white = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(1.0, 0.0, 0.0))
    tex = sa.add_texture_node(abi.TEXTURE_CHECKERBOARD, (black, white), frequency=math.pi)
    _texture_medium_slab(sa, sigma, tex)
    _emitter_facing_minus_z(sa, -3.0)
    sa.set_camera((0.5, 0.5, 5.0), (0.5, 0.5, -1.0), 0.5, 1.0)
    rs = abi.render_settings(8, 8, 256, 1, sampler=abi.SAMPLER_PCG, volume=1)
    return sa, rs, (1,), np.exp(-sigma * (2 * 0.25 + 2 * 1.0)), 0.03


def case_texture_density_shadow():
    """Shadow transmittance through a texture density: the grid-layer floor
    scene with a constant-density (0.5) texture slab, so every shadow ray
    crosses optical depth sigma 0.5 0.6 / cos."""
    sigma = 1.2
    sa = abi.SceneArrays()
    tex = sa.add_texture_node(abi.TEXTURE_CONSTANT_FLOAT, value=(0.5, 0.0, 0.0))
    _texture_medium_slab(sa, (sigma, sigma, sigma), tex, lo=(-50.0, 1.0, -50.0), hi=(50.0, 1.6, 50.0))
    sa.add_instance(sa.add_mesh(*scenes.quad((-60, 0, -60), (-60, 0, 60), (60, 0, 60), (60, 0, -60))[:2]),
                    sa.add_material(_lambert((0.5, 0.5, 0.5))))
    a, h = 1.0, 2.0
    lq = scenes.quad((-a, h, -a), (a, h, -a), (a, h, a), (-a, h, a))
    sa.add_instance(sa.add_mesh(*lq[:2]), sa.add_material(abi.light_material((1.0, 1.0, 1.0), 1.0)))
    sa.set_camera((0.0, 0.5, 0.0001), (0.0, 0.0, 0.0), 0.2, 1.0)
    rs = abi.render_settings(8, 8, 512, 1, sampler=abi.SAMPLER_PCG, volume=1)
    tau = sigma * 0.5 * 0.6
    x, w = np.polynomial.legendre.leggauss(64)
    u = a * x
    uu, vv = np.meshgrid(u, u)
    r = np.sqrt(uu * uu + vv * vv + h * h)
    integral = np.sum(np.outer(w, w) * a * a * (h * h / r ** 4) * np.exp(-tau * r / h))
    return sa, rs, (1,), np.full(3, 0.5 / math.pi * integral), 0.03


def test_texture_density_constant_is_homogeneous():
    _check_case(case_texture_density_constant)


def test_texture_density_checkerboard():
    _check_case(case_texture_density_checker)


def test_texture_density_shadow_transmittance():
    _check_case(case_texture_density_shadow)


def test_image_density_has_zero_majorant():
    """GetUpperBound returns 0 for an image texture (src/Texture.h:1183-1186),
    so a medium whose density is an image has majorant 0 and never
    attenuates: the emitter behind it keeps its full radiance."""
    sa = abi.SceneArrays()
    tex = sa.add_texture(np.full((4, 4), 255, np.uint8), color_space=abi.COLOR_LINEAR)
    _texture_medium_slab(sa, (2.0, 2.0, 2.0), tex)
    _emitter_facing_minus_z(sa, -3.0)
    sa.set_camera((0.0, 0.0, 5.0), (0.0, 0.0, -1.0), 1.0, 1.0)
    got, _ = _render_mean(sa, abi.render_settings(4, 4, 16, 1, sampler=abi.SAMPLER_PCG, volume=1))
    assert np.allclose(got, 1.0, rtol=1e-6), got


# --------------------------------------------------------- indirect clamp
def test_indirect_clamp_special_values():
    """indirect_light_clamp (src/Shading.cu:61-81) scales a contribution c of
    luminance Y (0.2126 r + 0.7152 g + 0.0722 b) at depth > 1 by
    (lim / (lim + Y))^p with p = min(0.05 + 0.25 (Y / lim)^1.5, 1):
    Y = lim gives 2^-0.3, Y = 4 lim gives p = 1 and a factor 1/5; depth 1,
    the setting off and Y <= 1e-10 leave c alone; a limit <= 0 clears it."""
    rs = abi.render_settings(1, 1, 1, 8)
    rs.enable_indirect_clamp = 1
    rs.indirect_limit = 2.0
    lum_w = np.array([0.2126, 0.7152, 0.0722])
    for col in (np.array([1.0, 1.0, 1.0]), np.array([3.0, 0.5, 7.0]), np.array([0.0, 2.0, 0.0])):
        unit = col / float(np.dot(lum_w, col))
        for ratio, factor in ((1.0, 2.0 ** -0.3), (4.0, 0.2)):
            c = (unit * ratio * 2.0).astype(np.float32)
            got = oracle.light_clamp(rs, c, 2)
            assert np.allclose(got, c * factor, rtol=2e-6), (col, ratio, got / c)
        c = (unit * 0.37).astype(np.float32)
        assert np.array_equal(oracle.light_clamp(rs, c, 1), c)         # camera depth
        # the factor falls as the luminance grows, and never raises a channel
        fs = [float(oracle.light_clamp(rs, (unit * y).astype(np.float32), 3)[np.argmax(col)] /
                    np.float32(unit * y)[np.argmax(col)]) for y in (0.01, 0.5, 1.0, 3.0, 9.0, 40.0)]
        assert all(1.0 >= a >= b > 0.0 for a, b in zip(fs, fs[1:])), fs
    c = np.array([1e-12, 0.0, 0.0], np.float32)
    assert np.array_equal(oracle.light_clamp(rs, c, 4), c)
    rs.indirect_limit = 0.0
    assert np.all(oracle.light_clamp(rs, np.array([1.0, 2.0, 3.0], np.float32), 2) == 0.0)
    rs.enable_indirect_clamp = 0
    c = np.array([100.0, 2.0, 3.0], np.float32)
    assert np.array_equal(oracle.light_clamp(rs, c, 5), c)
