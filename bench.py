#!/usr/bin/env python3
"""Throughput of the MI355X wavefront path tracer on the procedural interior.

Workload (one step = one progressive frame): 1920x1080, 64 samples per pixel,
path depth 8, Halton sampler, ~260k triangles in 40 instances with random
affine transforms, four 2048^2 RGBA8 value-noise textures, two quad lights,
homogeneous fog (sigma_t 0.05, albedo 0.8, g 0.3), FP32.  Scene arrays are
resident in HBM before timing starts.

Prints one JSON line: value = Mrays/s (camera + extension + shadow +
transmittance segments, counted on the device), plus 1080p samples/s, the
roofline of the dominant kernel (bytes it must move per launch / its HIP-event
time, against 8 TB/s) and the CPU oracle's throughput on a bounded sample of
the same frame.
"""
import argparse
import json
import re
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s & 1080p samples/s vs plain; adjoint/forward time at fixed RMSE"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per item of the traversal kernels.  extend, per
# closest-hit segment: queue id 4 + info 4 + ray 32 + last hit 8 + medium
# header 4 read; hit record 32 + nearby entry 8 + queue
# id 4 written.  Each visited wide node costs 128 B and each tested triangle
# 48 B; the kernel counts both on the device in every frame.
EXTEND_STATE_BYTES = 4 + 4 + 32 + 8 + 4 + 32 + 8 + 4
NODE_BYTES = 128
TRI_BYTES = 48
# shadow, per path with shadow rays: queue id 4 + flags 4 + vertex 16 + L 16
# read/write 32, per shadow ray o/d/f/li 64; plus visited nodes and triangles.
SHADOW_PATH_BYTES = 4 + 4 + 16 + 32
SHADOW_RAY_BYTES = 64


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--sampler", choices=["halton", "pcg"], default="halton")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-scene batch timing")
    ap.add_argument("--no-traversal", action="store_true", help="skip the 2 M-triangle traversal timing")
    ap.add_argument("--capacity", type=int, default=0, help="paths in flight per chunk (0: library default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before rendering, copy MIB MiB device-to-device once (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def adjoint_timing(ctx, frames=2):
    """The differentiable workload as configured: Cornell box at 512x512,
    16 spp, depth 6 with a 256^2 FP32 RGB albedo texture on the floor, 8
    metal spheres of their own GGX roughness and a 64^3 voxel-grid medium
    (sigma_max 2.0).  Forward: ymrt_render_frame; adjoint: ymrt_render_adjoint
    with every parameter a target (the forward frame plus the path replay),
    both timed with HIP events on the render stream over `frames` frames
    after a warm-up frame each."""
    import numpy as np
    import torch

    import ymrt_amd
    from ymrt_amd import abi, scenes
    sa, tg = scenes.cornell_inverse(512, 512, scenes.inverse_params())
    sc = ymrt_amd.Scene(ctx, sa)
    rs = abi.render_settings(512, 512, 16, 6, volume=1)
    adj = torch.from_numpy(np.random.default_rng(5).uniform(0.5, 1.5, (512, 512, 3)).astype(np.float32)).cuda()
    targets = dict(albedo=tg["albedo"], roughness=tg["roughness"], texture=tg["texture"], volume=tg["volume"])
    sc.render_frame(rs, 1)
    sc.render_adjoint(rs, adj, 1, **targets)
    torch.cuda.synchronize()
    t = {}
    stats = None
    for name in ("forward", "adjoint"):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for f in range(frames):
            if name == "forward":
                sc.render_frame(rs, 2 + f)
            else:
                _, _, stats = sc.render_adjoint(rs, adj, 2 + f, **targets)
        b.record()
        torch.cuda.synchronize()
        t[name] = a.elapsed_time(b) / frames
    sc.close()
    return {"workload": "Cornell box 512x512 16spp depth 6 with a 64^3 grid medium (sigma_max 2); gradients of a "
                        "256^2 RGB float albedo texture, 8 GGX roughnesses, 64^3 densities and 3 Lambert albedos",
            "forward_ms": round(t["forward"], 3), "adjoint_ms": round(t["adjoint"], 3),
            "adjoint_over_forward": round(t["adjoint"] / t["forward"], 3),
            "paths": int(stats.paths) if stats is not None else None}


def batch_timing(ctx, count=64, seed=11):
    """The multi-scene workload as configured: 64 scenes (Cornell box with a
    displaced floor grid, triangle counts log-uniform in [1k, 500k]),
    resolution uniform in {256^2, 512^2, 1024^2}, spp uniform in {4, 16, 64},
    depth 6.  Plain: ymrt_render_frame one frame after another, each timed
    with HIP events after a warm-up frame of its own scene, its path state
    released afterwards (64 scenes' buffers would not fit at once).
    Batched: one ymrt_render_batch call after a warm-up call.  Every batched
    image is compared with the plain one bit for bit (equal RMSE)."""
    import torch

    import ymrt_amd
    from ymrt_amd import abi, scenes
    work = scenes.batch_workload(count, seed)
    jobs, tris = [], 0
    t0 = time.time()
    for i, (t, res, spp, depth) in enumerate(work):
        sc = ymrt_amd.Scene(ctx, scenes.batch_scene(t, res, res, i))
        tris += sc.info().triangle_count
        jobs.append((sc, abi.render_settings(res, res, spp, depth), 2))
    gen_s = time.time() - t0
    plain_ms, plain_rays, plain_imgs, plain_each = 0.0, 0, [], []
    for sc, rs, frame in jobs:
        sc.render_frame(rs, 1)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        img, st = sc.render_frame(rs, frame)
        b.record()
        torch.cuda.synchronize()
        plain_ms += a.elapsed_time(b)
        plain_each.append(a.elapsed_time(b))
        plain_rays += st.extension_rays + st.shadow_rays + st.transmittance_rays
        plain_imgs.append(img)
        sc.release_buffers()
    ctx.render_batch(jobs)  # warm-up: the context's round buffers
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    out = ctx.render_batch(jobs)
    b.record()
    torch.cuda.synchronize()
    batch_ms = a.elapsed_time(b)
    rays = sum(st.extension_rays + st.shadow_rays + st.transmittance_rays for _, st in out)
    same = all(torch.equal(img, ref) for (img, _), ref in zip(out, plain_imgs))
    samples = sum(rs.width * rs.height * rs.spp for _, rs, _ in jobs)
    # the frames of at most 512^2 x 16 samples on their own: the share of the
    # workload where one frame cannot fill the GPU
    small = [j for j, (_, rs, _) in enumerate(jobs) if rs.width * rs.height * rs.spp <= 512 * 512 * 16]
    sub = [jobs[j] for j in small]
    ctx.render_batch(sub)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    ctx.render_batch(sub)
    b.record()
    torch.cuda.synchronize()
    small_batch_ms = a.elapsed_time(b)
    small_plain_ms = sum(plain_each[j] for j in small)
    for sc, _, _ in jobs:
        sc.close()
    return {"workload": "%d scenes, %d triangles in all (1k-500k each, log-uniform), 256^2/512^2/1024^2, 4/16/64 "
                        "spp, depth 6, %d samples" % (count, tris, samples),
            "plain_ms": round(plain_ms, 2), "batch_ms": round(batch_ms, 2),
            "speedup": round(plain_ms / batch_ms, 3),
            "plain_mrays_per_s": round(plain_rays / (plain_ms / 1e3) / 1e6, 1),
            "batch_mrays_per_s": round(rays / (batch_ms / 1e3) / 1e6, 1),
            "batch_samples_per_s": round(samples / (batch_ms / 1e3), 1),
            "images_bit_identical": bool(same), "scene_gen_s": round(gen_s, 1),
            "small_frames": {"frames": len(small), "plain_ms": round(small_plain_ms, 2),
                             "batch_ms": round(small_batch_ms, 2),
                             "speedup": round(small_plain_ms / small_batch_ms, 3),
                             "share_of_plain_time": round(small_plain_ms / plain_ms, 3)}}


def traversal_timing(ctx, reps=3):
    """The traversal workload: a 2 M-triangle heightfield (1001^2 vertices)
    and 4 M rays (origins uniform in its box, directions uniform on the
    sphere; any-hit segments to uniform points), through ymrt_trace_closest
    and ymrt_trace_any, timed with HIP events on the stream.  Algorithmic
    bytes per closest-hit ray: 32 B ray + 8 B hit + 128 B per visited wide
    node + 48 B per tested triangle (the counted variant gives the visits)."""
    import torch

    import ymrt_amd
    from ymrt_amd import scenes
    sa = scenes.heightfield(1001)
    sc = ymrt_amd.Scene(ctx, sa)
    closest, shadow = scenes.stress_rays(sa, 1 << 22)
    rc = ymrt_amd.Scene.rays_to_device(closest)
    rsh = ymrt_amd.Scene.rays_to_device(shadow)
    n = rc.shape[0]
    _, visits = sc.trace_closest(rc, counted=True)
    v = visits.double().sum(0).cpu().numpy()
    sc.trace_closest(rc)
    sc.trace_any(rsh)
    torch.cuda.synchronize()
    t = {}
    for name, fn in (("closest", lambda: sc.trace_closest(rc)), ("any", lambda: sc.trace_any(rsh))):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        t[name] = a.elapsed_time(b) / reps
    nodes, tris = v[0] / n, v[1] / n
    per_ray = 40.0 + nodes * 128.0 + tris * 48.0
    achieved = per_ray * n / (t["closest"] / 1e3) / 1e9
    return {"workload": "heightfield %d triangles, %d rays" % (sc.info().triangle_count, n),
            "closest_ms": round(t["closest"], 3), "any_ms": round(t["any"], 3),
            "closest_mrays_per_s": round(n / t["closest"] / 1e3, 1), "any_mrays_per_s": round(n / t["any"] / 1e3, 1),
            "nodes_per_ray": round(nodes, 2), "tris_per_ray": round(tris, 2),
            "closest_gb_per_s_algorithmic": round(achieved, 1), "frac_of_8tb": round(achieved / 8000.0, 4),
            "compulsory_frac": round(40.0 * n / (t["closest"] / 1e3) / 8e12, 4)}


def cpu_baseline(sa, rs_full, budget_s):
    """The oracle (the CPU restatement) rendering the same frame on the host's
    cores.  A first full frame at 1 spp (untimed for the result) sizes the
    timed sample: the whole frame at the spp that takes about budget_s."""
    import oracle
    from ymrt_amd import abi

    threads = max(1, min(16, os.cpu_count() or 1))
    sc = oracle.Scene(sa)
    w, h = rs_full.width, rs_full.height
    t0 = time.time()
    sc.render(abi.render_settings(w, h, 1, rs_full.ray_depth, volume=1), 2, threads=threads)
    frame_s = max(time.time() - t0, 1e-3)
    spp = int(max(1, min(rs_full.spp, round(budget_s / frame_s))))
    rs = abi.render_settings(w, h, spp, rs_full.ray_depth, volume=1)
    t0 = time.time()
    sc.render(rs, 1, threads=threads)
    dt = time.time() - t0
    c = sc.last_counts
    rays = int(c[1] + c[2] + c[3])
    samples = w * h * spp
    sc.close()
    return {
        "value": round(rays / dt / 1e6, 3),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "samples_per_s": round(samples / dt, 1),
        "sample": "oracle, the full %dx%d frame at %d spp (%d samples, %d rays), %.1f s on %d threads"
                  % (w, h, spp, samples, rays, dt, threads),
    }


def cornell_cpu_workload(ctx, budget_s):
    """The CPU workload as configured: Cornell box 256x256, 16 spp, depth 8,
    Halton (permutation seed 0), RR off.  The oracle builds its BVH on the
    host (timed) and renders on the host's cores; when the frame would take
    more than budget_s it renders the first rows only and the rate is taken
    over those.  The GPU renders the same frame for comparison."""
    import torch

    import oracle
    import ymrt_amd
    from ymrt_amd import abi, scenes
    threads = max(1, min(16, os.cpu_count() or 1))
    sa = scenes.cornell_box(256, 256, halton_seed=0)
    rs = abi.render_settings(256, 256, 16, 8, sampler=abi.SAMPLER_HALTON)
    t0 = time.time()
    osc = oracle.Scene(sa)
    build_ms = (time.time() - t0) * 1e3
    t0 = time.time()
    osc.render(rs, 1, rows=(0, 16), threads=threads)
    probe = time.time() - t0
    rows = 256 if probe * 16 <= budget_s else max(16, int(16 * budget_s / max(probe, 1e-3)) // 16 * 16)
    t0 = time.time()
    osc.render(rs, 1, rows=(0, rows), threads=threads)
    dt = time.time() - t0
    c = osc.last_counts
    cpu_rays = int(c[1] + c[2] + c[3])
    osc.close()
    sc = ymrt_amd.Scene(ctx, sa)
    sc.render_frame(rs, 1)
    torch.cuda.synchronize()
    _, st = sc.render_frame(rs, 2, timing=True)
    gpu_rays = st.extension_rays + st.shadow_rays + st.transmittance_rays
    return {"workload": "Cornell box 256x256 16spp depth 8, Halton seed 0, %d triangles" % sc.info().triangle_count,
            "cpu_kind": "port (oracle)", "cpu_threads": threads, "cpu_bvh_build_ms": round(build_ms, 1),
            "cpu_rows": rows, "cpu_render_s": round(dt, 3), "cpu_render_s_full_frame": round(dt * 256 / rows, 3),
            "cpu_mrays_per_s": round(cpu_rays / dt / 1e6, 3),
            "gpu_bvh_build_ms": round(sc.info().build_ms, 1), "gpu_frame_ms": round(st.ms_total, 3),
            "gpu_mrays_per_s": round(gpu_rays / (st.ms_total / 1e3) / 1e6, 1)}


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "k_extend_interior_1080p_pmc_traffic.md")


def measured_traffic(kernel, ext_rays, launches):
    """HBM bytes per k_extend launch: the calibrated FETCH_SIZE + WRITE_SIZE
    per extension ray of the committed rocprofv3 counter runs (same scene and
    kernels) times this run's extension rays per launch; None without it."""
    if kernel != "k_extend" or not launches or not os.path.exists(TRAFFIC_PROFILE):
        return None
    m = re.search(r"traffic bytes per extension ray \(k_extend, calibrated fetch \+ write\): ([0-9.]+)",
                  open(TRAFFIC_PROFILE).read())
    return int(float(m.group(1)) * ext_rays / launches) if m else None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the path tracer runs on one GPU; --gpus %d is not supported (the frame would be split "
              "by pixel tiles across GPUs with one reduction of the accumulation buffer)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import ymrt_amd
    from ymrt_amd import abi, scenes

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    t0 = time.time()
    sa = scenes.interior(args.width, args.height)
    gen_s = time.time() - t0
    ctx = ymrt_amd.Context(0)
    sc = ymrt_amd.Scene(ctx, sa)
    info = sc.info()
    rs = abi.render_settings(args.width, args.height, args.spp, args.depth, volume=1,
                             sampler=abi.SAMPLER_HALTON if args.sampler == "halton" else abi.SAMPLER_PCG,
                             path_capacity=args.capacity)
    noise = torch.empty((args.height, args.width, 4), dtype=torch.float32, device="cuda")
    print("scene: %d tris, %d wide nodes, %.1f MB in HBM, build %.0f ms (gen %.1f s); stack %d"
          % (info.triangle_count, info.node_count, info.device_bytes / 1e6, info.build_ms, gen_s, info.max_stack),
          file=sys.stderr, flush=True)

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        print("copy calibration: %d MiB" % args.copy_calibration, file=sys.stderr, flush=True)

    for w in range(args.warmup):
        sc.render_frame(rs, 1 + w, noise)
        torch.cuda.synchronize()
        print("warmup %d done" % w, file=sys.stderr, flush=True)

    # timed steps: events on the render stream around the whole frame
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    rays = 0
    stage = dict(extend=0.0, medium=0.0, shade=0.0, shadow=0.0, transmit=0.0, raygen=0.0, resolve=0.0)
    launches = dict(extend=0, medium=0, shade=0, shadow=0, transmit=0)
    visits = dict(extend_nodes=0, extend_tris=0, shadow_nodes=0, shadow_tris=0, transmit_nodes=0, transmit_tris=0)
    ext_rays = sh_rays = tr_rays = 0
    t_wall = time.time()
    start.record()
    for k in range(args.steps):
        _, st = sc.render_frame(rs, 1 + args.warmup + k, noise, timing=True)
        rays += st.extension_rays + st.shadow_rays + st.transmittance_rays
        ext_rays += st.extension_rays
        sh_rays += st.shadow_rays
        tr_rays += st.transmittance_rays
        for key in stage:
            stage[key] += getattr(st, "ms_" + key)
        for key in launches:
            launches[key] += getattr(st, "launches_" + key)
        for key in visits:
            visits[key] += getattr(st, key)
        print("step %d: %.1f ms" % (k, st.ms_total), file=sys.stderr, flush=True)
    stop.record()
    torch.cuda.synchronize()
    wall = time.time() - t_wall
    ms = start.elapsed_time(stop)
    ms_per_step = ms / args.steps
    mrays = rays / (ms / 1e3) / 1e6
    samples_per_s = args.width * args.height * args.spp * args.steps / (ms / 1e3)

    # Roofline of the dominant traversal kernel: algorithmic bytes (state +
    # visited nodes and triangles, counted on the device) over its HIP-event
    # time on the render stream.
    ext_bytes = (EXTEND_STATE_BYTES * ext_rays + NODE_BYTES * visits["extend_nodes"]
                 + TRI_BYTES * visits["extend_tris"])
    sh_bytes = (SHADOW_RAY_BYTES * sh_rays + SHADOW_PATH_BYTES * sh_rays // 2
                + NODE_BYTES * visits["shadow_nodes"] + TRI_BYTES * visits["shadow_tris"])
    if stage["shadow"] > stage["extend"]:
        dom, dom_bytes, dom_ms, dom_launch = "k_shadow", sh_bytes, stage["shadow"], launches["shadow"]
    else:
        dom, dom_bytes, dom_ms, dom_launch = "k_extend", ext_bytes, stage["extend"], launches["extend"]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    traffic = measured_traffic(dom, ext_rays, dom_launch)
    physical = traffic / (dom_ms / max(dom_launch, 1) / 1e3) / 1e9 if traffic else None
    per_step = {k: v / args.steps for k, v in stage.items()}
    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic procedural scene (no assets)",
        "config": {"workload": "interior 1920x1080 64spp depth 8, 260k tris / 40 instances, 4x2048^2 RGBA8 "
                               "textures, 2 quad lights, homogeneous fog",
                   "width": args.width, "height": args.height, "spp": args.spp, "depth": args.depth,
                   "triangles": int(info.triangle_count), "instances": len(sa.instances)},
        "samples_per_s": round(samples_per_s, 1),
        "rays_per_step": int(rays / args.steps),
        "stage_ms_per_step": {k: round(v, 2) for k, v in per_step.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "achieved_kind": "algorithmic bytes: path state + 128 B per visited wide node + 48 B per "
                                      "tested triangle, visits counted on the device in this run",
                     "traffic": traffic,
                     "traffic_source": "profiles/k_extend_interior_1080p_pmc_traffic.md: calibrated rocprofv3 "
                                       "FETCH_SIZE + WRITE_SIZE per extension ray (8 spp run of this scene), "
                                       "times this run's extension rays per launch" if traffic else None,
                     "physical_frac": round(physical / HBM_PEAK_GBS, 4) if physical else None,
                     "limiter": "not HBM: nodes and triangles come from L2/MALL (physical bytes are a sixth of "
                                "the algorithmic ones); the kernel is bound by VALU issue under divergence and "
                                "dependent-load latency (SQ counters in profiles/traversal_shading_sq_counters.md)",
                     "bytes_per_launch": int(dom_bytes / max(dom_launch, 1)),
                     "avg_launch_ms": round(dom_ms / max(dom_launch, 1), 4),
                     "launches": dom_launch},
        "mrays_per_s_without_transmittance": round((rays - tr_rays) / (ms / 1e3) / 1e6, 2),
        "traversal": {"extend_rays": ext_rays // args.steps,
                      "nodes_per_extend_ray": round(visits["extend_nodes"] / max(ext_rays, 1), 2),
                      "tris_per_extend_ray": round(visits["extend_tris"] / max(ext_rays, 1), 2),
                      "shadow_rays": sh_rays // args.steps, "transmittance_rays": tr_rays // args.steps,
                      "nodes_per_shadow_ray": round(visits["shadow_nodes"] / max(sh_rays, 1), 2),
                      "tris_per_shadow_ray": round(visits["shadow_tris"] / max(sh_rays, 1), 2),
                      "nodes_per_transmittance_ray": round(visits["transmit_nodes"] / max(tr_rays, 1), 2)},
        "wall_s": round(wall, 2),
    }
    if not args.no_adjoint:
        out["adjoint"] = adjoint_timing(ctx)
    if not args.no_batch:
        out["batch"] = batch_timing(ctx)
    if not args.no_traversal:
        out["traversal_c3"] = traversal_timing(ctx)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sa, rs, args.cpu_seconds)
        out["cornell_cpu_workload"] = cornell_cpu_workload(ctx, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
