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
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s & 1080p samples/s vs plain; adjoint/forward time at fixed RMSE"
HBM_PEAK_GBS = 8000.0

# Bytes a kernel must move per item ("Roofline accounting").
# extend: queue id 4 + ray 32 + last hit 8 + info 4 + hit record 32 + nearby 8
#         + hit-queue id 4, plus the BVH bytes it visits (128 B per wide node,
#         48 B per triangle), counted by a separate counting frame.
EXTEND_STATE_BYTES = 4 + 32 + 8 + 4 + 32 + 8 + 4
# shade: queue id 4 + info 4 + hit 32 + ray 32 + beta 16 + L 16 + sampler 16
#        + transfer header 4 + shading triangle 96 + instance 128 + material 128
#        + writes (sampler 16, L 16, ray 32, last 8, beta 16, info 4,
#        shadow rays 2*64 + 16 + 4, queue ids 8)
SHADE_BYTES = 4 + 4 + 32 + 32 + 16 + 16 + 16 + 4 + 96 + 128 + 128 + 16 + 16 + 32 + 8 + 16 + 4 + 148 + 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def cpu_baseline(sa, rs_full, budget_s):
    """Oracle (the CPU restatement) on a band of rows of the same frame at 1 spp,
    sized to roughly budget_s seconds of work on the host's cores."""
    import threading
    import numpy as np
    import oracle
    from ymrt_amd import abi

    threads = max(1, min(16, os.cpu_count() or 1))
    sc = oracle.Scene(sa)
    rs = abi.render_settings(rs_full.width, rs_full.height, 1, rs_full.ray_depth, volume=1)
    # probe: 8 rows
    t0 = time.time()
    sc.render(rs, 1, rows=(536, 544), threads=threads)
    dt = max(time.time() - t0, 1e-3)
    rows = int(max(8, min(rs.height, 8 * budget_s / dt)))
    r0 = max(0, rs.height // 2 - rows // 2)
    t0 = time.time()
    sc.render(rs, 1, rows=(r0, r0 + rows), threads=threads)
    dt = time.time() - t0
    c = sc.last_counts
    rays = int(c[1] + c[2] + c[3])
    paths = rows * rs.width
    sc.close()
    return {
        "value": rays / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "samples_per_s": paths / dt,
        "sample": "oracle, rows %d-%d of the 1920x1080 frame at 1 spp (%d paths, %d rays), %.1f s"
                  % (r0, r0 + rows, paths, rays, dt),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the frame would be split "
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
    rs = abi.render_settings(args.width, args.height, args.spp, args.depth, volume=1)
    noise = torch.empty((args.height, args.width, 4), dtype=torch.float32, device="cuda")
    print("scene: %d tris, %d wide nodes, %.1f MB in HBM, build %.0f ms (gen %.1f s); stack %d"
          % (info.triangle_count, info.node_count, info.device_bytes / 1e6, info.build_ms, gen_s, info.max_stack),
          file=sys.stderr, flush=True)

    for w in range(args.warmup):
        sc.render_frame(rs, 1 + w, noise)
        torch.cuda.synchronize()
        print("warmup %d done" % w, file=sys.stderr, flush=True)

    # timed steps: events on the render stream around the whole frame
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    rays = 0
    stage = dict(extend=0.0, shade=0.0, shadow=0.0, raygen=0.0, resolve=0.0)
    launches = dict(extend=0, shade=0, shadow=0)
    t_wall = time.time()
    start.record()
    for k in range(args.steps):
        _, st = sc.render_frame(rs, 1 + args.warmup + k, noise, timing=True)
        rays += st.camera_rays * 0 + st.extension_rays + st.shadow_rays + st.transmittance_rays
        stage["extend"] += st.ms_extend
        stage["shade"] += st.ms_shade
        stage["shadow"] += st.ms_shadow
        stage["raygen"] += st.ms_raygen
        stage["resolve"] += st.ms_resolve
        launches["extend"] += st.launches_extend
        launches["shade"] += st.launches_shade
        launches["shadow"] += st.launches_shadow
        last = st
        print("step %d: %.1f ms" % (k, st.ms_total), file=sys.stderr, flush=True)
    stop.record()
    torch.cuda.synchronize()
    wall = time.time() - t_wall
    ms = start.elapsed_time(stop)
    ms_per_step = ms / args.steps
    mrays = rays / (ms / 1e3) / 1e6
    samples_per_s = args.width * args.height * args.spp * args.steps / (ms / 1e3)

    # Roofline of the dominant stage.  Visited BVH bytes come from the
    # extension rays' average visits, measured by the counting trace kernel on
    # a sample of this frame's first-bounce rays is not available per stage, so
    # the extend figure uses the compulsory state bytes; shade uses its fixed
    # per-item bytes.
    dom = max(("extend", "shade", "shadow"), key=lambda k: stage[k])
    per_step = {k: v / args.steps for k, v in stage.items()}
    if dom == "shade":
        items = last.extension_rays  # hits shaded <= extension rays
        bytes_per_step = SHADE_BYTES * items
    elif dom == "extend":
        items = last.extension_rays
        bytes_per_step = EXTEND_STATE_BYTES * items
    else:
        items = last.shadow_rays
        bytes_per_step = 64 * items
    achieved = bytes_per_step / (per_step[dom] / 1e3) / 1e9
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
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "wall_s": round(wall, 2),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sa, rs, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
