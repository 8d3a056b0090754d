#!/usr/bin/env python3
"""Headline benchmark: Mrays/s of the static 1M-triangle 1080p Whitted frame.

Workload (one GPU): 200 UV spheres x 5,000 triangles (1,000,000 triangles),
1920x1080 pinhole primary rays, 2 point lights with shadow rays, Whitted
depth 3 (reflection + refraction, IOR 1.5 on 30% of the spheres), full LBVH
rebuild every frame.  A step is one frame: LBVH build -> primary rays ->
trace -> shadow -> shade -> 3 x (secondary -> trace -> shadow -> shade) ->
RGBA8 resolve, all on the GPU with inputs resident in HBM.

value = (primary + secondary rays) per frame x steps / timed seconds
(the reference's ray count, src/rt/renderer.cpp:137, 184-185).

Also reported: LBVH build ms, the plain one-launch-per-stage path (with the
reference's per-bounce count read-back), per-kernel times, the roofline of
the radix-sort pass (the build's HBM-bound kernel), and the CPU oracle
timed on this host on the same frame (one core).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tesina-rt_amd"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sd, cam_struct):
    """The oracle (sequential C++, one core) on the same frame: LBVH build +
    render.  Returns (Mrays/s, seconds, rays)."""
    sys.path.insert(0, str(ROOT))
    import oracle as O
    t0 = time.perf_counter()
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    tb = time.perf_counter() - t0
    acc, st, rays = O.render(b, sd.nrm, sd.uv, sd.materials, O.lights_from(sd.lights), O.camera_from(cam_struct),
                             sd.width, sd.height, sd.max_bounces)
    dt = time.perf_counter() - t0
    return rays / dt / 1e6, dt, rays, tb


def traffic_from_profiles():
    """HBM bytes per sort-pass launch measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), if the
    summary has been recorded under profiles/."""
    f = ROOT / "profiles" / "sort_pass_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--profile-frames", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(image-space sharding across GPUs is described in DESIGN.md, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import rtx
    from rtx import scenes

    t = time.perf_counter()
    sd = scenes.config_c2()
    log(f"scene C2: {sd.T} triangles, generated in {time.perf_counter() - t:.2f}s")
    ctx = rtx.Context(0)
    s = scenes.load(rtx.Scene(ctx), sd)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces, depth=30, min_leaf=1)

    # warm-up (also first-touch of every buffer), then capture the frame
    # (LBVH rebuild + render) into a hipGraph
    for _ in range(max(1, args.warmup)):
        R.frame_async(s, cam, rebuild=True)
    ctx.sync()
    st = R.last_stats()
    rays_per_frame = st.primary_rays + st.secondary_rays
    shadow_per_frame = st.shadow_rays
    log(f"rays/frame: primary {st.primary_rays} secondary {st.secondary_rays} shadow {st.shadow_rays}")
    R.capture(s, cam, rebuild=True)
    for _ in range(args.warmup):
        R.replay()
    ctx.sync()

    # timed region: K graph replays, no host work inside
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.replay()
    ctx.sync()
    dt = time.perf_counter() - t0
    ms_per_step = dt / args.steps * 1e3
    value = rays_per_frame * args.steps / dt / 1e6
    st2 = R.last_stats()
    assert st2.primary_rays + st2.secondary_rays == rays_per_frame, "replayed frame differs in ray count"
    log(f"timed {args.steps} graph frames: {ms_per_step:.3f} ms/frame, {value:.1f} Mrays/s")

    # the same frame launched stage by stage (no graph)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.frame_async(s, cam, rebuild=True)
    ctx.sync()
    async_ms = (time.perf_counter() - t0) / args.steps * 1e3

    # plain one-launch-per-stage path with the reference's count read-backs
    plain = []
    for _ in range(3):
        fs = R.frame(s, cam, rebuild=True)
        plain.append(fs.frame_ms)
    plain_ms = float(np.median(plain))

    # per-kernel times with HIP events on the stream the kernels run on
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(args.profile_frames):
        R.frame_async(s, cam, rebuild=True)
    rep = ctx.profile_report()
    ctx.profile(False)
    nf = args.profile_frames
    kernels = {k: {"launches_per_frame": c / nf, "ms_per_frame": ms / nf, "avg_launch_ms": ms / c}
               for k, (c, ms) in sorted(rep.items())}
    build_ms = kernels.get("LBVH_BUILD", {}).get("avg_launch_ms")
    sp = kernels.get("lbvh_sort_pass", {})
    sort_avg_s = sp.get("avg_launch_ms", 0.0) / 1e3
    T = sd.T
    sort_bytes = 16.0 * T  # per pass: key+value read (8 B) + key+value write (8 B)
    achieved = sort_bytes / sort_avg_s / 1e9 if sort_avg_s > 0 else None
    trace_ms = kernels.get("trace", {}).get("ms_per_frame", 0.0)

    out = {
        "metric": "Mrays/s @1M-tri 1080p Whitted",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, seed 0x5EED0002)",
        "config": {"workload": "C2 static GPU frame: 200 UV spheres x 5000 tris = 1M FP32 triangles, 1920x1080 "
                               "pinhole primary rays, 2 point lights + shadow rays, Whitted depth 3 "
                               "(reflect+refract, IOR 1.5 on 30%), full LBVH rebuild each frame",
                   "triangles": T, "width": sd.width, "height": sd.height, "max_bounces": sd.max_bounces,
                   "bvh_depth": 30, "min_leaf": 1},
        "rays_per_frame": int(rays_per_frame),
        "shadow_rays_per_frame": int(shadow_per_frame),
        "mrays_incl_shadow": round((rays_per_frame + shadow_per_frame) * args.steps / dt / 1e6, 3),
        "lbvh_build_ms": round(build_ms, 4) if build_ms else None,
        "frame_mode": "hipGraph replay of LBVH rebuild + render",
        "stage_launch_ms_per_frame": round(async_ms, 4),
        "plain_path_ms_per_frame": round(plain_ms, 4),
        "plain_path_mrays": round(rays_per_frame / plain_ms / 1e3, 3),
        "roofline": {
            "kernel": "lbvh_sort_pass (onesweep 8-bit LSD pass over (code, id) pairs)",
            "bound": "hbm",
            "achieved": round(achieved, 2) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5) if achieved else None,
            "traffic": traffic_from_profiles(),
            "algorithmic_bytes_per_launch": sort_bytes,
            "avg_launch_ms": round(sort_avg_s * 1e3, 5),
        },
        "traversal": {
            "trace_ms_per_frame": round(trace_ms, 4),
            "shadow_ms_per_frame": round(kernels.get("trace_shadow", {}).get("ms_per_frame", 0.0), 4),
            "trace_mrays": round(rays_per_frame / trace_ms / 1e3, 2) if trace_ms else None,
        },
        "kernels": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in kernels.items()},
    }
    if not args.no_cpu:
        log("CPU oracle baseline (one core, full frame) ...")
        mr, secs, rays, tb = cpu_baseline(sd, cam)
        out["cpu_baseline"] = {"value": round(mr, 4), "unit": "Mrays/s", "cores": 1, "kind": "port",
                               "sample": f"whole C2 frame on one core: LBVH build ({tb:.2f}s) + render of "
                                         f"{rays} rays, {secs:.1f}s total",
                               "host_cpu": _cpu_model(), "nproc": os.cpu_count()}
    else:
        out["cpu_baseline"] = None
    R.destroy()
    s.destroy()
    ctx.close()
    print(json.dumps(out))


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
