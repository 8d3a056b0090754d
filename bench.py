#!/usr/bin/env python3
"""Headline benchmark: Mrays/s of the static 1M-triangle 1080p Whitted frame.

Workload (one GPU): 200 UV spheres x 5,000 triangles (1,000,000 triangles),
1920x1080 pinhole primary rays, 2 point lights with shadow rays, Whitted
depth 3 (reflection + refraction, IOR 1.5 on 30% of the spheres), full LBVH
rebuild every frame.  A step is one frame: LBVH build -> primary rays ->
trace -> shadow -> shade -> 3 x (secondary -> trace -> shadow -> shade) ->
RGBA8 resolve, all on the GPU with inputs resident in HBM.  Frames are
pipelined as the reference's RendererT does (the next frame's build runs on a
second stream while this frame renders); the sequential frame is timed too.

value = (primary + secondary rays) per frame x steps / timed seconds
(the reference's ray count, src/rt/renderer.cpp:137, 184-185).

Also reported: LBVH build ms, the plain one-launch-per-stage path (with the
reference's per-bounce count read-back), per-kernel times (HIP events on the
stream each kernel runs on), the roofline of the dominant kernel (closest-hit
traversal) and of the build's HBM-bound radix-sort pass, with HBM traffic
per launch when recorded under profiles/, and the CPU oracle timed on this
host on the same frame (one core, median of 3 frames).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
    python bench.py --copy-calibration 256   # device copy for counter scaling
    python bench.py --c3-only --c3-frames 20 # the dynamic-scene leg alone
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
# Hardware queues: HIP's default (4 per process, or GPU_MAX_HW_QUEUES from the
# environment, recorded in the JSON).  16 queues measured the same on the
# frame pipeline and the server, and slower on pipelined graph replays.
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


def fp64_hit_error(ctx, s, sd, cam):
    """Hit distances of the frame's rays -- primary and every bounce, traced
    by the library on the built scene -- against the same hits recomputed in
    float64: Moller-Trumbore in float64 on the float32 ray (origin,
    direction) and the hit triangle's float32 vertices.  The library's
    distances are the float32 reference's bit for bit; this is their
    rounding error.  Returns max absolute and relative error and counts."""
    import numpy as np
    import rtx
    P = sd.width * sd.height
    pos = s.vertices()[0].reshape(-1, 3).astype(np.float64)
    idx = s.indices().astype(np.int64)
    worst_abs = worst_rel = 0.0
    hits_total = 0
    per_bounce = []
    bundles = []
    rays, cap = rtx.Rays(ctx, P), P
    rays.primary(cam, sd.width, sd.height)
    n = P
    for b in range(sd.max_bounces + 1):
        hb = rtx.Hits(ctx, cap)
        bundles += [rays, hb]
        rtx.trace(s, rays, hb, n)
        h = hb.read(n)
        o, d, _, _ = rays.read(n)
        m = h["id"] >= 0
        tri = idx[h["id"][m]]
        v0, v1, v2 = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
        oo, dd = o[m, :3].astype(np.float64), d[m, :3].astype(np.float64)
        e1, e2 = v1 - v0, v2 - v0
        hv = np.cross(dd, e2)
        a = np.einsum("ij,ij->i", e1, hv)
        q = np.cross(oo - v0, e1)
        t64 = np.einsum("ij,ij->i", e2, q) / a
        t32 = h["t"][m].astype(np.float64)
        err = np.abs(t32 - t64)
        rel = err / np.maximum(np.abs(t64), 1e-30)
        if err.size:
            worst_abs = max(worst_abs, float(err.max()))
            worst_rel = max(worst_rel, float(rel.max()))
        hits_total += int(m.sum())
        per_bounce.append({"rays": int(n), "hits": int(m.sum()),
                           "max_abs": float(err.max()) if err.size else 0.0,
                           "max_rel": float(rel.max()) if err.size else 0.0})
        if b == sd.max_bounces:
            break
        cap = (P << (b + 1)) + 2  # the renderer's bundle of this bounce
        nxt = rtx.Rays(ctx, cap)
        n = rtx.raygen_secondary(s, rays, hb, nxt, n)
        rays = nxt
        if n == 0:
            break
    for x in bundles + [rays]:
        x.destroy()
    return {"max_abs_error": worst_abs, "max_rel_error": worst_rel, "hits_checked": hits_total,
            "per_bounce": per_bounce,
            "what": "float32 hit distance of every hit of the frame (primary + bounces) minus the float64 "
                    "Moller-Trumbore distance on the same float32 ray and vertices"}


def cpu_visit_counts(sd, cam_struct, stride: int = 16):
    """Node visits and triangle tests per primary ray, counted by the oracle's
    traversal (deterministic) over every `stride`-th primary ray of the frame:
    the fetch counts behind the traversal's effective bytes."""
    sys.path.insert(0, str(ROOT))
    import oracle as O
    b = O.Build(sd.pos, sd.idx, sd.matmap, 30, 1)
    o4, d4, _, _ = O.raygen_primary(O.camera_from(cam_struct), sd.width, sd.height)
    o4, d4 = o4[::stride].copy(), d4[::stride].copy()
    h = O.trace(b, sd.nrm, sd.uv, o4, d4, count_visits=True)
    return float(h["visits"]) / len(o4), float(h["tri_tests"]) / len(o4), len(o4)


def visit_count_table(stride: int = 8):
    """The oracle's traversal work per ray (node visits, triangle tests, hit
    rate) of the C2 frame and of C3's first animated frame, for the primary
    rays and the first bounce's secondary rays (every `stride`-th ray):
    what a ray of each frame costs independent of the hardware."""
    sys.path.insert(0, str(ROOT))
    sys.path.insert(0, str(ROOT / "tesina-rt_amd"))
    import oracle as O
    from rtx import scenes
    out = {}
    for name in ("C2", "C3"):
        sd = scenes.config_c2() if name == "C2" else scenes.config_c3()
        pos = sd.pos
        if name == "C3":  # the animation's frame at t = 0
            pos = sd.pos.copy()
            pos[:, 1] = scenes.displace_host(sd, 0.0)[:, 1]
        b = O.Build(pos, sd.idx, sd.matmap, 30, 1)
        oo, od, op, oc = O.raygen_primary(O.camera_from(scenes.camera_of(sd)), sd.width, sd.height)
        o = O.trace(b, sd.nrm, sd.uv, oo, od)
        so, sdir, _, _, _ = O.raygen_secondary(sd.materials, b.matmap, oo, od, op, oc, o, 4 * len(oo))
        rows = {}
        for what, (a, d) in (("primary", (oo, od)), ("bounce_1", (so, sdir))):
            a, d = a[::stride].copy(), d[::stride].copy()
            h = O.trace(b, sd.nrm, sd.uv, a, d, count_visits=True)
            k = max(1, len(a))
            rows[what] = {"rays": len(oo) if what == "primary" else len(so), "sampled": len(a),
                          "node_visits_per_ray": round(float(h["visits"]) / k, 2),
                          "triangle_tests_per_ray": round(float(h["tri_tests"]) / k, 2),
                          "hit_rate": round(float((h["id"] >= 0).mean()) if len(a) else 0.0, 4)}
        out[name] = {"triangles": sd.T, "nodes": int(b.node_count), **rows}
    return out


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter reads for
    a device copy of known size), as recorded in profiles/hbm_traffic.json;
    None when not recorded."""
    f = ROOT / "profiles" / "hbm_traffic.json"
    if f.exists():
        try:
            return json.loads(f.read_text())["kernels"][kernel]["bytes_per_launch"]
        except Exception:
            return None
    return None


def copy_calibration(mib: int, reps: int = 3):
    """A plain device-to-device copy of `mib` MiB (read mib + write mib), run
    under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": mib << 20, "copy_calibration_bytes_written": mib << 20,
                      "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--profile-frames", type=int, default=3)
    ap.add_argument("--cpu-frames", type=int, default=3, help="oracle frames timed for the CPU baseline (median)")
    ap.add_argument("--c3-frames", type=int, default=300, help="frames of the dynamic-scene leg (0 = skip)")
    ap.add_argument("--c5-rates", type=str, default="50,100,200,500,1000,1100,1200,1300,1400,1500,2000",
                    help="arrival rates of the serving leg, per second ('' = skip)")
    ap.add_argument("--c5-max-requests", type=int, default=3000,
                    help="requests per stream at most (2 s of arrivals otherwise); the plain path takes the "
                         "first 1500 of the first stream")
    ap.add_argument("--c5-seeds", type=int, default=3,
                    help="request streams per rate from 1000/s up, their latencies pooled")
    ap.add_argument("--c4-scenes", type=int, default=64, help="scenes of the batched leg (0 = skip)")
    ap.add_argument("--inst-frames", type=int, default=50, help="frames of the instanced leg (0 = skip)")
    ap.add_argument("--accel-frames", type=int, default=10,
                    help="frames of the LBVH / SAH, stack / stackless leg (0 = skip)")
    ap.add_argument("--image-frames", type=int, default=10,
                    help="frames of the textured-frame leg; also times OBJ ingest (0 = skip)")
    ap.add_argument("--sort-sizes", type=str, default="16777216,67108864",
                    help="element counts of the sort-scaling leg ('' = skip)")
    ap.add_argument("--visit-counts", action="store_true",
                    help="only print the oracle's node visits and triangle tests per ray of the C2 and C3 frames "
                         "(primary rays and the first bounce; CPU only) and exit")
    ap.add_argument("--c3-only", action="store_true",
                    help="only run the dynamic-scene leg (--c3-frames frames) and print its JSON (profiling runs)")
    ap.add_argument("--c4-only", action="store_true",
                    help="only run the batched leg (--c4-scenes scenes) and print its JSON (profiling runs)")
    ap.add_argument("--c5-only", action="store_true",
                    help="only run the serving leg (--c5-rates) and print its JSON (profiling runs)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only run --profile-frames C2 frames with every launch on one stream (event-timed) and "
                         "print the two roofline objects: a run whose kernel trace holds only those launches")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(image-space sharding of a frame across GPUs is not built)", file=sys.stderr)
        sys.exit(2)
    if args.visit_counts:
        print(json.dumps(visit_count_table()), flush=True)
        return
    if args.c3_only or args.c4_only or args.c5_only:
        import rtx
        ctx = rtx.Context(0)
        out = {}
        if args.c3_only:
            out["c3_dynamic"] = bench_c3(ctx, max(1, args.c3_frames))
        if args.c4_only:
            out["c4_batched"] = bench_c4(ctx, max(1, args.c4_scenes))
        if args.c5_only:
            out["c5_serving"] = bench_c5(ctx, [int(x) for x in args.c5_rates.split(",") if x], args.c5_max_requests,
                                     seeds=args.c5_seeds)
        print(json.dumps(out), flush=True)
        return

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

    if args.roofline_only:
        print(json.dumps(roofline_only(ctx, s, R, cam, sd, args.profile_frames, args.warmup)), flush=True)
        return

    # the LBVH build's knobs, picked by back-to-back builds of this scene:
    # radix digit width and pairs per thread of the sort, persistent blocks
    # per CU of the level launches (every setting builds the same tree)
    build_tune = lbvh_autotune(ctx, s)
    log("LBVH build tuning: " + ", ".join(
        f"bits {t['sort_bits']} items {t['sort_items']} grid {t['build_grid']}: {t['ms_per_build']:.3f} ms"
        for t in build_tune["candidates"]) + f"; chosen {build_tune['chosen']}")

    # warm-up (also first-touch of every buffer), then capture the frame
    # (LBVH rebuild + render) into a hipGraph.  Both graph shapes are timed:
    # captured from one stream (linear), or with each bounce's shadow rays and
    # shading on a side stream (forked).  Which is faster depends on the
    # hardware queues: with a queue per stream HIP's graph executor turns every
    # fork and join into a cross-queue wait.
    for _ in range(max(1, args.warmup)):
        R.frame_async(s, cam, rebuild=True)
    ctx.device_sync()
    st = R.last_stats()
    rays_per_frame = st.primary_rays + st.secondary_rays
    shadow_per_frame = st.shadow_rays
    log(f"rays/frame: primary {st.primary_rays} secondary {st.secondary_rays} shadow {st.shadow_rays}")

    def graph_ms(pipelined: bool, side: int):
        ctx.tune("side_stream", side)
        R.set_pipelined(pipelined)
        R.frame_async(s, cam, rebuild=True)
        ctx.device_sync()
        R.capture(s, cam, rebuild=True)
        for _ in range(args.warmup):
            R.replay()
        # timed region: K graph replays, no host work inside
        ctx.device_sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            R.replay()
        ctx.device_sync()
        dt = time.perf_counter() - t0
        stg = R.last_stats()
        assert stg.primary_rays + stg.secondary_rays == rays_per_frame, "replayed frame differs in ray count"
        return dt / args.steps * 1e3

    graph_shapes = {}
    for pipelined in (False, True):
        for side in (0, 1):
            graph_shapes[("pipelined" if pipelined else "sequential", "forked" if side else "linear")] = \
                graph_ms(pipelined, side)
    for k, v in graph_shapes.items():
        log(f"timed {args.steps} {k[0]} graph frames, {k[1]}: {v:.3f} ms/frame, "
            f"{rays_per_frame / v / 1e3:.1f} Mrays/s")
    seq_shape = min(("linear", "forked"), key=lambda k: graph_shapes[("sequential", k)])
    pip_shape = min(("linear", "forked"), key=lambda k: graph_shapes[("pipelined", k)])
    seq_ms = graph_shapes[("sequential", seq_shape)]
    seq_value = rays_per_frame / seq_ms / 1e3
    pip_graph_ms = graph_shapes[("pipelined", pip_shape)]
    R.set_pipelined(True)
    # the same pipelined frames launched stage by stage (no graph)
    ctx.tune("side_stream", 1)
    for _ in range(args.warmup):
        R.frame_async(s, cam, rebuild=True)
    ctx.device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.frame_async(s, cam, rebuild=True)
    ctx.device_sync()
    pip_eager_ms = (time.perf_counter() - t0) / args.steps * 1e3
    log(f"pipelined stage-launched frames: {pip_eager_ms:.3f} ms/frame")
    R.set_pipelined(False)

    # frames in flight (frame pipeline): the next frames' builds overlap the
    # renders of the frames before, which overlap each other.  The traversal
    # knobs are tuned on this mode first (every setting gives the same
    # images): work-group size and cooperative-leaf size, untimed short runs.
    def flight_ms(depth, steps):
        pl = rtx.Pipeline(ctx, s, sd.width, sd.height, sd.max_bounces, depth=depth)
        for _ in range(depth + 2):
            pl.frame(cam, rebuild=True)
        ctx.device_sync()
        t0 = time.perf_counter()
        last = 0
        for _ in range(steps):
            last = pl.frame(cam, rebuild=True)
        pl.wait(last)
        ctx.device_sync()
        dt = time.perf_counter() - t0
        stf = pl.stats(last)
        pl.destroy()
        assert stf.primary_rays + stf.secondary_rays == rays_per_frame, "pipeline frame differs in ray count"
        return dt / steps * 1e3

    tune_runs = {}
    for tb, coop, sl in ((128, 8, 32), (64, 8, 32), (64, 4, 32), (128, 4, 32), (128, 8, 16)):
        ctx.tune("trace_block", tb)
        ctx.tune("coop_leaf", coop)
        ctx.tune("stack_lds", sl)
        tune_runs[(tb, coop, sl)] = flight_ms(3, 10)
    tb, coop, sl = min(tune_runs, key=tune_runs.get)
    ctx.tune("trace_block", tb)
    ctx.tune("coop_leaf", coop)
    ctx.tune("stack_lds", sl)
    log(f"frame pipeline tuning: {', '.join(f'block {k[0]} coop {k[1]} stack {k[2]}: {v:.3f} ms' for k, v in tune_runs.items())}"
        f"; chosen block {tb}, coop {coop}, stack_lds {sl}")
    flight_tune = [{"trace_block": k[0], "coop_leaf": k[1], "stack_lds": k[2], "ms_per_frame": round(v, 4)}
                   for k, v in tune_runs.items()]
    flights = {}
    for depth in (2, 3, 4, 5, 6, 7, 8):
        flights[depth] = flight_ms(depth, args.steps)
        log(f"frame pipeline, {depth} frames in flight: {flights[depth]:.3f} ms/frame, "
            f"{rays_per_frame / flights[depth] / 1e3:.1f} Mrays/s")
    # the headline is the fastest way of running the same frames (each timed
    # frame does one full LBVH rebuild and one full render in every mode)
    best_depth = min(flights, key=flights.get)
    modes = {
        "pipelined_graph": (pip_graph_ms, f"hipGraph replay ({pip_shape} capture) of the pipelined frame: LBVH "
                                          "rebuild of the next frame on a second stream while this frame renders "
                                          "(the reference's RendererT)"),
        "pipelined_stage_launch": (pip_eager_ms, "the pipelined frame (the reference's RendererT) launched stage by "
                                                 "stage, shadows and shading of each bounce on a side stream"),
        "frames_in_flight": (flights[best_depth], f"frame pipeline with {best_depth} frames in flight (stream "
                                                  "launches): the next frames' LBVH rebuilds and the renders of "
                                                  f"consecutive frames overlap; traversal work-group {tb}, "
                                                  f"cooperative leaves from {coop} triangles, {sl} LDS stack "
                                                  "entries"),
    }
    best_mode = min(modes, key=lambda k: modes[k][0])
    ms_per_step = modes[best_mode][0]
    value = rays_per_frame / ms_per_step / 1e3
    frame_mode = modes[best_mode][1] + "; every timed frame does one full rebuild and one full render"

    # the same frame launched stage by stage (no graph)
    ctx.device_sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        R.frame_async(s, cam, rebuild=True)
    ctx.device_sync()
    async_ms = (time.perf_counter() - t0) / args.steps * 1e3

    # plain one-launch-per-stage path with the reference's count read-backs
    plain = []
    for _ in range(3):
        fs = R.frame(s, cam, rebuild=True)
        plain.append(fs.frame_ms)
    plain_ms = float(np.median(plain))

    # per-kernel times with HIP events on the stream the kernels run on, over
    # frames launched on one stream (no kernel of the frame overlaps another)
    ctx.tune("side_stream", 0)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(args.profile_frames):
        R.frame_async(s, cam, rebuild=True)
    rep = ctx.profile_report()
    ctx.profile(False)
    ctx.tune("side_stream", 1)
    nf = args.profile_frames
    kernels = {k: {"launches_per_frame": c / nf, "ms_per_frame": ms / nf, "avg_launch_ms": ms / c}
               for k, (c, ms) in sorted(rep.items())}
    build_ms = kernels.get("LBVH_BUILD", {}).get("avg_launch_ms")
    T = sd.T

    def roofline(key, kernel, algo_bytes, note):
        k = kernels.get(key, {})
        avg_s = k.get("avg_launch_ms", 0.0) / 1e3
        ach = algo_bytes / avg_s / 1e9 if avg_s > 0 else None
        return {"kernel": kernel, "bound": "hbm", "achieved": round(ach, 2) if ach else None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 5) if ach else None,
                "traffic": traffic_from_profiles(kernel.split()[0]), "algorithmic_bytes_per_launch": algo_bytes,
                "avg_launch_ms": round(avg_s * 1e3, 5), "launches_per_frame": k.get("launches_per_frame"),
                "note": note}

    # dominant kernel: closest-hit traversal (4 launches per frame: primary
    # + 3 bounces); algorithmic bytes = ray in (32 B) + hit out (16 B) per ray
    trace_launches = kernels.get("trace", {}).get("launches_per_frame", 1.0) or 1.0
    rl_trace = roofline("trace", "k_trace (closest-hit LBVH traversal)", 48.0 * rays_per_frame / trace_launches,
                        "48 B/ray (32 B ray in + 16 B hit out), rays of a frame / launches per frame; node and "
                        "triangle fetches are cache traffic, not algorithmic bytes: the kernel is latency-bound")
    # the build's HBM-bound sort: 16 B/element/pass ideal (key+value read and write)
    rl_sort = roofline("lbvh_sort_pass", "k_sort_pass (onesweep 8-bit LSD pass over (code, id) pairs)", 16.0 * T,
                       "16 B per element per pass (8 B read + 8 B write)")
    trace_ms = kernels.get("trace", {}).get("ms_per_frame", 0.0)
    # the primary launch alone (2.07M rays), event-timed on its stream
    prim_rays, prim_hits = rtx.Rays(ctx, sd.width * sd.height), rtx.Hits(ctx, sd.width * sd.height)
    prim_rays.primary(cam, sd.width, sd.height)
    rtx.trace(s, prim_rays, prim_hits)
    ctx.device_sync()
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(5):
        rtx.trace(s, prim_rays, prim_hits)
    rep_p = ctx.profile_report()
    ctx.profile(False)
    primary_trace_ms = rep_p["trace"][1] / rep_p["trace"][0]
    prim_rays.destroy()
    prim_hits.destroy()
    # the LBVH build alone, back to back on one stream with no per-kernel
    # events (the profiled figure above adds an event pair per launch)
    for _ in range(3):
        s.build(30, 1)
    ctx.device_sync()
    t0 = time.perf_counter()
    for _ in range(20):
        s.build(30, 1)
    ctx.device_sync()
    build_wall_ms = (time.perf_counter() - t0) / 20 * 1e3
    log(f"LBVH build, back to back: {build_wall_ms:.3f} ms")
    fp64 = fp64_hit_error(ctx, s, sd, cam)
    log(f"float64 check of the frame's hit distances: max abs {fp64['max_abs_error']:.3e}, "
        f"max rel {fp64['max_rel_error']:.3e} over {fp64['hits_checked']} hits")

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
        "env": {"GPU_MAX_HW_QUEUES": os.environ.get("GPU_MAX_HW_QUEUES", "unset (HIP default: 4)")},
        "rays_per_frame": int(rays_per_frame),
        "shadow_rays_per_frame": int(shadow_per_frame),
        "mrays_incl_shadow": round((rays_per_frame + shadow_per_frame) / ms_per_step / 1e3, 3),
        "lbvh_build_ms": round(build_ms, 4) if build_ms else None,
        "lbvh_build_wall_ms": round(build_wall_ms, 4),
        "lbvh_autotune": build_tune,
        "frame_mode": frame_mode,
        "frames_in_flight_ms_per_frame": {str(k): round(v, 4) for k, v in flights.items()},
        "frames_in_flight_autotune": flight_tune,
        "pipelined_graph_ms_per_frame": round(pip_graph_ms, 4),
        "graph_ms_per_frame_by_shape": {f"{k[0]}_{k[1]}": round(v, 4) for k, v in graph_shapes.items()},
        "pipelined_stage_launch_ms_per_frame": round(pip_eager_ms, 4),
        "sequential_graph_ms_per_frame": round(seq_ms, 4),
        "sequential_graph_mrays": round(seq_value, 3),
        "stage_launch_ms_per_frame": round(async_ms, 4),
        "plain_path_ms_per_frame": round(plain_ms, 4),
        "plain_path_mrays": round(rays_per_frame / plain_ms / 1e3, 3),
        "roofline": rl_trace,
        "roofline_sort_pass": rl_sort,
        "build_stages": build_stage_rooflines(kernels, T, s.info()[0]),
        "traversal": {
            "trace_ms_per_frame": round(trace_ms, 4),
            "shadow_ms_per_frame": round(kernels.get("trace_shadow", {}).get("ms_per_frame", 0.0), 4),
            "trace_mrays": round(rays_per_frame / trace_ms / 1e3, 2) if trace_ms else None,
            "primary_trace_ms": round(primary_trace_ms, 4),
            "primary_trace_mrays": round(sd.width * sd.height / primary_trace_ms / 1e3, 1),
            "fp64_hit_distance": fp64,
        },
        "kernels": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in kernels.items()},
    }
    if args.c3_frames > 0:
        log(f"C3 dynamic scene, {args.c3_frames} graph frames ...")
        R.destroy()
        s.destroy()
        R = s = None
        out["c3_dynamic"] = bench_c3(ctx, args.c3_frames)
        log(f"C3: {out['c3_dynamic']['ms_per_frame']} ms/frame, {out['c3_dynamic']['mrays']} Mrays/s")
    if args.inst_frames > 0:
        out["instances"] = bench_instances(ctx, args.inst_frames)
    if args.accel_frames > 0:
        out["accelerators"] = bench_accelerators(ctx, args.accel_frames)
    if args.image_frames > 0:
        out["images_and_ingest"] = bench_images_and_ingest(ctx, args.image_frames)
    if args.sort_sizes:
        out["sort_scaling"] = bench_sort_scaling(ctx, [int(x) for x in args.sort_sizes.split(",") if x])
    if args.c4_scenes > 0:
        out["c4_batched"] = bench_c4(ctx, args.c4_scenes)
        c4 = out["c4_batched"]
        log(f"C4: {c4['ms_per_batch']} ms per batch, {c4['mrays']} Mrays/s (one by one: "
            f"{c4['one_by_one_mrays']} Mrays/s); chosen {c4['chosen']}")
    if args.c5_rates:
        rates = [int(x) for x in args.c5_rates.split(",") if x]
        out["c5_serving"] = bench_c5(ctx, rates, args.c5_max_requests, seeds=args.c5_seeds)
    if not args.no_cpu:
        log(f"CPU oracle baseline (one core, {args.cpu_frames} full frames) ...")
        runs = [cpu_baseline(sd, cam) for _ in range(max(1, args.cpu_frames))]
        runs.sort(key=lambda r: r[1])
        mr, secs, rays, tb = runs[len(runs) // 2]
        # fetch counts of the traversal (oracle counters), against the GPU's primary launch
        nv, nt, ns = cpu_visit_counts(sd, cam)
        P = sd.width * sd.height
        eff = P * (nv * 32.0 + nt * 48.0)
        out["traversal"]["oracle_fetches_per_primary_ray"] = {"node_visits": round(nv, 2),
                                                              "triangle_tests": round(nt, 2), "sample_rays": ns}
        out["traversal"]["effective_fetch_bytes_per_primary_launch"] = int(eff)
        out["traversal"]["effective_fetch_GBs"] = round(eff / (out["traversal"]["primary_trace_ms"] / 1e3) / 1e9, 1)
        out["traversal"]["effective_note"] = ("32 B per node visit and 48 B per triangle test (the oracle's counts), "
                                              "over the primary launch's time: cache traffic, mostly L2 hits")
        out["cpu_baseline"] = {"value": round(mr, 4), "unit": "Mrays/s", "cores": 1, "kind": "port",
                               "sample": f"median of {len(runs)} whole C2 frames on one core (sequential oracle): "
                                         f"LBVH build ({tb:.2f}s) + render of {rays} rays, {secs:.1f}s per frame",
                               "host_cpu": _cpu_model(), "nproc": os.cpu_count()}
    else:
        out["cpu_baseline"] = None
    if R is not None:
        R.destroy()
        s.destroy()
    ctx.close()
    print(json.dumps(out))


def build_stage_rooflines(kernels, T, nodes):
    """HBM roofline of the build's streaming stages, with the algorithmic
    bytes of the reference's stage definitions: prim AABB + Morton 76 B per
    triangle (12 B indices, 36 B positions, 24 B box, 4 B code); the
    permutation of the index buffer and material map 2 x (12 + 4) + 4 = 36 B
    per triangle; the refit 28 T + 32 N_L + 80 N_I bytes (per triangle 4 B id
    + 24 B box, per leaf 8 B range + 24 B box, per inner node two 24 B child
    boxes, 8 B ids and a 24 B box).  This build's stages move more in places
    (centroids, triangle records gathered with the permutation); the bytes
    are the reference's, the time is this build's."""
    ni = (nodes - 1) // 2
    nl = nodes - ni
    stages = {
        "prim_bounds_morton": (("lbvh_prim_bounds", "lbvh_morton"), 76.0 * T),
        "permutation": (("lbvh_rearrange",), 36.0 * T),
        "refit": (("lbvh_refit_level", "lbvh_refit_top", "lbvh_fat_leaves"), 28.0 * T + 32.0 * nl + 80.0 * ni),
        "sort_all_passes": (("lbvh_sort_pass",), 16.0 * T * 4),
        # the same passes against 4 + 16 P_s B per element (one key read for
        # the digit histograms, then key and value read and written per pass)
        "sort_with_histogram_read": (("lbvh_sort_pass",), (4.0 + 16.0 * 4) * T),
    }
    out = {}
    for name, (keys, algo) in stages.items():
        ms = sum(kernels.get(k, {}).get("ms_per_frame", 0.0) for k in keys)
        gbs = algo / (ms / 1e3) / 1e9 if ms > 0 else None
        out[name] = {"ms_per_build": round(ms, 4), "algorithmic_bytes": int(algo),
                     "achieved_GBs": round(gbs, 1) if gbs else None,
                     "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None, "kernels": list(keys)}
    return out


def bench_c3(ctx, frames: int):
    """Dynamic scene C3: 724x724 heightfield (1,045,458 triangles) displaced by
    8 travelling waves on the device every frame, LBVH rebuilt, 1280x720,
    1 point light, Whitted depth 2; the whole frame is one hipGraph whose
    device clock advances the animation by 1/60 s per replay."""
    import rtx
    from rtx import scenes
    sd = scenes.config_c3()
    s = scenes.load(rtx.Scene(ctx), sd)
    s.set_waves(sd.waves)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces, depth=30, min_leaf=1)
    R.set_animation(True, 0.0, 1.0 / 60.0)
    timing = {}
    ctx.tune("side_stream", 0)  # graphs are captured from one stream
    for mode in ("sequential", "pipelined"):
        R.set_pipelined(mode == "pipelined")
        s.set_clock(0)
        for _ in range(2):
            R.frame_async(s, cam, rebuild=True)
        ctx.device_sync()
        R.capture(s, cam, rebuild=True)
        # ray counts of every frame (frames are a function of the clock), untimed
        s.set_clock(0)
        rays = shadow = 0
        for _ in range(frames):
            R.replay()
            st = R.last_stats()
            rays += st.primary_rays + st.secondary_rays
            shadow += st.shadow_rays
        # the timed pass over the same frames
        s.set_clock(0)
        ctx.device_sync()
        t0 = time.perf_counter()
        for _ in range(frames):
            R.replay()
        ctx.device_sync()
        timing[mode] = (time.perf_counter() - t0, rays, shadow)
    ctx.tune("side_stream", 1)
    R.destroy()
    # frames in flight: the frame pipeline animates and rebuilds on its build
    # stream; frame k is the sequential frame k (same clock), so the ray
    # counts are the sequential pass's
    flight = {}
    for depth in (2, 3, 4, 5, 6, 8):
        pl = rtx.Pipeline(ctx, s, sd.width, sd.height, sd.max_bounces, depth=depth)
        pl.set_animation(True, 0.0, 1.0 / 60.0)
        s.set_clock(0)
        ctx.device_sync()
        t0 = time.perf_counter()
        last = 0
        for _ in range(frames):
            last = pl.frame(cam, rebuild=True)
        pl.wait(last)
        ctx.device_sync()
        flight[depth] = time.perf_counter() - t0
        pl.destroy()
    depth = min(flight, key=flight.get)
    timing["in_flight"] = (flight[depth],) + timing["sequential"][1:]
    best = min(timing, key=lambda k: timing[k][0])
    dt, rays, shadow = timing[best]
    sdt = timing["sequential"][0]
    desc = {"sequential": "one hipGraph per frame",
            "pipelined": "one hipGraph per frame, pipelined (next frame's displacement and build on a second stream)",
            "in_flight": f"frame pipeline with {depth} frames in flight (displacement and builds on a build stream, "
                         "renders of consecutive frames overlapping)"}
    out = {"workload": "C3 dynamic: 724x724 heightfield = 1,045,458 triangles displaced on the device every frame "
                       "(8 waves), full LBVH rebuild, 1280x720, 1 point light, Whitted depth 2; " + desc[best],
           "mode": best,
           "frames": frames, "ms_per_frame": round(dt / frames * 1e3, 4),
           "mrays": round(rays / dt / 1e6, 3), "mrays_incl_shadow": round((rays + shadow) / dt / 1e6, 3),
           "ms_per_frame_by_mode": {k: round(v[0] / frames * 1e3, 4) for k, v in timing.items()},
           "in_flight_ms_per_frame_by_depth": {str(k): round(v / frames * 1e3, 4) for k, v in flight.items()},
           "sequential_ms_per_frame": round(sdt / frames * 1e3, 4),
           "sequential_mrays": round(timing["sequential"][1] / sdt / 1e6, 3),
           "rays_per_frame_mean": round(rays / frames, 1), "triangles": sd.T}
    s.destroy()
    return out


def roofline_only(ctx, s, R, cam, sd, frames: int, warmup: int):
    """The roofline objects of the default bench line over `frames` C2 frames
    (rebuild + render) launched on one stream with per-kernel HIP events, so
    that a kernel trace of this run averages the same launches."""
    ctx.tune("side_stream", 0)
    for _ in range(max(1, warmup)):
        R.frame_async(s, cam, rebuild=True)
    ctx.device_sync()
    st = R.last_stats()
    rays = st.primary_rays + st.secondary_rays
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(frames):
        R.frame_async(s, cam, rebuild=True)
    rep = ctx.profile_report()
    ctx.profile(False)
    ctx.tune("side_stream", 1)
    out = {"frames": frames, "rays_per_frame": rays}
    for key, kernel, algo in (("trace", "k_trace", None), ("lbvh_sort_pass", "k_sort_pass", 16.0 * sd.T)):
        c, ms = rep[key]
        per_frame = c / frames
        if algo is None:
            algo = 48.0 * rays / per_frame  # ray in + hit out per ray, over the frame's launches
        avg_s = ms / c / 1e3
        ach = algo / avg_s / 1e9
        out[kernel] = {"launches": c, "avg_launch_ms": round(avg_s * 1e3, 5), "algorithmic_bytes_per_launch": algo,
                       "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": traffic_from_profiles(kernel)}
    return out


def lbvh_autotune(ctx, s, reps: int = 10):
    """Times `reps` back-to-back LBVH builds of scene `s` for each setting of
    the build's knobs and leaves the fastest set on the context."""
    cands = [(8, 16, 8), (8, 32, 8), (4, 16, 8), (4, 32, 8), (8, 16, 2), (8, 16, 4)]
    res = []
    for bits, items, grid in cands:
        ctx.tune("sort_bits", bits)
        ctx.tune("sort_items", items)
        ctx.tune("build_grid", grid)
        for _ in range(2):
            s.build(30, 1)
        ctx.device_sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            s.build(30, 1)
        ctx.device_sync()
        res.append({"sort_bits": bits, "sort_items": items, "build_grid": grid,
                    "ms_per_build": round((time.perf_counter() - t0) / reps * 1e3, 4)})
    best = min(res, key=lambda r: r["ms_per_build"])
    for k in ("sort_bits", "sort_items", "build_grid"):
        ctx.tune(k, best[k])
    return {"candidates": res, "chosen": {k: best[k] for k in ("sort_bits", "sort_items", "build_grid")}}


def bench_c5(ctx, rates, max_requests: int, workers: int = 128, arena_bytes: int = 32 << 30, groups: int = 2,
             seeds: int = 3, plain_requests: int = 1500):
    """C5 serving: 12 resident scenes (~20M triangles), Poisson arrivals per
    rate, Zipf(1.1) scene popularity, resolution uniform over 256^2, 512^2 and
    1280x720, depth uniform over 1..3, camera on a sphere of 2.5x the scene
    radius looking at its centre.  Latency = arrival -> image complete on the
    device.  The server admits requests by the bytes they take from a request
    arena of `arena_bytes`, at most `workers` in flight, and renders them with
    the wavefront renderer.

    Each rate runs a stream of min(max_requests, 2 s of arrivals) requests;
    from 1000/s up, `seeds` streams of independent seeds, whose latencies are
    pooled for p50/p99 and whose rays and spans are summed for Mrays/s (one
    stream's p99 moves by several ms with its bursts).  The plain
    per-request path runs the first `plain_requests` requests of the first
    stream of every rate; its throughput is flat once it saturates, and the
    largest it reaches over all rates is the reference for the ratio."""
    import numpy as np
    import rtx
    from rtx import scenes
    t0 = time.perf_counter()
    sds = scenes.config_c5()
    log(f"C5: generated {len(sds)} scenes, {sum(sd.T for sd in sds)} triangles in {time.perf_counter() - t0:.1f}s")
    scs = []
    for sd in sds:
        sc = scenes.load(rtx.Scene(ctx), sd)
        sc.build(30, 1)
        scs.append(sc)
    ctx.device_sync()
    classes = [(256, 256), (512, 512), (1280, 720)]
    res = {"workload": "C5 serving: 12 resident scenes, T ~ logU[10k, 4M] rescaled to ~20M triangles "
                       "(sphere sets and heightfields), Poisson arrivals, Zipf(1.1) scenes, resolution "
                       "U{256^2, 512^2, 1280x720}, depth U{1..3}, camera on a sphere of 2.5x the scene radius",
           "triangles": [int(sd.T) for sd in sds], "requests_in_flight_max": workers, "arena_bytes": arena_bytes,
           "groups": groups, "streams_per_rate_from_1000": seeds, "plain_requests_per_rate": plain_requests,
           "rates": list(rates), "server": [], "plain": []}
    for mode in ("server", "plain"):
        srv = rtx.Server(ctx, classes, workers=workers, max_bounces=3, plain=(mode == "plain"),
                         arena_bytes=arena_bytes, groups=groups)
        ids = [srv.add_scene(sc) for sc in scs]
        wb, sb = srv.memory()
        res["request_memory_bytes" if mode == "server" else "plain_renderer_bytes"] = wb
        res["resident_scene_bytes"] = sb
        for lam in rates:
            n = int(min(max_requests, max(40, lam * 2)))
            streams = seeds if (mode == "server" and lam >= 1000) else 1
            lat, rays, span, p99s = [], 0, 0.0, []
            for k in range(streams):
                reqs = scenes.c5_requests(sds, lam, n, int(lam) * 7 + 1 + 1000003 * k)
                if mode == "plain":
                    reqs = reqs[:plain_requests]
                start = rtx.now_ns() + 5_000_000
                for r in reqs:
                    r["scene"] = ids[r["scene"]]
                    r["arrival_ns"] += start
                srv.submit(reqs)
                srv.drain(600_000)
                done = srv.poll(len(reqs) + 16)
                assert len(done) == len(reqs) and all(c["status"] == 0 for c in done), "C5 request failed"
                ls = [(c["done_ns"] - c["arrival_ns"]) / 1e6 for c in done]
                lat += ls
                p99s.append(round(float(np.percentile(ls, 99)), 2))
                rays += sum(c["rays"] for c in done)
                span += (max(c["done_ns"] for c in done) - min(c["arrival_ns"] for c in done)) / 1e9
            lat = np.array(lat)
            row = {"rate_per_s": lam, "requests": int(lat.size), "streams": streams,
                   "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                   "mean_ms": round(float(lat.mean()), 3), "mrays": round(rays / span / 1e6, 2)}
            if streams > 1:
                row["p99_ms_per_stream"] = p99s
            res[mode].append(row)
            log(f"C5 {mode:6s} rate {lam:5d}/s: p50 {row['p50_ms']:8.2f} ms  p99 {row['p99_ms']:8.2f} ms  "
                f"{row['mrays']:8.1f} Mrays/s  ({row['requests']} requests, {streams} stream(s); "
                f"p99 per stream {p99s})")
        srv.destroy()
    for sc in scs:
        sc.destroy()
    # the served form against the plain path: the highest rate whose pooled
    # p99 stays within 50 ms, its throughput, and the plain path's saturated
    # throughput (the largest it reached)
    ok = [r for r in res["server"] if r["p99_ms"] <= 50.0]
    plain_sat = max((r["mrays"] for r in res["plain"]), default=0.0)
    if ok and plain_sat > 0:
        best = max(ok, key=lambda r: r["mrays"])
        res["p99_within_50ms"] = {"rate_per_s": best["rate_per_s"], "p99_ms": best["p99_ms"], "mrays": best["mrays"],
                                  "plain_saturated_mrays": plain_sat,
                                  "vs_plain_saturated": round(best["mrays"] / plain_sat, 3)}
    return res


def bench_c4(ctx, n_scenes: int = 64, reps: int = 5, streams: int = 8):
    """C4 batched: 64 scenes, T ~ logU[10k, 2M] (sphere sets and
    heightfields), resolution U{256^2, 512^2, 1280x720, 1920x1080}, depth
    U{1..4}; one frame of every scene per batch, rendered side by side on
    `streams` streams and captured into one hipGraph.  The traversal knobs
    (variant, work-group size, cooperative-leaf size) are autotuned on the
    batch first.  Compared with the same 64 frames rendered one after the
    other."""
    import numpy as np
    import rtx
    from rtx import scenes
    t0 = time.perf_counter()
    scs, rs, cams = [], [], []
    for fr in scenes.config_c4(n_scenes=n_scenes):
        sc = scenes.load(rtx.Scene(ctx), fr.sd)
        sc.build(30, 1)
        cams.append(scenes.orbit_camera(fr.sd, fr.direction, fr.width, fr.height))
        # bundles of 2 x pixels rays per bounce (the reference's capacity
        # rule drops rays beyond it), to keep 64 renderers within memory
        rs.append(rtx.Renderer(ctx, fr.width, fr.height, fr.bounces, ray_capacity=fr.ray_capacity))
        scs.append(sc)
    tris = np.array([sc.T for sc in scs])
    ctx.device_sync()
    log(f"C4: {n_scenes} scenes, {int(tris.sum())} triangles (target), set up in {time.perf_counter() - t0:.1f}s")

    def one_by_one():
        for R, sc, cam in zip(rs, scs, cams):
            R.frame_async(sc, cam, rebuild=False)
        ctx.device_sync()

    def timed(fn, n):
        ctx.device_sync()
        t = time.perf_counter()
        for _ in range(n):
            fn()
        ctx.device_sync()
        return (time.perf_counter() - t) / n

    # autotune the traversal knobs on the batch (every setting gives the same images)
    B = rtx.Batch(ctx, rs, scs, cams, streams=streams)
    tune = []
    for var, tb, coop in [(1, 128, 8), (1, 64, 8), (1, 256, 8), (0, 128, 8), (1, 128, 4), (1, 128, 16), (2, 128, 8),
                          (3, 128, 8)]:
        ctx.tune("trace_variant", var)
        ctx.tune("trace_block", tb)
        ctx.tune("coop_leaf", coop)
        B.render()
        dt = timed(B.render, 2)
        tune.append({"trace_variant": var, "trace_block": tb, "coop_leaf": coop, "ms": round(dt * 1e3, 3)})
    best = min(tune, key=lambda x: x["ms"])
    ctx.tune("trace_variant", best["trace_variant"])
    ctx.tune("trace_block", best["trace_block"])
    ctx.tune("coop_leaf", best["coop_leaf"])
    B.render()
    ctx.device_sync()
    rays, shadow = B.stats()
    dt_eager = timed(B.render, reps)
    B.capture()
    B.replay()
    dt_graph = timed(B.replay, reps)
    dt_seq = timed(one_by_one, reps)
    dt_best = min(dt_eager, dt_graph)
    B.destroy()
    wfr = bench_c4_wavefront(ctx, rs, scs, cams, rays, reps)
    for R in rs:
        R.destroy()
    for sc in scs:
        sc.destroy()
    ctx.tune("trace_variant", 1)
    ctx.tune("trace_block", 128)
    ctx.tune("coop_leaf", 8)
    return {"workload": "C4 batched: 64 resident scenes, T ~ logU[10k, 2M] (sphere sets and heightfields), "
                        "resolution U{256^2, 512^2, 1280x720, 1920x1080}, depth U{1..4}, one frame of each per "
                        f"batch, {streams} streams (stream launches or one hipGraph, the faster)",
            "scenes": n_scenes, "rays_per_batch": rays, "shadow_rays_per_batch": shadow,
            "ms_per_batch": round(dt_best * 1e3, 3), "mrays": round(rays / dt_best / 1e6, 2),
            "mode": "stream launches" if dt_eager <= dt_graph else "hipGraph replay",
            "stream_launch_ms_per_batch": round(dt_eager * 1e3, 3), "graph_ms_per_batch": round(dt_graph * 1e3, 3),
            "one_by_one_ms_per_batch": round(dt_seq * 1e3, 3), "one_by_one_mrays": round(rays / dt_seq / 1e6, 2),
            "autotune": tune, "chosen": best, "wavefront": wfr,
            "wavefront_vs_one_by_one": round(dt_seq / (wfr["ms_per_batch"] / 1e3), 2)}


def bench_c4_wavefront(ctx, rs, scs, cams, rays_expected, reps):
    """The C4 batch through the wavefront renderer: the same 64 frames (same
    scenes, cameras, sizes, depths and 2 x pixels + 2 ray capacity) admitted
    together and stepped to completion, one launch per stage per step.  The
    traversal knobs are autotuned on it first."""
    import rtx
    frames = [(sc, cam, R.p.width, R.p.height, R.p.max_bounces, 2 * R.p.width * R.p.height + 2)
              for R, sc, cam in zip(rs, scs, cams)]
    need = sum(rtx.Wavefront.frame_bytes(w, h, d, cap) for _, _, w, h, d, cap in frames)
    wf = rtx.Wavefront(ctx, arena_bytes=need + (256 << 20), max_frames=len(frames))
    last = {}

    def batch():
        ids = [wf.add(sc, cam, w, h, d, ray_capacity=cap) for sc, cam, w, h, d, cap in frames]
        wf.run()
        wf.sync()
        last["rays"] = sum(st.primary_rays + st.secondary_rays for st in map(wf.stats, ids))
        last["shadow"] = sum(wf.stats(i).shadow_rays for i in ids)
        for i in ids:
            wf.release(i)

    def timed(n):
        ctx.device_sync()
        t = time.perf_counter()
        for _ in range(n):
            batch()
        return (time.perf_counter() - t) / n

    tune = []
    keys = ("trace_variant", "trace_block", "coop_leaf", "wf_grain", "stack_lds")
    for cand in [(1, 128, 8, 1, 32), (1, 128, 8, 2, 32), (1, 128, 8, 4, 32), (1, 64, 8, 1, 32), (1, 256, 8, 1, 32),
                 (0, 128, 8, 1, 32), (2, 128, 8, 1, 32), (1, 128, 4, 1, 32), (1, 128, 16, 1, 32),
                 (1, 128, 8, 1, 16), (1, 256, 8, 1, 16), (1, 128, 8, 1, 8)]:
        for key, v in zip(keys, cand):
            ctx.tune(key, v)
        batch()
        tune.append({**dict(zip(keys, cand)), "ms": round(timed(2) * 1e3, 3)})
    best = min(tune, key=lambda x: x["ms"])
    for key in keys:
        ctx.tune(key, best[key])
    batch()
    assert last["rays"] == rays_expected, "the wavefront batch must trace the one-by-one frames' rays"
    ctx.profile(True)
    ctx.profile_reset()
    batch()
    stages = ctx.profile_report()
    ctx.profile(False)
    dt = timed(reps)
    wf.destroy()
    ctx.tune("trace_variant", 1)
    ctx.tune("trace_block", 128)
    ctx.tune("coop_leaf", 8)
    ctx.tune("wf_grain", 1)
    ctx.tune("stack_lds", 32)
    return {"what": "the 64 frames admitted together into the wavefront renderer and stepped to completion "
                    "(admission, steps, device completion and release timed)",
            "ms_per_batch": round(dt * 1e3, 3), "mrays": round(rays_expected / dt / 1e6, 2),
            "shadow_rays_per_batch": last["shadow"], "arena_bytes": need, "autotune": tune, "chosen": best,
            "stage_ms_per_batch": {k: round(v[1], 4) for k, v in stages.items()},
            "launches_per_batch": {k: v[0] for k, v in stages.items()}}


def bench_instances(ctx, frames: int = 50, reps: int = 3):
    """Instanced scene (multi-root traversal): 16 instances of 4 distinct
    meshes (jittered spheres of radius 5, 250k triangles each), 4M world
    triangles in a 4x4 grid covering about half the frame; every frame
    moves every instance rigidly (a new rotation)
    and traces 1920x1080 primary rays plus one shadow ray per hit and light
    (2 point lights), with no rebuild.  Beside it, the same frames over the
    aggregate mesh (the 16 copies' vertices moved on the host, as
    Scene::create_aggregate_mesh does): per frame a vertex upload and a full
    LBVH rebuild before the same rays."""
    import numpy as np
    import rtx
    from rtx import scenes
    t0 = time.perf_counter()
    meshes = []
    for k in range(4):
        rng = np.random.default_rng(0x5EED0010 + k)
        slices, stacks = 500, 250
        jit = rng.uniform(-0.02, 0.02, ((stacks + 1) * (slices + 1), 3))
        p, nn, uv, idx = scenes.uv_sphere(slices, stacks, (0.0, 0.0, 0.0), 5.0, jit)
        meshes.append(dict(pos=p.astype(np.float32), nrm=nn.astype(np.float32), uv=uv.astype(np.float32),
                           idx=idx.astype(np.uint32), T=len(idx)))
    W, H = 1920, 1080
    n_inst, spacing, scale = 16, 11.0, 1.0
    centres = [((i % 4 - 1.5) * spacing, (i // 4 - 1.5) * spacing, 0.0) for i in range(n_inst)]
    mesh_of = [i % 4 for i in range(n_inst)]
    spin = np.linspace(0.2, 1.0, n_inst)

    def pose(t):
        tr, tri = [], []
        for i in range(n_inst):
            a = spin[i] * t
            c, s_ = np.cos(a), np.sin(a)
            M = np.eye(4)
            M[:3, :3] = scale * np.array([[c, 0, s_], [0, 1, 0], [-s_, 0, c]])
            M[:3, 3] = centres[i]
            tr.append(M.astype(np.float32))
            tri.append(np.linalg.inv(M).astype(np.float32))
        return np.stack(tr), np.stack(tri)

    lights = scenes.config_c2(n_spheres=1).lights  # C2's two point lights
    scs = []
    for sd in meshes:
        sc = rtx.Scene(ctx)
        sc.set_geometry(sd["pos"], sd["idx"], sd["nrm"], sd["uv"], np.zeros(sd["T"], np.int32))
        sc.build(30, 1)
        scs.append(sc)
    tr, tri = pose(0.0)
    inst = rtx.Instances(ctx, [scs[m] for m in mesh_of], tr, tri)
    inst.set_lights(lights)
    cam = rtx.camera_set((0.0, 0.0, -48.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), np.pi / 4, W / H)
    n = W * H
    rays, hits = rtx.Rays(ctx, n), rtx.Hits(ctx, n)
    poses = [pose(0.05 * f) for f in range(frames)]
    ctx.device_sync()
    log(f"instances: {n_inst} instances of 4 meshes x {meshes[0]['T']} triangles, set up in "
        f"{time.perf_counter() - t0:.1f}s")

    def inst_frame(f):
        inst.set_transforms(*poses[f % frames])
        rays.primary(cam, W, H)
        inst.trace(rays, hits)
        inst.trace_shadow(rays, hits)

    def run(fn, nf):
        ctx.device_sync()
        t = time.perf_counter()
        for f in range(nf):
            fn(f)
        ctx.device_sync()
        return (time.perf_counter() - t) / nf

    for f in range(3):
        inst_frame(f)
    ctx.device_sync()
    g = hits.read(n)
    hit = (g["flags"] & 1) != 0
    n_hit = int(hit.sum())
    shadow_rays = n_hit * lights.count
    dt_inst = min(run(inst_frame, frames) for _ in range(reps))
    ctx.profile(True)
    ctx.profile_reset()
    for f in range(3):
        inst_frame(f)
    rep = ctx.profile_report()
    ctx.profile(False)
    kern = {k: {"launches": c, "avg_launch_ms": round(ms / c, 4)} for k, (c, ms) in rep.items()}
    trace_ms = kern.get("trace_instances", {}).get("avg_launch_ms")
    shadow_ms = kern.get("trace_shadow_instances", {}).get("avg_launch_ms")

    def schedule(fn, nf, mode):
        """fn timed under another root schedule: 1 each lane walks every root
        in one loop, 0 the wave-uniform root loop (every lane waits for the
        wave's slowest walk of each root)."""
        ctx.tune("inst_lane_roots", mode)
        try:
            fn(0)
            return min(run(fn, nf) for _ in range(reps))
        finally:
            ctx.tune("inst_lane_roots", 2)

    dt_inst_all = schedule(inst_frame, frames, 1)
    dt_inst_wave = schedule(inst_frame, frames, 0)
    log(f"instances: {dt_inst * 1e3:.3f} ms/frame, {(n + shadow_rays) / dt_inst / 1e6:.1f} Mrays/s incl. shadow; "
        f"trace {trace_ms} ms, shadow {shadow_ms} ms; lane roots without culling {dt_inst_all * 1e3:.3f} ms/frame, "
        f"wave-uniform root loop {dt_inst_wave * 1e3:.3f} ms/frame")

    # whole frames over the instances: Whitted depth 3, one material per mesh
    # (its first instance's), reflection and refraction in the mix
    rows = [(0.8, 0.3, 0.3, 0.5, 0.0, 0.0), (0.3, 0.8, 0.3, 1.0, 0.5, 0.0), (0.3, 0.3, 0.8, 0.2, 0.3, 1.5),
            (0.8, 0.8, 0.8, 1.0, 0.0, 0.0)]
    inst.set_materials(rtx.make_materials([rows[m] for m in mesh_of]))
    R = rtx.Renderer(ctx, W, H, 3)
    fst = {}

    def inst_render(f):
        inst.set_transforms(*poses[f % frames])
        fst["st"] = R.frame_instances(inst, cam)

    for f in range(3):
        inst_render(f)
    dt_inst_frame = min(run(inst_render, min(frames, 20)) for _ in range(reps))
    ist = fst["st"]
    ctx.profile(True)
    ctx.profile_reset()
    for f in range(3):
        inst_render(f)
    rep = ctx.profile_report()
    ctx.profile(False)
    frame_kern = {k: {"launches_per_frame": round(c / 3, 2), "ms_per_frame": round(ms / 3, 4)}
                  for k, (c, ms) in rep.items()}
    dt_inst_frame_all = schedule(inst_render, min(frames, 20), 1)
    dt_inst_frame_wave = schedule(inst_render, min(frames, 20), 0)
    inst_frame_rays = int(ist.primary_rays + ist.secondary_rays)
    log(f"instances, whole frames: {dt_inst_frame * 1e3:.3f} ms/frame, {inst_frame_rays} rays "
        f"({ist.secondary_rays} secondary, {ist.shadow_rays} shadow), {inst_frame_rays / dt_inst_frame / 1e6:.1f} "
        f"Mrays/s; lane roots without culling {dt_inst_frame_all * 1e3:.3f} ms/frame, wave-uniform root loop "
        f"{dt_inst_frame_wave * 1e3:.3f} ms/frame; stages {frame_kern}")

    # the aggregate mesh of the same frames: vertices moved on the host
    # (not timed), uploaded and rebuilt every frame
    def aggregate(t_pose):
        trs = t_pose[0]
        P, I, N, U = [], [], [], []
        base = 0
        for i in range(n_inst):
            sd = meshes[mesh_of[i]]
            M = trs[i].astype(np.float64)
            P.append((sd["pos"].astype(np.float64) @ M[:3, :3].T + M[:3, 3]).astype(np.float32))
            N.append((sd["nrm"].astype(np.float64) @ M[:3, :3].T).astype(np.float32))
            U.append(sd["uv"])
            I.append(sd["idx"] + base)
            base += len(sd["pos"])
        return np.concatenate(P), np.concatenate(I), np.concatenate(N), np.concatenate(U)

    P0, I0, N0, U0 = aggregate(poses[0])
    agg_poses = [P0, aggregate(poses[1])[0]]
    agg = rtx.Scene(ctx)
    agg_matmap = np.concatenate([np.full(meshes[mesh_of[i]]["T"], mesh_of[i], np.int32) for i in range(n_inst)])
    agg.set_geometry(P0, I0, N0, U0, agg_matmap)
    agg.set_materials(rtx.make_materials(rows))
    agg.set_lights(lights)
    agg.build(30, 1)

    def agg_frame(f):
        agg.update_positions(agg_poses[f % 2])
        agg.build(30, 1)
        rays.primary(cam, W, H)
        rtx.trace(agg, rays, hits)
        rtx.trace_shadow(agg, rays, hits)

    for f in range(2):
        agg_frame(f)
    dt_agg = min(run(agg_frame, min(frames, 20)) for _ in range(reps))
    ctx.profile(True)
    ctx.profile_reset()
    for f in range(3):
        agg_frame(f)
    rep = ctx.profile_report()
    ctx.profile(False)
    agg_kern = {k: round(ms / c, 4) for k, (c, ms) in rep.items()}
    log(f"aggregate mesh with rebuild: {dt_agg * 1e3:.3f} ms/frame; kernels {agg_kern}")

    def agg_render(f):
        agg.update_positions(agg_poses[f % 2])
        fst["st"] = R.frame(agg, cam, rebuild=True)

    for f in range(2):
        agg_render(f)
    dt_agg_frame = min(run(agg_render, min(frames, 20)) for _ in range(reps))
    ast = fst["st"]
    log(f"aggregate mesh, whole frames with rebuild: {dt_agg_frame * 1e3:.3f} ms/frame, "
        f"{int(ast.primary_rays + ast.secondary_rays)} rays")
    R.destroy()
    inst.destroy()
    agg.destroy()
    for sc in scs:
        sc.destroy()
    rays.destroy()
    hits.destroy()
    return {"workload": f"instanced scene: {n_inst} instances of 4 meshes (50-sphere C2-style clusters, "
                        f"{meshes[0]['T']} triangles each), {n_inst * meshes[0]['T']} world triangles, 1920x1080 primary "
                        "rays + shadow rays to 2 point lights; every frame moves every instance rigidly (new "
                        "transforms shipped on the stream), no rebuild",
            "triangles_world": int(n_inst * meshes[0]['T']), "triangles_unique": int(4 * meshes[0]['T']),
            "primary_rays": n, "hits": n_hit, "shadow_rays": shadow_rays,
            "ms_per_frame": round(dt_inst * 1e3, 4), "mrays": round(n / dt_inst / 1e6, 2),
            "lane_roots_no_culling_ms_per_frame": round(dt_inst_all * 1e3, 4),
            "wave_uniform_roots_ms_per_frame": round(dt_inst_wave * 1e3, 4),
            "root_schedule": "each lane walks, in one loop, the roots whose widened world box its ray meets "
                             "(inst_lane_roots 2); beside it every root in one loop (1) and the wave-uniform root "
                             "loop (0)",
            "mrays_incl_shadow": round((n + shadow_rays) / dt_inst / 1e6, 2),
            "trace_instances_ms": trace_ms, "trace_shadow_instances_ms": shadow_ms,
            "roofline": {"kernel": "k_trace_inst_lr (closest hit over 16 roots, culled by world boxes)", "bound": "hbm",
                         "achieved": round(48.0 * n / (trace_ms / 1e3) / 1e9, 2) if trace_ms else None,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(48.0 * n / (trace_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 5) if trace_ms else None,
                         "traffic": None, "algorithmic_bytes_per_launch": 48 * n,
                         "note": "48 B/ray (32 B ray in + 16 B hit out); node, triangle and root fetches are cache "
                                 "traffic: the kernel is latency-bound like the single-root traversal"},
            "whole_frame": {"what": "Renderer frame over the instances (rtx_render_instances): primary rays, then "
                                    "per bounce the multi-root trace and shadow kernels, secondary rays and shading "
                                    "through the shared material map, Whitted depth 3, one material per mesh "
                                    "(diffuse, reflective, refractive IOR 1.5, diffuse); ray counts read back per "
                                    "bounce; new transforms every frame, no rebuild",
                            "ms_per_frame": round(dt_inst_frame * 1e3, 4), "rays": inst_frame_rays,
                            "lane_roots_no_culling_ms_per_frame": round(dt_inst_frame_all * 1e3, 4),
                            "wave_uniform_roots_ms_per_frame": round(dt_inst_frame_wave * 1e3, 4),
                            "stages": frame_kern,
                            "secondary_rays": int(ist.secondary_rays), "shadow_rays": int(ist.shadow_rays),
                            "mrays": round(inst_frame_rays / dt_inst_frame / 1e6, 2),
                            "aggregate_rebuild_ms_per_frame": round(dt_agg_frame * 1e3, 4),
                            "aggregate_rays": int(ast.primary_rays + ast.secondary_rays),
                            "aggregate_note": "the same frames over the aggregate mesh (rtx_render_frame with a full "
                                              "LBVH rebuild after the vertex upload; normals of the first pose)"},
            "aggregate_rebuild_ms_per_frame": round(dt_agg * 1e3, 4),
            "aggregate_stage_avg_ms": agg_kern,
            "aggregate_note": "the same rays over one aggregate mesh of the 16 moved copies: vertex upload "
                              "(positions moved on the host beforehand, not timed) + full LBVH rebuild per frame"}


def bench_images_and_ingest(ctx, frames: int = 10, reps: int = 3):
    """Textured diffuse and the cubemap background, and OBJ ingest.

    Textures: the C2 frame without rebuild (1M triangles, 1920x1080, Whitted
    depth 3), plain, then with every material textured (four 512x512 RGBA8
    images, index m % 4) under a cubemap of six 512x512 faces; frame time and
    the shading stage's time (HIP events).  Ingest: a 250k-triangle OBJ with
    positions, uvs and normals and its MTL file, written at run time, read by
    the library's loader (host), then uploaded and built (device)."""
    import os
    import shutil
    import tempfile
    import numpy as np
    import rtx
    from rtx import scenes
    sd = scenes.config_c2()
    s = scenes.load(rtx.Scene(ctx), sd)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces)
    R.frame(s, cam, rebuild=True)

    def timed():
        ctx.device_sync()
        t = time.perf_counter()
        for _ in range(frames):
            R.frame(s, cam, rebuild=False)
        ctx.device_sync()
        return (time.perf_counter() - t) / frames

    def shade_ms():
        ctx.profile(True)
        ctx.profile_reset()
        R.frame(s, cam, rebuild=False)
        rep = ctx.profile_report()
        ctx.profile(False)
        c, ms = rep.get("shade", (1, 0.0))
        return round(ms, 4), int(c)

    dt_plain = min(timed() for _ in range(reps))
    shade_plain, launches = shade_ms()
    rng = np.random.default_rng(0x7E57)
    texs = [rng.integers(0, 256, (512, 512, 4), dtype=np.uint8) for _ in range(4)]
    faces = [rng.integers(0, 256, (512, 512, 4), dtype=np.uint8) for _ in range(6)]
    mats = np.ascontiguousarray(sd.materials).view(np.int32).reshape(-1, 8).copy()
    mats[:, 7] = np.arange(len(mats)) % 4
    s.set_materials(mats.view(np.uint8).reshape(-1))
    s.set_textures(texs)
    s.set_cubemap(faces)
    R.frame(s, cam, rebuild=False)
    dt_tex = min(timed() for _ in range(reps))
    shade_tex, _ = shade_ms()
    log(f"textures: C2 frame without rebuild {dt_plain * 1e3:.3f} ms plain, {dt_tex * 1e3:.3f} ms textured with "
        f"cubemap; shading {shade_plain} -> {shade_tex} ms per frame ({launches} launches)")
    R.destroy()
    s.destroy()

    # OBJ ingest
    tmp = tempfile.mkdtemp(prefix="rtx_obj_", dir="/tmp")
    try:
        p, nrm, uv, idx = scenes.uv_sphere(500, 250, (0.0, 0.0, 0.0), 1.0)
        path = os.path.join(tmp, "sphere.obj")
        with open(os.path.join(tmp, "sphere.mtl"), "w") as f:
            f.write("newmtl red\nKd 0.8 0.2 0.2\nKa 0.1 0.1 0.1\nKs 0.5 0.5 0.5\nNs 20\n")
        with open(path, "w") as f:
            f.write("mtllib sphere.mtl\n")
            np.savetxt(f, p, fmt="v %.6f %.6f %.6f")
            np.savetxt(f, uv, fmt="vt %.6f %.6f")
            np.savetxt(f, nrm, fmt="vn %.6f %.6f %.6f")
            f.write("usemtl red\n")
            i1 = idx.astype(np.int64) + 1
            np.savetxt(f, np.repeat(i1, 3, axis=1), fmt="f %d/%d/%d %d/%d/%d %d/%d/%d")
        size = os.path.getsize(path)
        t = time.perf_counter()
        m = rtx.ObjModel(path)
        dt_parse = time.perf_counter() - t
        sc = rtx.Scene(ctx)
        ctx.device_sync()
        t = time.perf_counter()
        sc.set_geometry(m.pos, m.idx, m.nrm, m.uv, np.zeros(len(m.idx), np.int32))
        sc.build(30, 1)
        ctx.device_sync()
        dt_up = time.perf_counter() - t
        sc.destroy()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    T = int(len(m.idx))
    log(f"OBJ ingest: {T} triangles, {size / 1e6:.1f} MB parsed in {dt_parse * 1e3:.1f} ms "
        f"({size / dt_parse / 1e6:.0f} MB/s), upload + LBVH build {dt_up * 1e3:.1f} ms")
    return {"textures": {"workload": "C2 frame without rebuild (1M triangles, 1920x1080, 2 point lights, Whitted "
                                     "depth 3); textured: every material samples one of four 512x512 RGBA8 images "
                                     "(OpenCL linear filter, repeat), misses sample a cubemap of six 512x512 faces "
                                     "(clamp to edge)",
                         "plain_ms_per_frame": round(dt_plain * 1e3, 4),
                         "textured_ms_per_frame": round(dt_tex * 1e3, 4),
                         "shade_ms_per_frame_plain": shade_plain, "shade_ms_per_frame_textured": shade_tex,
                         "shade_launches_per_frame": launches},
            "obj_ingest": {"workload": "one OBJ file with v, vt, vn and v/vt/vn faces plus its MTL file, written at "
                                       "run time (a 500x250 UV sphere), read by the library's loader on the host, "
                                       "then uploaded and built into an LBVH",
                           "triangles": T, "bytes": int(size), "parse_ms": round(dt_parse * 1e3, 2),
                           "parse_MBps": round(size / dt_parse / 1e6, 1),
                           "parse_mtris_per_s": round(T / dt_parse / 1e6, 3),
                           "upload_and_build_ms": round(dt_up * 1e3, 2)}}


def bench_accelerators(ctx, frames: int = 10):
    """The C2 frame (1M triangles, 1920x1080, Whitted depth 3) rendered
    without a rebuild over the two trees the reference can build -- the
    LBVH and the CPU SAH BVH (src/rt/bvh.cpp) -- each with the stack walk
    (form 1) and the stackless walk (form 4, trace-bvh-no-stack.cl); the SAH
    tree's host build time beside the LBVH's device build."""
    import numpy as np
    import rtx
    from rtx import scenes
    sd = scenes.config_c2()
    s = scenes.load(rtx.Scene(ctx), sd)
    cam = scenes.camera_of(sd)
    R = rtx.Renderer(ctx, sd.width, sd.height, sd.max_bounces)
    out = {"workload": "C2 frame without rebuild (1M triangles, 1920x1080, 2 point lights, Whitted depth 3) over "
                       "the LBVH and the CPU SAH BVH, stack walk (form 1) and stackless walk (form 4)"}
    old = ctx.tuned("trace_variant")

    def measure(tag):
        row = {}
        for form in (1, 4):
            ctx.tune("trace_variant", form)
            R.frame_async(s, cam, rebuild=False)
            ctx.device_sync()
            t = time.perf_counter()
            for _ in range(frames):
                R.frame_async(s, cam, rebuild=False)
            ctx.device_sync()
            ms = (time.perf_counter() - t) / frames * 1e3
            ctx.tune("side_stream", 0)
            ctx.profile(True)
            ctx.profile_reset()
            R.frame_async(s, cam, rebuild=False)
            rep = ctx.profile_report()
            ctx.profile(False)
            ctx.tune("side_stream", 1)
            st = R.last_stats()
            rays = st.primary_rays + st.secondary_rays
            row[f"form{form}"] = {"ms_per_frame": round(ms, 4), "mrays": round(rays / ms / 1e3, 1),
                                  "trace_ms": round(rep.get("trace", (0, 0.0))[1], 4),
                                  "shadow_ms": round(rep.get("trace_shadow", (0, 0.0))[1], 4)}
        ctx.tune("trace_variant", old)
        out[tag] = row

    s.build(30, 1)
    ctx.device_sync()
    out["lbvh_nodes"] = int(s.info()[0])
    measure("lbvh")
    t = time.perf_counter()
    s.build_sah()
    out["sah_host_build_s"] = round(time.perf_counter() - t, 3)
    out["sah_nodes"] = int(s.info()[0])
    measure("sah")
    log(f"accelerators: LBVH {out['lbvh']}, SAH {out['sah']} (host build {out['sah_host_build_s']} s)")
    R.destroy()
    s.destroy()
    return out


def bench_sort_scaling(ctx, sizes):
    """The build's radix-sort pass at sizes beyond the L2/MALL (random
    triangle soups): average pass time and GB/s at 16 B/element/pass."""
    import numpy as np
    import rtx
    rows = []
    for T in sizes:
        rng = np.random.default_rng(1)
        V = T // 2
        pos = rng.uniform(-100, 100, (V, 3)).astype(np.float32)
        idx = rng.integers(0, V, (T, 3), dtype=np.uint32)
        s = rtx.Scene(ctx)
        s.set_geometry(pos, idx, None, None, np.zeros(T, np.int32))
        del pos, idx
        s.build(30, 1)
        ctx.device_sync()
        ctx.profile(True)
        ctx.profile_reset()
        for _ in range(3):
            s.build(30, 1)
        rep = ctx.profile_report()
        ctx.profile(False)
        n, ms = rep["lbvh_sort_pass"]
        avg_s = ms / n / 1e3
        gbs = 16.0 * T / avg_s / 1e9
        rows.append({"elements": T, "avg_pass_ms": round(avg_s * 1e3, 4), "achieved_GBs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "lbvh_build_ms": round(rep["LBVH_BUILD"][1] / 3, 3)})
        log(f"sort pass at {T} elements: {avg_s * 1e3:.3f} ms, {gbs:.0f} GB/s; build {rows[-1]['lbvh_build_ms']} ms")
        s.destroy()
    return rows


def _cpu_model() -> str:
    """The host CPU's model name, with its family / model / stepping numbers
    (a virtualized host may report only the product line)."""
    fields = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping") and k not in fields:
                    fields[k] = v
            if not line.strip() and fields:
                break
    except OSError:
        return "unknown"
    name = fields.get("model name", "unknown")
    ids = ", ".join(f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields)
    return f"{name} ({ids})" if ids else name


if __name__ == "__main__":
    main()
