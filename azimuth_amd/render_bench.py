"""Render throughput of the wavefront path tracer on one GPU:

    python -m azimuth_amd.render_bench --workload cornell|interior|motion
        [--spp N] [--scale S] [--guide] [--env cosine|luminance] [--guide-max-depth D] [--autotune | --tuning B,W,L,C] [--cpu-seconds T]
        [--levels graph|host]
    python -m azimuth_amd.render_bench --workload motion --scenes 16 [--autotune]

Prints one JSON line: Msamples/s (width * height * spp_eff / render time,
spp_eff = floor(sqrt(spp))^2 as the reference renders), the render time from
HIP events around the whole render call (after one untimed warm-up render), and optionally the CPU oracle's rate
on a bounded strip of rows (the same camera, spp and seed), extrapolated
linearly to the full image.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import configs, native as N, workloads
from .scene import GpuContext


def build(name, scale, spp):
    if name == "cornell":
        sc, cam = configs.cornell_scene(int(256 * scale), int(256 * scale), spp=spp or 16, max_depth=8)
        filt = N.FilterType_Box
    elif name == "interior":
        sc, cam = workloads.interior_scene(int(1920 * scale), int(1080 * scale), spp=spp or 64, max_depth=6)
        filt = N.FilterType_Gaussian
    elif name == "motion":
        sc, cam = workloads.motion_scene(int(1280 * scale), int(720 * scale), spp=spp or 128, max_depth=10)
        filt = N.FilterType_Box
    else:
        raise SystemExit(f"unknown workload {name}")
    return sc, cam, filt


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="cornell")
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--guide", action="store_true", help="render with the learned path guide")
    ap.add_argument("--guide-max-depth", type=int, default=0, help="path depth up to which the guide is used (0: all)")
    ap.add_argument("--env", default="cosine", choices=("cosine", "luminance"),
                    help="environment next-event directions: the reference's cosine distribution or the map's "
                         "luminance distribution")
    ap.add_argument("--autotune", action="store_true", help="autotune chunk size and traversal variant first")
    ap.add_argument("--tuning", default="", help="fixed tuning instead: block,waves,loop,chunk (e.g. 128,6,1,4194304)")
    ap.add_argument("--cpu-seconds", type=float, default=0.0, help="time the CPU oracle on a strip of rows")
    ap.add_argument("--levels", default="graph", choices=("graph", "host"),
                    help="bounce levels replayed from a recorded graph per chunk, or launched from the host")
    ap.add_argument("--scenes", type=int, default=1,
                    help="motion workload only: render this many scenes (seeds 0..K-1) in one batched render")
    a = ap.parse_args(argv)
    if a.scenes > 1:
        return batch_main(a)
    sc, cam, filt = build(a.workload, a.scale, a.spp)
    W, H = cam.Resolution.x, cam.Resolution.y
    spp_eff = int(math.isqrt(cam.SampleCount)) ** 2
    ctx = GpuContext(0)
    ctx.set_level_mode(N.AZ_LEVELS_AUTO if a.levels == "graph" else N.AZ_LEVELS_HOST)
    t0 = time.time()
    ctx.upload_scene(sc)
    upload_s = time.time() - t0
    tuned = None
    if a.tuning:
        b, w, l, c = (int(x) for x in a.tuning.split(","))
        ctx.set_render_tuning(chunk=c, block=b, waves=w, loop=l)
        tuned = {"block": b, "waves": w, "chunk": c, "loop": "while-while" if l else "if-if", "candidates": 0}
    elif a.autotune:
        entries, best = ctx.autotune_render(cam, filt=filt, seed=a.seed)
        tuned = {"block": entries[best]["block"], "waves": entries[best]["waves"], "chunk": entries[best]["chunk"],
                 "loop": "while-while" if entries[best]["loop"] else "if-if", "candidates": len(entries)}
    env = N.AZ_ENV_LUMINANCE if a.env == "luminance" else N.AZ_ENV_COSINE
    # Warm-up: a full render, so that kernels are loaded and every queue
    # buffer has its steady-state size before the timed renders.
    ctx.render(cam, filt=filt, seed=a.seed, guide=a.guide, env_sampling=env, guide_max_depth=a.guide_max_depth)
    best = None
    for _ in range(a.repeat):
        ctx.timer_start()
        px = ctx.render(cam, filt=filt, seed=a.seed, guide=a.guide, env_sampling=env, guide_max_depth=a.guide_max_depth)
        ms = ctx.timer_stop()
        best = ms if best is None else min(best, ms)
    img = px["Color"] / np.maximum(px["TotalWeight"], 1e-30)[:, None]
    out = {"workload": a.workload, "width": W, "height": H, "spp": spp_eff, "triangles": sc.triangle_count,
           "render_ms": round(best, 3), "msamples_per_s": round(W * H * spp_eff / (best * 1e-3) / 1e6, 3),
           "upload_s": round(upload_s, 2), "mean_radiance": [round(float(x), 6) for x in img.mean(0)],
           "filter": "gaussian" if filt == N.FilterType_Gaussian else "box", "guide": bool(a.guide), "env": a.env, "guide_max_depth": a.guide_max_depth,
           "autotune": tuned,
           "levels": a.levels, "level_graphs": dict(zip(("recorded", "launched"), ctx.level_graph_counts()))}
    if a.cpu_seconds > 0:
        import oracle
        threads = max(1, min(16, os.cpu_count() or 1))
        rows = 1
        t0 = time.time()
        oracle.render(sc, cam, filt=filt, seed=a.seed, row_begin=H // 2, row_end=H // 2 + rows, threads=threads)
        dt = time.time() - t0
        rows = int(max(1, min(H, rows * a.cpu_seconds / max(dt, 1e-3))))
        t0 = time.time()
        oracle.render(sc, cam, filt=filt, seed=a.seed, row_begin=(H - rows) // 2, row_end=(H - rows) // 2 + rows,
                      threads=threads)
        dt = time.time() - t0
        rate = W * rows * spp_eff / dt / 1e6
        out["cpu_baseline"] = {"value": round(rate, 5), "unit": "Msamples/s", "cores": threads, "kind": "port",
                               "sample": f"{rows} middle rows x {W} px x {spp_eff} spp in {dt:.1f}s; a full "
                                         f"image extrapolates to {W * H * spp_eff / (rate * 1e6):.0f}s"}
        out["speedup_vs_cpu"] = round(out["msamples_per_s"] / rate, 1)
    print(json.dumps(out), flush=True)
    ctx.close()


def batch_main(a):
    """K motion-blur scenes from seeds 0..K-1 (each 2,000 moving spheres and a
    bumped icosphere), one context each, rendered by one batched render;
    Msamples/s over all of them, with the per-scene renders one after another
    timed beside it."""
    if a.workload != "motion":
        raise SystemExit("--scenes is for the motion workload")
    W, H = int(1280 * a.scale), int(720 * a.scale)
    ctxs, cams = [], []
    t0 = time.time()
    for k in range(a.scenes):
        sc, cam = workloads.motion_scene(W, H, spp=a.spp or 128, max_depth=10, seed=k)
        ctx = GpuContext(0)
        ctx.upload_scene(sc)
        ctxs.append(ctx)
        cams.append(cam)
        print(f"scene {k} built and uploaded ({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)
    tuned = None
    if a.autotune:
        entries, best = ctxs[0].autotune_render(cams[0], seed=a.seed)
        e = entries[best]
        tuned = {"block": e["block"], "waves": e["waves"], "chunk": e["chunk"],
                 "loop": "while-while" if e["loop"] else "if-if", "candidates": len(entries)}
    spp_eff = int(math.isqrt(cams[0].SampleCount)) ** 2
    if tuned:
        for ctx in ctxs:
            ctx.set_render_tuning(chunk=tuned["chunk"], block=tuned["block"], waves=tuned["waves"],
                                  loop=1 if tuned["loop"] == "while-while" else 0)
    # Both paths are warmed up by the same full render, untimed, so that
    # their queues have their steady-state sizes when the timed render starts.
    GpuContext.render_batch(ctxs, cams, seed=a.seed)
    _, t_batch = GpuContext.render_batch(ctxs, cams, seed=a.seed)
    print(f"batched render {t_batch:.2f}s", file=sys.stderr, flush=True)
    t_seq = 0.0
    for c in ctxs:
        c.trim()
    for ctx, cam in zip(ctxs, cams):
        ctx.render(cam, seed=a.seed)
        ctx.synchronize()
        t1 = time.perf_counter()
        ctx.render(cam, seed=a.seed)
        t_seq += time.perf_counter() - t1
        ctx.trim()                  # one context's full-size queues at a time
    total = a.scenes * W * H * spp_eff
    print(json.dumps({"workload": f"motion x{a.scenes} (seeds 0..{a.scenes - 1})", "width": W, "height": H,
                      "spp": spp_eff, "scenes": a.scenes, "batched_s": round(t_batch, 3),
                      "batched_msamples_per_s": round(total / t_batch / 1e6, 3),
                      "sequential_s": round(t_seq, 3), "sequential_msamples_per_s": round(total / t_seq / 1e6, 3),
                      "autotune": tuned}), flush=True)
    for c in ctxs:
        c.close()


if __name__ == "__main__":
    main()
