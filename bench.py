#!/usr/bin/env python3
"""Closest-hit traversal benchmark on one MI355X.

Workload (one step): the 1,310,720-triangle displaced icosphere (icosphere
subdivision 8, 3-octave gradient-noise displacement of amplitude 0.05, seed
1), traced with two batches of 4,194,304 rays already resident in HBM: the
coherent batch of a 2048x2048 pinhole camera and the incoherent batch with
origins uniform in the mesh box and directions uniform on the sphere.  The
value is closest-hit Mrays/s over both batches.

Before timing, the per-scene autotuner picks the traversal variant (LDS
stack entries, workgroup size, register target) on both batches.
Also reported:
- roofline: the traversal kernel's algorithmic bytes per launch (32 B ray in,
  16 B hit out, 128 B per node visited, 48 B per triangle tested, from an
  untimed counting launch) over the launch's HIP-event duration;
- any-hit rates;
- render_interior: the path tracer on the architectural interior
  (1920x1080, 64 spp, Gaussian filter) with its own roofline fraction from a
  counted render's algorithmic bytes; render_other: the Cornell box and one
  motion-blur scene;
- guide: the learned guide's packed inference (8.4M queries) and Adam
  training step (65,536 samples) on their own, algorithmic TFLOP/s against
  the 2.5 PFLOP/s dense BF16 peak;
- equal_time_mse: the metric's second half, MSE against a 4096-spp
  double-precision reference of the guided and the plain tracer at equal
  wall-clock on the guided interior, beside the environment luminance
  distribution (the strongest sampler that does not learn);
- cpu_baseline / cpu_renders: the oracle's reference traversal and recursive
  renderer timed on the host cores (the interior and motion renders
  extrapolated from a strip of rows at 1 spp).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RAY_BYTES, HIT_BYTES = 32, 16
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--block", type=int, default=0, help="workgroup size (0: library default)")
    ap.add_argument("--ring", type=int, default=0, help="LDS stack entries per lane (0: library default)")
    ap.add_argument("--waves", type=int, default=-1, help="register target in waves/SIMD (-1: library default)")
    ap.add_argument("--loop", type=int, default=-1, help="traversal loop form, 0 if-if, 1 while-while (-1: default)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile", action="store_true", help="short run for profilers: no CPU baseline, no any-hit")
    ap.add_argument("--no-render", action="store_true", help="skip the render lines")
    ap.add_argument("--no-mse", action="store_true", help="skip the equal-time MSE leg (about a minute)")
    ap.add_argument("--tune-build", action="store_true",
                    help="also autotune the BVH build (node width 2-4 x leaf size 2/4/8, one upload each)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; sharding across GPUs is not "
                                   "built"}))
        sys.exit(2)

    from azimuth_amd import configs, native as N
    from azimuth_amd.scene import GpuContext

    t0 = time.time()
    sc = configs.icosphere_scene(8, seed=1)
    coh = configs.coherent_rays(2048, 2048)
    inc = configs.incoherent_rays(4194304, sc.meta["bbox"], seed=2)
    n_coh, n_inc = coh.shape[0], inc.shape[0]
    log(f"scene {sc.triangle_count} triangles built in {time.time() - t0:.1f}s")

    ctx = GpuContext(0)
    t0 = time.time()
    ctx.upload_scene(sc)
    log(f"upload + device BVH build {time.time() - t0:.1f}s")
    d_coh, d_inc = ctx.alloc(coh.nbytes), ctx.alloc(inc.nbytes)
    d_coh.upload(coh)
    d_inc.upload(inc)
    d_hits = ctx.alloc(max(n_coh, n_inc) * HIT_BYTES)
    tuned = None
    if args.block or args.ring or args.waves >= 0 or args.loop >= 0:
        ctx.set_launch(block=args.block, ring=args.ring, waves=args.waves, loop=args.loop)
    else:
        # Per-scene autotune (untimed): every compiled traversal variant on
        # both batches; the fastest is kept.
        sample = np.concatenate([coh, inc])
        d_tune = ctx.alloc(sample.nbytes)
        d_tune.upload(sample)
        entries, best = ctx.autotune_trace(d_tune, sample.shape[0], repeats=3)
        tuned = {"candidates": len(entries), "ring": entries[best]["ring"], "block": entries[best]["block"],
                 "waves": entries[best]["waves"], "loop": "while-while" if entries[best]["loop"] else "if-if",
                 "ms_per_1m_rays": {f'{e["ring"]}/{e["block"]}/{e["waves"]}/{e["loop"]}':
                                    round(e["ms"] * 1e6 / sample.shape[0], 4) for e in entries}}
        log(f"autotune: ring {tuned['ring']} block {tuned['block']} waves {tuned['waves']} loop {tuned['loop']}")
        if args.tune_build:
            # One upload (host BVH build) per candidate, timed with the tuned
            # launch; the scene stays uploaded with the fastest build.
            t0 = time.time()
            bent, bbest = ctx.autotune_build(sc, d_tune, sample.shape[0], repeats=2)
            tuned["build"] = {"width": bent[bbest]["width"], "leaf": bent[bbest]["leaf"],
                              "ms_per_1m_rays": {f'w{e["width"]}/l{e["leaf"]}': round(e["ms"] * 1e6 / sample.shape[0], 4)
                                                 for e in bent}}
            log(f"build autotune ({time.time() - t0:.0f}s): width {bent[bbest]['width']} leaf {bent[bbest]['leaf']}")
        d_tune.free()

    # Kernel-counted work per batch (an untimed counting launch).
    ctx.set_counting(True)
    work = {}
    for name, d, n in (("coherent", d_coh, n_coh), ("incoherent", d_inc, n_inc)):
        ctx.trace_closest_device(d, n, d_hits)
        s = ctx.stats()
        work[name] = (s.NodesVisited, s.TrianglesTested, s.PrimitivesTested)
    ctx.set_counting(False)
    st = ctx.stats()
    node_b, tri_b = st.NodeBytes, st.TriangleBytes

    for _ in range(args.warmup):
        ctx.trace_closest_device(d_coh, n_coh, d_hits)
        ctx.trace_closest_device(d_inc, n_inc, d_hits)
    ctx.synchronize()

    ctx.record_launch_times(True)
    ctx.timer_start()
    for _ in range(args.steps):
        ctx.trace_closest_device(d_coh, n_coh, d_hits)
        ctx.trace_closest_device(d_inc, n_inc, d_hits)
    total_ms = ctx.timer_stop()
    lt = ctx.collect_launch_times()
    ctx.record_launch_times(False)
    ms_coh, ms_inc = float(np.mean(lt[0::2])), float(np.mean(lt[1::2]))
    rays_per_step = n_coh + n_inc
    value = rays_per_step * args.steps / (total_ms * 1e-3) / 1e6

    def alg_bytes(name, n, items=False):
        nodes, tris, prims = work[name]
        return n * (RAY_BYTES + HIT_BYTES) + nodes * node_b + tris * tri_b + (prims * 144 if items else 0)

    b_coh, b_inc = alg_bytes("coherent", n_coh), alg_bytes("incoherent", n_inc)
    achieved = (b_coh + b_inc) / ((ms_coh + ms_inc) * 1e-3) / 1e9
    # The same with the 144-byte top-level item record each walk reads (not
    # in the per-ray formula above).
    achieved_items = (alg_bytes("coherent", n_coh, True) + alg_bytes("incoherent", n_inc, True)) / \
        ((ms_coh + ms_inc) * 1e-3) / 1e9

    extra = {
        "coherent_mrays": n_coh / (ms_coh * 1e-3) / 1e6,
        "incoherent_mrays": n_inc / (ms_inc * 1e-3) / 1e6,
        "coherent_kernel_ms": ms_coh,
        "incoherent_kernel_ms": ms_inc,
        "nodes_per_ray": {k: work[k][0] / n for k, n in (("coherent", n_coh), ("incoherent", n_inc))},
        "tris_per_ray": {k: work[k][1] / n for k, n in (("coherent", n_coh), ("incoherent", n_inc))},
        "coherent_gbps": b_coh / (ms_coh * 1e-3) / 1e9,
        "incoherent_gbps": b_inc / (ms_inc * 1e-3) / 1e9,
        "achieved_gbps_with_item_records": achieved_items,
    }

    if args.profile:
        # One calibration copy of known size for the HBM counters.
        nb = 1 << 30
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.device_copy(b, a, nb)
        ctx.synchronize()
        a.free()
        b.free()
    else:
        # Any-hit rates (reported beside the headline; TMax = +inf).
        ctx.record_launch_times(True)
        for _ in range(max(1, args.steps // 4)):
            ctx.trace_any_device(d_coh, n_coh, d_hits)
            ctx.trace_any_device(d_inc, n_inc, d_hits)
        at = ctx.collect_launch_times()
        ctx.record_launch_times(False)
        extra["anyhit_coherent_mrays"] = n_coh / (float(np.mean(at[0::2])) * 1e-3) / 1e6
        extra["anyhit_incoherent_mrays"] = n_inc / (float(np.mean(at[1::2])) * 1e-3) / 1e6
        # HBM reference: a plain 1 GiB streaming copy.
        nb = 1 << 30
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.device_copy(b, a, nb)
        ctx.synchronize()
        ctx.timer_start()
        for _ in range(5):
            ctx.device_copy(b, a, nb)
        ms = ctx.timer_stop() / 5
        extra["copy_kernel_gbps"] = 2 * nb / (ms * 1e-3) / 1e9
        a.free()
        b.free()

    # HBM-side bytes per closest-hit launch, from the rocprofv3 FETCH_SIZE and
    # WRITE_SIZE runs summarized in profiles/ (scaled by a plain copy kernel).
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "trace_icosphere_closest.md")
    if os.path.exists(tpath):
        with open(tpath) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch"):
                    traffic = int(ln.split("=")[1])

    cpu = None
    if not (args.no_cpu_baseline or args.profile):
        cpu = cpu_baseline(sc, coh, inc, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")
    if not (args.profile or args.no_render):
        extra["render_interior"] = render_interior(ctx)
        extra["guide"] = guide_kernels(ctx)
        log(f"guide kernels: {extra['guide']}")
        log(f"render interior: {extra['render_interior']}")
        extra["render_other"] = render_other(ctx)
    if not (args.profile or args.no_render or args.no_mse):
        extra["equal_time_mse"] = equal_time_mse(ctx)
    if not (args.no_cpu_baseline or args.profile):
        extra["cpu_renders"] = cpu_renders()
        log(f"cpu renders: {extra['cpu_renders']}")

    line = {
        "metric": "Mrays/s closest-hit; equal-time MSE vs 4096spp ref, guided vs plain (1 GPU)",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "closest-hit traversal, displaced icosphere subdivision 8 (1,310,720 triangles), "
                        "4,194,304 coherent (2048x2048 pinhole) + 4,194,304 incoherent rays per step",
            "triangles": int(sc.triangle_count),
            "rays_per_step": int(rays_per_step),
            "bvh": f"{st.BvhWidth}-wide, {node_b} B nodes, {tri_b} B triangles, stack {st.StackDepth}",
            "block": int(st.BlockSize),
            "autotune": tuned,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            # The HBM side of the same launches: the profiled bytes per launch
            # over this run's mean launch time.  Algorithmic bytes above the
            # peak mean the L2 and the MALL serve most node fetches; the
            # kernel is then bound by VALU issue, not by HBM.
            "hbm_side_gbps": round(traffic / ((ms_coh + ms_inc) / 2 * 1e-3) / 1e9, 1) if traffic else None,
            "hbm_side_frac": round(traffic / ((ms_coh + ms_inc) / 2 * 1e-3) / 1e9 / HBM_PEAK_GBPS, 4)
            if traffic else None,
        },
        "cpu_baseline": cpu,
    }
    line.update({k: (round(v, 4) if isinstance(v, float) else v) for k, v in extra.items()})
    print(json.dumps(line), flush=True)
    ctx.close()


def cpu_threads():
    """Host threads for the CPU baselines: the cores this process may run
    on, at most the host share of a one-GPU box (OMP_NUM_THREADS there)."""
    n = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(n, cap) if cap > 0 else n), n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def render_bytes(st, node_b, tri_b, item_b=144):
    """Algorithmic bytes of a render from its counts (az_render_stats): camera
    rays written (64 B) with their accumulators (24 B); per traced path ray
    64 B in and a 32 B hit out plus the node, triangle and item records its
    walk fetched; per shading point the ray, the hit and 112 B of attributes
    in, 64 B per child and per shadow ray out; per shadow ray 64 B in, its
    walk's records and a 24 B accumulator add; the splat reads 24 B per
    sample (31 % more with the Gaussian filter's halo) and 32 B per pixel and
    chunk."""
    gen = 88 * st["CameraSamples"]
    ext = 96 * st["PathRays"] + node_b * st["ExtendNodes"] + tri_b * st["ExtendTriangles"] + item_b * st["ExtendItems"]
    shade = 208 * st["PathRays"] + 64 * (st["PathRays"] - st["CameraSamples"]) + 64 * st["ShadowRays"]
    conn = 88 * st["ShadowRays"] + node_b * st["ConnectNodes"] + tri_b * st["ConnectTriangles"] + \
        item_b * st["ConnectItems"]
    splat = 24 * st["CameraSamples"] * (1.31 if st["Gaussian"] else 1.0) + 32 * st["Pixels"] * st["Chunks"]
    return {"generate": gen, "extend": ext, "shade": shade, "connect": conn, "splat": splat,
            "total": gen + ext + shade + conn + splat}


def render_interior(ctx):
    """Beside the traversal line: the full path tracer on the architectural
    interior (299,722 triangles, 8 spot lights, HDR sky with a sun, 1920x1080,
    64 spp, depth 6, Gaussian filter), autotuned, one timed render after a
    warm-up render (HIP events around the render call); then an untimed
    counted render of the same samples gives the render's algorithmic bytes,
    whose rate over the timed render is its HBM roofline fraction."""
    from azimuth_amd import native as N, workloads
    sc, cam = workloads.interior_scene(1920, 1080, spp=64, max_depth=6)
    ctx.upload_scene(sc)
    entries, best = ctx.autotune_render(cam, filt=N.FilterType_Gaussian, seed=1)
    ctx.render(cam, filt=N.FilterType_Gaussian, seed=1)
    ctx.timer_start()
    ctx.render(cam, filt=N.FilterType_Gaussian, seed=1)
    ms = ctx.timer_stop()
    ctx.set_counting(True)
    ctx.render(cam, filt=N.FilterType_Gaussian, seed=1)
    ctx.set_counting(False)
    st = ctx.render_stats()
    g = ctx.stats()
    nb = render_bytes(st, g.NodeBytes, g.TriangleBytes)
    achieved = nb["total"] / (ms * 1e-3) / 1e9
    e = entries[best]
    return {"msamples_per_s": round(1920 * 1080 * 64 / (ms * 1e-3) / 1e6, 2), "render_ms": round(ms, 2),
            "tuning": f'{e["block"]} lanes, {e["waves"]} waves, {"while-while" if e["loop"] else "if-if"}, '
                      f'{e["chunk"]} samples per chunk',
            "work": {"path_rays": st["PathRays"], "shadow_rays": st["ShadowRays"],
                     "nodes_per_path_ray": round(st["ExtendNodes"] / max(1, st["PathRays"]), 3),
                     "tris_per_path_ray": round(st["ExtendTriangles"] / max(1, st["PathRays"]), 3),
                     "nodes_per_shadow_ray": round(st["ConnectNodes"] / max(1, st["ShadowRays"]), 3),
                     "tris_per_shadow_ray": round(st["ConnectTriangles"] / max(1, st["ShadowRays"]), 3)},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBPS, 4),
                         "bytes_by_stage_gb": {k: round(v / 1e9, 3) for k, v in nb.items()}}}


GUIDE_FLOPS_PER_QUERY = 2 * 23040                      # forward: 128-64-64-64-64-40 MACs x 2
# forward, the gradient of layers 1-4's inputs (64 x 64 x 3 + 64 x 40 MACs), weight gradients
GUIDE_FLOPS_PER_SAMPLE = 2 * (23040 + 14848 + 23040)
BF16_PEAK_TFLOPS = 2500.0


def guide_kernels(ctx, n_query=8388608, batch=65536, reps=10, steps=64):
    """The learned guide's MFMA kernels on their own, inputs in HBM: the
    renderer's packed inference over n_query valid queries (random positions,
    normals and roughness), and Adam training steps on batches of `batch`
    samples; algorithmic TFLOP/s against the 2.5 PFLOP/s dense BF16 peak.
    Wall-clock around the launches with the guide's stream synchronized on
    both sides."""
    import ctypes as C
    from azimuth_amd import native as N
    lib = N.load()
    g = C.c_void_p()
    N.check(lib.az_guide_create(C.byref(g), 0, 7), "az_guide_create")
    rng = np.random.default_rng(3)
    nrm = rng.normal(size=(n_query, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    q = np.zeros((n_query, 8), np.float32)
    q[:, 0:3], q[:, 3:6], q[:, 6] = rng.random((n_query, 3)), nrm, rng.random(n_query)
    dq, dl = ctx.alloc(q.nbytes), ctx.alloc(n_query * 64)
    dq.upload(q)
    smp = np.zeros((batch, 12), np.float32)
    d = rng.normal(size=(batch, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    smp[:, 0:3], smp[:, 3:6], smp[:, 6], smp[:, 7:10], smp[:, 10] = q[:batch, 0:3], q[:batch, 3:6], \
        q[:batch, 6], d, 1.0
    ds = ctx.alloc(smp.nbytes)
    ds.upload(smp)
    ctx.synchronize()
    chk = lambda st, what: N.check(st, what)
    chk(lib.az_guide_infer_packed_device(g, dq.ptr, n_query, dl.ptr), "az_guide_infer_packed_device")
    chk(lib.az_guide_synchronize(g), "az_guide_synchronize")
    t0 = time.perf_counter()
    for _ in range(reps):
        chk(lib.az_guide_infer_packed_device(g, dq.ptr, n_query, dl.ptr), "az_guide_infer_packed_device")
    chk(lib.az_guide_synchronize(g), "az_guide_synchronize")
    t_inf = (time.perf_counter() - t0) / reps
    chk(lib.az_guide_train_device(g, ds.ptr, batch, 1, None), "az_guide_train_device")
    chk(lib.az_guide_synchronize(g), "az_guide_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        chk(lib.az_guide_train_device(g, ds.ptr, batch, 1, None), "az_guide_train_device")
    chk(lib.az_guide_synchronize(g), "az_guide_synchronize")
    t_step = (time.perf_counter() - t0) / steps
    lib.az_guide_destroy(g)
    for b in (dq, dl, ds):
        b.free()
    inf_tf = n_query * GUIDE_FLOPS_PER_QUERY / t_inf / 1e12
    step_tf = batch * GUIDE_FLOPS_PER_SAMPLE / t_step / 1e12
    return {"inference": {"queries": n_query, "ms": round(t_inf * 1e3, 3), "tflops": round(inf_tf, 1),
                          "frac_of_bf16_peak": round(inf_tf / BF16_PEAK_TFLOPS, 4)},
            "train_step": {"batch": batch, "ms": round(t_step * 1e3, 4), "tflops": round(step_tf, 1),
                           "frac_of_bf16_peak": round(step_tf / BF16_PEAK_TFLOPS, 4)},
            "peak_tflops": BF16_PEAK_TFLOPS, "flops_per_query": GUIDE_FLOPS_PER_QUERY,
            "flops_per_training_sample": GUIDE_FLOPS_PER_SAMPLE}


def render_other(ctx):
    """The Cornell box (256x256, 16 spp, depth 8, box filter) and one motion-
    blur scene (2,000 moving spheres and 327,680 triangles, 1280x720, 121 spp,
    depth 10), default tuning, timed after a warm-up render."""
    from azimuth_amd import configs, native as N, workloads
    out = {}
    for name, (sc, cam), spp_eff in (("cornell", configs.cornell_scene(256, 256, spp=16, max_depth=8), 16),
                                     ("motion", workloads.motion_scene(1280, 720, spp=128, max_depth=10), 121)):
        ctx.upload_scene(sc)
        ctx.render(cam, seed=1)
        ctx.timer_start()
        ctx.render(cam, seed=1)
        ms = ctx.timer_stop()
        W, H = cam.Resolution.x, cam.Resolution.y
        out[name] = {"msamples_per_s": round(W * H * spp_eff / (ms * 1e-3) / 1e6, 2), "render_ms": round(ms, 2)}
        log(f"render {name}: {out[name]}")
    return out


def equal_time_mse(ctx, ref_spp=4096, spp=64, seeds=(1, 12)):
    """The guided-interior workload: MSE against a ref_spp render accumulated
    in double precision, of the plain tracer at spp and of the guided tracer
    (learned environment-direction guide, one-sample MIS at fraction 0.5
    against the reference's cosine sampling, trained on the first quarter of
    its spp) at the largest square spp whose render time stays within the
    plain render's; also the environment luminance distribution (the strongest
    sampler that does not learn) at equal time, and the guide over it, used up
    to path depth 1, 2 or 3 (deeper vertices take the luminance distribution
    alone; each depth at its own equal-time spp, the best reported with all
    three).  Besides the equal-time ratios, MSE x seconds ratios.  Wall-
    clock times around each render call, guide training included; each MSE is
    the mean over the given seeds.

    The reference is the plain tracer with the luminance distribution for the
    environment (unbiased, the same expectation as the cosine sampling, about
    60x less variance per sample on this scene), rendered as two halves of
    ref_spp/2 samples with distinct seeds: E[(a - b)^2] / 4 over the pixels
    estimates the variance left in their mean, which adds to every MSE
    measured against it; 'mse_net' subtracts it."""
    import math
    from azimuth_amd import native as N, workloads
    sc, cam = workloads.interior_scene(1920, 1080, spp=spp, max_depth=6)
    ctx.upload_scene(sc)
    filt = N.FilterType_Gaussian
    resolve = lambda px: px[:, :3] / np.maximum(px[:, 3:4], 1e-300)

    def run(seed, n, guide=False, env=N.AZ_ENV_COSINE, md=0):
        px = ctx.render_f64(cam, filt=filt, seed=seed, spp=n, guide=guide, env_sampling=env, guide_max_depth=md)
        return resolve(px), ctx.last_render_s

    t0 = time.time()
    halves = []
    for h in range(2):
        acc = None
        for i in range(ref_spp // 512):
            px = ctx.render_f64(cam, filt=filt, seed=5000 + 100 * h + i, spp=256, env_sampling=N.AZ_ENV_LUMINANCE)
            acc = px if acc is None else acc + px
        halves.append(acc)
    ref = resolve(halves[0] + halves[1])
    ref_noise = float(np.mean((resolve(halves[0]) - resolve(halves[1])) ** 2)) / 4.0
    t_ref = time.time() - t0
    log(f"reference {ref_spp} spp in {t_ref:.1f}s, its own noise {ref_noise:.3e}")

    def mse(n, guide=False, env=N.AZ_ENV_COSINE, md=0):
        vals, ts = [], []
        for sd in seeds:
            img, t = run(sd, n, guide, env, md)
            vals.append(float(np.mean((img - ref) ** 2)))
            ts.append(t)
        m, t = float(np.mean(vals)), float(np.mean(ts))
        return {"spp": n, "seconds": round(t, 3), "mse": m, "mse_net": m - ref_noise,
                "mse_net_x_seconds": (m - ref_noise) * t}

    def equal_spp(guide, env, t_budget, md=0):
        _, t_same = run(3, spp, guide, env, md)
        s = int(math.isqrt(max(1, int(spp * t_budget / t_same))))
        while s > 1 and run(3, s * s, guide, env, md)[1] > t_budget * 1.02:
            s -= 1
        if run(3, (s + 1) ** 2, guide, env, md)[1] <= t_budget * 1.02:
            s += 1
        return s * s

    run(1, spp)                                           # warm-up
    res = {"workload": "guided architectural interior, 1920x1080, depth 6, Gaussian filter",
           "reference": {"spp": ref_spp, "accumulation": "f64", "environment_sampling": "luminance",
                         "seconds": round(t_ref, 1), "noise_mse": ref_noise},
           "seeds": list(seeds), "guide_target": "environment NEE", "plain": mse(spp)}
    t_plain = res["plain"]["seconds"]
    # Over the cosine distribution the guide is used at every depth (limited
    # to depth 2 its MSE is 4.6x higher); over the luminance distribution it
    # pays only at the first vertices, so depths 1-3 are each measured.
    for name, guide, env, md in (("guided", True, N.AZ_ENV_COSINE, 0), ("luminance", False, N.AZ_ENV_LUMINANCE, 0)):
        res[name] = mse(equal_spp(guide, env, t_plain, md), guide, env, md)
        if guide:
            res[name]["guide_max_depth"] = md
        log(f"{name}: {res[name]}")
    by_depth = {}
    for md in (1, 2, 3):
        by_depth[md] = mse(equal_spp(True, N.AZ_ENV_LUMINANCE, t_plain, md), True, N.AZ_ENV_LUMINANCE, md)
        by_depth[md]["guide_max_depth"] = md
        log(f"guided_over_luminance to depth {md}: {by_depth[md]}")
    res["guided_over_luminance_by_depth"] = {str(md): r for md, r in by_depth.items()}
    res["guided_over_luminance"] = min(by_depth.values(), key=lambda r: r["mse_net"])
    ratio = lambda a, b, k="mse_net": round(res[a][k] / max(res[b][k], 1e-30), 3)
    res["mse_ratio_plain_over_guided"] = round(res["plain"]["mse"] / res["guided"]["mse"], 3)
    res["net_mse_ratio_plain_over_guided"] = ratio("plain", "guided")
    res["net_mse_ratio_luminance_over_guided_over_luminance"] = ratio("luminance", "guided_over_luminance")
    # MSE x seconds (unbiased estimators: MSE falls as 1 / time), free of the
    # square-spp steps of the equal-time renders; the best depth's.
    res["time_normalized_ratio_plain_over_guided"] = ratio("plain", "guided", "mse_net_x_seconds")
    best_t = min(by_depth.values(), key=lambda r: r["mse_net_x_seconds"])
    res["time_normalized_ratio_luminance_over_guided_over_luminance"] = round(
        res["luminance"]["mse_net_x_seconds"] / max(best_t["mse_net_x_seconds"], 1e-30), 3)
    res["time_normalized_guided_over_luminance_depth"] = best_t["guide_max_depth"]
    return res


def cpu_baseline(sc, coh, inc, budget_s):
    """The oracle's reference traversal (reference BVH, no pruning) on the host
    cores, on an equal number of coherent and incoherent rays sized to about
    budget_s seconds."""
    import oracle
    threads, nproc = cpu_threads()
    probe = 4096
    sel_c = coh[:: max(1, coh.shape[0] // probe)][:probe]
    sel_i = inc[:probe]
    t0 = time.time()
    oracle.trace(sc, sel_c, threads=threads)
    oracle.trace(sc, sel_i, threads=threads)
    rate = 2 * probe / max(1e-6, time.time() - t0)
    m = int(min(coh.shape[0], max(probe, rate * budget_s / 2)))
    sel_c = coh[:: max(1, coh.shape[0] // m)][:m]
    sel_i = inc[:m]
    t0 = time.time()
    oracle.trace(sc, sel_c, threads=threads)
    oracle.trace(sc, sel_i, threads=threads)
    dt = time.time() - t0
    return {"value": round(2 * m / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(), "nproc": nproc,
            "sample": f"{m} coherent (every {max(1, coh.shape[0] // m)}th) + {m} incoherent rays of the same "
                      f"batches, closest hit through the reference BVH without pruning, {dt:.1f}s"}


def cpu_renders(rows=8):
    """The oracle's recursive renderer (the reference's ProduceImageInLinearSpace
    over RenderScene's scanline bands, through the POSIX platform table) on
    the host cores: the full Cornell box, and 1 spp over `rows` rows of the
    interior and motion-blur workloads, extrapolated linearly in rows and spp
    to the full renders."""
    import oracle
    from azimuth_amd import configs, native as N, workloads
    threads, _ = cpu_threads()
    out = {}
    sc, cam = configs.cornell_scene(256, 256, spp=16, max_depth=8)
    t0 = time.time()
    oracle.render(sc, cam, seed=1, threads=threads)
    dt = time.time() - t0
    out["cornell"] = {"msamples_per_s": round(256 * 256 * 16 / dt / 1e6, 3), "seconds": round(dt, 2),
                      "sample": "full render, 256x256, 16 spp, depth 8"}
    for name, make, full_spp, filt in (
            ("interior", lambda: workloads.interior_scene(1920, 1080, spp=1, max_depth=6), 64, N.FilterType_Gaussian),
            ("motion", lambda: workloads.motion_scene(1280, 720, spp=1, max_depth=10), 121, N.FilterType_Box)):
        sc, cam = make()
        W, H = cam.Resolution.x, cam.Resolution.y
        r0 = H // 2 - rows // 2
        t0 = time.time()
        oracle.render(sc, cam, filt=filt, seed=1, spp=1, row_begin=r0, row_end=r0 + rows, threads=threads)
        dt = time.time() - t0
        rate = W * rows / dt
        out[name] = {"msamples_per_s": round(rate / 1e6, 4), "extrapolated": True,
                     "full_render_seconds": round(W * H * full_spp / rate, 1),
                     "sample": f"{rows} rows x {W} pixels x 1 spp in {dt:.2f}s, extrapolated to {W}x{H}x{full_spp}"}
    return {"cores": threads, "cpu": cpu_model(), "kind": "port", "renders": out}


if __name__ == "__main__":
    main()
