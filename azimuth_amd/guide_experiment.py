"""Equal-time comparison of the guided and the plain wavefront tracer on the
architectural interior (environment with a sun lobe, 8 spots):

    python -m azimuth_amd.guide_experiment [--scale S] [--ref-spp 4096]
        [--spp 64] [--seed 1] [--seeds 4]

The reference is a plain render with --ref-spp samples per pixel accumulated
in double precision (chunks of 256 spp with distinct seeds), its environment
directions drawn from the map's luminance distribution; its own remaining
noise, estimated from its two halves, is reported and subtracted in the
'net' ratios.  The plain tracer renders --spp samples per pixel (time
T); the guided tracer (environment directions from the learned guide mixed
50/50 with the reference's cosine sampling, trained online during its first
32 samples per pixel) renders the largest square spp whose measured time stays
within T.  Times are wall-clock around the render call (guide creation and
training included), after a device synchronize.  Because both estimators are unbiased, MSE scales as 1/time, and
the report also gives the time-normalized ratio (MSE_plain * T_plain) /
(MSE_guided * T_guided).  For comparison it also renders with the
environment map's luminance distribution, alone and as the guide's base
distribution, and at the same spp with the scatter guide (hemisphere scatter
directions) alone and beside the environment guide.  Every MSE is the mean over --seeds renders with distinct seeds
(the per-image MSE of this scene varies by tens of percent from seed to
seed), every time the mean of their times.  One JSON line is printed.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import native as N, workloads
from .scene import GpuContext


def resolve(px):
    return px[:, :3] / np.maximum(px[:, 3:4], 1e-300)


def timed(ctx, cam, filt, seed, spp, guide, env=N.AZ_ENV_COSINE, max_depth=0, target=N.AZ_GUIDE_ENV):
    px = ctx.render_f64(cam, filt=filt, seed=seed, spp=spp, guide=guide, env_sampling=env, guide_max_depth=max_depth,
                        guide_target=target)
    return resolve(px), ctx.last_render_s


def equal_time(ctx, cam, filt, seed, spp, t_plain, t_same, max_depth):
    """The largest square spp whose guided render fits in t_plain (2%)."""
    s = int(math.isqrt(max(1, int(spp * t_plain / t_same))))
    _, t = timed(ctx, cam, filt, seed, s * s, True, max_depth=max_depth)
    while t > t_plain * 1.02 and s > 1:
        s -= 1
        _, t = timed(ctx, cam, filt, seed, s * s, True, max_depth=max_depth)
    _, t_up = timed(ctx, cam, filt, seed, (s + 1) * (s + 1), True, max_depth=max_depth)
    return (s + 1) * (s + 1) if t_up <= t_plain * 1.02 else s * s


def measure(ctx, cam, filt, ref, seeds, spp, guide, env=N.AZ_ENV_COSINE, max_depth=0, target=N.AZ_GUIDE_ENV):
    """Mean MSE against ref and mean time over renders with the given seeds;
    also the per-seed MSEs."""
    mses, ts = [], []
    for s in seeds:
        img, t = timed(ctx, cam, filt, s, spp, guide, env, max_depth, target)
        mses.append(float(np.mean((img - ref) ** 2)))
        ts.append(t)
    return float(np.mean(mses)), float(np.mean(ts)), mses


def reference(ctx, cam, filt, seed, spp):
    """Plain render of spp samples per pixel in double precision with the
    environment's luminance distribution (unbiased, the same expectation as
    the cosine sampling, far less variance on this scene), as two halves of
    chunks of 256 spp with distinct seeds (sums of colour and weight add up).
    Returns the image and E[(a - b)^2] / 4 over the pixels of the two halves'
    images a and b: the variance left in the reference, which adds to every
    MSE measured against it."""
    chunk = min(spp // 2, 256)
    halves = []
    for h in range(2):
        acc = None
        for i in range(spp // (2 * chunk)):
            px = ctx.render_f64(cam, filt=filt, seed=seed + 100 * h + i, spp=chunk, env_sampling=N.AZ_ENV_LUMINANCE)
            acc = px if acc is None else acc + px
            print(f"reference half {h} chunk {i + 1}/{spp // (2 * chunk)} ({ctx.last_render_s:.1f}s)",
                  file=sys.stderr, flush=True)
        halves.append(acc)
    noise = float(np.mean((resolve(halves[0]) - resolve(halves[1])) ** 2)) / 4.0
    return resolve(halves[0] + halves[1]), noise


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--ref-spp", type=int, default=4096)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--seeds", type=int, default=4)
    a = ap.parse_args(argv)
    seeds = [a.seed + 11 * k for k in range(a.seeds)]
    W, H = int(1920 * a.scale), int(1080 * a.scale)
    sc, cam = workloads.interior_scene(W, H, spp=a.spp, max_depth=6)
    filt = N.FilterType_Gaussian
    ctx = GpuContext(0)
    ctx.upload_scene(sc)
    ctx.render(cam, filt=filt, spp=1, row_end=4)                      # warm-up
    t0 = time.time()
    ref, ref_noise = reference(ctx, cam, filt, 1000 * (a.seed + 1), a.ref_spp)
    t_ref = time.time() - t0
    print(f"reference {a.ref_spp} spp in {t_ref:.1f}s, its own noise {ref_noise:.3e}", file=sys.stderr, flush=True)
    timed(ctx, cam, filt, a.seed, a.spp, False)                        # warm-up at full size
    mse_plain, t_plain, per_plain = measure(ctx, cam, filt, ref, seeds, a.spp, False)
    print(f"plain {a.spp} spp {t_plain:.2f}s mse {mse_plain:.4e}", file=sys.stderr, flush=True)
    # Guided at the same spp gives its cost per sample; then the largest
    # square spp that fits in the plain time.
    mse_g_same, t_g_same, _ = measure(ctx, cam, filt, ref, seeds, a.spp, True)
    print(f"guided {a.spp} spp {t_g_same:.2f}s mse {mse_g_same:.4e}", file=sys.stderr, flush=True)
    spp_eq = equal_time(ctx, cam, filt, a.seed + 7, a.spp, t_plain, t_g_same, 0)
    mse_g_eq, t_g_eq, per_g_eq = measure(ctx, cam, filt, ref, seeds, spp_eq, True)
    print(f"guided {spp_eq} spp {t_g_eq:.2f}s mse {mse_g_eq:.4e}", file=sys.stderr, flush=True)
    # The guide up to path depth 3 only (deeper vertices: the reference's
    # cosine sampling alone), a cheaper setting of the same estimator.
    _, t_d3_same, _ = measure(ctx, cam, filt, ref, seeds[:1], a.spp, True, max_depth=3)
    spp_d3 = equal_time(ctx, cam, filt, a.seed + 7, a.spp, t_plain, t_d3_same, 3)
    mse_d3, t_d3, _ = measure(ctx, cam, filt, ref, seeds, spp_d3, True, max_depth=3)
    print(f"guided to depth 3: {spp_d3} spp {t_d3:.2f}s mse {mse_d3:.4e}", file=sys.stderr, flush=True)
    # Beside the learned guide: the environment map's luminance distribution
    # (no learning; samples the sun directly), alone and as the guide's base.
    mse_lum, t_lum, _ = measure(ctx, cam, filt, ref, seeds, a.spp, False, N.AZ_ENV_LUMINANCE)
    mse_glum, t_glum, _ = measure(ctx, cam, filt, ref, seeds, a.spp, True, N.AZ_ENV_LUMINANCE)
    mse_glum3, t_glum3, _ = measure(ctx, cam, filt, ref, seeds, a.spp, True, N.AZ_ENV_LUMINANCE, max_depth=3)
    print(f"luminance {a.spp} spp {t_lum:.2f}s mse {mse_lum:.4e}; guide+luminance {t_glum:.2f}s mse {mse_glum:.4e}",
          file=sys.stderr, flush=True)
    # The scatter guide (hemisphere scatter directions mixed with the
    # reference's cosine sampling), alone and beside the environment guide.
    mse_sc, t_sc, _ = measure(ctx, cam, filt, ref, seeds, a.spp, True, target=N.AZ_GUIDE_SCATTER)
    mse_both, t_both, _ = measure(ctx, cam, filt, ref, seeds, a.spp, True, target=N.AZ_GUIDE_BOTH)
    print(f"scatter guide {a.spp} spp {t_sc:.2f}s mse {mse_sc:.4e}; both guides {t_both:.2f}s mse {mse_both:.4e}",
          file=sys.stderr, flush=True)
    out = {
        "experiment": "equal-time MSE vs reference, guided vs plain",
        "workload": f"architectural interior {W}x{H} ({sc.triangle_count} triangles), Gaussian filter, depth 6",
        "reference": {"spp": a.ref_spp, "accumulation": "f64", "environment_sampling": "luminance",
                      "seconds": round(t_ref, 2), "noise_mse": ref_noise},
        "seeds": seeds,
        "plain": {"spp": a.spp, "seconds": round(t_plain, 3), "mse": mse_plain, "mse_per_seed": per_plain},
        "guided_same_spp": {"spp": a.spp, "seconds": round(t_g_same, 3), "mse": mse_g_same},
        "guided_equal_time": {"spp": spp_eq, "seconds": round(t_g_eq, 3), "mse": mse_g_eq, "mse_per_seed": per_g_eq},
        "mse_ratio_equal_time": round(mse_plain / mse_g_eq, 3),
        "net_mse_ratio_equal_time": round((mse_plain - ref_noise) / (mse_g_eq - ref_noise), 3),
        "mse_ratio_time_normalized": round((mse_plain * t_plain) / (mse_g_eq * t_g_eq), 3),
        "mse_ratio_same_spp": round(mse_plain / mse_g_same, 3),
        "guided_to_depth3_equal_time": {"spp": spp_d3, "seconds": round(t_d3, 3), "mse": mse_d3,
                                        "mse_ratio": round(mse_plain / mse_d3, 3)},
        "luminance_env_sampling": {"spp": a.spp, "seconds": round(t_lum, 3), "mse": mse_lum},
        "guided_over_luminance": {"spp": a.spp, "seconds": round(t_glum, 3), "mse": mse_glum},
        "net_mse_ratio_luminance_over_guided_over_luminance_same_spp":
            round((mse_lum - ref_noise) / (mse_glum - ref_noise), 3),
        "guided_over_luminance_to_depth3": {"spp": a.spp, "seconds": round(t_glum3, 3), "mse": mse_glum3},
        "time_normalized_ratio_luminance_over_guided_over_luminance_to_depth3":
            round((mse_lum - ref_noise) * t_lum / ((mse_glum3 - ref_noise) * t_glum3), 3),
        "scatter_guide_same_spp": {"spp": a.spp, "seconds": round(t_sc, 3), "mse": mse_sc},
        "both_guides_same_spp": {"spp": a.spp, "seconds": round(t_both, 3), "mse": mse_both},
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
