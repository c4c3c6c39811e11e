"""Timing of the output stage (az_gpu_output_pixels) on a device-resident
image:

    python -m azimuth_amd.output_bench [--width 1920] [--height 1080] [--repeats 20]

The image is a synthetic HDR pixel buffer (log-normal colours, weights in
(0.5, 4]).  Each configuration is timed with HIP events on the context stream
around one call, after a warm-up call; the report gives the mean time and the
algorithmic bytes per pixel over it: 16 read + 3 written without
tonemapping; with it 16 read + 8 written (luminance key and log), 4 read by
the log sum, 3 x 4 by the radix passes, 16 + 3 by the final pass.  One JSON
line is printed.
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from azimuth_amd import native as N
from azimuth_amd.scene import GpuContext


def hdr_pixels(w, h, seed=7):
    rng = np.random.default_rng(seed)
    n = w * h
    px = np.zeros(n, dtype=N.PIXEL_DTYPE)
    wgt = rng.uniform(0.5, 4.0, n).astype(np.float32)
    px["Color"] = np.exp(rng.normal(-1.5, 1.5, (n, 3))).astype(np.float32) * wgt[:, None]
    px["TotalWeight"] = wgt
    return px


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--repeats", type=int, default=20)
    a = ap.parse_args(argv)
    w, h = a.width, a.height
    n = w * h
    ctx = GpuContext(0)
    d_px, d_out = ctx.alloc(n * 16), ctx.alloc(n * 3)
    d_px.upload(hdr_pixels(w, h))
    tm = N.tonemap()
    tm.KeyValue, tm.BurnoutPercentage, tm.Saturation = 0.18, 1.0, 1.0
    tm.Operator = N.TonemapOperator_Photographic
    configs = {"photographic_srgb": (tm, 1.0 / 2.4, True, 16 + 8 + 4 + 12 + 16 + 3),
               "gamma_only": (None, 1.0 / 2.2, False, 16 + 3)}
    out = {"workload": f"output stage, {w}x{h} HDR pixel buffer in device memory", "pixels": n}
    for name, (t, g, srgb, bpp) in configs.items():
        ctx.output_pixels_device(d_px, w, h, d_out, t, g, srgb)
        ms = []
        for _ in range(a.repeats):
            ctx.timer_start()
            ctx.output_pixels_device(d_px, w, h, d_out, t, g, srgb)
            ms.append(ctx.timer_stop())
        m = float(np.mean(ms))
        out[name] = {"ms": round(m, 4), "mpixels_per_s": round(n / m / 1e3, 1),
                     "bytes_per_pixel": bpp, "gb_per_s": round(n * bpp / m / 1e6, 1)}
    d_px.free()
    d_out.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
