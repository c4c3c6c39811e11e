"""Upload timing with the host and the device BVH builders:

    python -m azimuth_amd.build_bench [--subdiv 8] [--repeats 3]

Scenes: the displaced icosphere of the traversal workload (subdivision 8:
1,310,720 triangles) and one motion-blur scene (2,000 moving spheres and a
327,680-triangle displaced icosphere).  Each upload (scene flattening, BVH
build, collapse into wide nodes, triangle packing, copies to the device) is
timed on the host clock around az_gpu_upload_scene after a warm-up upload;
both builders give the same tree, so the node counts are printed side by
side.  One JSON line is printed.
"""
from __future__ import annotations

import argparse
import json
import time

from azimuth_amd import configs, native as N, workloads
from azimuth_amd.scene import GpuContext


def timed_uploads(ctx, sc, mode, repeats):
    ctx.set_build_mode(mode)
    ctx.upload_scene(sc)
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        ctx.upload_scene(sc)
        ts.append(time.perf_counter() - t0)
    st = ctx.stats()
    ctx.set_build_mode(N.AZ_BUILD_AUTO)
    return min(ts), st.NodeCount


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--subdiv", type=int, default=8)
    ap.add_argument("--repeats", type=int, default=3)
    a = ap.parse_args(argv)
    ctx = GpuContext(0)
    scenes = {f"icosphere subdivision {a.subdiv}": configs.icosphere_scene(a.subdiv, seed=1),
              "motion-blur scene (seed 0)": workloads.motion_scene(64, 36, spp=1, max_depth=2)[0]}
    out = {"workload": "scene upload with the host and the device BVH builders"}
    for name, sc in scenes.items():
        th, nh = timed_uploads(ctx, sc, N.AZ_BUILD_HOST, a.repeats)
        td, nd = timed_uploads(ctx, sc, N.AZ_BUILD_DEVICE, a.repeats)
        out[name] = {"triangles": sc.triangle_count, "nodes_host": nh, "nodes_device": nd,
                     "upload_s_host": round(th, 4), "upload_s_device": round(td, 4)}
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
