"""Batched renders: several scenes, one per context, rendered together by
az_gpu_render_batch (one launch per wavefront stage over all of them) must
give each scene the image its own az_gpu_render gives: the same random
streams, the same per-sample fixed-point sums and, when both renders take
every sample in one chunk, the same float additions into the pixels, so
the images are equal bit for bit."""
import numpy as np
import pytest

from azimuth_amd import configs, native as N, workloads
from azimuth_amd.scene import GpuContext

pytestmark = pytest.mark.gpu


def rel_rmse(a, b):
    ia = a[:, :3] / np.maximum(a[:, 3:4], 1e-30)
    ib = b[:, :3] / np.maximum(b[:, 3:4], 1e-30)
    return float(np.sqrt(np.mean((ia - ib) ** 2)) / max(1e-6, float(np.mean(np.abs(ib)))))


def test_batch_matches_single_renders(gpu):
    scenes = [configs.cornell_scene(32, 32, spp=16, max_depth=6),
              configs.feature_scene(40, 40, spp=4, max_depth=4),
              workloads.motion_scene(40, 24, spp=9, max_depth=6, spheres=60, ico_subdiv=3)]
    ctxs = [gpu] + [GpuContext(0) for _ in scenes[1:]]
    try:
        for ctx, (sc, cam) in zip(ctxs, scenes):
            ctx.upload_scene(sc)
        cams = [cam for _, cam in scenes]
        got, seconds = GpuContext.render_batch(ctxs, cams, seed=4)
        for k, (ctx, cam) in enumerate(zip(ctxs, cams)):
            want = ctx.render(cam, seed=4)
            w = np.concatenate([want["Color"], want["TotalWeight"][:, None]], 1)
            assert np.array_equal(got[k][:, 3], w[:, 3]), k
            err = rel_rmse(got[k], w)
            print(f"scene {k}: batched vs single relative RMSE {err:.2e}", flush=True)
            assert np.array_equal(got[k], w), k
        # Gaussian filter and double accumulation through the batch as well.
        got64, _ = GpuContext.render_batch(ctxs[:2], cams[:2], filt=N.FilterType_Gaussian, seed=2, f64=True)
        for k in range(2):
            want = ctxs[k].render(cams[k], filt=N.FilterType_Gaussian, seed=2)
            w = np.concatenate([want["Color"], want["TotalWeight"][:, None]], 1).astype(np.float64)
            assert rel_rmse(got64[k], w) < 1e-5
    finally:
        for c in ctxs[1:]:
            c.close()


def test_full_motion_batch_matches_single_renders(gpu):
    """The batched motion-blur workload at its own size: 16 scenes (seeds
    0..15), 2,000 spheres and 327,680 triangles each, 1280x720, 121 spp,
    depth 10, in one batched render.  Scenes 0 and 15 rendered alone with
    the same samples per chunk are identical bit for bit."""
    from azimuth_amd.scene import GpuContext
    count = 16
    scenes = [workloads.motion_scene(1280, 720, spp=128, max_depth=10, seed=k) for k in range(count)]
    ctxs = [gpu] + [GpuContext(0) for _ in range(count - 1)]
    try:
        for ctx, (sc, cam) in zip(ctxs, scenes):
            ctx.upload_scene(sc)
        chunk = 32 << 20
        gpu.set_render_tuning(chunk=chunk)
        cams = [cam for _, cam in scenes]
        got, seconds = GpuContext.render_batch(ctxs, cams, seed=2)
        print(f"16-scene batch {seconds:.2f}s", flush=True)
        n_pix = 1280 * 720
        per_chunk = max(1, (chunk // count) // n_pix)        # samples per pixel per chunk in the batch
        for k in (0, count - 1):
            ctxs[k].set_render_tuning(chunk=per_chunk * n_pix)
            want = ctxs[k].render(cams[k], seed=2)
            w = np.concatenate([want["Color"], want["TotalWeight"][:, None]], 1)
            assert np.all(w[:, 3] == 121.0)
            assert np.array_equal(got[k], w), k
    finally:
        gpu.set_render_tuning(chunk=16 << 20)
        for c in ctxs[1:]:
            c.close()
