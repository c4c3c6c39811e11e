"""The autotuner picks launch parameters by timing; every candidate it can
pick must produce the same results: traversal variants bit-identical hits,
render variants (traversal workgroup size, register target, chunk size) the
oracle's image to the render tolerance."""
import numpy as np
import pytest

import oracle
from azimuth_amd import configs, native as N

pytestmark = pytest.mark.gpu


def test_trace_autotune_and_every_variant_identical(gpu):
    sc = configs.icosphere_scene(5, seed=7)
    rays = np.concatenate([configs.coherent_rays(64, 64), configs.incoherent_rays(4096, sc.meta["bbox"], seed=3)])
    want = oracle.trace(sc, rays)
    gpu.upload_scene(sc)
    d = gpu.alloc(rays.nbytes)
    d.upload(rays)
    entries, best = gpu.autotune_trace(d, rays.shape[0], repeats=2)
    print("trace candidates:", [(e["ring"], e["block"], e["waves"], e["loop"], round(e["ms"], 4)) for e in entries],
          "best", best, flush=True)
    assert len(entries) >= 2 and 0 <= best < len(entries)
    assert all(e["ms"] > 0 for e in entries)
    assert gpu.stats().BlockSize == entries[best]["block"]
    for e in entries:
        gpu.set_launch(block=e["block"], ring=e["ring"], waves=e["waves"], loop=e["loop"])
        got = gpu.trace(rays)
        assert got.tobytes() == want.tobytes(), e
    b = entries[best]
    gpu.set_launch(block=b["block"], ring=b["ring"], waves=b["waves"], loop=b["loop"])
    d.free()


def test_render_autotune_and_variants_match_oracle(gpu):
    sc, cam = configs.cornell_scene(48, 48, spp=4, max_depth=6)
    gpu.upload_scene(sc)
    entries, best = gpu.autotune_render(cam, seed=3)
    print("render candidates:", [(e["block"], e["waves"], e["loop"], e["chunk"], round(e["ms"], 3)) for e in entries],
          "best", best, flush=True)
    assert len(entries) >= 4 and 0 <= best < len(entries)
    want = oracle.render(sc, cam, seed=3)
    wr = want["Color"] / want["TotalWeight"][:, None]
    scale = float(np.mean(np.abs(wr)))
    seen = set()
    for e in entries:
        key = (e["block"], e["waves"], e["loop"])
        if key in seen:
            continue
        seen.add(key)
        # A small chunk forces several chunks per render.
        gpu.set_render_tuning(chunk=4096, block=e["block"], waves=e["waves"], loop=e["loop"])
        got = gpu.render(cam, seed=3)
        assert np.array_equal(got["TotalWeight"], want["TotalWeight"])
        g = got["Color"] / got["TotalWeight"][:, None]
        rmse = float(np.sqrt(np.mean((g - wr) ** 2))) / scale
        assert rmse < 1e-3, (e, rmse)
    gpu.set_render_tuning(chunk=16 << 20, block=128, waves=6, loop=1)


def test_trim_frees_queues_and_render_repeats(gpu):
    sc, cam = configs.cornell_scene(48, 48, spp=9, max_depth=6)
    gpu.upload_scene(sc)
    gpu.set_render_tuning(chunk=8192)
    first = gpu.render(cam, seed=5)
    gpu.trim()
    gpu.trim()                      # nothing left to free: still fine
    again = gpu.render(cam, seed=5)
    assert first.tobytes() == again.tobytes()
    gpu.set_render_tuning(chunk=16 << 20)


def test_set_render_tuning_rejects_unknown_variant(gpu):
    with pytest.raises(N.AzError):
        gpu.set_render_tuning(block=96, waves=6)


def test_bvh_width_and_leaf_size_builds_identical(gpu):
    """Every BVH build the build autotuner can choose (node width 2-4, leaf
    size 2/4/8) gives the oracle's closest hits bit for bit, on a mesh and on
    a two-level scene with moving spheres."""
    from azimuth_amd import workloads
    sc = configs.icosphere_scene(5, seed=7)
    rays = np.concatenate([configs.coherent_rays(48, 48), configs.incoherent_rays(4096, sc.meta["bbox"], seed=3)])
    want = oracle.trace(sc, rays)
    motion, cam = workloads.motion_scene(8, 8, spp=1, max_depth=2, spheres=60, ico_subdiv=3)
    mrays = configs.incoherent_rays(4096, (np.full(3, -10.0), np.full(3, 10.0)), seed=5)   # the spheres' box
    mwant = oracle.trace(motion, mrays)
    assert np.sum(mwant["Prim"] != 0xFFFFFFFF) > 100
    try:
        for width in (2, 3, 4):
            for leaf in (2, 4, 8):
                gpu.set_bvh_width(width)
                gpu.set_build(leaf=leaf)
                gpu.upload_scene(sc)
                assert gpu.stats().BvhWidth == width
                assert gpu.trace(rays).tobytes() == want.tobytes(), (width, leaf)
                gpu.upload_scene(motion)
                assert gpu.trace(mrays).tobytes() == mwant.tobytes(), ("motion", width, leaf)
        d = gpu.alloc(rays.nbytes)
        d.upload(rays)
        entries, best = gpu.autotune_build(sc, d, rays.shape[0], repeats=1)
        print("build candidates:", [(e["width"], e["leaf"], round(e["ms"], 4)) for e in entries], "best", best,
              flush=True)
        assert len(entries) == 9 and 0 <= best < 9
        assert gpu.stats().BvhWidth == entries[best]["width"]
        assert gpu.trace(rays).tobytes() == want.tobytes()
        d.free()
    finally:
        gpu.set_bvh_width(4)
        gpu.set_build(leaf=4)
