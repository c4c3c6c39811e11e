"""Bounce levels replayed from a recorded hipGraph (AZ_LEVELS_AUTO) against
the host-driven levels (AZ_LEVELS_HOST), through the C ABI: the same image,
bit for bit, whether a chunk replays the graph within its learned level
bounds or outgrows them and is redone host-driven.  Scenes with one child per shading point (a path-traced
Lambertian box), with mirrors and with dielectrics (several children per
point), several chunks per render (the graph replayed once per chunk), and
repeated renders with host-driven renders between the replays (the recorded
graph reused, every replay checked against the host-driven image)."""
import numpy as np
import pytest

from azimuth_amd import configs, native as N, workloads

pytestmark = pytest.mark.gpu

DEFAULT_CHUNK = 16 << 20


def render(gpu, cam, mode, **kw):
    gpu.set_level_mode(mode)
    try:
        return np.array(gpu.render(cam, **kw), copy=True)
    finally:
        gpu.set_level_mode(N.AZ_LEVELS_AUTO)


def same_bits(a, b):
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def check_replays(gpu, sc, cam, chunk, label, replays=3, **kw):
    """Host-driven render, then `replays` graph renders, each followed by a
    host-driven one; every image equal to the first."""
    gpu.upload_scene(sc)
    gpu.set_render_tuning(chunk=chunk)
    try:
        want = render(gpu, cam, N.AZ_LEVELS_HOST, **kw)
        assert np.isfinite(want["Color"]).all() and want["TotalWeight"].sum() > 0, label
        rec0, lau0 = gpu.level_graph_counts()
        for k in range(replays):
            got = render(gpu, cam, N.AZ_LEVELS_AUTO, **kw)
            assert same_bits(got, want), f"{label}: graph render {k} differs from the host-driven levels"
            again = render(gpu, cam, N.AZ_LEVELS_HOST, **kw)
            assert same_bits(again, want), f"{label}: host-driven render after graph replay {k} differs"
        rec1, lau1 = gpu.level_graph_counts()
    finally:
        gpu.set_render_tuning(chunk=DEFAULT_CHUNK)
    return rec1 - rec0, lau1 - lau0


def test_lambertian_box_one_child(gpu):
    # Path-traced box without mirror or glass spheres: one child per point,
    # 64x64 at 16 spp in chunks of 4 spp (4 chunks, one graph).
    sc, cam = configs.cornell_scene(64, 64, spp=16, max_depth=8, with_spheres=False)
    recorded, launched = check_replays(gpu, sc, cam, 64 * 64 * 4, "box")
    # The first graph render learns the level bounds on its first chunk
    # (host-driven); every later chunk replays the one recorded graph.
    assert recorded == 1, recorded
    assert launched == 3 * 4 - 1, launched


def test_cornell_spheres_gaussian(gpu):
    sc, cam = configs.cornell_scene(64, 64, spp=16, max_depth=8)
    recorded, launched = check_replays(gpu, sc, cam, 64 * 64 * 4, "cornell gaussian", filt=N.FilterType_Gaussian,
                                       seed=3)
    assert recorded >= 1 and launched >= 1


def test_interior_mirrors(gpu):
    # Mirrors: two children per point where a mirror is path traced.
    sc, cam = workloads.interior_scene(40, 24, spp=16, max_depth=4, shells=4, env_size=(128, 64))
    recorded, launched = check_replays(gpu, sc, cam, 40 * 24 * 4, "interior", filt=N.FilterType_Gaussian, seed=11)
    assert recorded >= 1 and launched >= 1


def test_motion_dielectrics(gpu):
    # Dielectrics: reflection and refraction children; levels may outgrow
    # 1.5 n, and such a chunk is redone with the host-driven levels.
    sc, cam = workloads.motion_scene(48, 32, spp=16, max_depth=8, spheres=80, ico_subdiv=3)
    check_replays(gpu, sc, cam, 48 * 32 * 8, "motion", replays=2, seed=5)


def test_rows_and_culling(gpu):
    sc, cam = configs.feature_scene(40, 40, spp=9, max_depth=5)
    check_replays(gpu, sc, cam, 40 * 3 * 9, "feature rows", replays=2, row_begin=7, row_end=31, cull=True)


def test_level_mode_rejects_unknown(gpu):
    with pytest.raises(Exception):
        gpu.set_level_mode(7)
