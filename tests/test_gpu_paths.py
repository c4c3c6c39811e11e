"""Paths of the device renderer that the workload tests leave out, each
against the oracle or an independent property, through the C ABI:

- the reference's default detailed camera (asymmetric near plane), the
  left-handed frame and a gaze given as a direction (main.cpp:1265-1323) on
  the feature scene;
- the background texture (main.cpp:1197-1201) and a 2D image bump map
  (main.cpp:323-350);
- the full 256x256, 16 spp, depth 8 Cornell box and the full 2,000-sphere,
  327,680-triangle motion-blur scene (at a small resolution);
- az_render_scene, the RenderScene-shaped drop-in (main.cpp:1799-1901);
- the out-of-memory path: a render memory cap halves the chunk, then fails
  with AZ_ERR_OUT_OF_MEMORY;
- the environment luminance sampler: its density integrates to one, its
  samples follow it (chi-square), and the density it reports with a sample
  is the density of that direction.
Render comparisons use the bar of test_gpu_render.py."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
import pin_scenes as S
from azimuth_amd import configs, native as N, workloads
from test_gpu_render import compare

pytestmark = pytest.mark.gpu


def feature_cameras(W, H, spp):
    """The feature scene seen through the three camera forms."""
    base = dict(position=(0.0, 3.0, 6.5), up=(0.0, 1.0, 0.0), near=1.2)
    cams = {
        "detailed, asymmetric near plane": dict(base, gaze=(0.0, 0.5, 0.0), gaze_is_point=True, left_handed=False,
                                                plane=(0.55, -0.35, -0.3, 0.45)),
        "detailed, left-handed, gaze direction": dict(base, gaze=(0.1, -2.5, -6.5), gaze_is_point=False,
                                                      left_handed=True, plane=(0.4, -0.5, -0.35, 0.3)),
        "simple, left-handed, gaze direction": dict(base, gaze=(-0.2, -2.4, -6.5), gaze_is_point=False,
                                                    left_handed=True, fovy=42.0),
    }
    return {k: S.make_camera(v, W, H, spp=spp) for k, v in cams.items()}


def test_feature_scene_camera_forms(gpu):
    sc, _ = configs.feature_scene(48, 36, spp=4, max_depth=4)
    gpu.upload_scene(sc)
    for name, cam in feature_cameras(48, 36, 4).items():
        got = gpu.render(cam, seed=13)
        want = oracle.render(sc, cam, seed=13)
        compare(got, want, f"features, {name}")


def test_background_texture_and_image_bump(gpu):
    sc, cam = configs.feature_scene(56, 40, spp=4, max_depth=4, env=False, background_texture=True, image_bump=True)
    gpu.upload_scene(sc)
    for filt in (N.FilterType_Box, N.FilterType_Gaussian):
        got = gpu.render(cam, filt=filt, seed=17)
        want = oracle.render(sc, cam, filt=filt, seed=17)
        compare(got, want, f"features, background texture and image bump, filter {filt}")


def test_cornell_box_full_size(gpu):
    sc, cam = configs.cornell_scene(256, 256, spp=16, max_depth=8)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=7)
    want = oracle.render(sc, cam, seed=7)
    compare(got, want, "cornell 256x256 16spp depth 8")


def test_motion_scene_full_geometry(gpu):
    # All 2,000 moving spheres (40 % dielectric) and the 327,680-triangle
    # bumped icosphere, depth 10, at 160x90.
    sc, cam = workloads.motion_scene(160, 90, spp=4, max_depth=10)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=5)
    want = oracle.render(sc, cam, seed=5)
    compare(got, want, "motion 2000 spheres + 327,680 triangles, 160x90 4spp depth 10")


def test_render_scene_drop_in(gpu):
    """az_render_scene(scene, camera, eps, threads, cull, pixels) gives the
    pixels az_gpu_render gives with seed 0 and the box filter, bit for bit,
    and matches the oracle; with culling too."""
    lib = N.load()
    sc, cam = configs.cornell_scene(64, 64, spp=16, max_depth=8)
    eps = N.epsilon_values(0.0, 0.000275)
    for cull in (0, 1):
        px = np.full(64 * 64, 7.0, dtype=N.PIXEL_DTYPE)      # the call clears the buffer itself
        assert lib.az_render_scene(sc.ptr, C.byref(cam), C.byref(eps), 8, cull, px.ctypes.data) == 1
        gpu.upload_scene(sc)
        want = gpu.render(cam, seed=0, cull=bool(cull))
        assert px.tobytes() == want.tobytes(), cull
        compare(px, oracle.render(sc, cam, seed=0, cull=bool(cull)), f"az_render_scene cull={cull}")
    bad = np.zeros(1, dtype=N.PIXEL_DTYPE)
    assert lib.az_render_scene(None, C.byref(cam), C.byref(eps), 1, 0, bad.ctypes.data) == 0


def test_render_memory_limit_halves_the_chunk_then_fails(gpu):
    # Depth 3: the path queues of the deepest level hold about twice the
    # camera rays (dielectric splits, mirror children), some 60 MB for the
    # render's single 16-spp chunk and under 4 MB at 1 spp.
    sc, cam = configs.cornell_scene(64, 64, spp=16, max_depth=3)
    gpu.upload_scene(sc)
    free = gpu.render(cam, seed=3)
    try:
        gpu.set_render_memory_limit(12 << 20)
        capped = gpu.render(cam, seed=3)
        assert np.array_equal(capped["TotalWeight"], free["TotalWeight"])
        np.testing.assert_allclose(capped["Color"], free["Color"], rtol=2e-6, atol=1e-7)
        gpu.set_render_memory_limit(64 << 10)
        with pytest.raises(N.AzError, match="AZ_ERR_OUT_OF_MEMORY"):
            gpu.render(cam, seed=3)
    finally:
        gpu.set_render_memory_limit(0)
    again = gpu.render(cam, seed=3)
    assert again.tobytes() == free.tobytes()


def env_uv(d):
    """The environment map's (u, v) of unit directions (main.cpp:726-733)."""
    u = (-np.arctan2(d[:, 2], d[:, 0]) + math.pi) / (2 * math.pi)
    v = np.arccos(np.clip(d[:, 1], -1, 1)) / math.pi
    return u, v


def test_environment_luminance_sampler_density_and_samples(gpu):
    sc, cam = workloads.interior_scene(32, 18, spp=1, max_depth=2, shells=2, env_size=(256, 128))
    gpu.upload_scene(sc)
    W, H = 256, 128
    k = 4                                         # quadrature points per cell side
    th = (np.arange(H * k) + 0.5) / (H * k) * math.pi
    ph = math.pi - (np.arange(W * k) + 0.5) / (W * k) * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    dirs = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pdf = gpu.env_luminance_pdf(dirs).astype(np.float64).reshape(H * k, W * k)
    dA = (math.pi / (H * k)) * (2 * math.pi / (W * k)) * np.sin(T)
    mass = pdf * dA
    total = float(mass.sum())
    print(f"integral of the density over the sphere: {total:.7f}", flush=True)
    assert abs(total - 1.0) < 1e-4
    assert pdf.min() > 0.0                        # every direction reachable
    cell = mass.reshape(H, k, W, k).sum(axis=(1, 3))
    # Samples at uniform inputs fall into the cells in proportion to the
    # density's mass there.
    M = 1 << 22
    u = np.random.default_rng(11).random((M, 2), dtype=np.float32)
    out = gpu.env_luminance_sample(u)
    d = out[:, :3].astype(np.float64)
    assert np.all(np.abs(np.linalg.norm(d, axis=1) - 1) < 1e-5)
    uu, vv = env_uv(d)
    c = np.minimum((uu * W).astype(np.int64), W - 1)
    r = np.minimum((vv * H).astype(np.int64), H - 1)
    counts = np.bincount(r * W + c, minlength=W * H).astype(np.float64)
    expect = cell.reshape(-1) * M
    big = expect >= 20
    rest_o, rest_e = counts[~big].sum(), expect[~big].sum()
    chi2 = float(np.sum((counts[big] - expect[big]) ** 2 / expect[big])) + \
        (float((rest_o - rest_e) ** 2 / rest_e) if rest_e > 0 else 0.0)
    dof = int(big.sum())
    print(f"chi-square {chi2:.0f} over {dof} cells ({chi2 / dof:.3f} per degree of freedom)", flush=True)
    assert chi2 / dof < 1.1
    # The density reported with each sample is the density of its direction
    # (up to samples on a cell border, where the lookup can pick the neighbour).
    q = gpu.env_luminance_pdf(d[: 1 << 18])
    rel = np.abs(out[: 1 << 18, 3] - q) / q
    assert float(np.mean(rel < 1e-3)) > 0.999


def test_environment_probe_needs_an_image_environment(gpu):
    sc, cam = configs.cornell_scene(8, 8, spp=1, max_depth=1)
    gpu.upload_scene(sc)
    with pytest.raises(N.AzError, match="AZ_ERR_UNSUPPORTED"):
        gpu.env_luminance_pdf(np.array([[0.0, 1.0, 0.0]]))
