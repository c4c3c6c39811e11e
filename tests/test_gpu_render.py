"""GPU wavefront renderer vs the oracle's recursive renderer
(ProduceImageInLinearSpace / TraceRay, main.cpp:966-1397) at matched spp and
random streams, through the C ABI.

Both draw every random number from the same (seed, pixel, sample, tree
node, draw) address and compute every path vertex with the same binary32
operations (sin, cos, atan2 and pow as double functions rounded to float on
both sides; the reference's own box test on each instance), so the paths are
the same bit for bit; the images differ only by the order in which a pixel's
contributions are summed (per-sample fixed-point sums on the device, float
atomics on the CPU), a few 1e-7 relative.  Bar, per image: relative RMSE over
pixels below 1e-4, and at least 99.9% of pixels within 1e-5 relative.  A path
can still part in rare cases, such as a hit within 1e-6 of a triangle edge
that only one of the two BVHs culls (the device does not walk the reference's
inner node boxes); at 256x256x16 spp, 2 pixels of 65,536 move by 3e-3."""
import numpy as np
import pytest

import oracle
from azimuth_amd import configs, native as N

pytestmark = pytest.mark.gpu

REL_RMSE = 1e-4
PIX_REL = 1e-5
PIX_FRAC = 0.999


def resolve(px):
    w = px["TotalWeight"][:, None]
    return px["Color"] / np.where(w == 0, 1, w)


def compare(gpu_px, ora_px, label, pix_frac=PIX_FRAC):
    assert np.array_equal(gpu_px["TotalWeight"] > 0, ora_px["TotalWeight"] > 0)
    np.testing.assert_allclose(gpu_px["TotalWeight"], ora_px["TotalWeight"], rtol=1e-5)
    g, o = resolve(gpu_px), resolve(ora_px)
    assert np.all(np.isfinite(g))
    scale = max(1e-6, float(np.mean(np.abs(o))))
    rmse = float(np.sqrt(np.mean((g - o) ** 2))) / scale
    rel = np.abs(g - o) / (np.abs(o) + 1e-2 * scale)
    frac = float(np.mean(np.all(rel <= PIX_REL, axis=1)))
    print(f"{label}: relative RMSE {rmse:.3e}, pixels within {PIX_REL}: {frac:.5f}, max rel {rel.max():.3e}",
          flush=True)
    assert rmse < REL_RMSE, f"{label}: relative RMSE {rmse}"
    assert frac >= pix_frac, f"{label}: only {frac} of pixels within {PIX_REL}"


def test_cornell_box_filter(gpu):
    sc, cam = configs.cornell_scene(64, 64, spp=16, max_depth=8)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=7)
    want = oracle.render(sc, cam, seed=7)
    compare(got, want, "cornell 64x64 16spp box")


def test_cornell_gaussian_filter(gpu):
    sc, cam = configs.cornell_scene(48, 40, spp=9, max_depth=8)
    gpu.upload_scene(sc)
    got = gpu.render(cam, filt=N.FilterType_Gaussian, seed=3)
    want = oracle.render(sc, cam, filt=N.FilterType_Gaussian, seed=3)
    compare(got, want, "cornell 48x40 9spp gaussian")


def test_feature_scene(gpu):
    sc, cam = configs.feature_scene(64, 64, spp=4, max_depth=5)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=11)
    want = oracle.render(sc, cam, seed=11)
    compare(got, want, "features 64x64 4spp")


def test_feature_scene_gaussian_rows(gpu):
    """A row range (the band of one reference thread) with the Gaussian filter."""
    sc, cam = configs.feature_scene(40, 40, spp=4, max_depth=4)
    gpu.upload_scene(sc)
    got = gpu.render(cam, filt=N.FilterType_Gaussian, seed=5, row_begin=10, row_end=27)
    want = oracle.render(sc, cam, filt=N.FilterType_Gaussian, seed=5, row_begin=10, row_end=27)
    compare(got, want, "features rows 10-27 gaussian")


def test_many_lights(gpu):
    """44 point, spot, directional and area lights and two emitters.  Point,
    spot and directional lights 0-31 go through the shading kernel's pass
    over marked (vertex, light) pairs, on floor-heavy waves more pairs than
    one LDS window holds; the rest through the per-lane loop, area and entity
    lights drawing their random numbers in light order between them."""
    sc, cam = configs.feature_scene(40, 40, spp=4, max_depth=3, extra_lights=40)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=13)
    want = oracle.render(sc, cam, seed=13)
    compare(got, want, "features with 44 lights 40x40 4spp")


def test_backface_culling_render(gpu):
    sc, cam = configs.cornell_scene(32, 32, spp=4, max_depth=4)
    gpu.upload_scene(sc)
    got = gpu.render(cam, cull=True, seed=9)
    want = oracle.render(sc, cam, cull=True, seed=9)
    compare(got, want, "cornell culled")


def test_spp_truncation_and_override(gpu):
    """spp is truncated to a square (main.cpp:1331-1336): 10 -> 9."""
    sc, cam = configs.cornell_scene(16, 16, spp=10, max_depth=3)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=1)
    assert np.all(got["TotalWeight"] == 9.0)
    got2 = gpu.render(cam, seed=1, spp=4)
    assert np.all(got2["TotalWeight"] == 4.0)


def test_renders_repeat_bit_for_bit(gpu):
    """Per-sample colour sums are fixed-point integers, so the order in which
    lanes add a sample's vertex and shadow-ray contributions does not matter:
    a second render with the same seed is identical bit for bit, here with
    dielectric splitting and several chunks."""
    from azimuth_amd import workloads
    sc, cam = workloads.motion_scene(48, 32, spp=16, max_depth=8, spheres=80, ico_subdiv=3)
    gpu.upload_scene(sc)
    gpu.set_render_tuning(chunk=8192)
    a = gpu.render(cam, seed=9)
    b = gpu.render(cam, seed=9)
    gpu.set_render_tuning(chunk=16 << 20)
    assert a.tobytes() == b.tobytes()
    assert np.all(a["TotalWeight"] > 0)
