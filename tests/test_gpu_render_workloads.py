"""Render parity on reduced versions of the benchmark workloads and on the
camera/sampling options the feature scene leaves out: the GPU wavefront
renderer against the oracle's recursive renderer at matched spp and random
streams, through the C ABI, to the same bar as test_gpu_render.py."""
import numpy as np
import pytest

import oracle
from azimuth_amd import configs, native as N, workloads

from test_gpu_render import compare

pytestmark = pytest.mark.gpu


def test_interior_reduced(gpu):
    # The architectural interior (spot lights, HDR sky with a sun, mirrors,
    # Blinn-Phong) with 4 arch shells at 40x24, Gaussian filter.
    sc, cam = workloads.interior_scene(40, 24, spp=4, max_depth=4, shells=4, env_size=(128, 64))
    gpu.upload_scene(sc)
    got = gpu.render(cam, filt=N.FilterType_Gaussian, seed=11)
    want = oracle.render(sc, cam, filt=N.FilterType_Gaussian, seed=11)
    compare(got, want, "interior 40x24 4spp gaussian")


def test_motion_reduced(gpu):
    # Moving spheres (40% dielectric) and the Perlin-bumped icosphere.
    sc, cam = workloads.motion_scene(40, 24, spp=4, max_depth=6, spheres=60, ico_subdiv=3)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=5)
    want = oracle.render(sc, cam, seed=5)
    compare(got, want, "motion 40x24 4spp")


def test_thin_lens_and_uniform_hemisphere(gpu):
    sc, cam = configs.feature_scene(40, 40, spp=4, max_depth=4, importance=False, aperture=0.15, focus=6.0)
    gpu.upload_scene(sc)
    got = gpu.render(cam, seed=9)
    want = oracle.render(sc, cam, seed=9)
    compare(got, want, "features thin lens, uniform hemisphere")
