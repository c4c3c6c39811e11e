"""The device renderer against the closed forms of analytic_pins.py: the
Gaussian pixel filter's weights, Fresnel and Beer through a glass slab, and
the pinhole camera's pixel-centre rays."""
import pytest

import analytic_pins as P
from azimuth_amd import native as N

pytestmark = pytest.mark.gpu


def make_render(gpu):
    def render(scene, camera, filt):
        gpu.upload_scene(scene)
        return gpu.render(camera, filt=filt, seed=1)
    return render


def test_gaussian_filter_weights(gpu):
    P.check_gaussian(make_render(gpu))


def test_dielectric_slab_fresnel_and_beer(gpu):
    P.check_slab(make_render(gpu))


def test_pinhole_camera_pixel_centres(gpu):
    P.check_camera(make_render(gpu))


def test_sphere_light_cone_sampling(gpu):
    P.check_sphere_light(make_render(gpu))


def test_hemisphere_scatter_samplers(gpu):
    P.check_scatter(make_render(gpu))


def test_glossy_scatter_both_samplers(gpu):
    """The BRDF lobe sampler (an option) and the reference's cosine sampler
    both converge to the quadrature of the reference BRDF over the emitter."""
    def render_opts(scene, camera, sampling):
        gpu.upload_scene(scene)
        return gpu.render(camera, filt=N.FilterType_Box, seed=1, scatter_sampling=sampling)
    P.check_glossy_scatter(render_opts)
