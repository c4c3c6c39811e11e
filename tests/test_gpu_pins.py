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


# ---- lights, mirrors, textures, bump maps, motion and cameras (pin_scenes.py)

import ctypes as C

import oracle
import pin_scenes as S


def test_area_light_parallelogram(gpu):
    S.check_area_light(make_render(gpu))


def test_directional_light_and_any_hit_occlusion(gpu):
    S.check_directional_light(make_render(gpu))


def test_spot_light_ramp_between_the_cones(gpu):
    S.check_spot_ramp(make_render(gpu))


def test_perfect_mirror(gpu):
    S.check_mirror(make_render(gpu))


def test_glossy_mirror_against_quadrature(gpu):
    S.check_glossy_mirror(make_render(gpu))


def test_texture_blend_modes_and_ambient(gpu):
    S.check_blend_modes_and_ambient(make_render(gpu))


def test_background_texture_filters_and_wrap(gpu):
    S.check_background_texture(make_render(gpu))


def test_texture_wrap_on_a_surface(gpu):
    S.check_texture_wrap_on_surface(make_render(gpu))


def test_image_bump_map(gpu):
    S.check_image_bump(make_render(gpu))


def test_perlin_bump_map(gpu):
    S.check_perlin_bump(make_render(gpu),
                        lambda t, x, y, z: oracle.load().azo_perlin_sample(C.byref(t), float(x), float(y), float(z)))


def test_motion_blur_shutter_fraction(gpu):
    S.check_motion_blur(make_render(gpu))


def test_detailed_left_handed_and_gaze_direction_cameras(gpu):
    S.check_cameras(make_render(gpu))


def test_thin_lens_focal_plane(gpu):
    S.check_thin_lens_focus(make_render(gpu))
