"""Closed forms that pin the oracle's pixel filter, dielectric and camera
(the analytic scenes and expectations are in analytic_pins.py)."""
import oracle
import analytic_pins as P


def render(scene, camera, filt):
    return oracle.render(scene, camera, filt=filt, seed=1)


def test_gaussian_filter_weights():
    P.check_gaussian(render)


def test_dielectric_slab_fresnel_and_beer():
    P.check_slab(render)


def test_pinhole_camera_pixel_centres():
    P.check_camera(render)


def test_sphere_light_cone_sampling():
    P.check_sphere_light(render)


def test_hemisphere_scatter_samplers():
    P.check_scatter(render)


def output(px, w, h, tm, inv_gamma, srgb):
    return oracle.output_pixels(px, w, h, tm, inv_gamma, srgb)


def test_output_bytes_and_srgb():
    P.check_output_bytes_and_srgb(output)


def test_photographic_operator():
    P.check_photographic_operator(output)


# ---- lights, mirrors, textures, bump maps, motion and cameras (pin_scenes.py)

import ctypes as C

import pin_scenes as S


def perlin_sample(t, x, y, z):
    return oracle.load().azo_perlin_sample(C.byref(t), float(x), float(y), float(z))


def test_area_light_parallelogram():
    S.check_area_light(render)


def test_directional_light_and_any_hit_occlusion():
    S.check_directional_light(render)


def test_spot_light_ramp_between_the_cones():
    S.check_spot_ramp(render)


def test_perfect_mirror():
    S.check_mirror(render)


def test_glossy_mirror_against_quadrature():
    S.check_glossy_mirror(render)


def test_texture_blend_modes_and_ambient():
    S.check_blend_modes_and_ambient(render)


def test_background_texture_filters_and_wrap():
    S.check_background_texture(render)


def test_texture_wrap_on_a_surface():
    S.check_texture_wrap_on_surface(render)


def test_image_bump_map():
    S.check_image_bump(render)


def test_perlin_bump_map():
    S.check_perlin_bump(render, perlin_sample)


def test_motion_blur_shutter_fraction():
    S.check_motion_blur(render)


def test_detailed_left_handed_and_gaze_direction_cameras():
    S.check_cameras(render)


def test_thin_lens_focal_plane():
    S.check_thin_lens_focus(render)
