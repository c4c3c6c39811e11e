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
