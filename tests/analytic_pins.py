"""Scenes whose rendered pixels have closed forms, shared by the CPU checks of
the oracle and the GPU checks of the device renderer (any `render(scene,
camera, filt)` callable returning pixel records).  Each expectation is
derived from the reference's definitions, not from either implementation:

- Gaussian pixel filter (ComputeSampleWeight, main.cpp:1209-1254): radius
  1.5, sigma = radius / 3, footprint one pixel each way, weight
  exp(-d^2 / 2 sigma^2) / (2 pi sigma^2) from a sample to each neighbouring
  pixel centre; stratified sample positions ((a + 0.5) / n, (b + 0.5) / n)
  inside a pixel (GetSampleCoord, main.cpp:1135-1153).
- Dielectric (CalculateTransparencyContribution, main.cpp:856-964): Schlick
  reflectance r0 + (1 - r0)(1 - cos)^5 with r0 = ((n1 - n2) / (n1 + n2))^2 and
  the air-side cosine, refraction weight 1 - reflectance, Beer attenuation
  exp((T - 1) d) over the distance d travelled inside.
- Pinhole camera (GetRayForPixel and camera set-up, main.cpp:1155-1178,
  1265-1323): a pixel centre's ray has direction
  f + (2 (i + 0.5) / W - 1) tan(fovy / 2) (W / H) r + (1 - 2 (j + 0.5) / H) tan(fovy / 2) u
  for the orthonormal frame f (forward), r = f x up, u = r x f; row 0 is the
  top of the image.
- Sphere light (SampleEntityLight, main.cpp:392-459, and
  CalculateLightContribution, main.cpp:571-603): directions uniform in the
  subtended cone, InvPDF 2 pi (1 - cos theta_max), light distance 1.
- Hemisphere scatter (main.cpp:1029-1083): cosine-weighted (InvPDF pi / cos)
  with importance sampling, else uniform (InvPDF 2 pi), weighted by the BRDF
  and cos.
"""
import math

import numpy as np

from azimuth_amd import configs, native as N
from azimuth_amd.scene import SceneBuilder, make_material

BACKGROUND = (0.25, 0.5, 0.75)


# ---------------------------------------------------------------- Gaussian filter

def background_scene():
    """Nothing but a floor far below a camera that looks up: every camera
    sample misses and returns the background colour."""
    b = SceneBuilder(max_depth=1, background=BACKGROUND)
    b.add_material(make_material())
    v = np.array([(-50, -100, -50), (50, -100, -50), (50, -100, 50), (-50, -100, 50)], np.float32)
    b.add_mesh(v, np.array([[0, 2, 1], [0, 3, 2]], np.uint32))
    return b.build()


def background_camera(width, height, spp):
    return configs.look_at_camera((0, 0, 0), (0, 10, 0.1), (0, 0, 1), width, height, 30.0, spp=spp)


def gaussian_total_weight(width, height, spp):
    """TotalWeight of every pixel (row-major) from the filter's definition."""
    n = int(math.isqrt(spp))
    sigma = 1.5 / 3.0
    norm = 1.0 / (2.0 * math.pi * sigma * sigma)
    off = (np.arange(n) + 0.5) / n
    w = np.zeros((height, width))
    for y in range(height):
        for x in range(width):
            for sy in range(y - 1, y + 2):
                for sx in range(x - 1, x + 2):
                    if not (0 <= sx < width and 0 <= sy < height):
                        continue
                    py = sy + off[:, None]
                    px = sx + off[None, :]
                    d2 = (px - (x + 0.5)) ** 2 + (py - (y + 0.5)) ** 2
                    w[y, x] += float(np.sum(norm * np.exp(-d2 / (2.0 * sigma * sigma))))
    return w.reshape(-1)


def check_gaussian(render):
    # One sample per pixel sits at the pixel centre: an inner pixel receives
    # the full centre weight plus four edge and four corner neighbours.
    centre = 1.0 / (2.0 * math.pi * 0.25)
    inner = centre * (1.0 + 4.0 * math.exp(-2.0) + 4.0 * math.exp(-4.0))
    sc = background_scene()
    for (W, H, spp) in ((6, 5, 1), (8, 8, 4), (7, 4, 9)):
        cam = background_camera(W, H, spp)
        px = render(sc, cam, N.FilterType_Gaussian)
        want = gaussian_total_weight(W, H, spp)
        np.testing.assert_allclose(px["TotalWeight"], want, rtol=2e-6)
        np.testing.assert_allclose(px["Color"], np.outer(want, BACKGROUND), rtol=2e-6)
        if spp == 1:
            got = px["TotalWeight"].reshape(H, W)
            np.testing.assert_allclose(got[1:-1, 1:-1], inner, rtol=2e-6)
            np.testing.assert_allclose(got[0, 0], centre * (1.0 + 2.0 * math.exp(-2.0) + math.exp(-4.0)), rtol=2e-6)


# ---------------------------------------------------------------- dielectric slab

SLAB_IOR = 1.5
SLAB_THICKNESS = 0.3
SLAB_TRANSPARENCY = (1.0, 0.8, 0.5)
EMITTER_RADIANCE = (2.0, 3.0, 4.0)


def _quad(z, s):
    return np.array([(-s, -s, z), (s, -s, z), (s, s, z), (-s, s, z)], np.float32)


def slab_scene(max_depth=2):
    """A glass slab (front face z = 0, back face z = -thickness) in front of
    an emitting quad at z = -5, black background."""
    b = SceneBuilder(max_depth=max_depth)
    glass = b.add_material(make_material(kd=(0, 0, 0), transparency=SLAB_TRANSPARENCY, ior=SLAB_IOR))
    dark = b.add_material(make_material(kd=(0, 0, 0)))
    faces = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    b.add_mesh(_quad(0.0, 10.0), faces, material=glass)
    b.add_mesh(_quad(-SLAB_THICKNESS, 10.0), faces, material=glass)
    b.add_mesh(_quad(-5.0, 50.0), faces, material=dark, emitter=True, radiance=EMITTER_RADIANCE)
    return b.build()


def slab_expected(cos_i):
    """Radiance through the slab for incidence cosine cos_i (max depth 2: the
    camera ray, the refracted ray inside, the refracted ray out to the
    emitter; every other branch ends in a miss or past the depth)."""
    r0 = ((1.0 - SLAB_IOR) / (1.0 + SLAB_IOR)) ** 2
    refl = r0 + (1.0 - r0) * (1.0 - cos_i) ** 5          # air-side cosine on both faces
    sin_t = math.sqrt(max(0.0, 1.0 - cos_i * cos_i)) / SLAB_IOR
    cos_t = math.sqrt(1.0 - sin_t * sin_t)
    inside = SLAB_THICKNESS / cos_t - 1e-4               # the refracted ray starts 1e-4 along its direction
    beer = np.exp((np.array(SLAB_TRANSPARENCY) - 1.0) * inside)
    return (1.0 - refl) ** 2 * beer * np.array(EMITTER_RADIANCE)


def check_slab(render):
    W = H = 5
    fovy = 30.0
    cam = configs.look_at_camera((0, 0, 4), (0, 0, 0), (0, 1, 0), W, H, fovy, spp=1)
    px = render(slab_scene(), cam, N.FilterType_Box)
    h = math.tan(math.radians(fovy) / 2.0)
    for j in range(H):
        for i in range(W):
            x = (2.0 * (i + 0.5) / W - 1.0) * h
            y = (1.0 - 2.0 * (j + 0.5) / H) * h
            cos_i = 1.0 / math.sqrt(1.0 + x * x + y * y)
            got = px["Color"][j * W + i] / px["TotalWeight"][j * W + i]
            np.testing.assert_allclose(got, slab_expected(cos_i), rtol=2e-5, err_msg=f"pixel {i},{j}")
    # At normal incidence (the centre pixel) the closed form is exact:
    # (1 - r0)^2 transmitted, r0 = 0.04 for n = 1.5.
    c = px["Color"][(H // 2) * W + W // 2] / px["TotalWeight"][(H // 2) * W + W // 2]
    beer = np.exp((np.array(SLAB_TRANSPARENCY) - 1.0) * (SLAB_THICKNESS - 1e-4))
    np.testing.assert_allclose(c, 0.96 ** 2 * beer * np.array(EMITTER_RADIANCE), rtol=2e-5)
    # One level deeper adds nothing on this path (the internal reflection only
    # reaches the emitter at depth 4), so depth 3 gives the same pixels.
    px3 = render(slab_scene(max_depth=3), cam, N.FilterType_Box)
    np.testing.assert_allclose(px3["Color"], px["Color"], rtol=2e-5)


# ---------------------------------------------------------------- pinhole camera

def camera_frame(position, target, up):
    f = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return f, r, u


def pixel_direction(position, target, up, W, H, fovy, i, j):
    f, r, u = camera_frame(position, target, up)
    h = math.tan(math.radians(fovy) / 2.0)
    w = h * W / H
    d = f + (2.0 * (i + 0.5) / W - 1.0) * w * r + (1.0 - 2.0 * (j + 0.5) / H) * h * u
    return d / np.linalg.norm(d)


CAMERA = dict(position=(1.0, 2.0, 3.0), target=(-2.0, 0.5, -4.0), up=(0.2, 1.0, 0.1))
CAMERA_SIZE = (32, 24, 40.0)
CAMERA_PIXELS = ((0, 0), (31, 0), (0, 23), (31, 23), (16, 12), (5, 17), (27, 3))


def camera_scene():
    """Tiny emitting spheres (0.2 of a pixel across) 10 units out along the
    hand-derived centre rays of CAMERA_PIXELS; black background."""
    W, H, fovy = CAMERA_SIZE
    b = SceneBuilder(max_depth=0)
    b.add_material(make_material(kd=(0, 0, 0)))
    p = np.asarray(CAMERA["position"], np.float64)
    pixel = 2.0 * math.tan(math.radians(fovy) / 2.0) / H
    for (i, j) in CAMERA_PIXELS:
        c = p + 10.0 * pixel_direction(CAMERA["position"], CAMERA["target"], CAMERA["up"], W, H, fovy, i, j)
        b.add_sphere(tuple(c), 0.1 * pixel * 10.0, emitter=True, radiance=(1.0, 1.0, 1.0))
    return b.build()


def check_camera(render):
    W, H, fovy = CAMERA_SIZE
    cam = configs.look_at_camera(CAMERA["position"], CAMERA["target"], CAMERA["up"], W, H, fovy, spp=1)
    px = render(camera_scene(), cam, N.FilterType_Box)
    lit = (px["Color"][:, 0] > 0.5).reshape(H, W)
    want = np.zeros((H, W), bool)
    for (i, j) in CAMERA_PIXELS:
        want[j, i] = True
    assert np.array_equal(lit, want), np.argwhere(lit ^ want)


# ---------------------------------------------------------------- sphere light (cone sampling)

FLOOR_KD = (0.6, 0.5, 0.4)


def _floor(b, material):
    v = np.array([(-50, 0, -50), (50, 0, -50), (50, 0, 50), (-50, 0, 50)], np.float32)
    b.add_mesh(v, np.array([[0, 2, 1], [0, 3, 2]], np.uint32), material=material)


SPHERE_LIGHT = dict(center=(0.0, 2.0, 0.0), radius=0.2, radiance=(5.0, 4.0, 3.0))


def sphere_light_scene():
    """A Kd/pi floor under an emitting sphere (an entity light), next-event
    estimation only."""
    b = SceneBuilder(max_depth=0)
    floor = b.add_material(make_material(kd=FLOOR_KD))
    dark = b.add_material(make_material(kd=(0, 0, 0)))
    _floor(b, floor)
    b.add_sphere(SPHERE_LIGHT["center"], SPHERE_LIGHT["radius"], material=dark, emitter=True,
                 radiance=SPHERE_LIGHT["radiance"])
    return b.build()


def check_sphere_light(render):
    """SampleEntityLight draws directions uniformly in the cone that the
    sphere subtends (InvPDF 2 pi (1 - cos theta_max)), with light distance 1
    in CalculateLightContribution: the mean of Kd/pi L cos InvPDF over the
    cone is Kd L sin^2(theta_max) (n . w), the irradiance of the sphere."""
    W = H = 3
    cam = configs.look_at_camera((0, 1, 0), (0, 0, 0), (0, 0, 1), W, H, 2.0, spp=64)
    px = render(sphere_light_scene(), cam, N.FilterType_Box)
    got = px["Color"] / px["TotalWeight"][:, None]
    h = math.tan(math.radians(1.0))
    c = np.asarray(SPHERE_LIGHT["center"])
    for j in range(H):
        for i in range(W):
            # Pixel centre ray from (0, 1, 0) down to the floor.
            f, r, u = camera_frame((0, 1, 0), (0, 0, 0), (0, 0, 1))
            d = f + (2.0 * (i + 0.5) / W - 1.0) * h * r + (1.0 - 2.0 * (j + 0.5) / H) * h * u
            p = np.array([0.0, 1.0, 0.0]) + d * (1.0 / -d[1])
            dist = np.linalg.norm(c - p)
            want = np.asarray(FLOOR_KD) * np.asarray(SPHERE_LIGHT["radiance"]) * \
                (SPHERE_LIGHT["radius"] / dist) ** 2 * ((c - p)[1] / dist)
            np.testing.assert_allclose(got[j * W + i], want, rtol=3e-3, err_msg=f"pixel {i},{j}")


# ---------------------------------------------------------------- hemisphere scatter

CEILING_RADIANCE = (3.0, 2.0, 1.0)


def scatter_scene(path_trace, importance):
    """A Kd/pi floor under an emitting 2 x 2 quad at height 1 (facing down),
    depth 1 so that a scatter ray ends at the quad or in the sky."""
    b = SceneBuilder(max_depth=1, path_trace=path_trace, importance_sampling=importance)
    floor = b.add_material(make_material(kd=FLOOR_KD))
    dark = b.add_material(make_material(kd=(0, 0, 0)))
    _floor(b, floor)
    q = np.array([(-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)], np.float32)
    b.add_mesh(q, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), material=dark, emitter=True,
               radiance=CEILING_RADIANCE)
    return b.build()


def view_factor_square(half=1.0, height=1.0, n=800):
    """(1/pi) int cos cos' / d^2 dA from the point below the centre of a
    square of side 2 half at the given height (midpoint rule)."""
    g = (np.arange(n) + 0.5) / n * 2.0 * half - half
    x, z = np.meshgrid(g, g)
    d2 = x * x + z * z + height * height
    return float(np.sum(height * height / d2 ** 2) * (2.0 * half / n) ** 2 / math.pi)


def check_scatter(render):
    """The hemisphere scatter (cosine-weighted with importance sampling, else
    uniform; main.cpp:1029-1083) reaches the quad with weight Kd for every
    cosine sample that hits it, 2 Kd cos for a uniform one: either way its
    mean is Kd L F, F the quad's view factor.  The scatter part is isolated
    by subtracting the same render without path tracing (the next-event
    draws come first in a vertex, so they are the same in both)."""
    W = H = 8
    cam = configs.look_at_camera((0, 0.5, 0), (0, 0, 0), (0, 0, 1), W, H, 2.0, spp=1024)
    want = np.asarray(FLOOR_KD) * np.asarray(CEILING_RADIANCE) * view_factor_square()
    for importance in (True, False):
        on = render(scatter_scene(True, importance), cam, N.FilterType_Box)
        off = render(scatter_scene(False, importance), cam, N.FilterType_Box)
        part = (on["Color"] - off["Color"]) / on["TotalWeight"][:, None]
        # 65,536 scatter samples: Bernoulli(F ~ 0.55) noise of 0.35 % (cosine),
        # about 0.5 % (uniform).
        np.testing.assert_allclose(part.mean(0), want, rtol=0.02, err_msg=f"importance={importance}")


# ---------------------------------------------------------------- glossy scatter (both samplers)

GLOSSY_KD, GLOSSY_KS, GLOSSY_EXP = (0.2, 0.15, 0.1), (0.6, 0.6, 0.6), 30.0


def glossy_scatter_scene(path_trace):
    """scatter_scene with a normalized modified Blinn-Phong floor."""
    b = SceneBuilder(max_depth=1, path_trace=path_trace, importance_sampling=True)
    floor = b.add_material(make_material(kd=GLOSSY_KD, ks=GLOSSY_KS, exponent=GLOSSY_EXP,
                                         brdf_type=N.BRDF_NormalizedModifiedBlinnPhong))
    dark = b.add_material(make_material(kd=(0, 0, 0)))
    _floor(b, floor)
    q = np.array([(-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)], np.float32)
    b.add_mesh(q, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), material=dark, emitter=True,
               radiance=CEILING_RADIANCE)
    return b.build()


def glossy_scatter_expected(n=600):
    """int_quad f(L) cos(theta) L dOmega at the origin, viewer straight up:
    f = Kd / pi + Ks (e + 8) / (8 pi) (N . H)^e (main.cpp:532-548)."""
    g = (np.arange(n) + 0.5) / n * 2.0 - 1.0
    x, z = np.meshgrid(g, g)
    d = np.stack([x, np.ones_like(x), z], -1)
    dist = np.linalg.norm(d, axis=-1)
    L = d / dist[..., None]
    H = L + np.array([0.0, 1.0, 0.0])
    H /= np.linalg.norm(H, axis=-1)[..., None]
    spec = (GLOSSY_EXP + 8.0) / (8.0 * math.pi) * H[..., 1] ** GLOSSY_EXP
    cos = L[..., 1]
    w = cos * cos / dist ** 2 * (2.0 / n) ** 2            # cos theta x dOmega (cos' dA / d^2)
    kd, ks = np.asarray(GLOSSY_KD), np.asarray(GLOSSY_KS)
    return np.asarray(CEILING_RADIANCE) * (kd * float(np.sum(w)) / math.pi + ks * float(np.sum(w * spec)))


def check_glossy_scatter(render_opts):
    """render_opts(scene, camera, sampling) renders with the given scatter
    sampler.  Each sampler's scatter part, isolated as in check_scatter,
    converges to the quadrature."""
    W = H = 8
    cam = configs.look_at_camera((0, 0.5, 0), (0, 0, 0), (0, 0, 1), W, H, 2.0, spp=1024)
    want = glossy_scatter_expected()
    off = render_opts(glossy_scatter_scene(False), cam, N.AZ_SCATTER_REFERENCE)
    for sampling in (N.AZ_SCATTER_REFERENCE, N.AZ_SCATTER_BRDF):
        on = render_opts(glossy_scatter_scene(True), cam, sampling)
        part = (on["Color"] - off["Color"]) / on["TotalWeight"][:, None]
        np.testing.assert_allclose(part.mean(0), want, rtol=0.03, err_msg=f"sampling={sampling}")


# ---- output stage (OutputPixelsToFile for the 8-bit formats) -------------------------------
# fn(pixels, width, height, tonemap, inv_gamma, srgb) -> (height, width, 3) uint8.

def _pixels(colors, weights=None):
    import numpy as np
    from azimuth_amd import native as N
    colors = np.asarray(colors, np.float32).reshape(-1, 3)
    px = np.zeros(colors.shape[0], dtype=N.PIXEL_DTYPE)
    px["Color"] = colors
    px["TotalWeight"] = 1.0 if weights is None else np.asarray(weights, np.float32)
    return px


def _tonemap(key, burnout, saturation):
    from azimuth_amd import native as N
    t = N.tonemap()
    t.KeyValue, t.BurnoutPercentage, t.Saturation = key, burnout, saturation
    t.Operator = N.TonemapOperator_Photographic
    return t


def check_output_bytes_and_srgb(fn):
    """PNG conversion (u8)(Clamp01(c) * 255) truncates, clamps, and maps NaN
    to 255 (MinOf(NaN, 1) is 1); the weight divides the colour; the sRGB
    curve (IEC 61966-2-1: 12.92 c below 0.0031308, else 1.055 c^(1/2.4) -
    0.055) at known points."""
    import numpy as np
    cols = [(0.5, 1.0, 2.0), (-1.0, 0.0, float("nan")), (1.5, 3.0, 0.75)]
    out = fn(_pixels(cols, [1.0, 1.0, 3.0]), 3, 1, None, 1.0, False)
    assert out.reshape(-1).tolist() == [127, 255, 255, 0, 0, 255, 127, 255, 63]
    # sRGB: linear 0.5 -> 0.735357 (187.5), 0.2 -> 0.484529 (123.6),
    # 0.003 (below the knee) -> 0.03876 (9.88); 1 -> 1.055f - 0.055f, which
    # binary32 rounds to 0.99999994 (254.99998): 254.
    assert np.float32(1.055) - np.float32(0.055) < np.float32(1.0)
    out = fn(_pixels([(0.5, 0.2, 0.003), (1.0, 0.0, 0.0031308)]), 2, 1, None, 1.0 / 2.4, True)
    assert out.reshape(-1).tolist()[:4] == [187, 123, 9, 254]
    assert out.reshape(-1)[4] == 0 and abs(int(out.reshape(-1)[5]) - 10) <= 1
    # A power gamma of 1/2.2: 0.5 -> 0.72974 (186.1).
    out = fn(_pixels([(0.5, 0.5, 0.5)]), 1, 1, None, 1.0 / 2.2, False)
    assert out.reshape(-1).tolist() == [186, 186, 186]


def check_photographic_operator(fn):
    """Reinhard et al. 2002 with a burnout white point, on a two-level grey
    image (half the pixels luminance a, half b): log-average sqrt((a + e)
    (b + e)) with e = 1e-5, scale key / log-average, white = the luminance at
    the (1 - burnout) percentile; Ld = Ls (1 + Ls / Lw^2) / (1 + Ls)."""
    import math
    import numpy as np
    a, b, key, n = 0.05, 0.8, 0.18, 64
    cols = [(a, a, a)] * (n // 2) + [(b, b, b)] * (n // 2)
    lum = lambda c: c * 0.27 + c * 0.67 + c * 0.06
    la, lb = lum(a), lum(b)
    scale = key / math.sqrt((la + 1e-5) * (lb + 1e-5))
    for burnout, white in ((0.0, lb), (60.0, la)):
        out = fn(_pixels(cols), n, 1, _tonemap(key, burnout, 1.0), 1.0, False).reshape(n, 3).astype(int)
        lw2 = (white * scale) ** 2
        for px, L in ((0, la), (n - 1, lb)):
            ls = L * scale
            want = min(1.0, ls * (1 + ls / lw2) / (1 + ls)) * 255
            assert np.all(np.abs(out[px] - want) <= 1.0), (burnout, px, out[px], want)
    # Saturation on a uniform colour image: the white is the colour's own
    # luminance, Ld = 1, and each channel is (c / L)^s.
    c = (0.5, 0.25, 0.125)
    L = 0.5 * 0.27 + 0.25 * 0.67 + 0.125 * 0.06
    for s in (0.5, 1.0):
        out = fn(_pixels([c] * 16), 4, 4, _tonemap(key, 1.0, s), 1.0, False).reshape(16, 3).astype(int)
        want = [min(1.0, (ch / L) ** s) * 255 for ch in c]
        assert np.all(np.abs(out - np.array(want)) <= 1.0), (s, out[0], want)
