"""Procedural workloads: the Cornell box (256x256, 16 spp, depth 8), the
displaced-icosphere traversal microbench (1,310,720 triangles, 4,194,304
coherent and incoherent rays), the architectural interior, the motion-blur
sphere field and the guided-interior case, built directly as reference
`scene` objects (there are no scene files).  Everything is seeded.
"""
from __future__ import annotations

import math

import numpy as np

from . import native as N
from .scene import (SceneBuilder, area_light, image_texture, make_material, perlin_texture, spot_light)

# ---------------------------------------------------------------------------
# geometry helpers


def icosphere(subdivisions):
    """Unit icosphere: 20*4^s faces, 10*4^s+2 vertices (float64)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    for _ in range(subdivisions):
        nv = v.shape[0]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 0).reshape(-1, 2)
        e.sort(axis=1)
        key = e[:, 0] * nv + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        i0, i1 = uniq // nv, uniq % nv
        mid = 0.5 * (v[i0] + v[i1])
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], 0)
        m = (inv + nv).reshape(3, -1)
        mab, mbc, mca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, mab, mca], 1), np.stack([b, mbc, mab], 1), np.stack([c, mca, mbc], 1),
                            np.stack([mab, mbc, mca], 1)], 0)
    return v, f


_GRAD3 = np.array([(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
                   (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)], dtype=np.float64)


def perlin3(p, perm):
    """Gradient noise (Perlin 2002 improved noise) at points p (n,3)."""
    pi = np.floor(p).astype(np.int64)
    f = p - pi
    u = f * f * f * (f * (f * 6 - 15) + 10)
    P = np.concatenate([perm, perm])
    res = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                h = P[P[P[(pi[:, 0] + dx) & 255] + ((pi[:, 1] + dy) & 255)] + ((pi[:, 2] + dz) & 255)] % 12
                g = _GRAD3[h]
                d = f - np.array([dx, dy, dz], dtype=np.float64)
                w = (u[:, 0] if dx else 1 - u[:, 0]) * (u[:, 1] if dy else 1 - u[:, 1]) * (u[:, 2] if dz else 1 - u[:, 2])
                res = res + w * np.einsum("ij,ij->i", g, d)
    return res


def fbm3(p, seed, octaves=3, base_freq=2.0):
    perm = np.random.default_rng(seed).permutation(256).astype(np.int64)
    s, amp, freq = 0.0, 1.0, base_freq
    for _ in range(octaves):
        s = s + amp * perlin3(p * freq, perm)
        amp *= 0.5
        freq *= 2.0
    return s


def displaced_icosphere(subdivisions, seed, amplitude=0.05):
    v, f = icosphere(subdivisions)
    n = fbm3(v, seed)
    v = v * (1.0 + amplitude * n)[:, None]
    return v.astype(np.float32), f.astype(np.uint32)


def look_at_camera(position, target, up, width, height, fovy_deg, spp=1, aperture=0.0, focus=0.0):
    cam = N.camera()
    cam.Type = N.CameraType_Simple
    cam.NearPlane = N.vec4(float(fovy_deg), 0.0, 0.0, 0.0)   # FovY aliases NearPlane.x
    cam.Position = N.vec3(*position)
    cam.Gaze = N.vec3(*target)
    cam.GazeIsPoint = 1
    cam.Up = N.vec3(*up)
    cam.Resolution = N.vec2u(int(width), int(height))
    cam.NearDistance = 1.0
    cam.FocusDistance = float(focus)
    cam.ApertureSize = float(aperture)
    cam.Gamma = cam.InvGamma = 1.0
    cam.SampleCount = int(spp)
    return cam


# ---------------------------------------------------------------------------
# traversal microbench (displaced icosphere)


def icosphere_scene(subdivisions=8, seed=1, amplitude=0.05):
    v, f = displaced_icosphere(subdivisions, seed, amplitude)
    b = SceneBuilder(max_depth=0)
    b.add_material(make_material())
    b.add_mesh(v, f, material=0, smooth=True)
    sc = b.build(meta={"vertices": v, "faces": f, "bbox": (v.min(0), v.max(0))})
    return sc


def coherent_rays(width=2048, height=2048, eye=(0.0, 0.0, 3.0), fovy_deg=45.0):
    """Pinhole camera rays through pixel centres, unit directions, looking at
    the origin along -z."""
    ys, xs = np.meshgrid(np.arange(height, dtype=np.float64), np.arange(width, dtype=np.float64), indexing="ij")
    h = math.tan(math.radians(fovy_deg) / 2.0)
    w = h * width / height
    px = ((xs + 0.5) / width * 2.0 - 1.0) * w
    py = (1.0 - (ys + 0.5) / height * 2.0) * h
    d = np.stack([px, py, -np.ones_like(px)], -1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros(d.shape[0], dtype=N.RAY_DTYPE)
    r["Origin"] = np.asarray(eye, dtype=np.float32)
    r["Direction"] = d.astype(np.float32)
    r["TMax"] = np.float32(np.inf)
    return r


def incoherent_rays(n, bbox, seed=2):
    """Origins uniform in the box, directions uniform on the sphere."""
    g = np.random.default_rng(seed)
    lo, hi = np.asarray(bbox[0], np.float64), np.asarray(bbox[1], np.float64)
    o = lo + (hi - lo) * g.random((n, 3))
    z = 1.0 - 2.0 * g.random(n)
    phi = 2.0 * math.pi * g.random(n)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    d = np.stack([s * np.cos(phi), s * np.sin(phi), z], -1)
    r = np.zeros(n, dtype=N.RAY_DTYPE)
    r["Origin"] = o.astype(np.float32)
    r["Direction"] = d.astype(np.float32)
    r["TMax"] = np.float32(np.inf)
    return r


# ---------------------------------------------------------------------------
# Cornell box


def cornell_scene(width=256, height=256, spp=16, max_depth=8, light_mode="mesh", with_spheres=True):
    """Unit-cube Cornell box: five walls as 10 triangles, a 0.25x0.25 ceiling
    light of radiance 15 as a 2-triangle emitting mesh, a dielectric sphere
    (r=0.18, IOR 1.5) and a Phong metal sphere (r=0.2).  Walls are Lambertian
    through NormalizedModifiedBlinnPhong with Ks=0 (diffuse Kd/pi)."""
    b = SceneBuilder(max_depth=max_depth, path_trace=True, importance_sampling=True)
    grey = b.add_material(make_material(kd=(0.73, 0.73, 0.73)))
    red = b.add_material(make_material(kd=(0.65, 0.05, 0.05)))
    green = b.add_material(make_material(kd=(0.12, 0.45, 0.15)))
    glass = b.add_material(make_material(kd=(0, 0, 0), transparency=(1, 1, 1), ior=1.5))
    metal = b.add_material(make_material(kd=(0.05, 0.05, 0.05), ks=(0.6, 0.6, 0.6), mirror=(0.4, 0.4, 0.4),
                                         exponent=50.0, brdf_type=N.BRDF_Phong))
    lightm = b.add_material(make_material(kd=(0, 0, 0)))

    def quad(p0, p1, p2, p3):
        return np.array([p0, p1, p2, p3], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.uint32)

    # floor (y=0, normal +y), ceiling (y=1, normal -y), back (z=0, normal +z)
    v, f = quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0))
    b.add_mesh(v, f, material=grey)
    v, f = quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1))
    b.add_mesh(v, f, material=grey)
    v, f = quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0))
    b.add_mesh(v, f, material=grey)
    v, f = quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1))   # left, normal +x
    b.add_mesh(v, f, material=red)
    v, f = quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0))   # right, normal -x
    b.add_mesh(v, f, material=green)
    y = 1.0 - 1e-3
    v, f = quad((0.375, y, 0.375), (0.625, y, 0.375), (0.625, y, 0.625), (0.375, y, 0.625))
    if light_mode == "mesh":
        b.add_mesh(v, f, material=lightm, emitter=True, radiance=(15, 15, 15))
    else:
        b.add_mesh(v, f, material=lightm, emitter=True, radiance=(15, 15, 15))
        b.add_light(area_light((0.375, y, 0.375), (15, 15, 15), (0.25, 0, 0), (0, 0, 0.25), (0, -1, 0)))
    if with_spheres:
        b.add_sphere((0.30, 0.18, 0.45), 0.18, material=glass)
        b.add_sphere((0.70, 0.20, 0.65), 0.20, material=metal)
    sc = b.build(meta={"name": "cornell"})
    cam = look_at_camera((0.5, 0.5, 2.0), (0.5, 0.5, 0.0), (0, 1, 0), width, height, 50.0, spp=spp)
    return sc, cam


# ---------------------------------------------------------------------------
# feature scene: every light type, every BRDF model, textures, bump maps,
# glossy mirrors, dielectrics with absorption, motion blur and an
# environment map, at a small resolution


def procedural_envmap(width=256, height=128, sun_peak=5000.0, sun_sigma_deg=0.5, sun_dir=(0.3, 0.8, 0.5)):
    """Sky gradient plus a Gaussian sun lobe, as an RGB F32 lat-long map in
    the reference's (u, v) = ((pi - atan2(z, x)) / 2pi, acos(y) / pi)
    parametrization; texel values are in 0..255 units (the reference divides
    by Normalizer = 255)."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi
    phi = np.pi - u * 2 * np.pi
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    d = np.stack([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)], -1)
    sky_t = np.clip(d[..., 1], 0, 1)[..., None]
    sky = (1 - sky_t) * np.array([0.9, 0.85, 0.8]) + sky_t * np.array([0.35, 0.55, 0.95])
    sky = sky * 0.6 * (d[..., 1:2] > -0.05) + (d[..., 1:2] <= -0.05) * np.array([0.15, 0.12, 0.1])
    sdir = np.asarray(sun_dir, np.float64)
    sdir /= np.linalg.norm(sdir)
    ang = np.arccos(np.clip(d @ sdir, -1, 1))
    sun = sun_peak * np.exp(-0.5 * (ang / np.radians(sun_sigma_deg)) ** 2)
    img = (sky + sun[..., None] * np.array([1.0, 0.95, 0.85])) * 255.0
    return img.astype(np.float32)


def feature_scene(width=48, height=48, spp=4, max_depth=5, seed=3, env=True, importance=True, aperture=0.0,
                  focus=0.0, background_texture=False, image_bump=False, extra_lights=0):
    """Every light type and BRDF model, textures, Perlin bump maps, glossy
    mirrors, dielectrics, motion blur and (env) an environment map; with
    background_texture (and env False) camera misses read a gradient
    texture, with image_bump a back wall carries a 2D image bump map;
    extra_lights more point, spot, directional and area lights (seeded)
    after the four of the scene."""
    from .scene import directional_light, point_light
    g = np.random.default_rng(seed)
    b = SceneBuilder(max_depth=max_depth, path_trace=True, importance_sampling=importance,
                     ambient=(0.02, 0.02, 0.03), background=(0.1, 0.1, 0.2))
    mats = []
    for bt in range(7):
        mats.append(b.add_material(make_material(kd=g.uniform(0.2, 0.8, 3), ks=g.uniform(0.1, 0.5, 3),
                                                 ka=(0.2, 0.2, 0.2), exponent=float(g.uniform(5, 80)), brdf_type=bt,
                                                 brdf_ior=1.5)))
    glossy = b.add_material(make_material(kd=(0.1, 0.1, 0.1), mirror=(0.6, 0.6, 0.6), glossiness=0.1))
    glass = b.add_material(make_material(kd=(0, 0, 0), transparency=(0.8, 0.9, 0.7), ior=1.45))
    lamp = b.add_material(make_material(kd=(0, 0, 0)))
    checker = (np.indices((16, 16)).sum(0) % 2).astype(np.float32)
    img = np.stack([checker * 200 + 30, checker * 60 + 60, (1 - checker) * 180 + 20], -1).astype(np.uint8)
    t2d, keep = image_texture(img, blend=N.TextureBlendMode_ReplaceDiffuse, filt=N.TextureFilter_Bilinear)
    tex_img = b.add_texture(t2d)
    tex_bump = b.add_texture(perlin_texture(seed + 1, scaling=3.0, bump=True, bump_multiplier=0.3,
                                            blend=N.TextureBlendMode_BlendDiffuse))
    tex_vein = b.add_texture(perlin_texture(seed + 2, scaling=5.0, appearance=N.TexturePerlinAppearance_Vein,
                                            blend=N.TextureBlendMode_ReplaceAll))
    # floor with uvs and the image texture
    fv = np.array([(-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, 4)], np.float32)
    ff = np.array([[0, 2, 1], [0, 3, 2]], np.uint32)
    uv = np.array([(0, 0), (2, 0), (2, 2), (0, 2)], np.float32)
    b.add_mesh(fv, ff, material=mats[5], uvs=uv, texture=tex_img)
    # spheres: one per BRDF, a glossy mirror, glass, bumped, veined, moving
    for k in range(7):
        ang = 2 * np.pi * k / 7
        b.add_sphere((2.2 * np.cos(ang), 0.45, 2.2 * np.sin(ang)), 0.45, material=mats[k])
    b.add_sphere((0, 0.6, 0), 0.6, material=glossy)
    b.add_sphere((0.9, 0.35, 0.9), 0.35, material=glass)
    b.add_sphere((-0.9, 0.35, -0.9), 0.35, material=mats[2], texture=tex_bump)
    b.add_sphere((-0.9, 0.35, 0.9), 0.35, material=mats[4], texture=tex_vein)
    b.add_sphere((0.9, 0.3, -0.9), 0.3, material=mats[6], motion=(0.4, 0.0, 0.0))
    # a smooth-shaded mesh with a transform and motion
    v, f = displaced_icosphere(2, seed)
    T = np.eye(4)
    T[:3, :3] *= 0.5
    T[:3, 3] = (0.0, 1.6, 1.6)
    b.add_mesh(v, f, material=mats[3], transform=T, motion=(0.0, 0.2, 0.0), smooth=True, texture=tex_bump)
    # lights
    b.add_light(point_light((2, 4, 2), (6, 6, 6)))
    b.add_light(area_light((-1, 3.5, -1), (4, 4, 4), (0.5, 0, 0), (0, 0, 0.5), (0, -1, 0)))
    b.add_light(directional_light((-0.3, -1, -0.2), (0.4, 0.4, 0.4)))
    b.add_light(spot_light((0, 4, 0), (20, 20, 20), (0, -1, 0), 60.0, 40.0))
    for k in range(extra_lights):
        pos = (float(g.uniform(-3, 3)), float(g.uniform(1.5, 4.5)), float(g.uniform(-3, 3)))
        col = tuple(float(x) for x in g.uniform(0.2, 1.5, 3))
        kind = k % 5
        if kind in (0, 3):
            b.add_light(point_light(pos, col))
        elif kind == 1:
            tgt = np.array([g.uniform(-2, 2), 0.0, g.uniform(-2, 2)])
            d = tgt - np.array(pos)
            b.add_light(spot_light(pos, tuple(4 * c for c in col), tuple(d / np.linalg.norm(d)),
                                   float(g.uniform(30, 90)), float(g.uniform(10, 30))))
        elif kind == 2:
            d = g.normal(size=3)
            d[1] = -abs(d[1]) - 0.2
            b.add_light(directional_light(tuple(d / np.linalg.norm(d)), tuple(0.1 * c for c in col)))
        else:
            b.add_light(area_light(pos, col, (0.3, 0, 0), (0, 0, 0.3), (0, -1, 0)))
    b.add_sphere((-2.5, 2.5, -2.5), 0.25, material=lamp, emitter=True, radiance=(8, 7, 6))
    lv = np.array([(2.5, 3.0, -2.5), (3.0, 3.0, -2.5), (3.0, 3.0, -2.0), (2.5, 3.0, -2.0)], np.float32)
    b.add_mesh(lv, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), material=lamp, emitter=True, radiance=(6, 6, 8))
    if env:
        envimg = procedural_envmap(64, 32)
        te, keep_env = image_texture(envimg, filt=N.TextureFilter_Nearest)
        b.set_environment_map(te)
    if background_texture:
        yy, xx = np.mgrid[0:24, 0:32]
        bg = np.stack([xx * 6 + 20, yy * 8 + 10, 200 - xx * 3 - yy * 2], -1).astype(np.uint8)
        tb, keep_bg = image_texture(bg, filt=N.TextureFilter_Bilinear, wrap=N.TextureWrapMode_Clamp)
        b.set_background_texture(tb)
    if image_bump:
        yy, xx = np.mgrid[0:32, 0:32]
        hgt = (np.sin(xx / 3.0) * 60 + np.cos(yy / 2.0) * 50 + 128).astype(np.float32)
        tbump, keep_bump = image_texture(np.stack([hgt] * 3, -1), filt=N.TextureFilter_Nearest,
                                         blend=N.TextureBlendMode_BlendDiffuse, bump=True, bump_multiplier=0.004)
        tex_wall = b.add_texture(tbump)
        wv = np.array([(-3.5, 0, -3), (3.5, 0, -3), (3.5, 3, -3), (-3.5, 3, -3)], np.float32)
        wuv = np.array([(0, 0), (2, 0), (2, 1), (0, 1)], np.float32)
        b.add_mesh(wv, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), material=mats[6], uvs=wuv, texture=tex_wall)
    sc = b.build(meta={"name": "features"})
    cam = look_at_camera((0.0, 3.0, 6.5), (0.0, 0.5, 0.0), (0, 1, 0), width, height, 45.0, spp=spp,
                         aperture=aperture, focus=focus)
    return sc, cam
