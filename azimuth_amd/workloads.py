"""The larger seeded workloads: the architectural interior (about 300k
triangles, HDR environment map with a sun lobe, 8 spot lights, 1920x1080, 64
spp, depth 6, Gaussian filter) and the motion-blur field (2,000 moving spheres,
40% dielectric, plus a 327,680-triangle displaced icosphere with a Perlin bump
map, 1280x720, 128 spp truncated to 121, depth 10).
"""
from __future__ import annotations

import math

import numpy as np

from . import native as N
from .configs import displaced_icosphere, look_at_camera, procedural_envmap
from .scene import SceneBuilder, directional_light, image_texture, make_material, perlin_texture, spot_light


def box_mesh(lo, hi):
    """Axis-aligned box as 12 outward-facing triangles."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
                  (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)
    return v, f


def half_cylinder_shell(center, radius, length, axis, segments=64, rings=32):
    """Tessellated half-cylinder (an arch vault) of segments x rings quads,
    2 * segments * rings triangles, open at both ends, normals inward."""
    cx, cy, cz = center
    th = np.linspace(0.0, math.pi, segments + 1)
    s = np.linspace(-length / 2, length / 2, rings + 1)
    T, S = np.meshgrid(th, s, indexing="ij")
    if axis == 0:
        P = np.stack([cx + S, cy + radius * np.sin(T), cz + radius * np.cos(T)], -1)
    else:
        P = np.stack([cx + radius * np.cos(T), cy + radius * np.sin(T), cz + S], -1)
    v = P.reshape(-1, 3).astype(np.float32)
    idx = np.arange((segments + 1) * (rings + 1)).reshape(segments + 1, rings + 1)
    a, b = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    c, d = idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)]).astype(np.uint32)
    return v, f


def _material_mix(b, g):
    """50% Lambertian (Kd/pi), 30% Blinn-Phong with exponent U[10, 200],
    20% perfect mirror."""
    u = g.random()
    if u < 0.5:
        return b.add_material(make_material(kd=g.uniform(0.2, 0.85, 3)))
    if u < 0.8:
        return b.add_material(make_material(kd=g.uniform(0.1, 0.6, 3), ks=g.uniform(0.1, 0.4, 3),
                                            exponent=float(g.uniform(10, 200)),
                                            brdf_type=N.BRDF_NormalizedModifiedBlinnPhong))
    return b.add_material(make_material(kd=(0, 0, 0), mirror=(0.95, 0.95, 0.95)))


def interior_scene(width=1920, height=1080, spp=64, max_depth=6, seed=0, env_size=(1024, 512), shells=72):
    """A 40 x 15 x 20 hall (floor and four walls, open to the sky) with a
    seeded 20 x 20 grid of 400 box columns and `shells` arch vaults of 64 x 32
    segments (4,096 triangles each), lit by an HDR sky with a sun lobe of
    peak 5000 and sigma 0.5 degrees, and by 8 spot lights."""
    g = np.random.default_rng(seed)
    b = SceneBuilder(max_depth=max_depth, path_trace=True, importance_sampling=True)
    hall = b.add_material(make_material(kd=(0.7, 0.68, 0.65)))
    W, H, D = 40.0, 15.0, 20.0
    walls = [((-W / 2, 0, -D / 2), (W / 2, 0, -D / 2), (W / 2, 0, D / 2), (-W / 2, 0, D / 2)),
             ((-W / 2, 0, -D / 2), (-W / 2, H, -D / 2), (W / 2, H, -D / 2), (W / 2, 0, -D / 2)),
             ((-W / 2, 0, D / 2), (W / 2, 0, D / 2), (W / 2, H, D / 2), (-W / 2, H, D / 2)),
             ((-W / 2, 0, -D / 2), (-W / 2, 0, D / 2), (-W / 2, H, D / 2), (-W / 2, H, -D / 2)),
             ((W / 2, 0, -D / 2), (W / 2, H, -D / 2), (W / 2, H, D / 2), (W / 2, 0, D / 2))]
    for q in walls:
        b.add_mesh(np.array(q, np.float32), np.array([[0, 2, 1], [0, 3, 2]], np.uint32), material=hall)
    # 400 columns on a jittered 20 x 20 grid.
    for i in range(20):
        for j in range(20):
            x = -W / 2 + (i + 0.5) * W / 20 + g.uniform(-0.3, 0.3)
            z = -D / 2 + (j + 0.5) * D / 20 + g.uniform(-0.2, 0.2)
            r = g.uniform(0.1, 0.25)
            h = g.uniform(2.0, 9.0)
            v, f = box_mesh((x - r, 0.0, z - r), (x + r, h, z + r))
            b.add_mesh(v, f, material=_material_mix(b, g))
    # Arch vaults between column rows.
    for k in range(shells):
        axis = k % 2
        length = g.uniform(1.0, 3.0)
        rad = g.uniform(0.5, 1.5)
        c = (g.uniform(-W / 2 + 2, W / 2 - 2), g.uniform(3.0, 11.0), g.uniform(-D / 2 + 2, D / 2 - 2))
        v, f = half_cylinder_shell(c, rad, length, axis)
        b.add_mesh(v, f, material=_material_mix(b, g), smooth=True)
    # 8 spot lights hanging below the open roof.
    for k in range(8):
        p = (g.uniform(-W / 2 + 3, W / 2 - 3), 13.5, g.uniform(-D / 2 + 3, D / 2 - 3))
        tgt = np.array([p[0] + g.uniform(-3, 3), 0.0, p[2] + g.uniform(-3, 3)])
        d = tgt - np.array(p)
        b.add_light(spot_light(p, (400.0, 380.0, 350.0), d / np.linalg.norm(d), 50.0, 30.0))
    env = procedural_envmap(env_size[0], env_size[1], sun_peak=5000.0, sun_sigma_deg=0.5, sun_dir=(0.4, 0.75, 0.3))
    t, keep = image_texture(env, filt=N.TextureFilter_Nearest)
    b.set_environment_map(t)
    sc = b.build(meta={"name": "interior", "seed": seed, "envmap": env})
    cam = look_at_camera((-17.0, 6.0, 8.5), (4.0, 3.0, -2.0), (0, 1, 0), width, height, 60.0, spp=spp)
    return sc, cam


def motion_scene(width=1280, height=720, spp=128, max_depth=10, seed=0, spheres=2000, ico_subdiv=7):
    """2,000 spheres with radii log-uniform in [0.05, 0.5], centres uniform
    in a 20^3 box and velocities uniform in [-0.5, 0.5]^3 over the shutter
    [0, 1); 40% dielectric with IOR U[1.33, 1.8].  A displaced icosphere at
    subdivision 7 (327,680 triangles) carries a Perlin bump map.  Lit by a
    sky with a sun and a directional light."""
    g = np.random.default_rng(seed)
    b = SceneBuilder(max_depth=max_depth, path_trace=True, importance_sampling=True, ambient=(0.01, 0.01, 0.01))
    mats = [b.add_material(make_material(kd=g.uniform(0.2, 0.9, 3))) for _ in range(12)]
    glass = [b.add_material(make_material(kd=(0, 0, 0), transparency=(1, 1, 1), ior=float(g.uniform(1.33, 1.8))))
             for _ in range(8)]
    for _ in range(spheres):
        r = float(math.exp(g.uniform(math.log(0.05), math.log(0.5))))
        c = g.uniform(-10, 10, 3)
        mv = g.uniform(-0.5, 0.5, 3)
        m = glass[g.integers(len(glass))] if g.random() < 0.4 else mats[g.integers(len(mats))]
        b.add_sphere(c, r, material=m, motion=mv)
    bump = b.add_texture(perlin_texture(seed + 100, scaling=4.0, bump=True, bump_multiplier=0.05))
    v, f = displaced_icosphere(ico_subdiv, seed + 7)
    T = np.eye(4)
    T[:3, :3] *= 3.0
    b.add_mesh(v, f, material=mats[0], transform=T, smooth=True, texture=bump)
    b.add_light(directional_light((-0.4, -1.0, -0.3), (1.5, 1.4, 1.3)))
    env = procedural_envmap(256, 128, sun_peak=800.0, sun_sigma_deg=2.0, sun_dir=(0.3, 0.8, 0.2))
    t, keep = image_texture(env)
    b.set_environment_map(t)
    sc = b.build(meta={"name": "motion", "seed": seed})
    cam = look_at_camera((0.0, 4.0, 24.0), (0.0, 0.0, 0.0), (0, 1, 0), width, height, 50.0, spp=spp)
    return sc, cam
