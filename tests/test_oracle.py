"""CPU checks that pin the oracle (the CPU restatement of the reference hot
path) to things other than itself: published known answers, brute force,
closed forms and invariants.  No GPU needed."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle
from azimuth_amd import configs, native as N
from azimuth_amd.scene import SceneBuilder, make_material, point_light, spot_light

# ---------------------------------------------------------------- random streams


def test_pcg32_known_answer():
    """pcg32_srandom_r(42, 54) -> the sequence printed by the PCG reference
    demo (pcg-c-basic, pcg32-demo)."""
    lib = oracle.load()
    got = [lib.azo_pcg32_at(42, 54, d) for d in range(6)]
    assert got == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_pcg32_jump_equals_stepping():
    """advance(d) is d single steps: the draw at d+1 after a jump to d equals
    the draw at d+1 directly, and jumps compose."""
    lib = oracle.load()
    for seed, seq in ((1, 2), (0xdeadbeef, 123456)):
        seqv = [lib.azo_pcg32_at(seed, seq, d) for d in range(300)]
        assert len(set(seqv)) > 290
        for d in (0, 1, 7, 63, 64, 65, 200, 299):
            assert lib.azo_pcg32_at(seed, seq, d) == seqv[d]
    # Large offsets: (sample, node, dim) addresses land where stepping would.
    big = (5 << 38) + (17 << 6)
    assert lib.azo_pcg32_at(9, 3, big + 2) != lib.azo_pcg32_at(9, 3, big + 1)


def test_rng_addressing():
    """azo_rng_draw(seed, pixel, sample, node, dim) = the (dim+1)-th output
    after seeding with (seed, pixel) and jumping (sample<<38) + (node<<6),
    as a 24-bit fraction."""
    lib = oracle.load()
    for (seed, pixel, sample, node, dim) in ((0, 0, 0, 0, 0), (7, 1234, 15, 1, 3), (99, 65535, 4095, 48828125, 10)):
        raw = lib.azo_pcg32_at(seed, pixel, (sample << 38) + (node << 6) + dim)
        want = np.float32(raw >> 8) * np.float32(1.0 / 16777216.0)
        assert oracle.rng_draw(seed, pixel, sample, node, dim) == want


def test_rng_uniformity():
    lib = oracle.load()
    x = np.array([lib.azo_pcg32_at(3, 5, d) >> 8 for d in range(20000)], np.float64) / 16777216.0
    hist, _ = np.histogram(x, bins=20, range=(0, 1))
    chi2 = float(np.sum((hist - 1000.0) ** 2 / 1000.0))
    assert chi2 < 50.0   # 19 dof, p ~ 1e-4
    assert x.min() >= 0.0 and x.max() < 1.0


# ---------------------------------------------------------------- intersection


@pytest.fixture(scope="module")
def ico3():
    return configs.icosphere_scene(3, seed=3)


def test_bvh_walk_equals_brute_force(ico3):
    rays = np.concatenate([configs.incoherent_rays(20000, ico3.meta["bbox"], 4), configs.coherent_rays(64, 64)])
    a = oracle.trace(ico3, rays)
    b = oracle.trace(ico3, rays, brute=True)
    assert np.array_equal(a["Prim"], b["Prim"])
    assert np.array_equal(a["t"], b["t"])
    assert np.array_equal(a["Beta"], b["Beta"])


def test_bvh_walk_equals_brute_force_culled(ico3):
    rays = configs.incoherent_rays(20000, ico3.meta["bbox"], 5)
    a = oracle.trace(ico3, rays, cull=True)
    b = oracle.trace(ico3, rays, cull=True, brute=True)
    assert np.array_equal(a["Prim"], b["Prim"])


def test_exact_ties_resolve_to_rightmost_leaf():
    """Rays through shared edges of the Cornell walls hit two triangles at
    the same t; the BVH walk and the brute force (later leaf wins) agree."""
    sc, cam = configs.cornell_scene(8, 8, spp=1, with_spheres=False)
    n = 64
    u = (np.arange(n) + 0.5) / n
    pts = np.concatenate([np.stack([u, np.zeros(n), np.zeros(n)], 1), np.stack([np.ones(n), np.zeros(n), u], 1)])
    eye = np.array([0.5, 0.5, 2.0], np.float32)
    r = np.zeros(pts.shape[0], dtype=N.RAY_DTYPE)
    r["Origin"] = eye
    r["Direction"] = pts.astype(np.float32) - eye
    a = oracle.trace(sc, r)
    b = oracle.trace(sc, r, brute=True)
    assert np.all(a["Prim"] != N.AZ_MISS)
    assert np.array_equal(a["Prim"], b["Prim"])


def test_triangle_barycentrics_reconstruct_hit_point(ico3):
    """P = (1-b-g) P0 + b P1 + g P2 equals o + t d: beta/gamma belong to P1/P2
    of the triangle in primitive-id order."""
    v, f = ico3.meta["vertices"].astype(np.float64), ico3.meta["faces"]
    rays = configs.incoherent_rays(5000, ico3.meta["bbox"], 6)
    h = oracle.trace(ico3, rays)
    m = h["Prim"] != N.AZ_MISS
    # Primitive ids follow the mesh's triangle-entity order, which the
    # reference BVH build permuted; recover positions through a brute match.
    o = rays["Origin"][m].astype(np.float64)
    d = rays["Direction"][m].astype(np.float64)
    P = o + h["t"][m][:, None] * d
    # Every triangle containing P within tolerance must reproduce b, g.
    tri = v[f]                                                # (F,3,3)
    b, g = h["Beta"][m].astype(np.float64), h["Gamma"][m].astype(np.float64)
    ok = 0
    for k in range(min(400, P.shape[0])):
        q = (1 - b[k] - g[k])[None, None] * tri[:, 0] + b[k] * tri[:, 1] + g[k] * tri[:, 2]
        err = np.linalg.norm(q - P[k], axis=1)
        ok += err.min() < 1e-4
    assert ok == min(400, P.shape[0])


def test_ray_sphere_closed_form():
    b = SceneBuilder()
    b.add_material(make_material())
    b.add_sphere((0.3, -0.2, 0.1), 0.7)
    sc = b.build()
    g = np.random.default_rng(1)
    n = 4000
    o = g.uniform(-3, 3, (n, 3))
    o = o[np.linalg.norm(o - [0.3, -0.2, 0.1], axis=1) > 0.8][:2000]
    tgt = np.array([0.3, -0.2, 0.1]) + g.normal(0, 0.4, (o.shape[0], 3))
    d = tgt - o
    r = np.zeros(o.shape[0], dtype=N.RAY_DTYPE)
    r["Origin"] = o.astype(np.float32)
    r["Direction"] = d.astype(np.float32)
    h = oracle.trace(sc, r)
    oo, dd = r["Origin"].astype(np.float64), r["Direction"].astype(np.float64)
    c = np.array([0.3, -0.2, 0.1]) - oo
    A = np.sum(dd * dd, 1)
    B = np.sum(dd * c, 1)
    Cc = np.sum(c * c, 1) - 0.49
    disc = B * B - A * Cc
    hit = disc > 1e-9 * A
    t = (B - np.sqrt(np.maximum(disc, 0))) / A
    both = hit & (h["Prim"] != N.AZ_MISS)
    assert np.mean(hit == (h["Prim"] != N.AZ_MISS)) > 0.999
    np.testing.assert_allclose(h["t"][both], t[both], rtol=2e-5)


# ---------------------------------------------------------------- BRDFs


def _brdf(kind, exponent=10.0, ior=1.5):
    b = N.brdf()
    b.Type, b.Exponent, b.RefractiveIndex = kind, exponent, ior
    return b


def _eval(b, L, Nn, V, cosI):
    lib = oracle.load()
    arr = lambda x: (C.c_float * 3)(*[float(t) for t in x])
    dif, spec = C.c_float(), C.c_float()
    lib.azo_evaluate_brdf(C.byref(b), arr(L), arr(Nn), arr(V), float(cosI), C.byref(dif), C.byref(spec))
    return dif.value, spec.value


def test_brdf_values_at_mirror_direction():
    """At the mirror direction cos(alpha) = 1: Phong -> 1/cos(theta_i),
    modified -> 1, normalized -> (n+2)/2pi and (n+8)/8pi; the non-normalized
    models keep Diffuse = 1 (main.cpp:588-589), the normalized ones 1/pi."""
    Nn = np.array([0.0, 0.0, 1.0])
    th = 0.6
    L = np.array([math.sin(th), 0, math.cos(th)])
    V = -np.array([-math.sin(th), 0, math.cos(th)])     # ray direction arriving at the surface
    n = 12.0
    d, s = _eval(_brdf(N.BRDF_Phong, n), L, Nn, V, math.cos(th))
    assert d == 1.0 and s == pytest.approx(1 / math.cos(th), rel=1e-5)
    d, s = _eval(_brdf(N.BRDF_ModifiedPhong, n), L, Nn, V, math.cos(th))
    assert d == 1.0 and s == pytest.approx(1.0, rel=1e-5)
    d, s = _eval(_brdf(N.BRDF_NormalizedModifiedPhong, n), L, Nn, V, math.cos(th))
    assert d == pytest.approx(1 / math.pi, rel=1e-6) and s == pytest.approx((n + 2) / (2 * math.pi), rel=1e-5)
    d, s = _eval(_brdf(N.BRDF_BlinnPhong, n), L, Nn, V, math.cos(th))
    assert s == pytest.approx(1 / math.cos(th), rel=1e-5)
    d, s = _eval(_brdf(N.BRDF_ModifiedBlinnPhong, n), L, Nn, V, math.cos(th))
    assert s == pytest.approx(1.0, rel=1e-5)
    d, s = _eval(_brdf(N.BRDF_NormalizedModifiedBlinnPhong, n), L, Nn, V, math.cos(th))
    assert d == pytest.approx(1 / math.pi, rel=1e-6) and s == pytest.approx((n + 8) / (8 * math.pi), rel=1e-5)


def test_normalized_modified_phong_lobe_integrates_to_one():
    """With the view along the normal, (n+2)/(2pi) cos^n(alpha) cos(theta)
    integrates to 1 over the hemisphere."""
    n = 20.0
    b = _brdf(N.BRDF_NormalizedModifiedPhong, n)
    Nn = np.array([0.0, 0.0, 1.0])
    V = np.array([0.0, 0.0, -1.0])
    ths = (np.arange(400) + 0.5) / 400 * (math.pi / 2)
    total = 0.0
    for th in ths:
        L = np.array([math.sin(th), 0.0, math.cos(th)])
        _, s = _eval(b, L, Nn, V, math.cos(th))
        total += s * math.cos(th) * math.sin(th) * (math.pi / 2 / 400) * 2 * math.pi
    assert total == pytest.approx(1.0, rel=2e-3)


def test_torrance_sparrow_fresnel_at_normal_incidence():
    """Light, view and normal aligned: D = (n+2)/2pi, G = 1, F = R0, so
    spec = D R0 / 4 (main.cpp:538-565)."""
    n, ior = 30.0, 1.5
    Nn = np.array([0.0, 0.0, 1.0])
    d, s = _eval(_brdf(N.BRDF_TorranceSparrow, n, ior), Nn, Nn, -Nn, 1.0)
    r0 = ((ior - 1) / (ior + 1)) ** 2
    assert s == pytest.approx((n + 2) / (2 * math.pi) * r0 / 4, rel=1e-5)


# ---------------------------------------------------------------- closed-form renders


def _plane_scene(**kw):
    b = SceneBuilder(max_depth=kw.get("depth", 0), path_trace=kw.get("path_trace", False),
                     importance_sampling=True, background=(0.25, 0.5, 0.75))
    b.add_material(make_material(kd=kw.get("kd", (0.6, 0.5, 0.4))))
    v = np.array([(-50, 0, -50), (50, 0, -50), (50, 0, 50), (-50, 0, 50)], np.float32)
    b.add_mesh(v, np.array([[0, 2, 1], [0, 3, 2]], np.uint32))
    for L in kw.get("lights", []):
        b.add_light(L)
    if "env" in kw:
        from azimuth_amd.scene import image_texture
        t, keep = image_texture(kw["env"])
        b.set_environment_map(t)
    return b.build()


def test_background_on_miss_and_weights():
    sc = _plane_scene()
    cam = configs.look_at_camera((0, 5, 0), (0, 10, 0.1), (0, 0, 1), 8, 8, 30.0, spp=4)   # looking up: all miss
    px = oracle.render(sc, cam)
    assert np.all(px["TotalWeight"] == 4.0)
    np.testing.assert_allclose(px["Color"] / 4.0, np.tile([0.25, 0.5, 0.75], (64, 1)), rtol=1e-6)
    pg = oracle.render(sc, cam, filt=N.FilterType_Gaussian)
    np.testing.assert_allclose(pg["Color"] / pg["TotalWeight"][:, None], np.tile([0.25, 0.5, 0.75], (64, 1)),
                               rtol=1e-5)


def test_lambertian_plane_under_uniform_environment():
    """Cosine-sampled environment NEE on a Lambertian (Kd/pi) plane under a
    constant environment of radiance Le: every sample contributes
    Kd/pi * Le * cos * pi/cos = Kd * Le exactly."""
    Le = np.array([0.8, 0.6, 0.3])
    env = np.tile((Le * 255.0).astype(np.float32), (8, 16, 1))
    sc = _plane_scene(env=env)
    cam = configs.look_at_camera((0, 3, 0.5), (0, 0, 0), (0, 0, 1), 8, 8, 20.0, spp=4)
    img = oracle.render(sc, cam, seed=3)
    col = img["Color"] / img["TotalWeight"][:, None]
    np.testing.assert_allclose(col, np.tile(np.array([0.6, 0.5, 0.4]) * Le, (64, 1)), rtol=2e-5)


def test_point_and_spot_light_irradiance():
    """Point light of intensity I at height h above a Kd/pi plane: the pixel
    looking straight down at the foot point sees Kd/pi * I / h^2; a spot light
    pointing at it (inside the inner cone) gives the same."""
    I, h = np.array([20.0, 10.0, 5.0]), 2.0
    for L in (point_light((0, h, 0), I), spot_light((0, h, 0), I, (0, -1, 0), 40.0, 20.0)):
        sc = _plane_scene(lights=[L])
        cam = configs.look_at_camera((0.0, 1.0, 0.0), (0, 0, 0), (0, 0, 1), 1, 1, 1.0, spp=1)
        px = oracle.render(sc, cam)
        col = px["Color"][0] / px["TotalWeight"][0]
        want = np.array([0.6, 0.5, 0.4]) / math.pi * I / h ** 2
        np.testing.assert_allclose(col, want, rtol=1e-4)


def test_mesh_light_estimator_matches_quadrature():
    """An emitting quad (entity light, uniform triangle choice with the
    total-area pdf, main.cpp:424-450) over a Kd/pi plane: the NEE estimate at
    the point below its centre converges to the irradiance integral
    Le * int cos cos' / d^2 dA computed by quadrature."""
    Le, s, hgt = 3.0, 0.5, 1.0
    b = SceneBuilder(max_depth=0, path_trace=False)
    b.add_material(make_material(kd=(1.0, 1.0, 1.0)))
    b.add_material(make_material(kd=(0, 0, 0)))
    v = np.array([(-50, 0, -50), (50, 0, -50), (50, 0, 50), (-50, 0, 50)], np.float32)
    b.add_mesh(v, np.array([[0, 2, 1], [0, 3, 2]], np.uint32), material=0)
    lv = np.array([(-s / 2, hgt, -s / 2), (s / 2, hgt, -s / 2), (s / 2, hgt, s / 2), (-s / 2, hgt, s / 2)], np.float32)
    b.add_mesh(lv, np.array([[0, 1, 2], [0, 2, 3]], np.uint32), material=1, emitter=True, radiance=(Le, Le, Le))
    sc = b.build()
    cam = configs.look_at_camera((0.0, 0.5, 0.0), (0, 0, 0), (0, 0, 1), 1, 1, 0.01, spp=4096)
    px = oracle.render(sc, cam, seed=5)
    est = float(px["Color"][0, 0] / px["TotalWeight"][0])
    m = 400
    xs = (np.arange(m) + 0.5) / m * s - s / 2
    X, Z = np.meshgrid(xs, xs)
    d2 = X ** 2 + Z ** 2 + hgt ** 2
    irr = float(np.sum(Le * hgt * hgt / d2 ** 2) * (s / m) ** 2)   # cos = cos' = h/d
    want = irr / math.pi                                           # Kd/pi
    assert est == pytest.approx(want, rel=0.02)


def test_perlin_lattice_points_and_range():
    """Gradient noise is 0 at lattice points, so the patch appearance gives
    (0+1)/2 = 0.5 there; values stay in [0, 1]."""
    from azimuth_amd.scene import perlin_texture
    lib = oracle.load()
    t = perlin_texture(4, scaling=1.0)
    for p in ((0, 0, 0), (3, -2, 5), (-7, 1, -1)):
        assert lib.azo_perlin_sample(C.byref(t), *p) == pytest.approx(0.5, abs=1e-7)
    g = np.random.default_rng(0).uniform(-10, 10, (2000, 3))
    vals = np.array([lib.azo_perlin_sample(C.byref(t), *q) for q in g])
    assert vals.min() >= -1e-3 and vals.max() <= 1.001
    grad = np.zeros(3, np.float32)
    for i in range(256):
        lib.azo_perlin_gradient(i, grad.ctypes.data)
        assert abs(np.linalg.norm(grad) - 1.0) < 1e-5


# ---------------------------------------------------------------- scene import


def _bvh_shape(tris, idx, depth=0):
    """ConstructBVHTrianglesOnly (bvh.cpp:223-328) on the triangles idx, in
    Python: the node box over all three vertices of each triangle
    (bvh.cpp:242-277), the split at the box centre on axis depth % 3, the
    partition loop from index 1 swapping left items to the front, a half
    split when a side is empty.  Returns the tree as nested (left, right)
    tuples with triangle indices at the leaves."""
    idx = list(idx)
    if len(idx) == 1:
        return idx[0]
    pts = tris[idx].reshape(-1, 3)
    mn, mx = pts.min(0), pts.max(0)
    centre = mn * np.float32(0.5) + mx * np.float32(0.5)
    dim = depth % 3
    split = centre[dim]
    left = 0
    for i in range(1, len(idx)):
        if tris[idx[i]][0][dim] < split:
            idx[left], idx[i] = idx[i], idx[left]
            left += 1
    if left == 0 or left == len(idx):
        left = len(idx) // 2
    return (_bvh_shape(tris, idx[:left], depth + 1), _bvh_shape(tris, idx[left:], depth + 1))


def _walk(e, tris_by_p0, boxes_ok):
    """The built entity tree as nested tuples of leaf triangle indices
    (matched by their vertices), checking that each node box holds its
    children."""
    import ref_structs as R
    if e.Type == R.EntityType_Triangle:
        t = e.u.tri
        key = (R.vec(t.P0), R.vec(t.P1), R.vec(t.P2))
        return tris_by_p0[key], np.array(key, np.float32)
    assert e.Type == R.EntityType_BVHNode
    box = e.u.boxed.AABB
    mn, mx = np.array(R.vec(box.Min), np.float32), np.array(R.vec(box.Max), np.float32)
    l, lp = _walk(e.u.boxed.p.node.Left.contents, tris_by_p0, boxes_ok)
    r, rp = _walk(e.u.boxed.p.node.Right.contents, tris_by_p0, boxes_ok)
    pts = np.concatenate([lp.reshape(-1, 3), rp.reshape(-1, 3)])
    boxes_ok.append(bool(np.all(pts >= mn) and np.all(pts <= mx)))
    return (l, r), pts


def test_reference_bvh_invariants():
    """The mesh BVH built like bvh.cpp:223-328: the same tree, leaf for leaf,
    as a Python run of the reference's construction over the mesh's
    triangles in input order; every node box holds every vertex below it;
    the leaves are exactly the mesh's triangles."""
    import ref_structs as R
    from azimuth_amd.scene import SceneBuilder
    v, f = configs.displaced_icosphere(2, 1)
    b = SceneBuilder()
    b.add_material(make_material())
    b.add_mesh(v, f)
    sc = b.build()
    assert sc.triangle_count == f.shape[0]
    tris = v[f].astype(np.float32)                       # (F, 3, 3) in input order
    keys = {tuple(map(tuple, t)): k for k, t in enumerate(tris)}
    assert len(keys) == f.shape[0]
    s = R.scene_of(sc)
    m = s.Meshes[0]
    boxes_ok = []
    built, _ = _walk(m.BVH.contents, {tuple(tuple(float(x) for x in p) for p in k): i for k, i in keys.items()},
                     boxes_ok)
    want = _bvh_shape(tris, range(f.shape[0]))
    assert built == want
    assert len(boxes_ok) == f.shape[0] - 1 and all(boxes_ok)

    def leaves(t):
        return [t] if not isinstance(t, tuple) else leaves(t[0]) + leaves(t[1])
    assert sorted(leaves(built)) == list(range(f.shape[0]))


def test_smooth_vertex_normals_are_radial():
    """xml.cpp:3632-3667: a vertex normal is the normalized sum of the
    unnormalized cross products (P1 - P0) x (P2 - P0) of its faces; on an
    icosphere they point along the radius.  Checked on the triangle entities
    the builder made (N0, N1, N2 of each face)."""
    import ref_structs as R
    v, f = configs.icosphere(3)
    v32 = v.astype(np.float32)
    sc = SceneBuilder()
    sc.add_material(make_material())
    sc.add_mesh(v32, f.astype(np.uint32), smooth=True)
    s = sc.build()
    acc = np.zeros((v.shape[0], 3))
    tri = v32[f].astype(np.float64)
    cr = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    for k in range(3):
        np.add.at(acc, f[:, k], cr)
    vn = acc / np.linalg.norm(acc, axis=1, keepdims=True)
    assert float(np.min(np.sum(vn * v / np.linalg.norm(v, axis=1, keepdims=True), 1))) > 0.9995
    mesh = R.scene_of(s).Meshes[0]
    # The mesh's triangle entities sit in the BVH leaves; walk them.
    seen = 0
    stack = [mesh.BVH.contents]
    by_pos = {tuple(np.round(p, 6)): k for k, p in enumerate(v32.astype(np.float64))}
    while stack:
        e = stack.pop()
        if e.Type == R.EntityType_BVHNode:
            stack += [e.u.boxed.p.node.Left.contents, e.u.boxed.p.node.Right.contents]
            continue
        t = e.u.tri
        for P, Nv in ((t.P0, t.N0), (t.P1, t.N1), (t.P2, t.N2)):
            k = by_pos[tuple(np.round(np.array(R.vec(P), np.float64), 6))]
            np.testing.assert_allclose(R.vec(Nv), vn[k], atol=2e-6)
        seen += 1
    assert seen == f.shape[0]


def test_builder_degamma_like_load_texture():
    """texture.h:391-413, 441-448: with Degamma set, 8-bit texels become
    (u8)(powf(c / 255, 2.2) * 255.001) and float texels powf(c, 2.2); the
    flag unset leaves them as they are."""
    import ref_structs as R
    from azimuth_amd.scene import image_texture
    u8 = np.arange(256, dtype=np.uint8).repeat(3).reshape(16, 16, 3)
    f32 = (np.arange(48, dtype=np.float32).reshape(4, 4, 3) * np.float32(0.37))
    for data in (u8, f32):
        for flag in (0, 1):
            t, keep = image_texture(data)
            t.Degamma = flag
            b = SceneBuilder()
            b.add_material(make_material())
            b.add_texture(t)
            b.add_sphere((0, 0, 0), 1.0)
            s = b.build()
            tex = R.scene_of(s).Textures[0]
            n = data.size
            ctype = C.c_uint8 if data.dtype == np.uint8 else C.c_float
            got = np.ctypeslib.as_array(C.cast(tex.u.img.Pixels, C.POINTER(ctype)), shape=(n,)).copy()
            src = data.reshape(-1).astype(np.float64)
            if not flag:
                assert np.array_equal(got, data.reshape(-1))
            elif data.dtype == np.uint8:
                exact = (src / 255.0) ** 2.2 * 255.001
                want = np.floor(exact)
                near = np.abs(exact - np.round(exact)) < 1e-4       # binary32 rounding may land either side
                assert np.all((got == want) | near), np.argwhere((got != want) & ~near)
                assert got[-1] == 255 and got[0] == 0 and got[3 * 128] == 55 and got[3 * 200] == 149
            else:
                np.testing.assert_allclose(got, src ** 2.2, rtol=2e-6)
            del keep


# ---------------------------------------------------------------- BRDFs off the mirror direction


def _textbook_brdf(kind, n, ior, L, Nn, Vray):
    """The seven models of EvaluateBRDF (main.cpp:461-569) in float64 from
    their definitions: Phong lobe (R . w_o)^n with R = reflect(L, N), Blinn
    lobe (N . H)^n, normalizations (n+2)/2pi and (n+8)/8pi, the unnormalized
    Phong and Blinn models divided by cos(theta_i), Torrance-Sparrow with the
    Blinn distribution, G = min(1, 2 (N.H) min(N.w_o, N.L) / (w_o.H)) and
    Schlick's F at w_o.H.  Returns (diffuse, specular) factors."""
    wo = -Vray
    cosI = float(Nn @ L)
    R = 2.0 * (L @ Nn) * Nn - L
    H = (wo + L) / np.linalg.norm(wo + L)
    cp, cb = float(wo @ R), float(Nn @ H)
    d = 1.0 / math.pi if kind in (N.BRDF_NormalizedModifiedPhong, N.BRDF_NormalizedModifiedBlinnPhong,
                                  N.BRDF_TorranceSparrow) else 1.0
    if kind == N.BRDF_Phong:
        return d, (cp ** n / cosI if cp > 0 else 0.0)
    if kind == N.BRDF_ModifiedPhong:
        return d, (cp ** n if cp > 0 else 0.0)
    if kind == N.BRDF_NormalizedModifiedPhong:
        return d, ((n + 2) / (2 * math.pi) * cp ** n if cp > 0 else 0.0)
    if kind == N.BRDF_BlinnPhong:
        return d, (cb ** n / cosI if cb > 0 else 0.0)
    if kind == N.BRDF_ModifiedBlinnPhong:
        return d, (cb ** n if cb > 0 else 0.0)
    if kind == N.BRDF_NormalizedModifiedBlinnPhong:
        return d, ((n + 8) / (8 * math.pi) * cb ** n if cb > 0 else 0.0)
    if cb <= 0:
        return d, 0.0
    D = (n + 2) / (2 * math.pi) * cb ** n
    nwo, woh = float(Nn @ wo), float(wo @ H)
    G = min(1.0, 2.0 * cb * min(nwo, cosI) / woh)
    r0 = (ior - 1) ** 2 / (ior + 1) ** 2
    F = r0 + (1 - r0) * (1 - woh) ** 5
    return d, D * F * G / (4 * cosI * nwo)


def _random_pair(g):
    Nn = g.normal(size=3)
    Nn /= np.linalg.norm(Nn)
    while True:
        L, wo = g.normal(size=3), g.normal(size=3)
        L /= np.linalg.norm(L)
        wo /= np.linalg.norm(wo)
        if L @ Nn > 0.05 and wo @ Nn > 0.05:
            return L, Nn, -wo


def test_brdfs_at_random_directions_against_their_definitions():
    g = np.random.default_rng(5)
    for kind in range(7):
        for _ in range(200):
            L, Nn, V = _random_pair(g)
            n = float(g.uniform(2, 60))
            d, s = _eval(_brdf(kind, n, 1.6), L, Nn, V, float(L @ Nn))
            wd, ws = _textbook_brdf(kind, n, 1.6, L, Nn, V)
            assert d == pytest.approx(wd, rel=1e-6)
            assert s == pytest.approx(ws, rel=3e-4, abs=1e-7), (kind, n)


def test_brdf_reciprocity():
    """Swapping the light and view directions leaves the reciprocal models'
    value unchanged (and the Phong and Blinn models' value times
    cos(theta_i), their lobe)."""
    g = np.random.default_rng(9)
    for kind in range(7):
        for _ in range(100):
            L, Nn, V = _random_pair(g)
            n = float(g.uniform(2, 60))
            _, a = _eval(_brdf(kind, n, 1.4), L, Nn, V, float(L @ Nn))
            _, b = _eval(_brdf(kind, n, 1.4), -V, Nn, -L, float(-V @ Nn))
            if kind in (N.BRDF_Phong, N.BRDF_BlinnPhong):
                a, b = a * float(L @ Nn), b * float(-V @ Nn)
            assert a == pytest.approx(b, rel=5e-4, abs=1e-7), kind


def test_torrance_sparrow_fresnel_goes_to_one_at_grazing():
    """Schlick's F = r0 + (1 - r0)(1 - w_o.H)^5 tends to 1 as w_o.H -> 0:
    with light and view at the mirror configuration near grazing (H = N,
    w_o.H = cos 89.5 deg), the specular value over D G / (4 cos_i cos_o)
    approaches 1 whatever the index."""
    Nn = np.array([0.0, 0.0, 1.0])
    for ior in (1.3, 2.5):
        th = math.radians(89.5)
        L = np.array([math.sin(th), 0.0, math.cos(th)])
        wo = np.array([-math.sin(th), 0.0, math.cos(th)])
        d, s = _eval(_brdf(N.BRDF_TorranceSparrow, 4.0, ior), L, Nn, -wo, float(L @ Nn))
        H = (wo + L) / np.linalg.norm(wo + L)
        cb, woh = float(Nn @ H), float(wo @ H)
        D = 6.0 / (2 * math.pi) * cb ** 4
        G = min(1.0, 2.0 * cb * min(float(Nn @ wo), float(Nn @ L)) / woh)
        F = s / (D * G / (4 * float(Nn @ L) * float(Nn @ wo)))
        r0 = (ior - 1) ** 2 / (ior + 1) ** 2
        assert F == pytest.approx(r0 + (1 - r0) * (1 - woh) ** 5, rel=1e-3)
        assert F > 0.9


# ---------------------------------------------------------------- Perlin noise


def test_perlin_gradient_at_lattice_points_follows_the_hash():
    """Near a lattice point Q the noise is dot(g_h, P - Q) to second order,
    g_h the gradient table entry at h = Perm[(I + Perm[(J + Perm[K]) % 256]) % 256]
    (PerlinHash, texture.h:134-144): finite differences of the patch value
    (raw + 1) / 2 at Q recover g_h / 2.  The vein appearance is |raw|."""
    from azimuth_amd.scene import perlin_texture
    lib = oracle.load()
    t = perlin_texture(21, scaling=1.0)
    tv = perlin_texture(21, scaling=1.0, appearance=N.TexturePerlinAppearance_Vein)
    perm = np.array(t.u.perlin.Perm[:], np.int64)
    grad = np.zeros(3, np.float32)
    eps = 1e-3
    for (I, J, K) in ((0, 0, 0), (3, 7, 1), (12, 5, 200), (100, 2, 33)):
        h = perm[(I + perm[(J + perm[K % 256]) % 256]) % 256]
        lib.azo_perlin_gradient(int(h), grad.ctypes.data)
        for k in range(3):
            P = np.array([I, J, K], np.float32)
            P[k] = np.float32(P[k] + eps)
            step = float(P[k]) - (I, J, K)[k]                 # the binary32 offset actually taken
            d = (lib.azo_perlin_sample(C.byref(t), *[float(x) for x in P]) - 0.5) * 2.0 / step
            assert d == pytest.approx(float(grad[k]), abs=3e-3), (I, J, K, k)
    q = np.random.default_rng(3).uniform(0, 50, (500, 3))
    for p in q:
        raw = (lib.azo_perlin_sample(C.byref(t), *p) - 0.5) * 2.0
        assert lib.azo_perlin_sample(C.byref(tv), *p) == pytest.approx(abs(raw), abs=2e-6)
