"""GPU closest-hit / any-hit parity with the oracle (reference BVH walk,
main.cpp:53-385), through the C ABI.

Bar: primitive ids bit-exact except ties (two primitives whose t differ by at
most 1e-6 relative); where ids agree, t, beta and gamma are compared bit for
bit (the stated tolerance is 1e-5 relative; the kernels use the reference's
operation order, so they are expected to be equal)."""
import numpy as np
import pytest

import oracle
from azimuth_amd import configs, native as N

pytestmark = pytest.mark.gpu

TIE_REL = 1e-6
VAL_REL = 1e-5


def compare_closest(got, want, label, allow_mismatch_frac=0.0):
    gp, wp = got["Prim"], want["Prim"]
    same = gp == wp
    diff = ~same
    # A differing id is a tie when both hit and their t agree within 1e-6.
    both = diff & (gp != N.AZ_MISS) & (wp != N.AZ_MISS)
    tie = both & (np.abs(got["t"] - want["t"]) <= TIE_REL * np.abs(want["t"]))
    bad = diff & ~tie
    nbad = int(bad.sum())
    n = got.shape[0]
    print(f"{label}: {n} rays, hits {int((wp != N.AZ_MISS).sum())}, id mismatches {int(diff.sum())} "
          f"(ties {int(tie.sum())}, non-tie {nbad})", flush=True)
    assert nbad <= allow_mismatch_frac * n, f"{label}: {nbad} non-tie id mismatches, first at {np.nonzero(bad)[0][:5]}"
    h = same & (wp != N.AZ_MISS)
    for f in ("t", "Beta", "Gamma"):
        a, b = got[f][h], want[f][h]
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-30)
        assert np.all((a == b) | (rel <= VAL_REL)), f"{label}: {f} outside 1e-5 relative"
    exact = np.mean((got["t"][h] == want["t"][h]) & (got["Beta"][h] == want["Beta"][h]) &
                    (got["Gamma"][h] == want["Gamma"][h])) if h.any() else 1.0
    return exact


@pytest.fixture(scope="module")
def ico5():
    return configs.icosphere_scene(5, seed=7)


def test_closest_icosphere_coherent_incoherent(gpu, ico5):
    gpu.upload_scene(ico5)
    rays = np.concatenate([configs.coherent_rays(256, 256), configs.incoherent_rays(65536, ico5.meta["bbox"], 11)])
    got = gpu.trace(rays)
    want = oracle.trace(ico5, rays)
    exact = compare_closest(got, want, "icosphere s5")
    assert exact == 1.0, f"t/beta/gamma bit-exact on only {exact:.6f} of matching hits"


def test_closest_matches_brute_force(gpu):
    sc = configs.icosphere_scene(3, seed=3)
    gpu.upload_scene(sc)
    rays = configs.incoherent_rays(8192, sc.meta["bbox"], 12)
    got = gpu.trace(rays)
    want = oracle.trace(sc, rays, brute=True)
    compare_closest(got, want, "brute force s3")


def test_any_hit_occlusion(gpu, ico5):
    gpu.upload_scene(ico5)
    rays = configs.incoherent_rays(65536, ico5.meta["bbox"], 13)
    g = np.random.default_rng(5)
    rays["TMax"] = g.uniform(0.0, 2.5, rays.shape[0]).astype(np.float32)
    got = gpu.trace(rays, any_hit=True)
    want = oracle.trace(ico5, rays, any_hit=True)
    occ_g = got["Prim"] != N.AZ_MISS
    occ_w = want["Prim"] != N.AZ_MISS
    flip = occ_g != occ_w
    # A flip is a boundary case when the closest t sits within 1e-6 of TMax.
    close = oracle.trace(ico5, rays)
    boundary = np.abs(close["t"] - rays["TMax"]) <= TIE_REL * np.abs(rays["TMax"])
    assert not np.any(flip & ~boundary), f"occlusion differs on {int((flip & ~boundary).sum())} rays"
    # Every reported occluder is a genuine hit below TMax.
    assert np.all(got["t"][occ_g] < rays["TMax"][occ_g])
    assert np.all(got["t"][occ_g] > 0.0)


def test_backface_culling(gpu, ico5):
    gpu.upload_scene(ico5)
    rays = configs.incoherent_rays(32768, ico5.meta["bbox"], 14)
    got = gpu.trace(rays, cull=True)
    want = oracle.trace(ico5, rays, cull=True)
    compare_closest(got, want, "culling")


def test_axis_aligned_and_degenerate_rays(gpu, ico5):
    gpu.upload_scene(ico5)
    dirs = np.array([(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (0.0, -0.0, 1),
                     (-0.0, 0.0, -1)], np.float32)
    g = np.random.default_rng(3)
    n = 4096
    r = np.zeros(n, dtype=N.RAY_DTYPE)
    r["Origin"] = g.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    r["Direction"] = dirs[g.integers(0, len(dirs), n)]
    r["TMax"] = np.inf
    got = gpu.trace(r)
    want = oracle.trace(ico5, r)
    compare_closest(got, want, "axis-aligned")


def test_ragged_batches_zero_and_far_rays(gpu, ico5):
    """Batch sizes that leave partial waves and small launches (one ray, one
    short of a wave, one past it, ...), with zero-length directions (the
    reference normalizes them to zero: no hit) and rays from far away."""
    gpu.upload_scene(ico5)
    g = np.random.default_rng(11)
    for n in (1, 63, 65, 257, 4099):
        r = np.zeros(n, dtype=N.RAY_DTYPE)
        r["Origin"] = g.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
        d = g.normal(size=(n, 3))
        d[::7] = 0.0                                         # zero-length directions
        r["Direction"] = (d / np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-30)).astype(np.float32)
        far = np.arange(n) % 5 == 3                          # from 1e4 away, aimed at the mesh
        r["Origin"][far] = (-1e4 * r["Direction"][far]).astype(np.float32)
        r["TMax"] = np.inf
        got = gpu.trace(r)
        want = oracle.trace(ico5, r)
        compare_closest(got, want, f"ragged n={n}")
        assert np.all(got["Prim"][::7][np.all(r["Direction"][::7] == 0, axis=1)] == N.AZ_MISS)
        ga, wa = gpu.trace(r, any_hit=True), oracle.trace(ico5, r, any_hit=True)
        assert np.array_equal(ga["Prim"] != N.AZ_MISS, wa["Prim"] != N.AZ_MISS), f"any-hit n={n}"


def test_empty_batch(gpu, ico5):
    gpu.upload_scene(ico5)
    r = np.zeros(0, dtype=N.RAY_DTYPE)
    got = gpu.trace(r)
    assert got.shape == (0,)


def test_spheres_instances_motion(gpu):
    """Spheres with transforms and motion, meshes with transforms: the
    two-level path (main.cpp:64-146, 244-281)."""
    from azimuth_amd.scene import SceneBuilder, make_material
    b = SceneBuilder()
    b.add_material(make_material())
    g = np.random.default_rng(21)
    for i in range(300):
        c = g.uniform(-5, 5, 3)
        T = np.eye(4)
        if i % 3 == 0:
            s = g.uniform(0.5, 2.0)
            T[:3, :3] *= s
            T[:3, 3] = g.uniform(-1, 1, 3)
        mv = g.uniform(-0.5, 0.5, 3) if i % 2 == 0 else np.zeros(3)
        b.add_sphere(c, float(np.exp(g.uniform(np.log(0.05), np.log(0.5)))), transform=T, motion=mv)
    v, f = configs.displaced_icosphere(3, 4)
    for k in range(3):
        T = np.eye(4)
        ang = 0.3 * (k + 1)
        T[:3, :3] = np.array([[np.cos(ang), 0, np.sin(ang)], [0, 1, 0], [-np.sin(ang), 0, np.cos(ang)]]) * (1.0 + k)
        T[:3, 3] = (k * 2.0 - 2.0, 0.5 * k, -k)
        b.add_mesh(v, f, transform=T, motion=(0.2 * k, 0, 0), smooth=True)
    sc = b.build()
    gpu.upload_scene(sc)
    r = configs.incoherent_rays(65536, (np.array([-6, -6, -6]), np.array([6, 6, 6])), 31)
    r["Time"] = g.random(r.shape[0]).astype(np.float32)
    got = gpu.trace(r)
    want = oracle.trace(sc, r)
    compare_closest(got, want, "spheres+instances+motion")
    r["TMax"] = g.uniform(0, 4, r.shape[0]).astype(np.float32)
    ga = gpu.trace(r, any_hit=True)
    wa = oracle.trace(sc, r, any_hit=True)
    close = oracle.trace(sc, r)
    boundary = np.abs(close["t"] - r["TMax"]) <= TIE_REL * np.abs(r["TMax"])
    flip = (ga["Prim"] != N.AZ_MISS) != (wa["Prim"] != N.AZ_MISS)
    assert not np.any(flip & ~boundary)


def test_full_size_icosphere_sampled(gpu):
    """The 1,310,720-triangle displaced icosphere with the full 4,194,304-ray
    coherent and incoherent batches on the GPU; every 16th ray is checked
    against the oracle, and every GPU hit is re-checked for t > 0."""
    sc = configs.icosphere_scene(8, seed=1)
    gpu.upload_scene(sc)
    for label, rays in (("coherent", configs.coherent_rays(2048, 2048)),
                        ("incoherent", configs.incoherent_rays(4194304, sc.meta["bbox"], 2))):
        got = gpu.trace(rays)
        sub = slice(0, None, 16)
        want = oracle.trace(sc, rays[sub])
        exact = compare_closest(got[sub], want, f"full {label}")
        assert exact == 1.0
        h = got["Prim"] != N.AZ_MISS
        assert np.all(got["t"][h] > 0)


def test_exact_ties_follow_reference_order(gpu):
    """Rays through the exact seams of the Cornell box (camera on the box's
    symmetry axis): equal t on two triangles resolves to the right-most leaf
    of the reference BVH, as its recursive minimum does."""
    sc, cam = configs.cornell_scene(64, 64, spp=1, with_spheres=False)
    gpu.upload_scene(sc)
    n = 64
    u = (np.arange(n) + 0.5) / n
    # Targets on the four seams of the back wall and the floor/right-wall seam.
    pts = np.concatenate([np.stack([u, np.zeros(n), np.zeros(n)], 1), np.stack([np.zeros(n), u, np.zeros(n)], 1),
                          np.stack([np.ones(n), u, np.zeros(n)], 1), np.stack([np.ones(n), np.zeros(n), u], 1)])
    eye = np.array([0.5, 0.5, 2.0], np.float32)
    r = np.zeros(pts.shape[0], dtype=N.RAY_DTYPE)
    r["Origin"] = eye
    r["Direction"] = (pts.astype(np.float32) - eye)
    r["TMax"] = np.inf
    got = gpu.trace(r)
    want = oracle.trace(sc, r)
    brute = oracle.trace(sc, r, brute=True)
    assert np.array_equal(want["Prim"], brute["Prim"])
    assert np.array_equal(got["Prim"], want["Prim"]), f"{int((got['Prim'] != want['Prim']).sum())} tie mismatches"
    assert np.array_equal(got["t"], want["t"])


def test_degenerate_triangles(gpu):
    """A displaced icosphere in which every fifth face is collapsed to a point
    and every seventh made collinear (zero area, as meshes exported with
    welded vertices carry them): the zero-area faces never hit (the
    reference's Cramer's rule divides by a zero determinant), their padded
    boxes do not disturb the walk, and the other faces' hits are the
    oracle's, ids and t, beta, gamma, closest and any hit."""
    from azimuth_amd.scene import SceneBuilder, make_material
    v, f = configs.displaced_icosphere(4, 5)
    v = v.astype(np.float32).copy()
    f = f.astype(np.uint32).copy()
    f[::5, 1] = f[::5, 0]
    f[::5, 2] = f[::5, 0]                                  # a point
    col = np.arange(3, f.shape[0], 7)
    v = np.concatenate([v, 0.5 * (v[f[col, 0]] + v[f[col, 1]])]).astype(np.float32)
    f[col, 2] = np.arange(v.shape[0] - col.size, v.shape[0], dtype=np.uint32)   # on the edge (0, 1)
    b = SceneBuilder(max_depth=1)
    b.add_material(make_material())
    b.add_mesh(v, f)
    sc = b.build()
    gpu.upload_scene(sc)
    lo, hi = v.min(0), v.max(0)
    rays = np.concatenate([configs.coherent_rays(192, 192), configs.incoherent_rays(32768, (lo, hi), 21)])
    got = gpu.trace(rays)
    want = oracle.trace(sc, rays)
    exact = compare_closest(got, want, "degenerate faces")
    assert exact == 1.0
    assert int((want["Prim"] != N.AZ_MISS).sum()) > 10000
    anyg = gpu.trace(rays, any_hit=True)
    anyw = oracle.trace(sc, rays, any_hit=True)
    assert np.array_equal(anyg["Prim"] != N.AZ_MISS, anyw["Prim"] != N.AZ_MISS)
