"""Guided rendering and environment luminance sampling: both replace the
reference's cosine-distributed environment directions by another
distribution (the guide in a one-sample MIS mixture; the map's luminance
CDF), so both must stay unbiased (their mean image converges to the plain
tracer's) and should reduce variance on a scene lit by a small, bright sun."""
import numpy as np
import pytest

from azimuth_amd import native as N, workloads

pytestmark = pytest.mark.gpu


def resolve(px):
    return px[:, :3] / np.maximum(px[:, 3:4], 1e-300)


def test_guided_interior_unbiased_and_lower_variance(gpu):
    # The architectural interior at 1/8 of its 1920x1080 resolution, depth 6:
    # big enough that the online training sees the whole scene (at a few
    # thousand pixels the gain varies run to run with the training batches).
    sc, cam = workloads.interior_scene(240, 135, spp=64, max_depth=6)
    gpu.upload_scene(sc)
    filt = N.FilterType_Box
    ref = resolve(gpu.render_f64(cam, filt=filt, seed=100, spp=2304))
    plain = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64))
    guided = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64, guide=True))
    mse_p = float(np.mean((plain - ref) ** 2))
    mse_g = float(np.mean((guided - ref) ** 2))
    # Unbiasedness: the image means agree within their Monte Carlo error,
    # estimated from the per-pixel errors against the reference.
    m_ref, m_g = float(ref.mean()), float(guided.mean())
    se = float(np.sqrt(mse_g / guided.size)) * 4 + float(np.sqrt(mse_p / plain.size))
    print(f"mean ref {m_ref:.5f} plain {plain.mean():.5f} guided {m_g:.5f} (se {se:.2e}); "
          f"mse plain {mse_p:.4e} guided {mse_g:.4e} ratio {mse_p / mse_g:.2f}", flush=True)
    assert np.all(np.isfinite(guided))
    assert abs(m_g - m_ref) < 5 * se + 1e-3 * m_ref
    assert mse_g < mse_p
    # The training schedule is a caller's choice (16 spp in 2-spp chunks
    # here); the estimator stays unbiased whatever the network learned.
    g2 = resolve(gpu.render_f64(cam, filt=filt, seed=2, spp=64, guide=True, guide_train_spp=16, guide_chunk_spp=2))
    mse_2 = float(np.mean((g2 - ref) ** 2))
    se2 = float(np.sqrt(mse_2 / g2.size)) * 4 + float(np.sqrt(mse_p / plain.size))
    print(f"guided (train 16 spp, 2-spp chunks) mean {g2.mean():.5f} mse {mse_2:.4e}", flush=True)
    assert np.all(np.isfinite(g2))
    assert abs(float(g2.mean()) - m_ref) < 5 * se2 + 1e-3 * m_ref
    assert mse_2 < mse_p


def test_luminance_environment_sampling_unbiased_and_lower_variance(gpu):
    sc, cam = workloads.interior_scene(96, 54, spp=64, max_depth=3, shells=12)
    gpu.upload_scene(sc)
    filt = N.FilterType_Box
    ref = resolve(gpu.render_f64(cam, filt=filt, seed=100, spp=2304))
    plain = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64))
    lum = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64, env_sampling=N.AZ_ENV_LUMINANCE))
    both = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64, guide=True, env_sampling=N.AZ_ENV_LUMINANCE))
    mse_p = float(np.mean((plain - ref) ** 2))
    mse_l = float(np.mean((lum - ref) ** 2))
    mse_b = float(np.mean((both - ref) ** 2))
    m_ref = float(ref.mean())
    se_ref = float(np.sqrt(mse_p / plain.size)) * (64 / 2304) ** 0.5 * 4
    print(f"mean ref {m_ref:.5f} luminance {lum.mean():.5f} guide+luminance {both.mean():.5f}; "
          f"mse plain {mse_p:.4e} luminance {mse_l:.4e} guide+luminance {mse_b:.4e}", flush=True)
    assert np.all(np.isfinite(lum)) and np.all(np.isfinite(both))
    # The reference itself carries Monte Carlo error (se_ref); the luminance
    # renders' own error is small.
    for img, mse in ((lum, mse_l), (both, mse_b)):
        se = float(np.sqrt(mse / img.size)) * 4
        assert abs(float(img.mean()) - m_ref) < 5 * (se + se_ref) + 1e-3 * m_ref
    assert mse_l < mse_p


def test_luminance_sampling_without_environment_changes_nothing(gpu):
    from azimuth_amd import configs
    sc, cam = configs.cornell_scene(24, 24, spp=4, max_depth=3)
    gpu.upload_scene(sc)
    a = gpu.render(cam, seed=2)
    b = gpu.render(cam, seed=2, env_sampling=N.AZ_ENV_LUMINANCE)
    assert a.tobytes() == b.tobytes()


def emitter_floor_scene(max_depth, path_trace=True):
    """A diffuse floor lit by a small, bright emitting sphere: the scatter
    rays that happen to hit the sphere carry most of the noise."""
    from azimuth_amd.scene import SceneBuilder, make_material
    b = SceneBuilder(max_depth=max_depth, path_trace=path_trace, importance_sampling=True)
    floor = b.add_material(make_material(kd=(0.5, 0.5, 0.5)))
    dark = b.add_material(make_material(kd=(0, 0, 0)))
    v = np.array([(-20, 0, -20), (20, 0, -20), (20, 0, 20), (-20, 0, 20)], np.float32)
    b.add_mesh(v, np.array([[0, 2, 1], [0, 3, 2]], np.uint32), material=floor)
    b.add_sphere((0.0, 1.0, -1.0), 0.15, material=dark, emitter=True, radiance=(40.0, 40.0, 40.0))
    return b.build()


def test_scatter_guide_unbiased_and_lower_variance(gpu):
    """The scatter guide (GuideTarget AZ_GUIDE_SCATTER) learns, from the
    radiance each scatter ray's path delivers, to aim the floor's scatter
    rays at the emitter; mixed with the reference's cosine sampling (and,
    second, with BRDF lobe sampling) it stays unbiased and lowers the noise."""
    from azimuth_amd import configs
    sc = emitter_floor_scene(3)
    cam = configs.look_at_camera((0, 1, 1), (0, 0, 0), (0, 1, 0), 96, 96, 40.0, spp=64)
    gpu.upload_scene(sc)
    ref = resolve(gpu.render_f64(cam, seed=100, spp=4096))
    plain = resolve(gpu.render_f64(cam, seed=1, spp=64))
    mse_p = float(np.mean((plain - ref) ** 2))
    m_ref = float(ref.mean())
    for base in (N.AZ_SCATTER_REFERENCE, N.AZ_SCATTER_BRDF):
        img = resolve(gpu.render_f64(cam, seed=1, spp=64, guide=True, guide_target=N.AZ_GUIDE_SCATTER,
                                     scatter_sampling=base))
        mse = float(np.mean((img - ref) ** 2))
        se = float(np.sqrt(mse / img.size)) * 4 + float(np.sqrt(mse_p / plain.size)) * (64 / 4096) ** 0.5 * 4
        print(f"base {base}: mean ref {m_ref:.5f} plain {plain.mean():.5f} scatter-guided {img.mean():.5f} "
              f"(se {se:.2e}); mse plain {mse_p:.4e} guided {mse:.4e} ratio {mse_p / mse:.2f}", flush=True)
        assert np.all(np.isfinite(img))
        assert abs(float(img.mean()) - m_ref) < 5 * se + 1e-3 * m_ref
        assert mse < 0.5 * mse_p


def test_both_guides_unbiased_on_interior(gpu):
    """Both networks at once (environment NEE and scatter), and the scatter
    guide alone, on the interior at 1/8 of 1080p: unbiased.  Their noise is
    not asserted here: with two networks trained online the MSE varies from
    run to run (0.15-0.23 in five of six runs against 1.09 plain, once 1.10);
    the guide experiment measures it over several seeds."""
    sc, cam = workloads.interior_scene(240, 135, spp=64, max_depth=6)
    gpu.upload_scene(sc)
    filt = N.FilterType_Box
    ref = resolve(gpu.render_f64(cam, filt=filt, seed=100, spp=2304))
    plain = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64))
    both = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64, guide=True, guide_target=N.AZ_GUIDE_BOTH))
    scat = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64, guide=True, guide_target=N.AZ_GUIDE_SCATTER))
    mse_p = float(np.mean((plain - ref) ** 2))
    m_ref = float(ref.mean())
    se_ref = float(np.sqrt(mse_p / plain.size)) * (64 / 2304) ** 0.5 * 4
    for name, img in (("both", both), ("scatter", scat)):
        mse = float(np.mean((img - ref) ** 2))
        se = float(np.sqrt(mse / img.size)) * 4
        print(f"{name}: mean ref {m_ref:.5f} guided {img.mean():.5f}; mse plain {mse_p:.4e} guided {mse:.4e}",
              flush=True)
        assert np.all(np.isfinite(img))
        assert abs(float(img.mean()) - m_ref) < 5 * (se + se_ref) + 1e-3 * m_ref


def test_scatter_guide_without_path_tracing_changes_nothing(gpu):
    """Without hemisphere scatter there is nothing for the scatter guide to
    propose: the render equals the plain one."""
    from azimuth_amd import configs
    sc = emitter_floor_scene(3, path_trace=False)
    cam = configs.look_at_camera((0, 1, 1), (0, 0, 0), (0, 1, 0), 24, 24, 40.0, spp=4)
    gpu.upload_scene(sc)
    a = gpu.render(cam, seed=2)
    b = gpu.render(cam, seed=2, guide=True, guide_target=N.AZ_GUIDE_SCATTER)
    assert a.tobytes() == b.tobytes()


def test_guided_interior_full_size(gpu):
    """The guided-interior case at its own size: 1920x1080, 64 spp, depth 6,
    guide fraction 0.5, against a 1,024-spp plain render accumulated in
    double precision.  The guided render is unbiased (image means agree
    within Monte Carlo error) and has lower MSE than the plain one at equal
    spp."""
    sc, cam = workloads.interior_scene(1920, 1080, spp=64, max_depth=6)
    gpu.upload_scene(sc)
    filt = N.FilterType_Gaussian
    ref = resolve(gpu.render_f64(cam, filt=filt, seed=100, spp=1024))
    plain = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64))
    t_plain = gpu.last_render_s
    guided = resolve(gpu.render_f64(cam, filt=filt, seed=1, spp=64, guide=True))
    t_guided = gpu.last_render_s
    mse_p = float(np.mean((plain - ref) ** 2))
    mse_g = float(np.mean((guided - ref) ** 2))
    m_ref, m_g = float(ref.mean()), float(guided.mean())
    se = float(np.sqrt(mse_g / guided.size)) * 4 + float(np.sqrt(mse_p / plain.size))
    print(f"1080p: mean ref {m_ref:.5f} plain {plain.mean():.5f} guided {m_g:.5f} (se {se:.2e}); "
          f"mse plain {mse_p:.4e} ({t_plain:.2f}s) guided {mse_g:.4e} ({t_guided:.2f}s) "
          f"ratio {mse_p / mse_g:.2f}", flush=True)
    assert np.all(np.isfinite(guided))
    assert abs(m_g - m_ref) < 5 * se + 1e-3 * m_ref
    assert mse_g < 0.5 * mse_p


def test_guided_renders_repeat_bit_for_bit(gpu):
    """Training batches are drawn from the records in the order of their path
    vertex (sample, random-stream node), not in the order the waves appended
    them, and the gradients are summed in a fixed order, so a guided render
    with the same seed gives the same pixels bit for bit: with the
    environment guide, and with both guides (the scatter guide's records and
    their radiance links too)."""
    sc, cam = workloads.interior_scene(160, 90, spp=16, max_depth=5)
    gpu.upload_scene(sc)
    for target in (N.AZ_GUIDE_ENV, N.AZ_GUIDE_BOTH):
        a = gpu.render(cam, seed=4, guide=True, guide_target=target)
        b = gpu.render(cam, seed=4, guide=True, guide_target=target)
        c = gpu.render(cam, seed=5, guide=True, guide_target=target)
        assert np.all(a["TotalWeight"] > 0)
        assert a.tobytes() == b.tobytes(), target
        assert a.tobytes() != c.tobytes(), target
