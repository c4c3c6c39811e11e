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
