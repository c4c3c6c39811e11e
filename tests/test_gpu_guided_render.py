"""Guided rendering: the one-sample MIS estimator with the learned guide must
stay unbiased (its mean image converges to the plain tracer's) and should
reduce variance on a scene lit by a small, bright sun."""
import numpy as np
import pytest

from azimuth_amd import native as N, workloads

pytestmark = pytest.mark.gpu


def resolve(px):
    return px[:, :3] / np.maximum(px[:, 3:4], 1e-300)


def test_guided_interior_unbiased_and_lower_variance(gpu):
    sc, cam = workloads.interior_scene(96, 54, spp=64, max_depth=3, shells=12)
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
