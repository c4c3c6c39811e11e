"""CPU pins of the float64 guide reference (oracle/guide_ref.py): finite
differences, the vMF normalization and moments, and the Adam update."""
import math

import numpy as np
import pytest

from oracle import guide_ref as R


def test_gradients_match_finite_differences():
    g = np.random.default_rng(0)
    Ws, bs = R.init_params(1)
    Ws = [w.astype(np.float64) for w in Ws]
    bs = [b.astype(np.float64) + g.normal(0, 0.1, b.shape) for b in bs]
    n = 12
    x = R.encode(g.random((n, 3)), g.normal(size=(n, 3)), g.random(n))
    w = g.normal(size=(n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    wgt = g.random(n) + 0.1
    _, dWs, dbs = R.loss_and_grads(Ws, bs, x, w, wgt)
    eps = 1e-6
    for l in range(5):
        for _ in range(4):
            i, j = g.integers(Ws[l].shape[0]), g.integers(Ws[l].shape[1])
            Wp = [a.copy() for a in Ws]
            Wm = [a.copy() for a in Ws]
            Wp[l][i, j] += eps
            Wm[l][i, j] -= eps
            fd = (R.loss_and_grads(Wp, bs, x, w, wgt)[0] - R.loss_and_grads(Wm, bs, x, w, wgt)[0]) / (2 * eps)
            assert dWs[l][i, j] == pytest.approx(fd, rel=1e-4, abs=1e-9)
        j = g.integers(bs[l].shape[0])
        bp = [a.copy() for a in bs]
        bm = [a.copy() for a in bs]
        bp[l][j] += eps
        bm[l][j] -= eps
        fd = (R.loss_and_grads(Ws, bp, x, w, wgt)[0] - R.loss_and_grads(Ws, bm, x, w, wgt)[0]) / (2 * eps)
        assert dbs[l][j] == pytest.approx(fd, rel=1e-4, abs=1e-9)


@pytest.mark.parametrize("kappa", [0.05, 1.0, 30.0, 3000.0])
def test_vmf_pdf_normalization(kappa):
    """kappa / (4 pi sinh kappa) exp(kappa mu.w) integrates to 1 on S^2."""
    m = 200000
    th = (np.arange(m) + 0.5) / m * np.pi
    w = np.stack([np.sin(th), 0 * th, np.cos(th)], 1)
    lp = R.vmf_log_pdf(np.tile([[[0.0, 0.0, 1.0]]], (m, 1, 1)), np.full((m, 1), kappa), w)[:, 0]
    assert np.sum(np.exp(lp) * np.sin(th)) * (np.pi / m) * 2 * np.pi == pytest.approx(1.0, rel=1e-6)


@pytest.mark.parametrize("kappa", [0.5, 5.0, 200.0])
def test_vmf_sampler_mean_cosine(kappa):
    """E[mu.w] = coth(kappa) - 1/kappa for the Jakob 2012 sampler."""
    g = np.random.default_rng(3)
    n = 200000
    mu = np.tile([0.3, -0.4, np.sqrt(1 - 0.25)], (n, 1))
    d = R.sample_vmf(mu, np.full(n, kappa), g.random(n), g.random(n))
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-9)
    want = 1 / math.tanh(kappa) - 1 / kappa
    assert np.mean(np.sum(d * mu, 1)) == pytest.approx(want, abs=4e-3)


def test_adam_first_step():
    p = [np.array([1.0, -2.0, 3.0])]
    gr = [np.array([0.5, -1e-3, 2.0])]
    m, v = [np.zeros(3)], [np.zeros(3)]
    out = R.adam_step(p, gr, m, v, 1)
    np.testing.assert_allclose(out[0], p[0] - 1e-3 * np.sign(gr[0]), rtol=0, atol=1e-7)


def test_encoding_layout():
    x = R.encode(np.array([[0.25, 0.5, 0.75]]), np.array([[0.0, 1.0, 0.0]]), np.array([0.3]))
    assert x.shape == (1, 128)
    # frequency 0, coordinate 1 (y = 0.5): sin(pi/2) = 1, cos = 0
    assert x[0, 2] == pytest.approx(1.0) and abs(x[0, 3]) < 1e-12
    # frequency 2 of x = 0.25: sin(pi) = 0, cos(pi) = -1 at features 48, 49;
    # frequency 15 of z = 0.75: cos(24576 pi) = 1 at feature 95
    assert abs(x[0, 48]) < 1e-12 and x[0, 49] == pytest.approx(-1.0)
    assert x[0, 95] == pytest.approx(1.0)
    j, c, cs = R.feature_map()
    assert len(set(zip(j.tolist(), c.tolist(), cs.tolist()))) == 96
    assert list(x[0, 96:100]) == [0.0, 1.0, 0.0, pytest.approx(0.3)]
    assert np.all(x[0, 100:] == 0.0)
