"""The guide's BF16 MFMA kernels against the float64 numpy reference
(oracle/guide_ref.py), through the C ABI.

Tolerances: inference is compared with a reference that rounds features,
weights and activations as the kernels do (float32 angle-doubling features,
bf16 operands), so that only the fp32 summation order and the device's
sinpi/cospi ulps differ: 2e-3 relative for 98% of the samples, 2e-2 for all
(0.1 rad for mean directions; a feature or activation that lands on the
other side of a bf16 rounding boundary moves its sample by one bf16 ulp of
that input).  Gradients are compared with a reference that rounds where the
kernels round (relative L2 error below 5e-3 per layer) and, as a sanity
bound, with the unrounded float64 reference: bf16 rounding of activations
and back-propagated gradients, summed over 4096 samples with strong
cancellation in the first layers, leaves a relative error of up to about
0.25 there (the bf16-rounded reference itself shows the same), with the
gradient direction kept (cosine above 0.95)."""
import math

import numpy as np
import pytest

from oracle import guide_ref as R
from azimuth_amd.guide import Guide, pack_params, unpack_params

pytestmark = pytest.mark.gpu


def bf16(x):
    a = np.asarray(x, np.float32).copy()
    u = a.view(np.uint32)
    u += (np.uint32(0x7FFF) + ((u >> 16) & 1)).astype(np.uint32)
    u &= np.uint32(0xFFFF0000)
    return u.view(np.float32).astype(np.float64)


def encode_f32(p01, nrm, rough):
    """The features as the kernels round them: float32 sin/cos at the lowest
    frequency 2^(4q) of each operand-lane quarter q, then float32 angle
    doubling (sin 2a = 2 sin a cos a, cos 2a = (cos a - sin a)(cos a + sin a))
    for the next three; laid out as the oracle's encode."""
    x = R.encode(p01, nrm, rough)
    p = np.asarray(p01, np.float32)
    j, c, cs = R.feature_map()
    for q in range(4):
        th = (np.float32(2.0 ** (4 * q)) * p).astype(np.float64)
        sn = [np.sin(np.pi * np.mod(th, 2.0)).astype(np.float32)]
        co = [np.cos(np.pi * np.mod(th, 2.0)).astype(np.float32)]
        for _ in range(3):
            a, b = sn[-1], co[-1]
            sn.append(np.float32(2.0) * a * b)
            co.append((b - a) * (b + a))
        for f in range(96):
            if j[f] // 4 == q:
                x[:, f] = (co if cs[f] else sn)[j[f] - 4 * q][:, c[f]]
    return x


def forward_bf16(Ws, bs, x):
    h = bf16(x)
    for l in range(4):
        h = bf16(np.maximum(h @ bf16(Ws[l]) + bs[l].astype(np.float32), 0.0))
    return h @ bf16(Ws[4]) + bs[4]


@pytest.fixture(scope="module")
def guide():
    g = Guide(0, seed=3)
    yield g
    g.close()


def random_batch(n, seed):
    g = np.random.default_rng(seed)
    p = g.random((n, 3)).astype(np.float32)
    nrm = g.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rough = g.random(n).astype(np.float32)
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return p, nrm.astype(np.float32), rough, d.astype(np.float32), (g.random(n) + 0.2).astype(np.float32)


def test_mfma_lane_layout(guide):
    g = np.random.default_rng(0)
    A, B = g.normal(size=(16, 16)), g.normal(size=(16, 16))
    D = guide.mfma_selftest(A, B)
    np.testing.assert_allclose(D, bf16(A) @ bf16(B), rtol=1e-5, atol=1e-5)


def test_inference_matches_reference(guide):
    Ws, bs = R.init_params(7)
    g = np.random.default_rng(1)
    bs = [b + g.normal(0, 0.2, b.shape).astype(np.float32) for b in bs]
    guide.set_params(pack_params(Ws, bs))
    p, nrm, rough, _, _ = random_batch(1000, 2)     # not a multiple of 16 or 64
    lobes = guide.infer(p, nrm, rough)
    x = R.encode(p, nrm, rough)
    o = forward_bf16(Ws, bs, encode_f32(p, nrm, rough))
    pi, mu, kappa, _, _ = R.mixture(o)
    # Per sample: relative error of the weights and concentrations, angle
    # between the mean directions.  A device sinpi/cospi ulp, amplified by
    # the doubling, can still move a feature across a bf16 rounding boundary
    # and change that sample by a bf16 ulp of one input: a few samples, small.
    e_pi = np.max(np.abs(lobes[:, :, 0] - pi) / (np.abs(pi) + 1e-4), axis=1)
    e_mu = np.max(np.arccos(np.clip(np.sum(lobes[:, :, 1:4] * mu, axis=2), -1.0, 1.0)), axis=1)
    e_ka = np.max(np.abs(lobes[:, :, 4] - kappa) / kappa, axis=1)
    bad = (e_pi > 2e-3) | (e_mu > 2e-3) | (e_ka > 2e-3)
    print(f"inference: samples outside 2e-3 {int(bad.sum())} of {len(bad)}; max pi {e_pi.max():.2e} "
          f"mu angle {e_mu.max():.2e} kappa {e_ka.max():.2e}", flush=True)
    assert bad.mean() <= 0.02
    # (a mean direction is the normalized raw output: its angle moves by
    # the raw change over the raw length, larger for short raw vectors)
    assert e_pi.max() < 2e-2 and e_ka.max() < 2e-2 and e_mu.max() < 0.1
    # and the float64 network agrees to bf16 precision
    o64, _ = R.forward(Ws, bs, x)
    pi64 = R.mixture(o64)[0]
    assert np.max(np.abs(lobes[:, :, 0] - pi64)) < 0.05


def half_to_float(h):
    return np.asarray(h, np.uint16).view(np.float16).astype(np.float64)


def octahedral_decode(u, v):
    """Meyer et al. 2010: the point (x, y) of the unit octahedron's map, with
    z = 1 - |x| - |y| and the lower half folded over the diagonals."""
    x = u.astype(np.float64) * (2.0 / 65535.0) - 1.0
    y = v.astype(np.float64) * (2.0 / 65535.0) - 1.0
    z = 1.0 - np.abs(x) - np.abs(y)
    fold = z < 0
    fx = (1.0 - np.abs(y)) * np.where(x >= 0, 1.0, -1.0)
    fy = (1.0 - np.abs(x)) * np.where(y >= 0, 1.0, -1.0)
    x, y = np.where(fold, fx, x), np.where(fold, fy, y)
    d = np.stack([x, y, z], -1)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def test_packed_lobes_match_the_float_lobes(guide):
    """The renderer's packed lobes (half-precision weight and kappa,
    octahedral mean direction) decode to the float lobes of the same queries
    within the formats' rounding, and queries of negative roughness (misses,
    emitters) are skipped: their words keep the caller's fill."""
    Ws, bs = R.init_params(11)
    g = np.random.default_rng(5)
    bs = [b + g.normal(0, 0.3, b.shape).astype(np.float32) for b in bs]
    guide.set_params(pack_params(Ws, bs))
    p, nrm, rough, _, _ = random_batch(3001, 6)
    skip = g.random(3001) < 0.4
    skip[:70] = True                  # a whole first chunk of 64 skipped, and more
    skip[-5:] = False
    rough_q = np.where(skip, -1.0, rough).astype(np.float32)
    want = guide.infer(p, nrm, rough)
    fill = 0xDEADBEEF
    got = guide.infer_packed(p, nrm, rough_q, fill=fill)
    assert np.all(got[skip] == fill)
    keep = ~skip
    w = got[keep][:, :, 0]
    pi = half_to_float(w & 0xFFFF)
    kappa = half_to_float(w >> 16)
    mu = octahedral_decode(got[keep][:, :, 1] & 0xFFFF, got[keep][:, :, 1] >> 16)
    ref = want[keep]
    # half precision: 2^-11 relative (weights below the half normal range
    # are absolute 2^-25); the mean within about 1e-4 rad.
    np.testing.assert_allclose(pi, ref[:, :, 0], rtol=5e-4, atol=6e-8)
    np.testing.assert_allclose(kappa, ref[:, :, 4], rtol=5e-4)
    # (the float lobes' means are normalized with a 1-ulp reciprocal square
    # root: renormalized here, or their 1e-7 length error alone reads as
    # 4.5e-4 rad through arccos)
    mref = ref[:, :, 1:4].astype(np.float64)
    mref /= np.linalg.norm(mref, axis=-1, keepdims=True)
    ang = np.arccos(np.clip(np.sum(mu * mref, -1), -1.0, 1.0))
    assert ang.max() < 1e-4, ang.max()
    assert np.all(np.abs(np.sum(pi, axis=1) - 1.0) < 4e-3)


def grads_bf16(Ws, bs, x, w, wgt):
    """The float64 reference's backward with bf16 rounding where the kernels
    round: activations, the output gradient and each back-propagated
    gradient before it enters the next product; bias gradients are sums of
    the rounded output gradients, as the weight-gradient kernel forms them."""
    n = x.shape[0]
    hs = [bf16(x)]
    for l in range(4):
        hs.append(bf16(np.maximum(hs[-1] @ bf16(Ws[l]) + bs[l].astype(np.float32), 0.0)))
    o = hs[-1] @ bf16(Ws[4]) + bs[4]
    pi, mu, kappa, raw, rn = R.mixture(o)
    lp = R.vmf_log_pdf(mu, kappa, w) + np.log(pi)
    m = lp.max(1, keepdims=True)
    lq = (m + np.log(np.exp(lp - m).sum(1, keepdims=True)))[:, 0]
    r = np.exp(lp - lq[:, None])
    gq = (-wgt / n)[:, None]
    do = np.zeros_like(o)
    do[:, :8] = gq * (r - pi)
    c = np.einsum("nkd,nd->nk", mu, w)
    do[:, 8:32] = (gq[:, :, None] * (r * kappa)[:, :, None] * (w[:, None, :] - c[:, :, None] * mu) /
                   np.maximum(rn, 1e-12)).reshape(n, 24)
    kr = o[:, 32:40]
    inside = (kr > R.KMIN) & (kr < R.KMAX)
    do[:, 32:40] = gq * r * (1 / kappa - 2 / np.expm1(2 * kappa) + (c - 1)) * kappa * inside
    dWs, dbs = [None] * 5, [None] * 5
    dy = do
    for l in range(4, -1, -1):
        dyb = bf16(dy)
        dbs[l] = dyb.sum(0)
        dWs[l] = hs[l].T @ dyb
        if l > 0:
            dy = (dyb @ bf16(Ws[l]).T) * (hs[l] > 0)
    return dWs, dbs


def test_gradients_match_reference(guide):
    Ws, bs = R.init_params(11)
    g = np.random.default_rng(4)
    bs = [b + g.normal(0, 0.1, b.shape).astype(np.float32) for b in bs]
    guide.set_params(pack_params(Ws, bs))
    p, nrm, rough, d, w = random_batch(4096, 5)
    guide.zero_grad()
    loss_gpu = guide.train(p, nrm, rough, d, w, update=False)
    gg = unpack_params(guide.grad())
    x = R.encode(p, nrm, rough)
    d64, w64 = d.astype(np.float64), w.astype(np.float64)
    loss, dWs, dbs = R.loss_and_grads([W.astype(np.float64) for W in Ws], [b.astype(np.float64) for b in bs], x,
                                      d64, w64)
    eWs, ebs = grads_bf16(Ws, bs, x, d64, w64)
    assert loss_gpu == pytest.approx(loss, rel=2e-2)
    for l in range(5):
        for got, want, emu in ((gg[0][l], dWs[l], eWs[l]), (gg[1][l], dbs[l], ebs[l])):
            err = np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-20)
            cos = float(np.dot(got.ravel(), want.ravel()) / (np.linalg.norm(got) * np.linalg.norm(want) + 1e-30))
            err_emu = np.linalg.norm(got - emu) / max(np.linalg.norm(emu), 1e-20)
            print(f"layer {l}: vs float64 rel err {err:.4f} cos {cos:.5f}; vs bf16-rounded reference {err_emu:.2e}",
                  flush=True)
            assert err_emu < 5e-3
            assert err < 0.35 and cos > 0.95
    guide.zero_grad()


def test_adam_step_closed_form(guide):
    """One update from zero moments moves every parameter by
    -lr * g / (|g| + eps') (bias-corrected first Adam step)."""
    Ws, bs = R.init_params(2)
    p0 = pack_params(Ws, bs)
    guide.set_params(p0)
    p, nrm, rough, d, w = random_batch(1024, 6)
    guide.zero_grad()
    guide.train(p, nrm, rough, d, w, update=False)
    gr = guide.grad().astype(np.float64)
    guide.set_params(p0)                  # resets moments and gradient
    guide.train(p, nrm, rough, d, w, update=True)
    p1 = guide.params().astype(np.float64)
    want = p0 - 1e-3 * gr / (np.abs(gr) + 1e-8)
    big = np.abs(gr) > 1e-6
    np.testing.assert_allclose(p1[big], want[big], rtol=0, atol=2e-5)


def test_learns_a_position_dependent_lobe(guide):
    """Directions drawn from vMF(mu(p), kappa=20) with unit weights: the
    trained mixture's loss approaches the target's entropy and its dominant
    lobe points along mu(p)."""
    guide.set_params(pack_params(*R.init_params(5)))
    g = np.random.default_rng(9)

    def batch(n):
        p = g.random((n, 3)).astype(np.float32)
        mu = np.stack([p[:, 0] - 0.5, np.ones(n), p[:, 2] - 0.5], 1)
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        d = R.sample_vmf(mu, np.full(n, 20.0), g.random(n), g.random(n))
        nrm = np.tile([0.0, 1.0, 0.0], (n, 1))
        return p, nrm.astype(np.float32), np.full(n, 0.5, np.float32), d.astype(np.float32), np.ones(n, np.float32), mu

    losses = []
    for it in range(400):
        p, nrm, rough, d, w, _ = batch(4096)
        losses.append(guide.train(p, nrm, rough, d, w, update=True))
    entropy = -(math.log(20 / (2 * math.pi)) + 20 * (1 / math.tanh(20) - 1 / 20 - 1))
    print(f"loss {losses[0]:.3f} -> {np.mean(losses[-20:]):.3f}, target entropy {entropy:.3f}", flush=True)
    assert np.mean(losses[-20:]) < entropy + 0.25
    p, nrm, rough, d, w, mu = batch(512)
    lobes = guide.infer(p, nrm, rough)
    k = np.argmax(lobes[:, :, 0], 1)
    dom = lobes[np.arange(512), k, 1:4]
    assert np.mean(np.sum(dom * mu, 1)) > 0.95
