"""TEST INFRASTRUCTURE ONLY: float64 numpy restatement of the learned path
guide, used to check the GPU's BF16 MFMA kernels.  The guide has no
counterpart in the reference renderer.

Network (the layout the GPU kernels use):
  input   128 = 16 frequencies x 3 coordinates x (sin, cos) of the position
                normalized to the scene box (feature 6j + 2c + s, argument
                pi * 2^j * x), then the normal (96..98), the roughness (99),
                zero padding (100..127)
  hidden  4 x 64, ReLU
  output  40 = 8 mixture logits, 8 x 3 raw mean directions, 8 log-concentrations
vMF lobe (Jakob 2012, "Numerically stable sampling of the von Mises-Fisher
distribution on S^2"):
  p(w) = k / (2 pi (1 - exp(-2k))) * exp(k (mu.w - 1)), mu = raw/|raw|,
  k = exp(clamp(raw_k, -6, 11))
Loss per training sample (direction w drawn with pdf p_s, target t >= 0):
  -(t / p_s) * log q(w), q the mixture; weights are normalized by the batch
  mean of t/p_s.  Adam (Kingma & Ba 2015) with lr 1e-3, b1 0.9, b2 0.999,
  eps 1e-8.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, LOBES, FREQS = 128, 64, 40, 8, 16
LAYERS = [(IN, HID), (HID, HID), (HID, HID), (HID, HID), (HID, OUT)]
KMIN, KMAX = -6.0, 11.0


def init_params(seed=0):
    """He-normal weights, zero biases; W[l] has shape (in, out)."""
    g = np.random.default_rng(seed)
    Ws, bs = [], []
    for i, o in LAYERS:
        Ws.append(g.normal(0.0, np.sqrt(2.0 / i), (i, o)).astype(np.float32))
        bs.append(np.zeros(o, np.float32))
    return Ws, bs


def feature_map():
    """Frequency j, coordinate c and cos flag of the 96 sinusoidal features.

    Feature f = 16 t + 4 q + i (t < 6) belongs to the q-th quarter of the
    MFMA operand lanes; with m = 4 t + i, it is frequency j = 4 q + m // 6,
    coordinate (m % 6) // 2, sin for even m % 6 and cos for odd.  A lane
    thus holds the four frequencies 4q..4q+3 of all three coordinates.
    """
    f = np.arange(96)
    t, q, i = f // 16, (f % 16) // 4, f % 4
    m = 4 * t + i
    return 4 * q + m // 6, (m % 6) // 2, (m % 6) % 2


def encode(p01, nrm, rough):
    """(n,3) positions in [0,1], (n,3) normals, (n,) roughness -> (n,128)."""
    p01 = np.asarray(p01, np.float32)
    n = p01.shape[0]
    x = np.zeros((n, IN), np.float64)
    j, c, cosf = feature_map()
    v = (np.float32(2.0) ** j.astype(np.float32))[None, :] * p01[:, c]      # exact in float32
    r = np.mod(v.astype(np.float64), 2.0)
    x[:, :96] = np.where(cosf[None, :] == 1, np.cos(np.pi * r), np.sin(np.pi * r))
    x[:, 96:99] = nrm
    x[:, 99] = rough
    return x


def forward(Ws, bs, x):
    hs = [x]
    h = x
    for l in range(4):
        h = np.maximum(h @ Ws[l].astype(np.float64) + bs[l], 0.0)
        hs.append(h)
    o = h @ Ws[4].astype(np.float64) + bs[4]
    return o, hs


def mixture(o):
    logits = o[:, :8]
    m = logits.max(1, keepdims=True)
    e = np.exp(logits - m)
    pi = e / e.sum(1, keepdims=True)
    raw = o[:, 8:32].reshape(-1, 8, 3)
    rn = np.linalg.norm(raw, axis=2, keepdims=True)
    mu = raw / np.maximum(rn, 1e-12)
    kr = np.clip(o[:, 32:40], KMIN, KMAX)
    kappa = np.exp(kr)
    return pi, mu, kappa, raw, rn


def vmf_log_pdf(mu, kappa, w):
    """log p for lobes (n,8,3),(n,8) at directions (n,3)."""
    c = np.einsum("nkd,nd->nk", mu, w)
    return np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa)) + kappa * (c - 1.0)


def log_q(o, w):
    pi, mu, kappa, _, _ = mixture(o)
    lp = vmf_log_pdf(mu, kappa, w) + np.log(pi)
    m = lp.max(1, keepdims=True)
    return (m + np.log(np.exp(lp - m).sum(1, keepdims=True)))[:, 0]


def loss_and_grads(Ws, bs, x, w, wgt):
    """Mean over the batch of -wgt * log q(w); returns the loss and the
    gradients of every W and b (float64)."""
    n = x.shape[0]
    o, hs = forward(Ws, bs, x)
    pi, mu, kappa, raw, rn = mixture(o)
    lp = vmf_log_pdf(mu, kappa, w) + np.log(pi)
    m = lp.max(1, keepdims=True)
    lq = (m + np.log(np.exp(lp - m).sum(1, keepdims=True)))[:, 0]
    loss = float(np.mean(-wgt * lq))
    r = np.exp(lp - lq[:, None])                                  # responsibilities
    g = (-wgt / n)[:, None]                                       # dL/dlog q per sample
    do = np.zeros_like(o)
    do[:, :8] = g * (r - pi)
    c = np.einsum("nkd,nd->nk", mu, w)
    dmu_raw = (w[:, None, :] - c[:, :, None] * mu) / np.maximum(rn, 1e-12)
    do[:, 8:32] = (g[:, :, None] * (r * kappa)[:, :, None] * dmu_raw).reshape(n, 24)
    dlogp_dk = 1.0 / kappa - 2.0 / np.expm1(2.0 * kappa) + (c - 1.0)
    kr = o[:, 32:40]
    inside = (kr > KMIN) & (kr < KMAX)
    do[:, 32:40] = g * r * dlogp_dk * kappa * inside
    dWs, dbs = [None] * 5, [None] * 5
    dh = do
    for l in range(4, -1, -1):
        dWs[l] = hs[l].T @ dh
        dbs[l] = dh.sum(0)
        if l > 0:
            dh = (dh @ Ws[l].astype(np.float64).T) * (hs[l] > 0)
    return loss, dWs, dbs


def adam_step(params, grads, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    out = []
    for p, gr, mm, vv in zip(params, grads, m, v):
        mm[...] = b1 * mm + (1 - b1) * gr
        vv[...] = b2 * vv + (1 - b2) * gr * gr
        mh = mm / (1 - b1 ** t)
        vh = vv / (1 - b2 ** t)
        out.append(p - lr * mh / (np.sqrt(vh) + eps))
    return out


def sample_vmf(mu, kappa, u1, u2):
    """Jakob 2012 sampler for one lobe per row: (n,3), (n,), uniforms."""
    wz = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    s = np.sqrt(np.maximum(0.0, 1.0 - wz * wz))
    phi = 2 * np.pi * u2
    a = np.where(np.abs(mu[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(mu, a)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(mu, t)
    return wz[:, None] * mu + s[:, None] * (np.cos(phi)[:, None] * t + np.sin(phi)[:, None] * b)
