"""CPU reference of the path-guiding network (float64 numpy).

TEST INFRASTRUCTURE ONLY: the gradient check of the device training kernels
and the unit tests import it; the product never does.

The network and loss are this build's own (the reference renderer has no
learned guiding); this module restates them independently of the device
code:

* input: 64 features of (position in the scene box, shading normal,
  outgoing direction) -- frequency encoding sin/cos(2^k pi p), k = 0..7, then
  2p - 1, n, wo, sin(pi n), sin(pi wo), n . wo;
* 3 hidden layers of 64 with ReLU, 40 outputs = 8 lobes x (mean direction,
  log concentration clamped to [-2, 7], weight logit), softmax weights;
* density p(w) = sum_k w_k kappa_k / (2 pi (1 - e^{-2 kappa_k}))
  exp(kappa_k (mu_k . w - 1));
* loss L = -sum_s c_s log p(wi_s), c_s = weight_s / sum weight;
* selection learning: output 40 is the logit of the guide-selection
  probability a = sigmoid(o_40), fitted through a alone to the one-sample
  mixture's loss -sum_s c_s log(a p(wi_s) + (1 - a) pbsdf_s) (p held fixed);
  the reported loss stays L.

`bf16=True` mirrors the device's precision split: inputs, hidden activations,
weights and back-propagated deltas rounded to BF16, sums in float64 (the
device accumulates in FP32).  `bf16=False` is the exact float64 network whose
analytic gradient the tests check against finite differences.
"""

from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT, N_PAD, N_LOBES = 64, 64, 40, 48, 8
SEL = 40  # selection logit
LAYERS = [(64, 64), (64, 64), (64, 64), (64, 48)]
LOGK_MIN, LOGK_MAX = -2.0, 7.0


def offsets():
    """(weight offset, bias offset) of each layer in the parameter vector."""
    out, o = [], 0
    for i, j in LAYERS:
        out.append((o, o + i * j))
        o += i * j + j
    return out, o


OFFS, N_PARAMS = offsets()


def unpack(params):
    ws, bs = [], []
    for (i, j), (wo, bo) in zip(LAYERS, OFFS):
        ws.append(np.asarray(params[wo:wo + i * j], np.float64).reshape(i, j))
        bs.append(np.asarray(params[bo:bo + j], np.float64))
    return ws, bs


def pack(ws, bs):
    p = np.zeros(N_PARAMS)
    for w, b, (wo, bo) in zip(ws, bs, OFFS):
        p[wo:wo + w.size] = w.reshape(-1)
        p[bo:bo + b.size] = b
    return p


def bf16(x):
    """Round to bfloat16 (round to nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(p, n, wo, bmin, bmax):
    """(S, 64) input features; p, n, wo (S, 3)."""
    p, n, wo = (np.asarray(a, np.float64) for a in (p, n, wo))
    pn = (p - np.asarray(bmin, np.float64)) / (np.asarray(bmax, np.float64) - np.asarray(bmin, np.float64))
    S = p.shape[0]
    f = np.zeros((S, 64))
    for idx in range(48):
        c, q = idx % 3, idx // 3
        k = q >> 1
        a = pn[:, c] * (np.pi * (1 << k))
        f[:, idx] = np.cos(a) if q & 1 else np.sin(a)
    f[:, 48:51] = 2 * pn - 1
    f[:, 51:54] = n
    f[:, 54:57] = wo
    f[:, 57:60] = np.sin(np.pi * n)
    f[:, 60:63] = np.sin(np.pi * wo)
    f[:, 63] = np.sum(n * wo, axis=1)
    return f


def decode(o):
    """(S, 40) raw outputs -> mu (S,8,3), raw length (S,8), kappa (S,8), log c (S,8), weights (S,8)."""
    o = np.asarray(o, np.float64).reshape(-1, N_LOBES, 5)
    raw = o[..., :3]
    ln = np.sqrt(np.sum(raw * raw, axis=-1)) + 1e-6
    mu = raw / ln[..., None]
    lk = np.clip(o[..., 3], LOGK_MIN, LOGK_MAX)
    kappa = np.exp(lk)
    logc = np.log(kappa) - np.log(2 * np.pi) - np.log(-np.expm1(-2 * kappa))
    a = o[..., 4]
    w = np.exp(a - a.max(axis=1, keepdims=True))
    w /= w.sum(axis=1, keepdims=True)
    return mu, ln, kappa, logc, w


def pdf(o, wi):
    mu, _, kappa, logc, w = decode(o)
    wi = np.asarray(wi, np.float64)
    lf = logc + kappa * (np.einsum("skc,sc->sk", mu, wi) - 1.0)
    return np.sum(w * np.exp(lf), axis=1)


def loss(o, wi, weight):
    c = np.asarray(weight, np.float64) / np.sum(weight)
    return float(-np.sum(c * np.log(np.maximum(pdf(o, wi), 1e-300))))


def loss_grad(o, wi, weight):
    """d loss / d raw outputs (S, 40), from the closed-form derivatives."""
    o = np.asarray(o, np.float64).reshape(-1, N_OUT)
    mu, ln, kappa, logc, w = decode(o)
    wi = np.asarray(wi, np.float64)
    c = (np.asarray(weight, np.float64) / np.sum(weight))[:, None]
    mw = np.einsum("skc,sc->sk", mu, wi)
    f = np.exp(logc + kappa * (mw - 1.0))
    p = np.maximum(np.sum(w * f, axis=1, keepdims=True), 1e-300)
    r = w * f / p
    d = np.zeros((o.shape[0], N_LOBES, 5))
    d[..., :3] = (-c * r * kappa / ln)[..., None] * (wi[:, None, :] - mw[..., None] * mu)
    lk = o.reshape(-1, N_LOBES, 5)[..., 3]
    q = 2 * kappa / np.expm1(2 * kappa)
    inside = (lk > LOGK_MIN) & (lk < LOGK_MAX)
    d[..., 3] = np.where(inside, -c * r * (1.0 - q + kappa * (mw - 1.0)), 0.0)
    d[..., 4] = -c * (r - w)
    return d.reshape(-1, N_OUT)


def sel_grad(o_sel, p, pbsdf, weight):
    """d L_sel / d o_40 per sample; 0 where pbsdf <= 0."""
    c = np.asarray(weight, np.float64) / np.sum(weight)
    a = 1.0 / (1.0 + np.exp(-np.asarray(o_sel, np.float64)))
    pb = np.asarray(pbsdf, np.float64)
    q = a * p + (1.0 - a) * pb
    return np.where(pb > 0, -c * a * (1.0 - a) * (p - pb) / np.where(q > 0, q, 1.0), 0.0)


def sel_loss(o_sel, p, pbsdf, weight):
    c = np.asarray(weight, np.float64) / np.sum(weight)
    a = 1.0 / (1.0 + np.exp(-np.asarray(o_sel, np.float64)))
    pb = np.asarray(pbsdf, np.float64)
    return float(-np.sum(np.where(pb > 0, c * np.log(a * p + (1.0 - a) * pb), 0.0)))


def forward(params, x, bf=True):
    """Activations X_0..X_3 (inputs of each layer) and the raw outputs."""
    ws, bs = unpack(params)
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    xs = [rnd(x)]
    for l in range(4):
        z = xs[-1] @ rnd(ws[l]) + bs[l]
        if l < 3:
            xs.append(rnd(np.maximum(z, 0.0)))
        else:
            out = z
    return xs, out


def step(params, samples, bmin, bmax, bf=True, learn_sel=True):
    """Raw outputs (S, 40), gradient (N_PARAMS,) and loss of one training
    step over `samples` (S, 16): p, n, wo, wi, weight, pbsdf."""
    s = np.asarray(samples, np.float64)
    x = features(s[:, 0:3], s[:, 3:6], s[:, 6:9], bmin, bmax)
    xs, out = forward(params, x, bf)
    o = out[:, :N_OUT]
    L = loss(o, s[:, 9:12], s[:, 12])
    d = np.zeros((s.shape[0], N_PAD))
    d[:, :N_OUT] = loss_grad(o, s[:, 9:12], s[:, 12])
    if learn_sel:
        d[:, SEL] = sel_grad(out[:, SEL], np.maximum(pdf(o, s[:, 9:12]), 1e-300), s[:, 13], s[:, 12])
    ws, _ = unpack(params)
    rnd = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    dy = [None] * 4
    dy[3] = rnd(d)
    for l in range(3, 0, -1):
        dx = dy[l] @ rnd(ws[l]).T
        dy[l - 1] = rnd(np.where(xs[l] > 0, dx, 0.0))
    gws, gbs = [], []
    for l in range(4):
        gw = xs[l].T @ dy[l]
        gb = dy[l].sum(axis=0)
        if l == 3:
            gw[:, SEL + 1:] = 0.0
            gb[SEL + 1:] = 0.0
        gws.append(gw)
        gbs.append(gb)
    return o, pack(gws, gbs), L


def init_params(seed=0, scale_out=0.1):
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for l, (i, j) in enumerate(LAYERS):
        lim = np.sqrt(6.0 / i) * (scale_out if l == 3 else 1.0)
        w = rng.uniform(-lim, lim, (i, j))
        b = rng.uniform(-0.1, 0.1, j)
        if l == 3:
            w[:, N_OUT:] = 0.0
            b[N_OUT:] = 0.0
        ws.append(w)
        bs.append(b)
    return pack(ws, bs)
