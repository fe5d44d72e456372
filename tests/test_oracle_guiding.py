"""The CPU reference of the guiding network (oracle/guide_ref.py), pinned by
mathematics rather than by restating its formulas:

* its closed-form gradient (vMF mean, clamped log concentration, softmax
  weights, back-propagation through the ReLU layers) equals central finite
  differences of the float64 loss;
* the decoded mixture is a density: it integrates to 1 over the sphere
  (Fibonacci-lattice quadrature), for concentrations from 0.14 to 50;
* the BF16 mode only rounds: with exactly representable inputs it agrees
  with the float64 mode to BF16 precision;
* selection learning: the gradient with the selection term equals central
  finite differences of L + L_sel, the mixture loss taken with the guide's
  density held at its value (the selection logit is fitted through a alone),
  on every selection parameter and on hidden-layer parameters; its sign
  moves a towards the technique that has the larger density at the samples.
"""

import numpy as np
import pytest

import guide_ref as G


def _samples(rng, S, box=((-1, -1, -1), (1, 1, 1))):
    p = rng.uniform(box[0], box[1], (S, 3))
    n = rng.normal(size=(S, 3)); n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.normal(size=(S, 3)); wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(S, 3)); wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    w = rng.uniform(0.1, 2.0, S)
    return np.concatenate([p, n, wo, wi, w[:, None], np.zeros((S, 3))], axis=1)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    params = G.init_params(seed=4, scale_out=1.0)
    smp = _samples(rng, 24)
    box = ((-1, -1, -1), (1, 1, 1))
    _, grad, _ = G.step(params, smp, *box, bf=False)
    idx = []
    for (wo, bo), (i, j) in zip(G.OFFS, G.LAYERS):
        jj = min(j, 40)
        idx += [wo + r * j + c for r, c in zip(rng.integers(0, i, 6), rng.integers(0, jj, 6))]
        idx += [bo + c for c in rng.integers(0, jj, 3)]
    h = 1e-6
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        lp = G.step(pp, smp, *box, bf=False)[2]
        lm = G.step(pm, smp, *box, bf=False)[2]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 + 1e-4 * abs(fd), (k, fd, grad[k])


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


@pytest.mark.parametrize("logk", [-2.0, 0.0, 2.0, np.log(50.0)])
def test_mixture_integrates_to_one(logk):
    rng = np.random.default_rng(5)
    o = rng.normal(size=(1, 40))
    o.reshape(8, 5)[:, 3] = logk
    w = _fib_sphere(400_000)
    p = G.pdf(np.repeat(o, len(w), 0), w)
    assert abs(p.mean() * 4 * np.pi - 1.0) < 2e-3


def test_bf16_mode_is_rounding_only():
    rng = np.random.default_rng(6)
    params = G.bf16(G.init_params(seed=7))
    smp = _samples(rng, 64)
    box = ((-1, -1, -1), (1, 1, 1))
    o1, g1, l1 = G.step(params, smp, *box, bf=True)
    o2, g2, l2 = G.step(params, smp, *box, bf=False)
    assert np.max(np.abs(o1 - o2)) < 0.05 * (np.max(np.abs(o2)) + 1e-3)
    assert abs(l1 - l2) < 0.02 * abs(l2)
    assert np.linalg.norm(g1 - g2) < 0.05 * np.linalg.norm(g2)


def test_selection_gradient_matches_finite_differences():
    rng = np.random.default_rng(13)
    params = G.init_params(seed=14, scale_out=1.0)
    ws, bs = G.unpack(params)
    ws[3][:, G.SEL] = rng.uniform(-0.3, 0.3, 64)  # a non-trivial selection column
    bs[3][G.SEL] = 0.4
    params = G.pack(ws, bs)
    smp = _samples(rng, 24)
    smp[:, 13] = rng.uniform(0.02, 1.5, 24)
    smp[::5, 13] = 0.0  # no sampler density recorded: no selection term
    box = ((-1, -1, -1), (1, 1, 1))
    _, grad, _ = G.step(params, smp, *box, bf=False)
    x = G.features(smp[:, 0:3], smp[:, 3:6], smp[:, 6:9], *box)
    p0 = G.pdf(G.forward(params, x, bf=False)[1][:, :G.N_OUT], smp[:, 9:12])

    def total(pr):
        _, out = G.forward(pr, x, bf=False)
        return (G.loss(out[:, :G.N_OUT], smp[:, 9:12], smp[:, 12])
                + G.sel_loss(out[:, G.SEL], p0, smp[:, 13], smp[:, 12]))

    wo3, bo3 = G.OFFS[3]
    idx = [wo3 + r * 48 + G.SEL for r in range(0, 64, 7)] + [bo3 + G.SEL]
    for (wo, bo), (i, j) in zip(G.OFFS[:3], G.LAYERS[:3]):
        idx += [wo + r * j + c for r, c in zip(rng.integers(0, i, 4), rng.integers(0, j, 4))]
    h = 1e-6
    for k in idx:
        pp, pm = params.copy(), params.copy()
        pp[k] += h
        pm[k] -= h
        fd = (total(pp) - total(pm)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-5 + 1e-4 * abs(fd), (k, fd, grad[k])
    # the padding columns 41..47 get no gradient
    assert not grad[wo3:bo3].reshape(64, 48)[:, G.SEL + 1:].any() and not grad[bo3 + G.SEL + 1:bo3 + 48].any()
    # a guide density above the sampler's at every sample pushes a up (gradient < 0 on the logit)
    one = smp[:1].copy()
    one[0, 13] = 1e-3
    o = G.forward(params, G.features(one[:, 0:3], one[:, 3:6], one[:, 6:9], *box), bf=False)[1]
    assert G.pdf(o[:, :G.N_OUT], one[:, 9:12])[0] > 1e-3
    assert G.sel_grad(o[:, G.SEL], G.pdf(o[:, :G.N_OUT], one[:, 9:12]), one[:, 13], one[:, 12])[0] < 0
