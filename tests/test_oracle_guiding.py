"""The CPU reference of the guiding network (oracle/guide_ref.py), pinned by
mathematics rather than by restating its formulas:

* its closed-form gradient (vMF mean, clamped log concentration, softmax
  weights, back-propagation through the ReLU layers) equals central finite
  differences of the float64 loss;
* the decoded mixture is a density: it integrates to 1 over the sphere
  (Fibonacci-lattice quadrature), for concentrations from 0.14 to 50;
* the BF16 mode only rounds: with exactly representable inputs it agrees
  with the float64 mode to BF16 precision.
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
