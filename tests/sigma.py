"""Three-sigma comparison of two Monte Carlo estimates of the same image.

Used by the GPU tests that compare pixel means of the HIP tracer and the
oracle over independent seed streams.  For every pixel channel

    z = (m_a - m_b) / sqrt(v_a / n_a + v_b / n_b)

with v the per-sample variances.  Two bounds are checked:

* the fraction of channels with |z| > 3 stays below the rate of a Student t
  with n_a + n_b - 2 degrees of freedom plus four binomial standard errors,
  counted over pixels (the channels of one pixel are strongly correlated, so
  the pixel count is the conservative number of independent trials);
* the image-mean signed z stays within 4 standard errors of 0: with equal
  sample counts z is antisymmetric under exchanging the two estimates, so its
  mean is exactly 0 under the null hypothesis, skewed pixel distributions
  included.  A bias confined to a few percent of the pixels fails this bound
  long before it moves the fraction of |z| > 3.

Channels whose two variance estimates are both zero (deterministic pixels,
e.g. camera rays that only see the environment) must agree to 1e-6
relative instead.

Calibrated on the CPU with the oracle against itself (two seed offsets):
closed box, glass/lights box and icosphere scenes at 64-256 spp all give a
fraction of |z| > 3 at or below the t rate and std(z) within 1.00-1.04.
"""

from __future__ import annotations

import math

import numpy as np


def t_tail(df: int) -> float:
    """P(|T| > 3) for Student's t with df degrees of freedom."""
    from scipy import stats

    return float(2.0 * stats.t.sf(3.0, df))


def check(m_a, v_a, n_a, m_b, v_b, n_b, what: str = "", mean_z: bool = True) -> dict:
    """Asserts both bounds (the mean-z one only when `mean_z`: it needs the
    two estimates to come from the same estimator); returns the figures."""
    m_a, v_a, m_b, v_b = (np.asarray(x, np.float64).reshape(-1, np.shape(x)[-1]) for x in (m_a, v_a, m_b, v_b))
    se = np.sqrt(v_a / n_a + v_b / n_b)
    det = se < 1e-9
    if det.any():
        gap = np.abs(m_a - m_b)[det]
        assert (gap <= 1e-6 * (1.0 + np.abs(m_b[det]))).all(), f"{what}: deterministic channels differ"
    z = ((m_a - m_b) / np.where(det, 1.0, se))[~det]
    m_pix = int((~det).any(axis=1).sum())
    out = {"n_pixels": m_pix, "frac": 0.0, "bound": 1.0, "mean_z": 0.0, "z_bound": math.inf}
    if m_pix == 0:
        return out
    p0 = t_tail(int(n_a + n_b - 2))
    out["frac"] = float(np.mean(np.abs(z) > 3.0))
    out["bound"] = p0 + 4.0 * math.sqrt(p0 * (1.0 - p0) / m_pix) + 1.0 / m_pix
    out["mean_z"] = float(z.mean())
    out["z_bound"] = 4.0 * float(z.std()) / math.sqrt(m_pix)
    assert out["frac"] <= out["bound"], f"{what}: {out['frac']:.4f} of channels beyond 3 sigma (bound {out['bound']:.4f})"
    if mean_z:
        assert abs(out["mean_z"]) <= out["z_bound"], f"{what}: image-mean z {out['mean_z']:+.4f} (bound {out['z_bound']:.4f})"
    return out


def from_sums(s, s2, n):
    """(mean, unbiased per-sample variance) from per-channel sums and sums of squares."""
    s, s2 = np.asarray(s, np.float64), np.asarray(s2, np.float64)
    m = s / n
    return m, np.maximum(s2 / n - m * m, 0.0) * n / (n - 1)
