"""Checks of one optimizer step against a float64 gradient known to within a
per-element tolerance.

After torch.optim.Adam's first step (no weight decay) from zero moments,
m = (1 - b1) g, v = (1 - b2) g^2 and the update is
p1 - p0 = -lr m_hat / (sqrt(v_hat) + eps) = -lr g / (|g| + eps), a monotone
function of g.  With the exact gradient known to lie in [g - t, g + t], each
output must lie between its values at the two ends (plus the float32 rounding
of the update itself); where the interval holds 0 the update may be anything
in [-lr, lr], and nothing else."""
import numpy as np


def _ulp(x):
    return np.spacing(np.abs(x).astype(np.float32)).astype(np.float64)


def adam_first_step(p0, p1, m, v, g, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, what="param"):
    """p0, p1, m, v: float32 arrays (before, after, moments); g, t: float64
    gradient and its tolerance.  Asserts every element; returns the number of
    elements whose update is fixed to better than 1% by the gradient (the
    ones away from g = 0)."""
    p0, p1, m, v = (np.asarray(a).astype(np.float64) for a in (p0, p1, m, v))
    lo, hi = g - t, g + t
    f = lambda x: -lr * x / (np.abs(x) + eps)
    d_lo, d_hi = f(hi), f(lo)  # f is decreasing
    slack = 2e-6 * lr + 2 * _ulp(p1) + _ulp(p0)
    d = p1 - p0
    bad = (d < d_lo - slack) | (d > d_hi + slack)
    assert not bad.any(), (f"{what}: {int(bad.sum())} of {bad.size} updates outside the gradient's interval; "
                           f"first {np.argwhere(bad)[:3].tolist()}")
    m_lo, m_hi = (1 - b1) * lo, (1 - b1) * hi
    ms = 1e-6 * np.abs(m_hi - m_lo) + 1e-6 * (1 - b1) * np.abs(g) + 1e-45
    assert ((m >= m_lo - ms) & (m <= m_hi + ms)).all(), f"{what}: exp_avg outside the interval"
    sq_hi = np.maximum(lo * lo, hi * hi)
    sq_lo = np.where((lo <= 0) & (hi >= 0), 0.0, np.minimum(lo * lo, hi * hi))
    vs = 1e-6 * (1 - b2) * sq_hi + 1e-45
    assert ((v >= (1 - b2) * sq_lo - vs) & (v <= (1 - b2) * sq_hi + vs)).all(), f"{what}: exp_avg_sq outside"
    return int(((d_hi - d_lo) < 1e-2 * lr).sum())
