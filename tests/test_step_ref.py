"""tests/step_ref.py's interval check of Adam's first step, on the CPU:
torch.optim.Adam's own float32 step from a gradient known to within a
tolerance passes, and a step with a planted mistake (twice the learning
rate, a missing bias correction, the sign flipped) fails."""
import numpy as np
import pytest
import torch

from step_ref import adam_first_step


def run_adam(p0, g32, lr=1e-3, bias_fix=True):
    p = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([p], lr=lr, foreach=False)
    p.grad = torch.tensor(g32)
    opt.step()
    st = opt.state[p]
    m, v = st["exp_avg"].numpy(), st["exp_avg_sq"].numpy()
    p1 = p.detach().numpy()
    if not bias_fix:
        p1 = (p0.astype(np.float64) - lr * m / (np.sqrt(v) + 1e-8)).astype(np.float32)
    return p1, m, v


def case(seed=0, n=20000):
    rng = np.random.default_rng(seed)
    p0 = rng.normal(0, 0.05, n).astype(np.float32)
    g = rng.normal(0, 1e-5, n) * rng.choice([1e-6, 1e-3, 1.0, 1e3], n)
    t = 1e-5 * np.abs(g) * rng.uniform(1, 50, n)  # a sum|terms| bound above |g|
    g32 = (g + rng.uniform(-1, 1, n) * t).astype(np.float32)  # a computed gradient inside the tolerance
    return p0, g, t, g32


def test_torch_adam_step_passes():
    p0, g, t, g32 = case()
    p1, m, v = run_adam(p0, g32)
    assert adam_first_step(p0, p1, m, v, g, t) > 0.5 * p0.size


@pytest.mark.parametrize("mistake", ["lr", "bias", "sign"])
def test_planted_mistakes_fail(mistake):
    p0, g, t, g32 = case(1)
    if mistake == "lr":
        p1, m, v = run_adam(p0, g32, lr=2e-3)
    elif mistake == "bias":
        p1, m, v = run_adam(p0, g32, bias_fix=False)
    else:
        p1, m, v = run_adam(p0, -g32)
    with pytest.raises(AssertionError):
        adam_first_step(p0, p1, m, v, g, t)
