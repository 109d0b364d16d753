"""The float64 L1 + SSIM reference (tests/ssim_ref.py) against torch autograd of
the pytorch_ssim.py:17-66 expression itself (conv2d with the 11x11 window,
zero padding), and its magnitude bound against a float32 evaluation of that
expression: the bound is what the GPU gradient is held to, so it must cover
float32 rounding (with margin) and not be vacuous."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from ssim_ref import l1_ssim_ref, window_taps


def direct_loss(x, y, lam=0.2):
    g = torch.tensor(window_taps(), dtype=x.dtype)
    c = x.shape[0]
    w = (g[:, None] @ g[None, :]).expand(c, 1, 11, 11).contiguous()
    conv = lambda a: F.conv2d(a[None], w, padding=5, groups=c)[0]
    mu1, mu2 = conv(x), conv(y)
    s1, s2, s12 = conv(x * x) - mu1 * mu1, conv(y * y) - mu2 * mu2, conv(x * y) - mu1 * mu2
    C1 = float(np.float32(0.01 ** 2))
    C2 = float(np.float32(0.03 ** 2))
    m = ((2 * mu1 * mu2 + C1) * (2 * s12 + C2)) / ((mu1 * mu1 + mu2 * mu2 + C1) * (s1 + s2 + C2))
    return (1 - lam) * (x - y).abs().mean() + lam * (1 - m.mean())


@pytest.mark.parametrize("shape,seed", [((3, 37, 53), 0), ((3, 64, 80), 1), ((1, 20, 20), 2)])
def test_reference_gradient_equals_autograd(shape, seed):
    rng = np.random.default_rng(seed)
    tgt = rng.uniform(0, 1, shape).astype(np.float32)
    img = np.clip(tgt + rng.normal(0, 0.1, shape), 0, 1).astype(np.float32)
    img[:, ::5, ::3] = tgt[:, ::5, ::3]  # exact ties: sign(0) = 0
    img[:, :6, :6] = 0.5  # flat patch with a flat target: zero variances
    tgt[:, :6, :6] = 0.5
    loss, grad, mag, _ = l1_ssim_ref(img, tgt)
    x = torch.tensor(img, dtype=torch.float64, requires_grad=True)
    L = direct_loss(x, torch.tensor(tgt, dtype=torch.float64))
    L.backward()
    assert loss == pytest.approx(float(L), rel=1e-13, abs=1e-15)
    ref = x.grad.numpy()
    assert np.abs(grad - ref).max() <= 1e-12 * np.abs(ref).max()
    assert (mag >= np.abs(grad) * (1 - 1e-12)).all()


def _inputs(kind, shape, seed):
    rng = np.random.default_rng(seed)
    if kind == "random":
        tgt = rng.uniform(0, 1, shape).astype(np.float32)
        img = np.clip(tgt + rng.normal(0, 0.1, shape), 0, 1).astype(np.float32)
    else:  # smooth image and a slightly perturbed copy: small variances on large means (cancellation)
        base = rng.uniform(0.3, 0.9, (shape[0], 1, 1)) + 0.05 * np.sin(np.arange(shape[2]) / 7.0)[None, None, :]
        tgt = np.broadcast_to(base, shape).astype(np.float32).copy()
        img = (tgt + rng.normal(0, 0.01, shape)).astype(np.float32)
    return img, tgt


@pytest.mark.parametrize("kind", ["random", "cancellation"])
def test_bound_covers_float32(kind):
    """A float32 evaluation (torch on the CPU, another operation order than
    the GPU kernels) lies far inside 1e-5 of the bound (below 1e-6 of it)."""
    img, tgt = _inputs(kind, (3, 64, 96), 3)
    L, grad, mag, Lmag = l1_ssim_ref(img, tgt)
    x = torch.tensor(img, requires_grad=True)
    l32 = direct_loss(x, torch.tensor(tgt))
    l32.backward()
    assert abs(float(l32.detach()) - L) <= 1e-6 * Lmag
    err = np.abs(x.grad.numpy().astype(np.float64) - grad)
    assert (err <= 1e-6 * mag).all(), float(np.max(err / mag))


def test_bound_not_vacuous():
    """On ordinary images the bound is within two orders of magnitude of the
    gradient, and a gradient with one term dropped (the y_p (w * dS/de12)
    term) falls outside 1e-5 of it: the bar catches such a mistake."""
    img, tgt = _inputs("random", (3, 64, 96), 5)
    _, grad, mag, _ = l1_ssim_ref(img, tgt)
    assert np.median(mag / np.maximum(np.abs(grad), 1e-300)) < 1000
    from ssim_ref import _blur
    x, y = (torch.from_numpy(a.astype(np.float64)) for a in (img, tgt))
    m1, m2 = _blur(x, window_taps()), _blur(y, window_taps())
    s1 = _blur(x * x, window_taps()) - m1 * m1
    s2 = _blur(y * y, window_taps()) - m2 * m2
    C1, C2 = float(np.float32(1e-4)), float(np.float32(9e-4))
    dC = 2 * (2 * m1 * m2 + C1) / ((m1 * m1 + m2 * m2 + C1) * (s1 + s2 + C2))
    dropped = grad + 0.2 / x.numel() * (y * _blur(dC, window_taps())).numpy()
    assert ((np.abs(dropped - grad) > 1e-5 * mag).mean()) > 0.5
