"""L1 + SSIM image loss of the finetune-step workload.

SSIM follows gsplat/pytorch_ssim.py:17-66 (11x11 Gaussian window, sigma 1.5,
C1 = (0.01 r)^2, C2 = (0.03 r)^2, mean over the map); the window is applied as
two separable 1-D passes, which is the same window (it is an outer product).
The combined loss is (1 - lambda) * L1 + lambda * (1 - SSIM), lambda = 0.2
(vanilla_3dgs/gsplat/pytorch_ssim.py:64-67).
"""
import math

import torch
import torch.nn.functional as F

_WIN = {}


def _window(size, sigma, channels, device, dtype):
    key = (size, sigma, channels, device, dtype)
    if key not in _WIN:
        g = torch.tensor([math.exp(-(x - size // 2) ** 2 / float(2 * sigma ** 2)) for x in range(size)],
                         dtype=torch.float32)
        g = (g / g.sum()).to(device=device, dtype=dtype)
        _WIN[key] = (g.view(1, 1, 1, size).repeat(channels, 1, 1, 1),
                     g.view(1, 1, size, 1).repeat(channels, 1, 1, 1))
    return _WIN[key]


def ssim(img1, img2, range_=1.0, window_size=11):
    """Mean SSIM of two (C,H,W) or (B,C,H,W) images."""
    if img1.dim() == 3:
        img1, img2 = img1.unsqueeze(0), img2.unsqueeze(0)
    c = img1.shape[1]
    x = torch.cat([img1, img2, img1 * img1, img2 * img2, img1 * img2], dim=1)
    wx, wy = _window(window_size, 1.5, 5 * c, x.device, x.dtype)
    p = window_size // 2
    y = F.conv2d(F.conv2d(x, wx, padding=(0, p), groups=5 * c), wy, padding=(p, 0), groups=5 * c)
    mu1, mu2, e11, e22, e12 = torch.split(y, c, dim=1)
    mu1_sq, mu2_sq, mu12 = mu1 * mu1, mu2 * mu2, mu1 * mu2
    s1, s2, s12 = e11 - mu1_sq, e22 - mu2_sq, e12 - mu12
    C1, C2 = (0.01 * range_) ** 2, (0.03 * range_) ** 2
    m = ((2 * mu12 + C1) * (2 * s12 + C2)) / ((mu1_sq + mu2_sq + C1) * (s1 + s2 + C2))
    return m.mean()


def l1_ssim_loss(image, target, lam=0.2):
    return (1.0 - lam) * torch.abs(image - target).mean() + lam * (1.0 - ssim(image, target))


def window_taps(size=11, sigma=1.5):
    """The normalised 1-D taps of create_window (pytorch_ssim.py:17-29), as
    float32 values."""
    import numpy as np
    g = np.array([math.exp(-(x - size // 2) ** 2 / float(2 * sigma ** 2)) for x in range(size)], np.float32)
    return (g / g.sum()).astype(np.float32)


class L1SSIMLoss(torch.autograd.Function):
    """(1 - lam) * mean|img - target| + lam * (1 - SSIM(img, target)) on the
    fused gfx950 kernels; the target is a constant."""

    @staticmethod
    def forward(ctx, img, target, lam=0.2, range_=1.0):
        import ctypes

        from ._lib import call, f32c, ptr, require_cuda, stream
        require_cuda(img, target)
        x, y = f32c(img), f32c(target)
        if x.dim() == 4:
            assert x.shape[0] == 1
            x, y = x[0], y[0]
        C, H, W = x.shape
        taps = window_taps()
        ctx.taps = taps
        abc = torch.empty((3, C, H, W), dtype=torch.float32, device=x.device)
        sums = torch.zeros(128, dtype=torch.float64, device=x.device)
        loss = torch.empty((), dtype=torch.float32, device=x.device)
        call("rgp_l1_ssim_forward", C, H, W, ptr(x), ptr(y), taps.ctypes.data_as(ctypes.c_void_p),
             (0.01 * range_) ** 2, (0.03 * range_) ** 2, float(lam), ptr(abc), ptr(sums), ptr(loss), stream(x.device))
        ctx.lam = float(lam)
        ctx.shape = img.shape
        ctx.save_for_backward(x, y, abc)
        return loss

    @staticmethod
    def backward(ctx, gout):
        import ctypes

        from ._lib import call, ptr, stream
        x, y, abc = ctx.saved_tensors
        C, H, W = x.shape
        g = gout.detach().float().contiguous().reshape(1)
        dimg = torch.empty_like(x)
        call("rgp_l1_ssim_backward", C, H, W, ptr(x), ptr(y), ptr(abc), ctx.taps.ctypes.data_as(ctypes.c_void_p),
             ctx.lam, ptr(g), ptr(dimg), stream(x.device))
        return dimg.reshape(ctx.shape), None, None, None


def l1_ssim_loss_fused(image, target, lam=0.2):
    return L1SSIMLoss.apply(image, target, lam)
