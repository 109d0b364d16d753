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

        from ._lib import call, f32c, lib, ptr, require_cuda, stream
        require_cuda(img, target)
        x, y = f32c(img), f32c(target)
        if x.dim() == 4:
            assert x.shape[0] == 1
            x, y = x[0], y[0]
        C, H, W = x.shape
        taps = window_taps()
        ctx.taps = taps
        abc = torch.empty((3, C, H, W), dtype=torch.float32, device=x.device)
        sums = torch.empty(int(lib().rgp_l1_ssim_sums(C, H, W)), dtype=torch.float64, device=x.device)
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


# ---------------------------------------------------------------------------
# barycentric palette loss of the finetune step (gsplat/gsmodel.py:219-285,
# finetune.py:42-76 and 353-392)
# ---------------------------------------------------------------------------
def _bary_ws(Kc, device):
    from ._lib import lib
    return torch.empty(int(lib().rgp_bary_workspace_bytes(int(Kc))), dtype=torch.uint8, device=device)


class BaryLoss(torch.autograd.Function):
    """L_geo + lambda_pos L_pos [+ lambda_alpha L_alpha] against the barycentric
    target weights of the frozen render C (finetune.py:353-392): one kernel
    pass (rgp_bary_loss) forms the loss, dL/dW_tilde and dL/dpalette; the
    backward only scales them.  Returns (loss, parts) with parts = (L_geo,
    L_pos, L_alpha), not differentiable."""

    @staticmethod
    def forward(ctx, W_tilde, palette_rgb, C, gt_alpha=None, clamp=False, lambda_pos=0.5, lambda_alpha=1.0,
                tau=1e-9):
        from ._lib import call, f32c, ptr, require_cuda, stream
        require_cuda(W_tilde, palette_rgb, C, gt_alpha)
        Wt, P, Ci = f32c(W_tilde), f32c(palette_rgb), f32c(C)
        Kc = P.shape[0]
        Kt, H, W = Wt.shape
        if Kt != Kc + 2 or Ci.shape != (3, H, W):
            raise ValueError(f"W_tilde {tuple(Wt.shape)}, C {tuple(Ci.shape)} and palette {tuple(P.shape)} disagree")
        a = None if gt_alpha is None else f32c(gt_alpha).reshape(H, W)
        dW, dP = torch.empty_like(Wt), torch.empty_like(P)
        out = torch.empty(4, dtype=torch.float32, device=Wt.device)
        call("rgp_bary_loss", W, H, Kc, ptr(Ci), ptr(Wt), ptr(a), ptr(P), float(tau), int(bool(clamp)),
             float(lambda_pos), float(lambda_alpha), None, ptr(dW), ptr(dP), ptr(out), ptr(_bary_ws(Kc, Wt.device)),
             stream(Wt.device))
        ctx.save_for_backward(dW, dP)
        ctx.shapes = (W_tilde.shape, palette_rgb.shape)
        parts = out[1:]
        ctx.mark_non_differentiable(parts)
        return out[0], parts

    @staticmethod
    def backward(ctx, gloss, _):
        dW, dP = ctx.saved_tensors
        g = gloss.to(torch.float32)
        return ((dW * g).reshape(ctx.shapes[0]), (dP * g).reshape(ctx.shapes[1]), None, None, None, None, None,
                None)


def bary_loss(W_tilde, palette_rgb, C, gt_alpha=None, clamp=False, lambda_pos=0.5, lambda_alpha=1.0, tau=1e-9):
    return BaryLoss.apply(W_tilde, palette_rgb, C, gt_alpha, clamp, lambda_pos, lambda_alpha, tau)


class BaryCoords(torch.autograd.Function):
    """barycentric_coords_3d (gsmodel.py:219-285) on the gfx950 kernels: the
    (K+2, H, W) weights, differentiable in the palette (C is a constant)."""

    @staticmethod
    def forward(ctx, C, P, tau=1e-9):
        from ._lib import call, f32c, ptr, require_cuda, stream
        require_cuda(C, P)
        Ci, Pc = f32c(C), f32c(P)
        Kc = Pc.shape[0]
        _, H, W = Ci.shape
        out = torch.empty((Kc + 2, H, W), dtype=torch.float32, device=Ci.device)
        call("rgp_bary_coords", W, H, Kc, ptr(Ci), ptr(Pc), float(tau), ptr(out), stream(Ci.device))
        ctx.save_for_backward(Ci, Pc)
        ctx.tau = float(tau)
        return out

    @staticmethod
    def backward(ctx, dWb):
        from ._lib import call, f32c, ptr, stream
        Ci, Pc = ctx.saved_tensors
        Kc = Pc.shape[0]
        _, H, W = Ci.shape
        dP = torch.empty_like(Pc)
        call("rgp_bary_coords_backward", W, H, Kc, ptr(Ci), ptr(Pc), ctx.tau, ptr(f32c(dWb)), ptr(dP),
             ptr(_bary_ws(Kc, Ci.device)), stream(Ci.device))
        return None, dP, None


def barycentric_coords_3d(C, P, tau=1e-9, eps=1e-12):
    """Drop-in for gsmodel.py:219: the kernels use eps = 1e-12, the reference default."""
    if eps != 1e-12:
        raise ValueError("barycentric_coords_3d: only eps = 1e-12 is implemented")
    return BaryCoords.apply(C, P, tau)


def palette_compactness_loss(palette_rgb):
    """finetune.py:42-63: the hull volume of black, white and the K chromatic
    colours as K tetrahedra (black, white, c_k, c_{k+1}) fanned round the grey
    axis; |det| / 6 each."""
    K = palette_rgb.shape[0]
    if K < 2:
        raise ValueError("palette_compactness_loss needs at least 2 chromatic colours")
    nxt = torch.roll(palette_rgb, -1, 0)
    white = torch.ones_like(palette_rgb)
    # det of the rows (white - black, c_k - black, c_{k+1} - black) as a triple product
    det = (white * torch.linalg.cross(palette_rgb, nxt, dim=1)).sum(1)
    return (torch.abs(det) / 6.0).sum()


def hue_separation_loss(gap_raw, min_gap_fraction=0.5):
    """finetune.py:66-76: squared shortfall of each hue gap below
    min_gap_fraction of the even spacing, averaged."""
    K = gap_raw.shape[0]
    g = F.softplus(gap_raw)
    gaps = (2 * math.pi) * g / g.sum()
    return (F.relu(min_gap_fraction * (2 * math.pi) / K - gaps) ** 2).mean()
