"""The fused L1 + SSIM loss kernels against the same loss written with torch
conv2d in float64 on the CPU (gsplat/pytorch_ssim.py:17-66 semantics)."""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from rgp.losses import L1SSIMLoss

pytestmark = pytest.mark.gpu


def ref_loss(img, tgt, lam=0.2):
    g = torch.tensor([math.exp(-(x - 5) ** 2 / (2 * 1.5 ** 2)) for x in range(11)], dtype=torch.float64)
    g = g / g.sum()
    w = (g[:, None] @ g[None, :]).expand(3, 1, 11, 11).contiguous()
    x, y = img[None], tgt[None]
    conv = lambda a: F.conv2d(a, w, padding=5, groups=3)
    mu1, mu2 = conv(x), conv(y)
    s1, s2, s12 = conv(x * x) - mu1 ** 2, conv(y * y) - mu2 ** 2, conv(x * y) - mu1 * mu2
    C1, C2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu1 * mu2 + C1) * (2 * s12 + C2)) / ((mu1 ** 2 + mu2 ** 2 + C1) * (s1 + s2 + C2))
    return (1 - lam) * (x - y).abs().mean() + lam * (1 - m.mean())


@pytest.mark.parametrize("shape", [(3, 97, 131), (3, 64, 64), (3, 1080, 1920)])
def test_fused_l1_ssim_matches_torch(device, shape):
    print(f"case: l1+ssim {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    tgt = rng.uniform(0, 1, shape)
    img = np.clip(tgt + rng.normal(0, 0.1, shape), 0, 1)
    img[:, ::7, ::5] = tgt[:, ::7, ::5]  # exact ties: sign(0) = 0
    xi = torch.tensor(img, dtype=torch.float64, requires_grad=True)
    L = ref_loss(xi, torch.tensor(tgt))
    L.backward()
    x = torch.tensor(img, dtype=torch.float32, device=device, requires_grad=True)
    y = torch.tensor(tgt, dtype=torch.float32, device=device)
    l = L1SSIMLoss.apply(x, y, 0.2)
    (2.5 * l).backward()
    assert float(l) == pytest.approx(float(L), rel=1e-5, abs=1e-6)
    g = x.grad.double().cpu()
    ref = 2.5 * xi.grad
    assert (g - ref).abs().max() <= 1e-4 * ref.abs().max()
