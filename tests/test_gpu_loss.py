"""The fused L1 + SSIM loss kernels (gsplat/pytorch_ssim.py:17-66 semantics,
zero padding, lambda = 0.2) against the float64 reference of tests/ssim_ref.py:
the loss within 1e-5 relative, and every element of dL/dimg within 1e-5 of
its sum of |terms| (the magnitude bound that covers float32 rounding of the
expression, including the cancellation in the windowed variances)."""
import numpy as np
import pytest
import torch

from rgp.losses import L1SSIMLoss
from ssim_ref import l1_ssim_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,kind", [((3, 97, 131), "random"), ((3, 64, 64), "random"),
                                        ((3, 1080, 1920), "random"), ((3, 200, 328), "smooth")])
def test_fused_l1_ssim_matches_reference(device, shape, kind):
    print(f"case: l1+ssim {shape} {kind}", flush=True)
    rng = np.random.default_rng(sum(shape))
    if kind == "random":
        tgt = rng.uniform(0, 1, shape).astype(np.float32)
        img = np.clip(tgt + rng.normal(0, 0.1, shape), 0, 1).astype(np.float32)
    else:  # small local variances on large means
        base = rng.uniform(0.3, 0.9, (shape[0], 1, 1)) + 0.05 * np.sin(np.arange(shape[2]) / 7.0)[None, None, :]
        tgt = np.broadcast_to(base, shape).astype(np.float32).copy()
        img = (tgt + rng.normal(0, 0.01, shape)).astype(np.float32)
    img[:, ::7, ::5] = tgt[:, ::7, ::5]  # exact ties: sign(0) = 0
    L, ref, mag, Lmag = l1_ssim_ref(img, tgt)
    x = torch.tensor(img, device=device, requires_grad=True)
    y = torch.tensor(tgt, device=device)
    loss = L1SSIMLoss.apply(x, y, 0.2)
    (2.5 * loss).backward()
    assert abs(float(loss.detach()) - L) <= 1e-5 * Lmag
    err = np.abs(x.grad.double().cpu().numpy() - 2.5 * ref)
    lim = 1e-5 * 2.5 * mag
    assert (err <= lim).all(), f"max err / (1e-5 sum|terms|) = {float(np.max(err / lim)):.3g}"


def test_fused_l1_ssim_loss_is_deterministic(device):
    """The per-tile partial sums are added in a fixed order: the same inputs
    give the same loss bits on every call."""
    rng = np.random.default_rng(11)
    tgt = rng.uniform(0, 1, (3, 540, 960)).astype(np.float32)
    img = np.clip(tgt + rng.normal(0, 0.1, tgt.shape), 0, 1).astype(np.float32)
    x, y = torch.tensor(img, device=device), torch.tensor(tgt, device=device)
    vals = {float(L1SSIMLoss.apply(x, y, 0.2)) for _ in range(4)}
    assert len(vals) == 1
