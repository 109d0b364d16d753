"""The reference's two coefficient leaves without their concatenation
(WeightGSFunctionSplit, preprocess mode 2, rgp_sh_backward_adam_split).

* forward: layer and mask bit for bit equal to WeightGSFunctionTest on
  torch.cat(low, high) (gsplat/utils.py:157-158), for a high leaf staged in
  LDS (16-byte aligned) and one read per lane (unaligned);
* backward: the two gradients, against the oracle's dL/dshs split at K
  (within 1e-5 of the sum of |terms|);
* the fused update of the pair equals rgp_sh_backward_adam_full on the
  concatenated tensor and moments, bit for bit, for the same dL/dw;
* PaletteAdam with the pair in fused_sh gives the update of the
  concatenated tensor, also when a leaf carries a second gradient.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel, ops
from rgp._lib import I64, call, ptr, stream
from rgp.optim import PaletteAdam
from rgp.scene import workload_c0

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.detach().cpu().numpy()


def leaves(shs, K, dev, aligned=True):
    low = T(shs[:, :K], dev)
    hi = np.ascontiguousarray(shs[:, K:])
    if aligned:
        return low, T(hi, dev)
    flat = torch.empty(hi.size + 1, dtype=torch.float32, device=dev)
    high = flat[1:].view(hi.shape)
    high.copy_(T(hi, dev))
    assert high.is_contiguous() and high.data_ptr() % 16 != 0
    return low, high


@pytest.mark.parametrize("aligned", [True, False])
def test_split_forward_backward(device, aligned):
    print(f"case: split leaves, aligned={aligned}", flush=True)
    scene, cam = workload_c0()
    K = scene.K
    tc = cam.to_torch(device)
    args = (T(scene.alphas, device), T(scene.scales, device), T(scene.rots, device), tc)
    pws = T(scene.pws, device)
    ref_layer, ref_mask = gsmodel.WeightGSFunctionTest.apply(pws, T(scene.shs, device), *args, K)
    low, high = leaves(scene.shs, K, device, aligned)
    low.requires_grad_(True)
    high.requires_grad_(True)
    layer, mask = gsmodel.WeightGSFunctionSplit.apply(pws, low, high, *args, K)
    assert np.array_equal(N(layer), N(ref_layer))
    assert np.array_equal(N(mask), N(ref_mask))
    o = O.forward_weight(scene, cam)
    dl = np.random.default_rng(2).normal(size=o["layer"].shape).astype(np.float32)
    layer.backward(T(dl, device))
    ob = O.backward_weight(scene, o, dl, replay=True)
    for name, g, sl in (("low", low.grad, slice(0, K)), ("high", high.grad, slice(K, None))):
        ref, ref_abs = ob["dshs"][:, sl], ob["dshs_abs"][:, sl]
        err = np.abs(N(g).astype(np.float64) - ref)
        assert (err <= 1e-5 * ref_abs + 1e-30).all(), (name, float(np.max(err / (1e-5 * ref_abs + 1e-30))))


def test_split_adam_kernel_equals_concatenated(device):
    print("case: split adam kernel vs concatenated", flush=True)
    scene, cam = workload_c0()
    K, n = scene.K, 4000 + 21
    rng = np.random.default_rng(5)
    pws = T(scene.pws[:n], device)
    camd = ops.camera_struct(cam.to_torch(device), device)
    p = T(rng.normal(0, 0.1, (n, 16 * K)).astype(np.float32), device)
    m, v = torch.zeros_like(p), torch.zeros_like(p)
    lo, hi = p[:, :K].contiguous(), p[:, K:].contiguous()
    mlo, mhi, vlo, vhi = (torch.zeros_like(x) for x in (lo, hi, lo, hi))
    for step in range(1, 4):
        dw = T(rng.normal(0, 1e-3, (n, K)).astype(np.float32), device)
        call("rgp_sh_backward_adam_full", n, K, 16, ptr(camd), ptr(pws), ptr(dw), ptr(p), ptr(m), ptr(v),
             I64(step), 3e-3, 0.9, 0.999, 1e-8, 1e-2, stream(device))
        call("rgp_sh_backward_adam_split", n, K, ptr(camd), ptr(pws), ptr(dw), ptr(lo), ptr(mlo), ptr(vlo), ptr(hi),
             ptr(mhi), ptr(vhi), I64(step), 3e-3, 0.9, 0.999, 1e-8, 1e-2, stream(device))
    for a, b, what in ((p, torch.cat([lo, hi], 1), "param"), (m, torch.cat([mlo, mhi], 1), "exp_avg"),
                       (v, torch.cat([vlo, vhi], 1), "exp_avg_sq")):
        assert np.array_equal(N(a), N(b)), what


@pytest.mark.parametrize("second", [False, True])
def test_palette_adam_split_pair(device, second):
    """One step of PaletteAdam on the pair equals one on the concatenated
    tensor from the same dL/dw; second=True adds an ordinary gradient to the
    high leaf (the pair then falls back to the materialised gradient)."""
    print(f"case: PaletteAdam split pair, second gradient={second}", flush=True)
    scene, cam = workload_c0()
    K, n = scene.K, scene.n
    tc = cam.to_torch(device)
    pws = T(scene.pws, device)
    args = (T(scene.alphas, device), T(scene.scales, device), T(scene.rots, device), tc)
    low, high = (x.requires_grad_(True) for x in leaves(scene.shs, K, device))
    opt = PaletteAdam([low, high], lr=3e-3, fused_sh=[(low, high)])
    layer, _ = gsmodel.WeightGSFunctionSplit.apply(pws, low, high, *args, K)
    dl = T(np.random.default_rng(6).normal(size=tuple(layer.shape)).astype(np.float32), device)
    loss = (layer * dl).sum() + (0.5 * (high ** 2).sum() if second else 0.0)
    opt.zero_grad()
    loss.backward()
    # the dL/dw the split backward handed over, applied to the concatenated tensor
    pend = opt._pending[id(low)]
    dw, camd = pend[2], pend[3]
    shs = T(scene.shs, device)
    m, v = torch.zeros_like(shs), torch.zeros_like(shs)
    if second:
        g = torch.empty_like(shs)
        call("rgp_sh_backward_full", n, K, 16, ptr(camd), ptr(pws), ptr(dw), ptr(g), stream(device))
        g[:, K:] += high.grad
        call("rgp_adam", I64(shs.numel()), ptr(shs), ptr(g), ptr(m), ptr(v), I64(1), 3e-3, 0.9, 0.999, 1e-8, 0.0,
             stream(device))
    else:
        call("rgp_sh_backward_adam_full", n, K, 16, ptr(camd), ptr(pws), ptr(dw), ptr(shs), ptr(m), ptr(v), I64(1),
             3e-3, 0.9, 0.999, 1e-8, 0.0, stream(device))
    opt.step()
    assert np.array_equal(N(torch.cat([low, high], 1)), N(shs))
    assert np.array_equal(N(torch.cat([opt.state[low]["exp_avg"], opt.state[high]["exp_avg"]], 1)), N(m))
