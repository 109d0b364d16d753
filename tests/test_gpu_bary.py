"""Barycentric palette loss kernels (csrc/rgp_bary.hip) against the FP32
oracle (oracle/rgp_oracle_bary.cpp, pinned in tests/test_oracle_bary.py).

Bar: W_bary and dL/dW_tilde bit-exact (the kernel and the FP32 oracle perform
the same IEEE operations in the same order, near-tie pixels of the tau = 1e-9
softmax included); the loss terms within 1e-6 relative (sums of the same FP32
per-pixel values in another order); dL/dpalette within 1e-5 * sum|terms| of
the oracle's float64 sum of the same FP32 per-pixel terms."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import losses
from rgp._lib import call, lib, ptr, stream

pytestmark = pytest.mark.gpu


def palette(Kc, seed):
    rng = np.random.default_rng(seed)
    h = np.sort(rng.uniform(0, 2 * np.pi, Kc)) if Kc > 1 else np.array([0.3])
    u, v = np.array([1, -1, 0]) / np.sqrt(2), np.array([1, 1, -2]) / np.sqrt(6)
    P = 0.5 + rng.uniform(0.25, 0.45, (Kc, 1)) * (np.cos(h)[:, None] * u + np.sin(h)[:, None] * v)
    return np.clip(P + rng.uniform(-0.1, 0.1, (Kc, 1)), 0.02, 0.98).astype(np.float32)


def inputs(Kc, H, W, seed):
    rng = np.random.default_rng(seed)
    P = palette(Kc, seed)
    C = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    C[:, 0, :Kc] = P.T  # pixels equal to palette colours
    Wt = rng.normal(0.15, 0.35, (Kc + 2, H, W)).astype(np.float32)
    a = rng.uniform(0, 1, (H, W)).astype(np.float32)
    a[0, :4] = [0.0, 1e-4, 1e-3, 1.0]
    return P, C, Wt, a


def T(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def gpu_loss(dev, C, Wt, P, a, clamp, tau=1e-9, lam_pos=0.5, lam_alpha=1.0):
    Kc = P.shape[0]
    Kt, H, W = Wt.shape
    Ct, Wtt, Pt = T(C, dev), T(Wt, dev), T(P, dev)
    at = None if a is None else T(a, dev)
    wb, dW = torch.empty_like(Wtt), torch.empty_like(Wtt)
    dP, out = torch.empty_like(Pt), torch.empty(4, dtype=torch.float32, device=dev)
    ws = torch.empty(int(lib().rgp_bary_workspace_bytes(Kc)), dtype=torch.uint8, device=dev)
    call("rgp_bary_loss", W, H, Kc, ptr(Ct), ptr(Wtt), ptr(at), ptr(Pt), tau, int(clamp), lam_pos, lam_alpha, ptr(wb),
         ptr(dW), ptr(dP), ptr(out), ptr(ws), stream(dev))
    torch.cuda.synchronize()
    return wb.cpu().numpy(), dW.cpu().numpy(), dP.cpu().numpy(), out.cpu().numpy()


def check(dev, Kc, H, W, clamp, with_alpha, seed):
    P, C, Wt, a = inputs(Kc, H, W, seed)
    a = a if with_alpha else None
    print(f"case: bary loss Kc={Kc} {W}x{H} clamp={clamp} alpha={with_alpha}", flush=True)
    wb, dW, dP, out = gpu_loss(dev, C, Wt, P, a, clamp)
    o = O.bary_loss(C, Wt, P, alpha=a, clamp=clamp)
    assert np.array_equal(wb, o["wbary"]), np.argwhere(wb != o["wbary"])[:5]
    assert np.array_equal(dW, o["dW"]), np.argwhere(dW != o["dW"])[:5]
    nk, hw = (Kc + 2) * H * W, H * W
    want = [o["loss"], o["sums"][0] / nk, o["sums"][1] / nk, o["sums"][2] / hw if with_alpha else 0.0]
    np.testing.assert_allclose(out, want, rtol=1e-6, atol=1e-12)
    err = np.abs(dP.astype(np.float64) - o["dP"])
    lim = 1e-5 * o["dP_abs"] + 1e-30
    assert (err <= lim).all(), float(np.max(err / lim))


@pytest.mark.parametrize("Kc", [1, 2, 3, 4, 5, 6])
def test_loss_bitexact_plain(device, Kc):
    check(device, Kc, 37, 53, False, False, Kc)


@pytest.mark.parametrize("clamp,with_alpha", [(True, False), (False, True), (True, True)])
@pytest.mark.parametrize("Kc", [3, 5])
def test_loss_bitexact_variants(device, Kc, clamp, with_alpha):
    check(device, Kc, 64, 96, clamp, with_alpha, 10 + Kc)


def test_loss_1080p(device):
    """The finetune-step workload: K = 5 (3 chromatic colours), 1920x1080."""
    check(device, 3, 1080, 1920, False, False, 99)


def test_coords_and_backward(device):
    Kc, H, W = 4, 40, 50
    P, C, _, _ = inputs(Kc, H, W, 5)
    print("case: barycentric_coords_3d forward/backward", flush=True)
    Ct, Pt = T(C, device), T(P, device).requires_grad_(True)
    Wb = losses.barycentric_coords_3d(Ct, Pt)
    o = O.bary_loss(C, np.zeros((Kc + 2, H, W), np.float32), P)
    assert np.array_equal(Wb.detach().cpu().numpy(), o["wbary"])
    G = np.random.default_rng(3).normal(size=(Kc + 2, H, W)).astype(np.float32)
    (Wb * T(G, device)).sum().backward()
    dP, dPa = O.bary_coords_backward(C, P, G)
    err = np.abs(Pt.grad.cpu().numpy().astype(np.float64) - dP)
    assert (err <= 1e-5 * dPa + 1e-30).all(), float(np.max(err / (1e-5 * dPa + 1e-30)))


def test_autograd_loss(device):
    Kc, H, W = 3, 48, 64
    P, C, Wt, a = inputs(Kc, H, W, 8)
    print("case: BaryLoss autograd", flush=True)
    Wtt, Pt = T(Wt, device).requires_grad_(True), T(P, device).requires_grad_(True)
    loss, parts = losses.bary_loss(Wtt, Pt, T(C, device), T(a, device), clamp=True)
    (3.0 * loss).backward()
    _, dW, dP, out = gpu_loss(device, C, Wt, P, a, True)
    assert float(loss) == float(out[0]) and np.array_equal(parts.cpu().numpy(), out[1:])
    assert np.array_equal(Wtt.grad.cpu().numpy(), (torch.from_numpy(dW) * 3.0).numpy())
    assert np.array_equal(Pt.grad.cpu().numpy(), (torch.from_numpy(dP) * 3.0).numpy())


def test_finetune_step_runs(device):
    """One step of the reference's finetune loop (finetune.py:346-399) on the
    kernels: FinetuneModel, barycentric loss, compactness and separation
    terms, backward, Adam; the palette parameters receive gradients."""
    from rgp import gsmodel
    from rgp.scene import make_scene, orbit_cameras
    print("case: palette finetune step (20k Gaussians, 256x256)", flush=True)
    sc = make_scene(20_000, seed=1, extent=1.0, K=5)
    cam = orbit_cameras(4, 4.0, 256, 256)[0].to_torch(device)
    frozen = {k: T(getattr(sc, k), device) for k in ("pws", "alphas", "scales", "rots")}
    params = {"low_shs_w": T(sc.shs[:, :5], device).requires_grad_(True),
              "high_shs_w": T(sc.shs[:, 5:], device).requires_grad_(True), "low_rank": False}
    for k, v in (("hue_base", [0.3]), ("gap_raw", [0.0, 0.1, -0.1]), ("sat_raw", [0.5, 0.2, 0.8]),
                 ("lit_raw", [0.0, 0.1, -0.2])):
        params[k] = torch.tensor(v, dtype=torch.float32, device=device, requires_grad=True)
    model = gsmodel.FinetuneModel(num_palette=3, use_fast=False)
    opt = torch.optim.Adam([v for v in params.values() if torch.is_tensor(v)], lr=1e-3)
    C = torch.rand((3, 256, 256), device=device)
    W_tilde, pal = model(frozen["pws"], params, frozen["alphas"], frozen["scales"], frozen["rots"], cam)
    loss, parts = losses.bary_loss(W_tilde, pal, C)
    loss = loss + 0.001 * losses.palette_compactness_loss(pal) + losses.hue_separation_loss(params["gap_raw"], 0.3)
    opt.zero_grad()
    loss.backward()
    for k in ("low_shs_w", "high_shs_w", "hue_base", "sat_raw", "lit_raw"):
        assert params[k].grad is not None and torch.isfinite(params[k].grad).all(), k
    assert float(params["sat_raw"].grad.abs().sum()) > 0
    opt.step()
    assert torch.isfinite(loss)
