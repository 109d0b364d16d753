"""The fused low-rank workload at its larger size: 3,000,000 Gaussians in
[-3,3]^3, K = 5, rank-2 factors A (N,5,2), B (N,2,15) ~ N(0, 0.1) (P = 5,
Q = 15) reconstructed inside the preprocess (sh2Weight_fast,
gsplatcu/kernel.cu:884-987, and its backward sh2Weight_fast_B, :994-1032),
1920x1080, the 8-camera orbit of bench.py's rank-2 step.
(The 1M size of this workload is checked by float64 finite differences in
test_gpu_fd64.py.)

Against the FP32 oracle over the whole frame of one view:
* depths, mask, every tile's range, the whole sorted id list, bit for bit;
* layer, contrib and final_tau of every pixel, bit for bit;
* dL/d(dc, A, B) of every Gaussian for L = sum(G * layer) within
  1e-5 * sum|terms| of the reference's backward (division form, then
  gsmodel.py:194-198 and sh2Weight_fast_B).
One step (mix, L1 + SSIM, backward, PaletteAdam with dc, A, B updated by the
kernel fused with their backward): the Adam step of every factor inside the
interval its float64 gradient fixes (tests/step_ref.py).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel, ops
from rgp.losses import L1SSIMLoss
from rgp.optim import PaletteAdam
from rgp.scene import make_scene, orbit_cameras
from step_ref import adam_first_step

pytestmark = pytest.mark.gpu
N_G, K, R, P, Q, W, H = 3_000_000, 5, 2, 5, 15, 1920, 1080
VIEW = 3


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.detach().cpu().numpy()


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b))) if a.dtype.kind == "f" else (a != b)
    assert not bad.any(), f"{what}: {int(bad.sum())} of {a.size} differ; first at {np.argwhere(bad)[:3].tolist()}"


def within(g, ref, ref_abs, what, tol=1e-5):
    err = np.abs(np.asarray(g, np.float64).reshape(ref.shape) - ref)
    lim = tol * ref_abs + 1e-30
    assert (err <= lim).all(), f"{what}: max err / lim {float(np.max(err / lim)):.3g}"


@pytest.fixture(scope="module")
def c2():
    print("case: C2 3M Gaussians K=5 rank-2 P=5 Q=15 1920x1080: scene", flush=True)
    sc = make_scene(N_G, seed=0, extent=3.0, K=K, lowrank=True, rank=R, P=P, Q=Q)
    cam = orbit_cameras(8, 6.0, W, H)[VIEW]
    print("oracle forward", flush=True)
    return sc, cam, O.forward_weight(sc, cam)


def factors(sc, device):
    return [T(x, device).requires_grad_(True) for x in (sc.dc, sc.A, sc.B)]


def test_c2_3m_full_frame(device, c2):
    sc, cam, o = c2
    tc = cam.to_torch(device)
    args = [T(x, device) for x in (sc.pws, sc.alphas, sc.scales, sc.rots)]
    dc, A, B = factors(sc, device)
    print("GPU forward + backward", flush=True)
    layer, mask = gsmodel.WeightGSFunctionFast.apply(args[0], dc, A, B, args[1], args[2], args[3], tc, K, Q)
    with torch.no_grad():
        _, depths, st = gsmodel.render_weights(args[0], args[1], args[2], args[3], tc, K, dc=dc.detach(),
                                               A=A.detach(), B=B.detach(), Q=Q, return_state=True)
    G = np.random.default_rng(33).standard_normal((K, H, W)).astype(np.float32)
    (layer * T(G, device)).sum().backward()
    torch.cuda.synchronize()
    m = st["m"]
    assert m == len(o["ids"]) and m > 8_000_000
    same(N(depths), o["depths"], "depths")
    same(N(mask), o["mask"], "mask")
    same(N(st["ranges"]), o["ranges"], "ranges")
    same(N(st["ids"][:m]), o["ids"], "ids")
    same(N(layer), o["layer"], "layer")
    same(N(st["contrib"]), o["contrib"], "contrib")
    same(N(st["final_tau"]), o["final_tau"], "final_tau")
    print(f"{m} pairs, E = {o['n_eval']}, C = {o['n_contrib']}; oracle backward", flush=True)
    ob = O.backward_weight(sc, o, G, replay=False)
    within(N(dc.grad), ob["ddc"], ob["ddc_abs"], "ddc")
    within(N(A.grad), ob["dA"], ob["dA_abs"] * 1.0001 + 1e-12, "dA")
    within(N(B.grad), ob["dB"], ob["dB_abs"] * 1.0001 + 1e-12, "dB")
    assert (np.abs(ob["ddc"]) > 0).any(axis=1).sum() > 100_000  # visible Gaussians with a gradient


def test_c2_3m_finetune_step(device, c2):
    sc, cam, o = c2
    tc = cam.to_torch(device)
    lr = 1e-3
    target = O.mix(np.clip(sc.palette + 0.05, 0, 1).astype(np.float32), o["layer"])
    args = [T(x, device) for x in (sc.pws, sc.alphas, sc.scales, sc.rots)]
    dc, A, B = factors(sc, device)
    pal = T(sc.palette, device)
    opt = PaletteAdam([dc, A, B], lr=lr, fused_lowrank=[(dc, A, B)])
    seen = {}
    print("step: GPU", flush=True)
    layer, _ = gsmodel.WeightGSFunctionFast.apply(args[0], dc, A, B, args[1], args[2], args[3], tc, K, Q)
    layer.register_hook(lambda g: seen.__setitem__("dlayer", g.detach().clone()))
    loss = L1SSIMLoss.apply(ops.Mix.apply(layer, pal), T(target, device), 0.2)
    opt.zero_grad(set_to_none=True)
    loss.backward()
    assert dc.grad is None and A.grad is None and B.grad is None  # went to the fused update
    opt.step()
    torch.cuda.synchronize()
    same(N(layer), o["layer"], "layer")
    print("step: oracle backward of the step's dL/dlayer", flush=True)
    ob = O.backward_weight(sc, o, N(seen["dlayer"]), replay=False)
    n_fixed = 0
    for name, t, p0, ref, ref_abs in (("dc", dc, sc.dc, ob["ddc"], ob["ddc_abs"]),
                                      ("A", A, sc.A, ob["dA"], ob["dA_abs"] * 1.0001 + 1e-12),
                                      ("B", B, sc.B, ob["dB"], ob["dB_abs"] * 1.0001 + 1e-12)):
        s = opt.state[t]
        n_fixed += adam_first_step(p0, N(t), N(s["exp_avg"]), N(s["exp_avg_sq"]), ref.reshape(p0.shape),
                                   1e-5 * ref_abs.reshape(p0.shape) + 1e-30, lr=lr, what=name)
    assert n_fixed > 1_000_000, n_fixed
