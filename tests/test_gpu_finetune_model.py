"""FinetuneModel (gsplat/gsmodel.py:292-329) on one set of rank-1 low-rank
parameters, the reference's own fused / unfused pair (gsmodel.py:303-316):
use_fast=True renders through WeightGSFunctionFast (the factors reconstructed
inside the preprocess), use_fast=False reconstructs the high bands with
reconstruct_high_shs (gsmodel.py:90-97) and renders through
WeightGSFunctionTest.

* The two layers agree within 1e-5 of the per-pixel sum of |terms|
  (sum_i tau_i alpha'_i (C0 |dc_k| + sum_lm |Y_lm| |A_p B_q|), rendered by the
  oracle): they differ only in the rounding of A B Y.
* Each path's dL/d(dc, A, B) lies within 1e-5 * sum|terms| of the oracle's
  float64 low-rank backward (division form, then sh2Weight_fast_B).
* The palette is get_palette_rgb of the HSL parameters in both.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp.gsmodel import FinetuneModel
from rgp.scene import make_scene, orbit_cameras

pytestmark = pytest.mark.gpu
K, P, Q = 5, 5, 15


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def within(g, ref, ref_abs, what):
    err = np.abs(np.asarray(g, np.float64).reshape(ref.shape) - ref)
    lim = 1e-5 * ref_abs + 1e-30
    assert (err <= lim).all(), f"{what}: max err / lim {float(np.max(err / lim)):.3g}"


def test_use_fast_pair(device):
    print("case: FinetuneModel use_fast True / False, rank 1, 6000 Gaussians 256x192", flush=True)
    sc = make_scene(6000, seed=31, extent=1.0, K=K, lowrank=True, rank=1, P=P, Q=Q)
    cam = orbit_cameras(5, 4.0, 256, 192)[2]
    tc = cam.to_torch(device)
    geo = [T(x, device) for x in (sc.pws, sc.alphas, sc.scales, sc.rots)]
    A1, B1 = sc.A[:, :, 0], sc.B[:, 0, :]  # the reference's rank-1 shapes (N,P) and (N,Q)
    hsl = dict(hue_base=[0.4], gap_raw=[0.1, -0.2, 0.05], sat_raw=[0.6, 0.3, 0.9], lit_raw=[0.1, -0.1, 0.2])
    G = np.random.default_rng(4).standard_normal((K, cam.height, cam.width)).astype(np.float32)
    out = {}
    for fast in (True, False):
        params = {"low_rank": True, "reshape_Q": Q, "low_shs_w": T(sc.dc, device).requires_grad_(True),
                  "high_shs_A": T(A1, device).requires_grad_(True), "high_shs_B": T(B1, device).requires_grad_(True)}
        for k, v in hsl.items():
            params[k] = torch.tensor(v, dtype=torch.float32, device=device)
        model = FinetuneModel(num_palette=3, use_fast=fast)
        W_tilde, pal = model(geo[0], params, geo[1], geo[2], geo[3], tc)
        (W_tilde * T(G, device)).sum().backward()
        out[fast] = dict(layer=W_tilde.detach().cpu().numpy(), pal=pal.detach().cpu().numpy(),
                         mask=model.mask.cpu().numpy(),
                         g=[params[k].grad.cpu().numpy() for k in ("low_shs_w", "high_shs_A", "high_shs_B")])
    assert np.array_equal(out[True]["pal"], out[False]["pal"])
    assert np.array_equal(out[True]["mask"], out[False]["mask"])

    o = O.forward_weight(sc, cam)  # the fused path's oracle (sh2Weight_fast on the rank-1 factors)
    assert np.array_equal(out[True]["layer"], o["layer"])
    C = np.abs(A1.astype(np.float64))[:, :, None] * np.abs(B1.astype(np.float64))[:, None, :]
    wabs = (np.float64(0.28209479177387814) * np.abs(sc.dc.astype(np.float64)) +
            np.einsum("nl,nlk->nk", np.abs(o["ylm"].astype(np.float64)), C.reshape(sc.n, 15, K)))
    M, _, _, _, _ = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"], sc.alphas,
                           wabs.astype(np.float32))
    diff = np.abs(out[True]["layer"].astype(np.float64) - out[False]["layer"])
    assert (diff <= 1e-5 * M.astype(np.float64) + 1e-30).all(), float(np.max(diff / np.maximum(M, 1e-30)))

    ob = O.backward_weight(sc, o, G, replay=False)
    for fast in (True, False):
        gdc, gA, gB = out[fast]["g"]
        within(gdc, ob["ddc"], ob["ddc_abs"], f"ddc fast={fast}")
        within(gA, ob["dA"].reshape(A1.shape), ob["dA_abs"].reshape(A1.shape) * 1.0001 + 1e-12, f"dA fast={fast}")
        within(gB, ob["dB"].reshape(B1.shape), ob["dB_abs"].reshape(B1.shape) * 1.0001 + 1e-12, f"dB fast={fast}")
