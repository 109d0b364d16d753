"""The headline finetune workload at full size: 1,000,000 Gaussians in
[-3,3]^3, K = 5 at full rank (16 SH coefficients per channel), 1920x1080, the
32-camera orbit at radius 6 (the scene and cameras of bench.py).  The
reference chain is WeightGSFunctionTest (gsplat/gsmodel.py:104-144) inside
the step of finetune.py:346-399 with the L1 + SSIM loss of
gsplat/pytorch_ssim.py:17-66.

For two of the 32 views, against the FP32 oracle over the whole frame:
* depths, the visibility mask, every tile's range and the whole sorted id
  list, bit for bit;
* layer, contrib and final_tau of every pixel, bit for bit, and the
  forward's per-entry masks of row strips with a contribution (which the
  replay backward evaluates), bit for bit wherever the forward writes them;
* dL/dshs of every Gaussian for L = sum(G * layer), G ~ N(0,1), within
  1e-5 * sum|terms| of the reference's own backward (tau recovered by
  division from final_tau, kernel.cu:1291-1379; the GPU runs the replay form,
  which differs from it by well under that bar).

One full step (palette mix, L1 + SSIM against a target rendered with a
perturbed palette, backward, PaletteAdam on the coefficients and the
chromatic palette):
* rgb bit for bit against the oracle's mix of the oracle's layer;
* the loss and every element of dL/drgb within 1e-5 of their sums of |terms|
  (tests/ssim_ref.py, float64);
* dL/dlayer bit for bit against the mix backward's float32 formula;
* dL/dpalette within 1e-5 * sum|terms| of the float64 sum;
* the Adam step of all 80M coefficients and of the palette inside the
  interval their float64 gradients (from the oracle's backward of this
  dL/dlayer) and tolerances fix (tests/step_ref.py).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel, ops
from rgp.losses import L1SSIMLoss
from rgp.optim import PaletteAdam
from rgp.scene import make_scene, orbit_cameras
from ssim_ref import l1_ssim_ref
from step_ref import adam_first_step
from test_gpu_parity import check_strips

pytestmark = pytest.mark.gpu
N_G, K, W, H = 1_000_000, 5, 1920, 1080
VIEWS = (0, 13)


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
    err = np.abs(np.asarray(g, np.float64) - ref)
    lim = tol * ref_abs + 1e-30
    assert (err <= lim).all(), f"{what}: max err / lim {float(np.max(err / lim)):.3g}"


@pytest.fixture(scope="module")
def c1():
    print("case: C1 1M Gaussians K=5 full rank 1920x1080: scene", flush=True)
    sc = make_scene(N_G, seed=0, extent=3.0, K=K)
    cams = orbit_cameras(32, 6.0, W, H)
    return sc, cams, {}


def oracle_forward(c1, v):
    sc, cams, cache = c1
    if v not in cache:
        print(f"view {v}: oracle forward", flush=True)
        cache[v] = O.forward_weight(sc, cams[v])
    return cache[v]


@pytest.mark.parametrize("v", VIEWS)
def test_c1_view_full_frame(device, c1, v):
    sc, cams, _ = c1
    o = oracle_forward(c1, v)
    tc = cams[v].to_torch(device)
    args = [T(x, device) for x in (sc.pws, sc.alphas, sc.scales, sc.rots)]
    shs = T(sc.shs, device).requires_grad_(True)
    print(f"view {v}: GPU forward + backward", flush=True)
    layer, mask = gsmodel.WeightGSFunctionTest.apply(args[0], shs, args[1], args[2], args[3], tc, K)
    with torch.no_grad():
        _, depths, st = gsmodel.render_weights(args[0], args[1], args[2], args[3], tc, K, shs=shs.detach(),
                                               return_state=True, strips=True)
    G = np.random.default_rng(100 + v).standard_normal((K, H, W)).astype(np.float32)
    (layer * T(G, device)).sum().backward()
    torch.cuda.synchronize()
    m = st["m"]
    assert m == len(o["ids"]) and m > 3_000_000
    same(N(depths), o["depths"], "depths")
    same(N(mask), o["mask"], "mask")
    same(N(st["ranges"]), o["ranges"], "ranges")
    same(N(st["ids"][:m]), o["ids"], "ids")
    same(N(layer), o["layer"], "layer")
    same(N(st["contrib"]), o["contrib"], "contrib")
    same(N(st["final_tau"]), o["final_tau"], "final_tau")
    assert check_strips(st, o, sc, cams[v]) > 1_000_000
    print(f"view {v}: {m} pairs, E = {o['n_eval']}, C = {o['n_contrib']}; oracle backward", flush=True)
    ob = O.backward_weight(sc, o, G, replay=False)
    within(N(shs.grad), ob["dshs"], ob["dshs_abs"], "dshs")
    assert (np.abs(ob["dshs"]) > 0).sum() > 5_000_000  # visible Gaussians with a gradient


def test_c1_finetune_step(device, c1):
    sc, cams, _ = c1
    v = VIEWS[0]
    o = oracle_forward(c1, v)
    tc = cams[v].to_torch(device)
    lr = 1e-3
    rng = np.random.default_rng(7)
    pal_t = np.clip(sc.palette + rng.normal(0, 0.1, sc.palette.shape), 0, 1).astype(np.float32)
    pal_t[:2] = sc.palette[:2]
    target = O.mix(pal_t, o["layer"])
    args = [T(x, device) for x in (sc.pws, sc.alphas, sc.scales, sc.rots)]
    shs = T(sc.shs, device).requires_grad_(True)
    pal = T(sc.palette, device)
    chroma = pal[2:].clone().requires_grad_(True)
    opt = PaletteAdam([shs, chroma], lr=lr, fused_sh=[shs])
    seen = {}
    print("step: GPU", flush=True)
    layer, _ = gsmodel.WeightGSFunctionTest.apply(args[0], shs, args[1], args[2], args[3], tc, K)
    layer.register_hook(lambda g: seen.__setitem__("dlayer", g.detach().clone()))
    rgb = ops.Mix.apply(layer, torch.cat([pal[:2], chroma], 0))
    rgb.register_hook(lambda g: seen.__setitem__("drgb", g.detach().clone()))
    loss = L1SSIMLoss.apply(rgb, T(target, device), 0.2)
    opt.zero_grad(set_to_none=True)
    loss.backward()
    assert shs.grad is None  # the SH gradient went to the fused update
    shs0 = sc.shs
    opt.step()
    torch.cuda.synchronize()

    print("step: checks", flush=True)
    rgb_n = N(rgb)
    same(rgb_n, O.mix(sc.palette, o["layer"]), "rgb")
    L, dref, dmag, Lmag = l1_ssim_ref(rgb_n, target)
    assert abs(float(loss.detach()) - L) <= 1e-5 * Lmag
    drgb = N(seen["drgb"])
    within(drgb, dref, dmag, "dL/drgb")
    P = sc.palette
    dl_ref = (P[:, 0, None, None] * drgb[0] + P[:, 1, None, None] * drgb[1]) + P[:, 2, None, None] * drgb[2]
    dlayer = N(seen["dlayer"])
    same(dlayer, dl_ref.astype(np.float32), "dL/dlayer")
    terms = o["layer"].reshape(K, 1, -1).astype(np.float64) * drgb.reshape(1, 3, -1).astype(np.float64)
    dpal, dpal_abs = terms.sum(-1), np.abs(terms).sum(-1)
    within(N(chroma.grad), dpal[2:], dpal_abs[2:], "dL/dchroma")

    ob = O.backward_weight(sc, o, dlayer, replay=False)
    st = opt.state[shs]
    fixed = adam_first_step(shs0, N(shs), N(st["exp_avg"]), N(st["exp_avg_sq"]), ob["dshs"],
                            1e-5 * ob["dshs_abs"] + 1e-30, lr=lr, what="shs")
    assert fixed > 0.3 * shs0.size, fixed
    sp = opt.state[chroma]
    adam_first_step(sc.palette[2:], N(chroma), N(sp["exp_avg"]), N(sp["exp_avg_sq"]), dpal[2:],
                    1e-5 * dpal_abs[2:] + 1e-30, lr=lr, what="palette")
