"""Multi-view batches (render_weights_views, WeightGSFunction*Views): several
cameras of one resolution in one preprocess launch and one compositing launch
per direction.  Each view must give, bit for bit, what the single-view path
and the oracle give for its camera (layer, contrib, final_tau, ranges, the
sorted ids, the culling mask), and the coefficient gradients must be the sum
over the views of the oracle's per-view gradients, within the gradient bar
(1e-5 of the sum of |terms|).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel
from rgp.scene import make_scene, orbit_cameras

pytestmark = pytest.mark.gpu
GRAD_TOL = 1e-5


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.detach().cpu().numpy()


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else (a == b))
    assert not bad.any(), f"{what}: {bad.sum()} of {a.size} differ"


def grad_close(g, ref, ref_abs, what):
    err = np.abs(g.astype(np.float64) - ref)
    lim = GRAD_TOL * ref_abs + 1e-30
    assert (err <= lim).all(), f"{what}: max err/lim {np.max(err / lim):.3g}"


def check_views_forward(scene, cams, layers, depths, st, dev, **coef):
    for v, cam in enumerate(cams):
        o = O.forward_weight(scene, cam)
        m = len(o["ids"])
        same(N(layers[v]), o["layer"], f"view {v} layer")
        same(N(st["contrib"][v]), o["contrib"], f"view {v} contrib")
        same(N(st["final_tau"][v]), o["final_tau"], f"view {v} final_tau")
        same(N(st["ranges"][v]), o["ranges"], f"view {v} ranges")
        same(N(st["ids"][v][:m]), o["ids"], f"view {v} ids")
        assert st["m"][v] == m
        same(N(depths[v]), o["depths"], f"view {v} depths")
        # and the single-view path for the same camera
        l1, d1 = gsmodel.render_weights(T(scene.pws, dev), T(scene.alphas, dev), T(scene.scales, dev),
                                        T(scene.rots, dev), cam.to_torch(dev), scene.K, **coef)
        same(N(l1), N(layers[v]), f"view {v} vs single-view layer")


@pytest.mark.parametrize("views", [1, 3])
def test_views_full_rank(device, views):
    print(f"case: multi-view full rank, {views} views, 6000 Gaussians 160x128", flush=True)
    scene = make_scene(6000, seed=21, extent=1.0, K=5)
    cams = orbit_cameras(5, 4.0, 160, 128)[:views]
    coef = dict(shs=T(scene.shs, device))
    layers, depths, st = gsmodel.render_weights_views(T(scene.pws, device), T(scene.alphas, device),
                                                      T(scene.scales, device), T(scene.rots, device), cams,
                                                      scene.K, return_state=True, **coef)
    check_views_forward(scene, cams, layers, depths, st, device, **coef)
    # without the state the views' compositing forms no contrib images; the layers are the same
    lay0, _ = gsmodel.render_weights_views(T(scene.pws, device), T(scene.alphas, device), T(scene.scales, device),
                                           T(scene.rots, device), cams, scene.K, **coef)
    same(N(lay0), N(layers), "layers without contrib")
    shs = T(scene.shs, device).requires_grad_(True)
    lay, mask = gsmodel.WeightGSFunctionTestViews.apply(T(scene.pws, device), shs, T(scene.alphas, device),
                                                         T(scene.scales, device), T(scene.rots, device), cams,
                                                         scene.K)
    rng = np.random.default_rng(7)
    dl = rng.normal(size=tuple(lay.shape)).astype(np.float32)
    lay.backward(T(dl, device))
    ref = np.zeros(scene.shs.shape)
    ref_abs = np.zeros(scene.shs.shape)
    for v, cam in enumerate(cams):
        o = O.forward_weight(scene, cam)
        same(N(mask[v]), o["mask"], f"view {v} mask")
        ob = O.backward_weight(scene, o, dl[v], replay=True)
        ref += ob["dshs"]
        ref_abs += ob["dshs_abs"]
    grad_close(N(shs.grad), ref, ref_abs * 1.0001 + 1e-12, "dshs")


@pytest.mark.parametrize("rank", [1, 2])
def test_views_low_rank(device, rank):
    print(f"case: multi-view low rank r={rank}, 4 views, 5000 Gaussians K=7 192x96", flush=True)
    scene = make_scene(5000, seed=40 + rank, extent=1.0, K=7, lowrank=True, rank=rank, P=7, Q=15)
    cams = orbit_cameras(4, 4.0, 192, 96)
    A_np = scene.A if rank > 1 else scene.A[:, :, 0]
    B_np = scene.B if rank > 1 else scene.B[:, 0, :]
    coef = dict(dc=T(scene.dc, device), A=T(A_np, device), B=T(B_np, device), Q=scene.Q)
    layers, depths, st = gsmodel.render_weights_views(T(scene.pws, device), T(scene.alphas, device),
                                                      T(scene.scales, device), T(scene.rots, device), cams,
                                                      scene.K, return_state=True, **coef)
    check_views_forward(scene, cams, layers, depths, st, device, **coef)
    dc = T(scene.dc, device).requires_grad_(True)
    A = T(A_np, device).requires_grad_(True)
    B = T(B_np, device).requires_grad_(True)
    lay, _ = gsmodel.WeightGSFunctionFastViews.apply(T(scene.pws, device), dc, A, B, T(scene.alphas, device),
                                                      T(scene.scales, device), T(scene.rots, device), cams,
                                                      scene.K, scene.Q)
    dl = np.random.default_rng(8).normal(size=tuple(lay.shape)).astype(np.float32)
    lay.backward(T(dl, device))
    acc = {k: 0.0 for k in ("ddc", "dA", "dB", "ddc_abs", "dA_abs", "dB_abs")}
    for v, cam in enumerate(cams):
        ob = O.backward_weight(scene, O.forward_weight(scene, cam), dl[v], replay=True)
        for k in acc:
            acc[k] = acc[k] + ob[k]
    grad_close(N(dc.grad), acc["ddc"], acc["ddc_abs"] * 1.0001 + 1e-12, "ddc")
    grad_close(N(A.grad).reshape(acc["dA"].shape), acc["dA"], acc["dA_abs"] * 1.0001 + 1e-12, "dA")
    grad_close(N(B.grad).reshape(acc["dB"].shape), acc["dB"], acc["dB_abs"] * 1.0001 + 1e-12, "dB")


def test_views_overflow_and_resolution_check(device):
    """A capacity far below the pair count overflows on some view, grows and
    reruns with the same results; cameras of two resolutions are refused."""
    from rgp import ops
    scene = make_scene(3000, seed=44, extent=1.0, K=5)
    cams = orbit_cameras(3, 4.0, 128, 96)
    key = (scene.n, 128, 96)
    ops.CAPACITY.cap[key] = 16
    layers, depths, st = gsmodel.render_weights_views(T(scene.pws, device), T(scene.alphas, device),
                                                      T(scene.scales, device), T(scene.rots, device), cams,
                                                      scene.K, shs=T(scene.shs, device), return_state=True)
    assert ops.CAPACITY.cap[key] > max(st["m"])
    for v, cam in enumerate(cams):
        same(N(layers[v]), O.forward_weight(scene, cam)["layer"], f"view {v} layer after regrowth")
    other = orbit_cameras(1, 4.0, 96, 96)[0]
    with pytest.raises(ValueError):
        gsmodel.render_weights_views(T(scene.pws, device), T(scene.alphas, device), T(scene.scales, device),
                                     T(scene.rots, device), [cams[0], other], scene.K, shs=T(scene.shs, device))
