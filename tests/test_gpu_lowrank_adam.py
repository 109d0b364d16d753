"""The low-rank factor update fused with its backward
(rgp_sh_backward_adam_lowrank, PaletteAdam's fused_lowrank).

* the kernel equals rgp_sh_backward_lowrank followed by rgp_adam on each of
  dc, A and B, bit for bit, over three steps, for rank 1 and 2, K = 5 and 7,
  with a partial last block of Gaussians and weight decay;
* PaletteAdam with the triple in fused_lowrank gives, after a
  WeightGSFunctionFast forward and backward, the same dc, A, B and moments
  as the gradients materialised from the same dL/dw followed by rgp_adam,
  also when A carries a second gradient the usual way;
* unsupported shapes are refused by the entry point.
"""
import numpy as np
import pytest
import torch

from rgp import gsmodel, ops
from rgp._lib import I64, RgpError, call, ptr, stream
from rgp.optim import PaletteAdam
from rgp.scene import make_scene, orbit_cameras

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("K,rank", [(5, 1), (5, 2), (7, 2)])
def test_kernel_equals_backward_then_adam(device, K, rank):
    print(f"case: fused low-rank adam kernel K={K} rank={rank}", flush=True)
    n = 4000 + 37
    rng = np.random.default_rng(K * 10 + rank)
    sc = make_scene(n, seed=3, extent=1.0, K=K, lowrank=True, rank=rank, P=K, Q=15)
    cam = orbit_cameras(3, 4.0, 96, 64)[1]
    camd = ops.camera_struct(cam.to_torch(device), device)
    pws = T(sc.pws, device)
    fused = [T(sc.dc, device), T(sc.A, device), T(sc.B, device)]
    ref = [t.clone() for t in fused]
    mf = [torch.zeros_like(t) for t in fused]
    vf = [torch.zeros_like(t) for t in fused]
    mr = [torch.zeros_like(t) for t in fused]
    vr = [torch.zeros_like(t) for t in fused]
    for step in range(1, 4):
        dw = T(rng.normal(0, 1e-2, (n, K)).astype(np.float32), device)
        g = [torch.empty_like(t) for t in ref]
        call("rgp_sh_backward_lowrank", n, K, K, 15, rank, ptr(camd), ptr(pws), ptr(ref[1]), ptr(ref[2]), ptr(dw),
             ptr(g[0]), ptr(g[1]), ptr(g[2]), stream(device))
        for t, gt, m, v in zip(ref, g, mr, vr):
            call("rgp_adam", I64(t.numel()), ptr(t), ptr(gt), ptr(m), ptr(v), I64(step), 3e-3, 0.9, 0.999, 1e-8,
                 1e-2, stream(device))
        call("rgp_sh_backward_adam_lowrank", n, K, K, 15, rank, ptr(camd), ptr(pws), ptr(dw), ptr(fused[0]),
             ptr(mf[0]), ptr(vf[0]), ptr(fused[1]), ptr(mf[1]), ptr(vf[1]), ptr(fused[2]), ptr(mf[2]), ptr(vf[2]),
             I64(step), 3e-3, 0.9, 0.999, 1e-8, 1e-2, stream(device))
    for name, a, b in (("dc", fused[0], ref[0]), ("A", fused[1], ref[1]), ("B", fused[2], ref[2]),
                       ("m_B", mf[2], mr[2]), ("v_A", vf[1], vr[1]), ("m_dc", mf[0], mr[0])):
        assert np.array_equal(N(a), N(b)), name
    assert not np.array_equal(N(fused[1]), sc.A)  # the update moved the factors


@pytest.mark.parametrize("views,K,rank", [(3, 5, 2), (4, 7, 2), (2, 5, 1)])
def test_views_kernel_equals_backward_then_adam(device, views, K, rank):
    print(f"case: fused low-rank adam, {views} views K={K} rank={rank}", flush=True)
    n = 3000 + 11
    rng = np.random.default_rng(views * 100 + K * 10 + rank)
    sc = make_scene(n, seed=6, extent=1.0, K=K, lowrank=True, rank=rank, P=K, Q=15)
    camsd = ops.camera_structs(orbit_cameras(views, 4.0, 96, 64), device)
    pws = T(sc.pws, device)
    fused = [T(sc.dc, device), T(sc.A, device), T(sc.B, device)]
    ref = [t.clone() for t in fused]
    mf, vf, mr, vr = ([torch.zeros_like(t) for t in fused] for _ in range(4))
    for step in range(1, 4):
        dw = T(rng.normal(0, 1e-2, (views, n, K)).astype(np.float32), device)
        g = [torch.empty_like(t) for t in ref]
        call("rgp_sh_backward_lowrank_views", views, n, K, K, 15, rank, ptr(camsd), ptr(pws), ptr(ref[1]),
             ptr(ref[2]), ptr(dw), ptr(g[0]), ptr(g[1]), ptr(g[2]), stream(device))
        for t, gt, m, v in zip(ref, g, mr, vr):
            call("rgp_adam", I64(t.numel()), ptr(t), ptr(gt), ptr(m), ptr(v), I64(step), 3e-3, 0.9, 0.999, 1e-8,
                 0.0, stream(device))
        call("rgp_sh_backward_adam_lowrank_views", views, n, K, K, 15, rank, ptr(camsd), ptr(pws), ptr(dw),
             ptr(fused[0]), ptr(mf[0]), ptr(vf[0]), ptr(fused[1]), ptr(mf[1]), ptr(vf[1]), ptr(fused[2]), ptr(mf[2]),
             ptr(vf[2]), I64(step), 3e-3, 0.9, 0.999, 1e-8, 0.0, stream(device))
    for name, a, b in (("dc", fused[0], ref[0]), ("A", fused[1], ref[1]), ("B", fused[2], ref[2]),
                       ("m_A", mf[1], mr[1]), ("v_B", vf[2], vr[2]), ("v_dc", vf[0], vr[0])):
        assert np.array_equal(N(a), N(b)), name


@pytest.mark.parametrize("second", [False, True])
def test_palette_adam_fused_lowrank(device, second):
    """Two steps of PaletteAdam with the triple in fused_lowrank, each
    against the unfused update of copies from the same dL/dw (the weight
    gradient itself is summed with float atomics, so it is taken from the
    run, not recomputed); second=True adds 0.5 |A|^2 to the loss, so A also
    gets a gradient the usual way and the step materialises the fused one."""
    print(f"case: PaletteAdam fused_lowrank, second gradient={second}", flush=True)
    K, rank, n = 5, 2, 3000
    sc = make_scene(n, seed=8, extent=1.0, K=K, lowrank=True, rank=rank, P=K, Q=15)
    cam = orbit_cameras(4, 4.0, 160, 128)[2]
    geo = (T(sc.alphas, device), T(sc.scales, device), T(sc.rots, device), cam.to_torch(device))
    pws = T(sc.pws, device)
    params = [T(x, device).requires_grad_(True) for x in (sc.dc, sc.A, sc.B)]
    opt = PaletteAdam(params, lr=3e-3, weight_decay=1e-3, fused_lowrank=[tuple(params)])
    mr = [torch.zeros_like(p) for p in params]
    vr = [torch.zeros_like(p) for p in params]
    dl = None
    for it in range(2):
        layer, _ = gsmodel.WeightGSFunctionFast.apply(pws, *params, *geo, K, 15)
        if dl is None:
            dl = T(np.random.default_rng(5).normal(size=tuple(layer.shape)).astype(np.float32), device)
        loss = (layer * dl).sum() + (0.5 * (params[1] ** 2).sum() if second else 0.0)
        opt.zero_grad()
        loss.backward()
        pend = opt._pending[id(params[0])]
        assert pend[0] == "lowrank"
        dw, camd = pend[3], pend[4]
        ref = [p.detach().clone() for p in params]
        g = [torch.empty_like(p) for p in ref]
        call("rgp_sh_backward_lowrank", n, K, K, 15, rank, ptr(camd), ptr(pws), ptr(ref[1]), ptr(ref[2]), ptr(dw),
             ptr(g[0]), ptr(g[1]), ptr(g[2]), stream(device))
        if second:
            g[1] = params[1].grad + g[1]
        for t, gt, m, v in zip(ref, g, mr, vr):
            call("rgp_adam", I64(t.numel()), ptr(t), ptr(gt), ptr(m), ptr(v), I64(it + 1), 3e-3, 0.9, 0.999, 1e-8,
                 1e-3, stream(device))
        opt.step()
        for i, what in enumerate(("dc", "A", "B")):
            assert np.array_equal(N(params[i]), N(ref[i])), (it, what)
            assert np.array_equal(N(opt.state[params[i]]["exp_avg"]), N(mr[i])), (it, what, "exp_avg")
            assert np.array_equal(N(opt.state[params[i]]["exp_avg_sq"]), N(vr[i])), (it, what, "exp_avg_sq")


@pytest.mark.parametrize("second,wd", [(False, 1e-3), (True, 1e-3), (False, 0.0)])
def test_palette_adam_fused_lowrank_views(device, second, wd):
    """The multi-view path through Python: WeightGSFunctionFastViews over 3
    views hands the (views, N, K) weight gradient to PaletteAdam, whose step
    runs rgp_sh_backward_adam_lowrank_views; two steps against copies updated
    by the view-summed backward (rgp_sh_backward_lowrank_views) then rgp_adam,
    bit for bit, params and both moments.  second=True adds 0.5 |B|^2, so B
    also gets a gradient the usual way and the step materialises the fused
    one (the V > 1 branch)."""
    print(f"case: PaletteAdam fused_lowrank over 3 views, second={second} wd={wd}", flush=True)
    K, rank, n, V = 5, 2, 3000, 3
    sc = make_scene(n, seed=18, extent=1.0, K=K, lowrank=True, rank=rank, P=K, Q=15)
    cams = orbit_cameras(6, 4.0, 160, 128)[1:1 + V]
    geo = (T(sc.alphas, device), T(sc.scales, device), T(sc.rots, device))
    pws = T(sc.pws, device)
    params = [T(x, device).requires_grad_(True) for x in (sc.dc, sc.A, sc.B)]
    opt = PaletteAdam(params, lr=3e-3, weight_decay=wd, fused_lowrank=[tuple(params)])
    mr, vr = [torch.zeros_like(p) for p in params], [torch.zeros_like(p) for p in params]
    dl = None
    for it in range(2):
        layers, _ = gsmodel.WeightGSFunctionFastViews.apply(pws, *params, *geo, cams, K, 15)
        if dl is None:
            dl = T(np.random.default_rng(6).normal(size=tuple(layers.shape)).astype(np.float32), device)
        loss = (layers * dl).sum() + (0.5 * (params[2] ** 2).sum() if second else 0.0)
        opt.zero_grad()
        loss.backward()
        pend = opt._pending[id(params[0])]
        assert pend[0] == "lowrank" and pend[-1] == V
        dw, camsd = pend[3], pend[4]
        ref = [p.detach().clone() for p in params]
        g = [torch.empty_like(p) for p in ref]
        call("rgp_sh_backward_lowrank_views", V, n, K, K, 15, rank, ptr(camsd), ptr(pws), ptr(ref[1]), ptr(ref[2]),
             ptr(dw), ptr(g[0]), ptr(g[1]), ptr(g[2]), stream(device))
        if second:
            g[2] = params[2].grad + g[2]
        for t, gt, m, v in zip(ref, g, mr, vr):
            call("rgp_adam", I64(t.numel()), ptr(t), ptr(gt), ptr(m), ptr(v), I64(it + 1), 3e-3, 0.9, 0.999, 1e-8,
                 wd, stream(device))
        opt.step()
        for i, what in enumerate(("dc", "A", "B")):
            assert np.array_equal(N(params[i]), N(ref[i])), (it, what)
            assert np.array_equal(N(opt.state[params[i]]["exp_avg"]), N(mr[i])), (it, what, "exp_avg")
            assert np.array_equal(N(opt.state[params[i]]["exp_avg_sq"]), N(vr[i])), (it, what, "exp_avg_sq")


def test_palette_adam_fused_lowrank_unequal_steps(device):
    """A step that updates A alone (a regulariser without a render) leaves the
    triple with different step counts; the next fused step then falls back to
    the materialised gradients and each tensor's own bias correction, equal
    to the unfused update bit for bit."""
    print("case: PaletteAdam fused_lowrank with unequal step counts", flush=True)
    K, rank, n = 5, 2, 2500
    sc = make_scene(n, seed=28, extent=1.0, K=K, lowrank=True, rank=rank, P=K, Q=15)
    cam = orbit_cameras(4, 4.0, 160, 128)[3]
    geo = (T(sc.alphas, device), T(sc.scales, device), T(sc.rots, device), cam.to_torch(device))
    pws = T(sc.pws, device)
    params = [T(x, device).requires_grad_(True) for x in (sc.dc, sc.A, sc.B)]
    opt = PaletteAdam(params, lr=3e-3, fused_lowrank=[tuple(params)])
    ref = [p.detach().clone() for p in params]
    mr, vr = [torch.zeros_like(p) for p in params], [torch.zeros_like(p) for p in params]
    opt.zero_grad()
    (0.5 * (params[1] ** 2).sum()).backward()
    opt.step()
    call("rgp_adam", I64(ref[1].numel()), ptr(ref[1]), ptr(ref[1].clone()), ptr(mr[1]), ptr(vr[1]), I64(1), 3e-3,
         0.9, 0.999, 1e-8, 0.0, stream(device))
    assert [opt.state[p].get("step", 0) for p in params] == [0, 1, 0]
    layer, _ = gsmodel.WeightGSFunctionFast.apply(pws, *params, *geo, K, 15)
    dl = T(np.random.default_rng(8).normal(size=tuple(layer.shape)).astype(np.float32), device)
    opt.zero_grad()
    (layer * dl).sum().backward()
    dw, camd = opt._pending[id(params[0])][3], opt._pending[id(params[0])][4]
    g = [torch.empty_like(p) for p in ref]
    call("rgp_sh_backward_lowrank", n, K, K, 15, rank, ptr(camd), ptr(pws), ptr(ref[1]), ptr(ref[2]), ptr(dw),
         ptr(g[0]), ptr(g[1]), ptr(g[2]), stream(device))
    for t, gt, m, v, step in zip(ref, g, mr, vr, (1, 2, 1)):
        call("rgp_adam", I64(t.numel()), ptr(t), ptr(gt), ptr(m), ptr(v), I64(step), 3e-3, 0.9, 0.999, 1e-8, 0.0,
             stream(device))
    opt.step()
    assert [opt.state[p]["step"] for p in params] == [1, 2, 1]
    for i, what in enumerate(("dc", "A", "B")):
        assert np.array_equal(N(params[i]), N(ref[i])), what
        assert np.array_equal(N(opt.state[params[i]]["exp_avg"]), N(mr[i])), (what, "exp_avg")
        assert np.array_equal(N(opt.state[params[i]]["exp_avg_sq"]), N(vr[i])), (what, "exp_avg_sq")


def test_unsupported_shapes_refused(device):
    z = torch.zeros(64, dtype=torch.float32, device=device)
    with pytest.raises(RgpError):
        # Q = 9: only the Q = 15, P = K register path is fused
        call("rgp_sh_backward_adam_lowrank", 4, 5, 5, 9, 2, None, *([ptr(z)] * 11), I64(1), 1e-3, 0.9, 0.999, 1e-8,
             0.0, stream(device))
