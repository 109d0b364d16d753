"""Parity of the gfx950 kernels against the CPU oracle, through the
gsplatcu-compatible module and the fused autograd wrappers.

Bar: bit-exact for every integer output (areas, tile lists, sort order,
ranges, contrib) and for every FP32 output of the forward (the kernels and
the FP32 oracle perform the same IEEE operations in the same order).  The
backward sums per-pixel terms with atomics in an order that varies from run
to run; each gradient must lie within 1e-5 * sum|terms| of the exact sum
(the oracle sums the same FP32 terms in float64).
"""
import numpy as np
import pytest
import torch

import gsplatcu as gsc
from oracle import oracle as O
from rgp import gsmodel, ops
from rgp.scene import make_scene, orbit_cameras, workload_c0

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
    assert not bad.any(), f"{what}: {bad.sum()} of {a.size} differ; first at {np.argwhere(bad)[:3].tolist()}"


def grad_close(g, ref, ref_abs, what):
    err = np.abs(g.astype(np.float64) - ref)
    lim = GRAD_TOL * ref_abs + 1e-30
    assert (err <= lim).all(), f"{what}: max err/lim {np.max(err / lim):.3g}"


@pytest.fixture(scope="module")
def c0():
    print("case: C0 20k Gaussians 256x256 K=5", flush=True)
    return workload_c0()


def gpu_chain(scene, cam, dev, fast=False):
    """The reference chain of WeightGSFunctionTest/Fast.forward on gsplatcu."""
    tc = cam.to_torch(dev)
    pws = T(scene.pws, dev)
    us, pcs, depths = gsc.project(pws, tc.Rcw, tc.tcw, cam.fx, cam.fy, cam.cx, cam.cy, False)
    c3, = gsc.computeCov3D(T(scene.rots, dev), T(scene.scales, dev), depths, False)
    c2, = gsc.computeCov2D(c3, pcs, tc.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height, False)
    if fast:
        w, dwdc, ylm = gsc.sh2Weight_fast(T(scene.dc, dev), T(scene.A, dev), T(scene.B, dev), pws, tc.twc, True,
                                          scene.K, scene.Q)
        extra = dict(ylm=ylm, dwdc=dwdc)
    else:
        w, basis, _ = gsc.sh2Weight(T(scene.shs, dev), pws, tc.twc, True, scene.K)
        extra = dict(basis=basis)
    cinv, areas = gsc.inverseCov2D(c2, depths, False)
    layer, contrib, final_tau, ranges, ids = gsc.splat_weight(cam.height, cam.width, scene.K, us, cinv,
                                                              T(scene.alphas, dev), depths, w, areas)
    return dict(us=us, pcs=pcs, depths=depths, cov3ds=c3, cov2ds=c2, weights=w, cinv=cinv, areas=areas,
                layer=layer, contrib=contrib, final_tau=final_tau, ranges=ranges, ids=ids, **extra)


def check_forward(g, o, fast):
    for k in ("us", "pcs", "cov3ds", "cov2ds", "weights", "cinv", "areas", "depths", "ranges", "ids", "layer",
              "contrib", "final_tau"):
        same(N(g[k]), o[k], k)
    if fast:
        same(N(g["ylm"]), o["ylm"], "ylm")
    else:
        same(N(g["basis"]).reshape(o["basis"].shape), o["basis"], "basis")


def test_forward_chain_bitexact_c0(device, c0):
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    g = gpu_chain(scene, cam, device)
    check_forward(g, o, fast=False)
    assert len(o["ids"]) > 10000


def test_backward_chain_c0(device, c0):
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    g = gpu_chain(scene, cam, device)
    dl = np.random.default_rng(1).normal(size=o["layer"].shape).astype(np.float32)
    ob = O.backward_weight(scene, o, dl)
    dw, = gsc.splat_weight_B(cam.height, cam.width, scene.K, g["us"], g["cinv"], T(scene.alphas, device),
                             g["depths"], g["weights"], g["contrib"], g["final_tau"], g["ranges"], g["ids"],
                             T(dl, device))
    assert dw.shape == (scene.n, 1, scene.K)
    grad_close(N(dw).reshape(scene.n, scene.K), ob["dweights"], ob["dweights_abs"], "dweights")


@pytest.mark.parametrize("mode", ["replay", "reference"])
def test_fused_wrapper_full_rank_c0(device, c0, mode):
    scene, cam = c0
    gsmodel.BACKWARD_MODE = mode
    o = O.forward_weight(scene, cam)
    tc = cam.to_torch(device)
    shs = T(scene.shs, device).requires_grad_(True)
    layer, mask = gsmodel.WeightGSFunctionTest.apply(T(scene.pws, device), shs, T(scene.alphas, device),
                                                     T(scene.scales, device), T(scene.rots, device), tc, scene.K)
    same(N(layer), o["layer"], "layer")
    same(N(mask), o["mask"], "mask")
    dl = np.random.default_rng(2).normal(size=o["layer"].shape).astype(np.float32)
    layer.backward(T(dl, device))
    gsmodel.BACKWARD_MODE = "replay"
    ob = O.backward_weight(scene, o, dl, replay=(mode == "replay"))
    grad_close(N(shs.grad), ob["dshs"], ob["dshs_abs"], "dshs")


def test_fused_forward_unaligned_coefficients(device, c0):
    """Coefficient rows that are not 16-byte aligned are read per lane from
    global memory instead of through the LDS staging; same bits."""
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    flat = torch.empty(scene.shs.size + 1, dtype=torch.float32, device=device)
    shs = flat[1:].view(scene.shs.shape)
    shs.copy_(T(scene.shs, device))
    assert shs.is_contiguous() and shs.data_ptr() % 16 != 0
    layer, mask = gsmodel.WeightGSFunctionTest.apply(T(scene.pws, device), shs, T(scene.alphas, device),
                                                     T(scene.scales, device), T(scene.rots, device),
                                                     cam.to_torch(device), scene.K)
    same(N(layer), o["layer"], "layer")
    same(N(mask), o["mask"], "mask")


@pytest.mark.parametrize("rank", [1, 2])
def test_fused_wrapper_low_rank(device, rank):
    print(f"case: low rank r={rank} 4000 Gaussians 192x160", flush=True)
    scene = make_scene(4000, seed=3 + rank, extent=1.0, K=5, lowrank=True, rank=rank, P=5, Q=15)
    cam = orbit_cameras(3, 4.0, 192, 160)[1]
    o = O.forward_weight(scene, cam)
    g = gpu_chain(scene, cam, device, fast=True)
    check_forward(g, o, fast=True)
    tc = cam.to_torch(device)
    dc = T(scene.dc, device).requires_grad_(True)
    A = T(scene.A if rank > 1 else scene.A[:, :, 0], device).requires_grad_(True)
    B = T(scene.B if rank > 1 else scene.B[:, 0, :], device).requires_grad_(True)
    layer, mask = gsmodel.WeightGSFunctionFast.apply(T(scene.pws, device), dc, A, B, T(scene.alphas, device),
                                                     T(scene.scales, device), T(scene.rots, device), tc,
                                                     scene.K, scene.Q)
    same(N(layer), o["layer"], "layer")
    dl = np.random.default_rng(5).normal(size=o["layer"].shape).astype(np.float32)
    layer.backward(T(dl, device))
    ob = O.backward_weight(scene, o, dl, replay=True)
    grad_close(N(dc.grad), ob["ddc"], ob["ddc_abs"], "ddc")
    grad_close(N(A.grad).reshape(ob["dA"].shape), ob["dA"], ob["dA_abs"] * 1.0001 + 1e-12, "dA")
    grad_close(N(B.grad).reshape(ob["dB"].shape), ob["dB"], ob["dB_abs"] * 1.0001 + 1e-12, "dB")


def test_lowrank_backward_op_bitexact(device):
    """sh2Weight_fast_B sums in loop order, like the reference's thread-private
    atomics: bit-exact against the FP32 oracle."""
    scene = make_scene(3000, seed=9, K=5, lowrank=True, rank=2, P=5, Q=15)
    cam = orbit_cameras(1, 4.0, 64, 64)[0]
    w, ylm = O.sh2weight_fast(scene.dc, scene.A, scene.B, scene.pws, cam.twc, 5, 15)
    dw = np.random.default_rng(0).normal(size=(scene.n, 5)).astype(np.float32)
    oA, oB = O.sh2weight_fast_backward(scene.A, scene.B, ylm, dw, 5, 15)
    gA, gB = gsc.sh2Weight_fast_B(T(scene.A, device), T(scene.B, device), T(ylm, device), T(dw, device), 5, 15)
    same(N(gA), oA, "dA")
    same(N(gB), oB, "dB")


def test_colour_path_bitexact(device):
    scene = make_scene(5000, seed=11, K=5)
    cam = orbit_cameras(2, 4.0, 200, 120)[0]
    shs = np.random.default_rng(3).normal(0, 0.3, (scene.n, 48)).astype(np.float32)
    img = gsmodel.render_vanilla(T(scene.pws, device), T(shs, device), T(scene.alphas, device),
                                 T(scene.scales, device), T(scene.rots, device), cam.to_torch(device))
    us, pcs, depths = O.project(scene.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    c3 = O.cov3d(scene.rots, scene.scales, depths)
    c2 = O.cov2d(c3, pcs, cam.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height)
    col = O.sh2color(shs, scene.pws, cam.twc)
    cinv, areas = O.inverse_cov2d(c2, depths)
    rects, counts = O.get_rects(us, areas, depths, cam.width, cam.height)
    keys, ids, ranges = O.bin_sort(depths, rects, counts, cam.width, cam.height)
    ref, _, _, _, _ = O.draw(cam.width, cam.height, ranges, ids, us, cinv, scene.alphas, col, 1.0)
    same(N(img), ref, "image")


def strips_written(ranges, contrib, final_tau, width, height):
    """The list positions whose strip masks the forward writes: every entry of
    each batch of 64 a tile stages, i.e. while some pixel of the tile is live
    at the batch start (a pixel stops at its contrib when final_tau < 1e-4,
    else runs through the list; pixels outside the image never start)."""
    gx = (width + 15) // 16
    keep = np.zeros(int(ranges[:, 1].max(initial=0)), bool)
    for t in np.flatnonzero(ranges[:, 1] > ranges[:, 0]):
        a, b = ranges[t]
        ty, tx = divmod(int(t), gx)
        c = contrib[16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16]
        f = final_tau[16 * ty:16 * ty + 16, 16 * tx:16 * tx + 16]
        stop = np.where(f < np.float32(1e-4), c, b - a).max()
        keep[a:a + min(b - a, 64 * ((int(stop) + 63) // 64))] = True
    return keep


def check_strips(st, o, sc, cam):
    _, _, _, _, _, ref = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"], sc.alphas,
                                o["weights"], strips=True)
    keep = strips_written(o["ranges"], o["contrib"], o["final_tau"], cam.width, cam.height)
    got = N(st["strips"][:len(o["ids"])])
    same(got[keep], ref[keep], "strips")
    assert (ref[~keep] == 0).all()  # nothing contributes after every pixel has stopped
    return int((ref != 0).sum())


def test_forward_strip_masks_c0(device, c0):
    """The forward's per-entry masks of row strips with a contribution, bit
    for bit against the oracle; the replay backward with them equals the
    oracle's gradient, as without them."""
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    tc = cam.to_torch(device)
    args = [T(x, device) for x in (scene.pws, scene.alphas, scene.scales, scene.rots)]
    layer, depths, st = gsmodel.render_weights(*args, tc, scene.K, shs=T(scene.shs, device), return_state=True,
                                               strips=True)
    same(N(layer), o["layer"], "layer")
    assert check_strips(st, o, scene, cam) > 1000
    dl = np.random.default_rng(3).normal(size=o["layer"].shape).astype(np.float32)
    ob = O.backward_weight(scene, o, dl, replay=True)
    for strips in (st["strips"], None):
        dw = ops.draw_backward_replay(cam.width, cam.height, scene.K, st["ranges"], st["ids"], st["records"],
                                      T(dl, device), scene.n, strips)
        grad_close(N(dw), ob["dweights"], ob["dweights_abs"], f"dweights strips={strips is not None}")


def test_forward_without_contrib_c0(device, c0):
    """The training forward (strip masks, no contrib image) and the frame
    forward (images only) write the same layer and final_tau as the forward
    that also forms contrib (the training forward the same masks too), and
    the autograd wrapper's gradient in replay mode equals the oracle's."""
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    tc = cam.to_torch(device)
    args = [T(x, device) for x in (scene.pws, scene.alphas, scene.scales, scene.rots)]
    shs = T(scene.shs, device)
    _, _, a = gsmodel.render_weights(*args, tc, scene.K, shs=shs, return_state=True, strips=True)
    layer, _, b = gsmodel.render_weights(*args, tc, scene.K, shs=shs, return_state=True, strips=True, contrib=False)
    assert b["contrib"] is None
    same(N(layer), o["layer"], "layer")
    same(N(b["final_tau"]), N(a["final_tau"]), "final_tau")
    assert check_strips(b, o, scene, cam) > 1000
    l2, c2, f2 = ops.draw_forward(cam.width, cam.height, scene.K, a["ranges"], a["ids"], a["records"], 0.0, None,
                                  False)
    assert c2 is None
    same(N(l2), o["layer"], "frame layer")
    same(N(f2), N(a["final_tau"]), "frame final_tau")
    same(N(gsmodel.render_weights(*args, tc, scene.K, shs=shs)[0]), o["layer"], "render_weights layer")
    x = shs.clone().requires_grad_(True)
    out, _ = gsmodel.WeightGSFunctionTest.apply(*args[:1], x, *args[1:], tc, scene.K)
    dl = np.random.default_rng(5).normal(size=o["layer"].shape).astype(np.float32)
    out.backward(T(dl, device))
    ob = O.backward_weight(scene, o, dl, replay=True)
    grad_close(N(x.grad), ob["dshs"], ob["dshs_abs"], "dshs")


def test_mix_matches_bake(device, c0):
    """mix(render(w), P) against the oracle mix, and the linear identity with
    the bake-then-render colour path (gsmodel.py:488-509, 13-26)."""
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    pal = scene.palette
    rgb = ops.mix(T(o["layer"], device), T(pal, device))
    same(N(rgb), O.mix(pal, o["layer"]), "mix")
    # bake: colour SH = weight SH @ P, rendered by the colour path
    shs_w = scene.shs.reshape(scene.n, 16, scene.K)
    baked = np.einsum("nlk,kc->nlc", shs_w.astype(np.float64), pal.astype(np.float64)).astype(np.float32)
    img = gsmodel.render_vanilla(T(scene.pws, device), T(baked.reshape(scene.n, 48), device),
                                 T(scene.alphas, device), T(scene.scales, device), T(scene.rots, device),
                                 cam.to_torch(device))
    # Equal in exact arithmetic (both paths see the same alpha' and tau per entry; empty tiles give 0 in
    # both).  Bound: 1e-5 of sum_k |P_kc| sum_i tau_i alpha'_i sum_lm |Y_lm| |w_lm,k|, the magnitude of every
    # term either side sums (rendered by the oracle with per-Gaussian weights sum_lm |Y_lm||w_lm,k|).
    wabs = np.einsum("nl,nlk->nk", np.abs(o["basis"].astype(np.float64)),
                     np.abs(shs_w.astype(np.float64))).astype(np.float32)
    MW, _, _, _, _ = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"], scene.alphas, wabs)
    M = np.einsum("kc,khw->chw", np.abs(pal.astype(np.float64)), MW.astype(np.float64))
    diff = np.abs(N(img).astype(np.float64) - N(rgb))
    assert (diff <= 1e-5 * M).all(), float(np.max(diff / np.maximum(M, 1e-300)))
    assert M.max() > 0.5


@pytest.mark.parametrize("shape,w_off,rgb_off", [((5, 48, 64), 0, 0), ((5, 47, 63), 0, 0), ((5, 48, 64), 1, 0),
                                                  ((5, 48, 64), 0, 3), ((7, 47, 63), 1, 2), ((5, 1080, 1920), 0, 0)])
def test_mix_forward(device, shape, w_off, rgb_off):
    """rgp_mix and rgp_mix_batched bit for bit against the oracle's mix
    (gsmodel.py:488-509 semantics), with W and rgb at 4-byte offsets and odd
    H*W: the float4 path only when H*W % 4 == 0 and both planes are 16-byte
    aligned, the scalar one otherwise."""
    from rgp._lib import I64, call, ptr, stream
    print(f"case: mix forward {shape} W offset {w_off} rgb offset {rgb_off}", flush=True)
    K, H, W = shape
    rng = np.random.default_rng(5 + w_off + 7 * rgb_off)
    Wn = rng.random(shape, dtype=np.float32)
    P = rng.random((K, 3), dtype=np.float32)
    ref = O.mix(P, Wn)
    Wflat = torch.empty(Wn.size + w_off, dtype=torch.float32, device=device)
    Wt = Wflat[w_off:].view(shape)
    Wt.copy_(T(Wn, device))
    outs = []
    for _ in range(2):
        oflat = torch.full((3 * H * W + rgb_off,), float("nan"), dtype=torch.float32, device=device)
        outs.append(oflat[rgb_off:].view(3, H, W))
    Pt = T(P, device)
    call("rgp_mix", K, I64(H * W), ptr(Wt), ptr(Pt), ptr(outs[0]), stream(device))
    # the batched form: the same image twice (second output), plus a 1-pixel image
    tiny_w = T(Wn[:, :1, :1], device)
    tiny_o = torch.empty((3, 1, 1), dtype=torch.float32, device=device)
    args = torch.tensor([Wt.data_ptr(), tiny_w.data_ptr(), outs[1].data_ptr(), tiny_o.data_ptr(), H * W, 1],
                        dtype=torch.int64, device=device)
    pals = T(np.concatenate([P, P]), device)
    base = args.data_ptr()
    call("rgp_mix_batched", 2, K, base, base + 16, base + 32, ptr(pals), I64(H * W), stream(device))
    torch.cuda.synchronize()
    same(N(outs[0]), ref, "rgp_mix")
    same(N(outs[1]), ref, "rgp_mix_batched")
    same(N(tiny_o), O.mix(P, Wn[:, :1, :1].copy()), "rgp_mix_batched 1 pixel")


@pytest.mark.parametrize("shape,offset", [((5, 48, 64), 0), ((5, 47, 63), 0), ((3, 48, 64), 1),
                                          ((5, 1080, 1920), 0)])
def test_mix_backward(device, shape, offset):
    """Mix.backward: dL/dW_k = (P_k0 d_0 + P_k1 d_1) + P_k2 d_2 bit for bit
    against the same float32 operations in torch (the float4 path when
    H*W % 4 == 0 and the planes are 16-byte aligned, else the scalar one:
    odd sizes, and W read at a 4-byte offset), and dL/dP_kc = sum_p W_k d_c
    within 1e-5 of sum_p |W_k d_c| against float64."""
    print(f"case: mix backward {shape} offset {offset}", flush=True)
    K, H, W = shape
    rng = np.random.default_rng(11)
    Wn = rng.random(shape, dtype=np.float32)
    P = rng.random((K, 3), dtype=np.float32)
    d = rng.normal(size=(3, H, W)).astype(np.float32)
    if offset:
        flat = torch.empty(Wn.size + offset, dtype=torch.float32, device=device)
        Wt = flat[offset:].view(shape)
        Wt.copy_(T(Wn, device))
    else:
        Wt = T(Wn, device)
    Wt.requires_grad_(True)
    Pt = T(P, device).requires_grad_(True)
    rgb = ops.Mix.apply(Wt, Pt)
    rgb.backward(T(d, device))
    dt, Pc = T(d, device), T(P, device)
    ref_dW = (Pc[:, 0, None, None] * dt[0] + Pc[:, 1, None, None] * dt[1]) + Pc[:, 2, None, None] * dt[2]
    same(N(Wt.grad), N(ref_dW), "dW")
    terms = Wn.reshape(K, 1, -1).astype(np.float64) * d.reshape(1, 3, -1).astype(np.float64)
    ref_dP, ref_abs = terms.sum(-1), np.abs(terms).sum(-1)
    err = np.abs(N(Pt.grad).astype(np.float64) - ref_dP)
    assert (err <= 1e-5 * ref_abs).all(), float(np.max(err / ref_abs))


def test_binning_many_tiles(device):
    """1080p with a dense scene: two tile bytes and two depth bytes in the
    sort, hundreds of sort blocks."""
    print("case: 200k Gaussians 1920x1080", flush=True)
    scene = make_scene(200_000, seed=21, extent=3.0, K=5)
    cam = orbit_cameras(4, 6.0, 1920, 1080)[1]
    o = O.forward_weight(scene, cam)
    g = gpu_chain(scene, cam, device)
    for k in ("depths", "areas", "ranges", "ids", "contrib", "final_tau", "layer"):
        same(N(g[k]), o[k], k)
    assert len(o["ids"]) > 4096 * 50


def test_edge_cases(device):
    """M = 0 (everything behind the camera), M = 1 (the reference leaves
    every range (0,0)), N = 1, and a Gaussian over the whole image."""
    cam = orbit_cameras(1, 4.0, 96, 80)[0]
    tc = cam.to_torch(device)
    for n, tweak in ((5, "behind"), (1, "single"), (3, "huge")):
        print(f"case: edge {tweak}", flush=True)
        sc = make_scene(n, seed=31, K=3)
        if tweak == "behind":
            sc.pws[:] = cam.twc - 2.0 * (cam.Rcw[2] * 1.0)  # 2 units behind the camera
        if tweak == "single":
            # a tiny splat in the middle of one tile
            # pixel (40, 40): its 2-pixel box stays inside tile (2, 2)
            pc = np.array([(40.0 - cam.cx) * 4.0 / cam.fx, (40.0 - cam.cy) * 4.0 / cam.fy, 4.0])
            sc.pws[:] = (np.linalg.inv(cam.Rcw.astype(np.float64)) @ (pc - cam.tcw)).astype(np.float32)
            sc.scales[:] = 1e-4
        if tweak == "huge":
            sc.scales[:] = 2.0
        o = O.forward_weight(sc, cam)
        g = gpu_chain(sc, cam, device)
        for k in ("depths", "areas", "ranges", "ids", "contrib", "final_tau", "layer"):
            same(N(g[k]), o[k], f"{tweak}:{k}")
        if tweak == "behind":
            assert len(o["ids"]) == 0
        if tweak == "single":
            assert len(o["ids"]) == 1 and (o["ranges"] == 0).all()


@pytest.mark.parametrize("mode", ["replay", "reference"])
def test_extreme_opacities(device, mode):
    """Opacities far outside (0, 1): alpha up to 3e38, inf and NaN give
    cutoffs above 170 (or NaN), where exp(-m/2) is taken by the range-checked
    path down to denormal results; alpha 1e20 stays on the short path with
    distances near 100.  Forward bit-exact, both backward forms within the
    gradient bar."""
    print(f"case: extreme opacities, {mode} backward, 3000 Gaussians 128x96", flush=True)
    scene = make_scene(3000, seed=17, extent=1.0, K=5)
    cam = orbit_cameras(1, 4.0, 128, 96)[0]
    special = np.array([5e35, 3e38, np.inf, np.nan, 1e20, 0.0, -0.5], dtype=np.float32)
    scene.alphas[::7] = special[np.arange(len(scene.alphas[::7])) % len(special)]
    o = O.forward_weight(scene, cam)
    g = gpu_chain(scene, cam, device)
    for k in ("depths", "areas", "ranges", "ids", "contrib", "final_tau", "layer"):
        same(N(g[k]), o[k], k)
    dl = np.random.default_rng(4).normal(size=o["layer"].shape).astype(np.float32)
    if mode == "reference":
        ob = O.backward_weight(scene, o, dl)
        dw, = gsc.splat_weight_B(cam.height, cam.width, scene.K, g["us"], g["cinv"], T(scene.alphas, device),
                                 g["depths"], g["weights"], g["contrib"], g["final_tau"], g["ranges"], g["ids"],
                                 T(dl, device))
        grad_close(N(dw).reshape(scene.n, scene.K), ob["dweights"], ob["dweights_abs"], "dweights")
    else:
        ob = O.backward_weight(scene, o, dl, replay=True)
        shs = T(scene.shs, device).requires_grad_(True)
        layer, _ = gsmodel.WeightGSFunctionTest.apply(T(scene.pws, device), shs, T(scene.alphas, device),
                                                      T(scene.scales, device), T(scene.rots, device),
                                                      cam.to_torch(device), scene.K)
        same(N(layer), o["layer"], "layer")
        layer.backward(T(dl, device))
        grad_close(N(shs.grad), ob["dshs"], ob["dshs_abs"], "dshs")


def test_forward_counts_match_oracle(device, c0):
    """The instrumented forward counts exactly the evaluations and
    contributions of the oracle's per-pixel loop (the roofline's work)."""
    scene, cam = c0
    o = O.forward_weight(scene, cam)
    layer, depths, st = gsmodel.render_weights(T(scene.pws, device), T(scene.alphas, device), T(scene.scales, device),
                                               T(scene.rots, device), cam.to_torch(device), scene.K,
                                               shs=T(scene.shs, device), return_state=True)
    lay, contrib, ft, e, c = ops.draw_forward_counts(cam.width, cam.height, scene.K, st["ranges"], st["ids"],
                                                     st["records"], 0.0)
    same(N(lay), o["layer"], "layer")
    assert (e, c) == (o["n_eval"], o["n_contrib"])


def test_deferred_overflow_check_matches_exact(device, c0):
    """exact=False launches sort and draw before reading the overflow flag;
    a capacity far below M must overflow, grow and give the exact-mode
    results (same M, ids[:M], ranges and images)."""
    scene, cam = c0
    args = (T(scene.pws, device), T(scene.alphas, device), T(scene.scales, device), T(scene.rots, device),
            cam.to_torch(device), scene.K)
    kw = dict(shs=T(scene.shs, device), return_state=True)
    la, da, sa = gsmodel.render_weights(*args, exact=True, **kw)
    key = (scene.n, cam.width, cam.height)
    ops.CAPACITY.cap[key] = 1024
    lb, db, sb = gsmodel.render_weights(*args, exact=False, **kw)
    m = sa["m"]
    assert m > 4096 and sb["m"] == m and ops.CAPACITY.cap[key] > m
    same(N(lb), N(la), "layer")
    same(N(db), N(da), "depths")
    for k in ("contrib", "final_tau", "ranges"):
        same(N(sb[k]), N(sa[k]), k)
    same(N(sb["ids"][:m]), N(sa["ids"][:m]), "ids")


def gpu_bin(us, areas, depths, W, H, device, cap):
    """getRects -> scan -> sort on the GPU from given us / areas / depths."""
    from rgp._lib import I64, call, ptr, stream
    n = len(depths)
    ws = ops.new_workspace(n, cap, W, H, device)
    a, d = T(areas, device), T(depths, device)
    call("rgp_bin_rects", n, ptr(T(us, device)), ptr(a), ptr(d), W, H, ptr(ws), I64(cap), stream(device))
    call("rgp_bin_scan", n, W, H, ptr(ws), I64(cap), stream(device))
    ids, ranges, keys = ops.bin_sort(n, d, W, H, ws, cap, device, keys=True)
    m = int(ops.header(ws, 1).item())
    return N(keys[:m]), N(ids[:m]), N(ranges), m


@pytest.mark.parametrize("case", ["long_lists_ties", "mid_lists", "two_runs", "4k", "8k_radix", "huge_1080p",
                                  "huge_4k", "far_depths"])
def test_binning_sort_paths(device, case):
    """Keys, ids and ranges against the oracle's stable sort: tile lists far
    longer than one LDS sort (merged through global memory) with many equal
    quantised depths, lists of two to four register runs of 1024 keys
    merged in LDS (2049-4096 entries; 1018-1408 entries, on both sides of
    the one-wavefront limit), a 3840x2160 grid (32,400 tiles, the largest
    grid of the tile-bucket path) and a 7680x4320 grid (the LSD radix
    path); splats whose rects reach more than 32 super-tiles (walked by a
    whole wavefront in the coarse passes) at 1080p (4 x 4-tile super-tiles)
    and 4K (8 x 8); lists where some keys' quantised depths reach
    0x7FEFFFFF and above (the in-register sort's integer form, next to runs
    sorted in the double form)."""
    print(f"case: binning {case}", flush=True)
    rng = np.random.default_rng(51)
    if case in ("long_lists_ties", "mid_lists", "two_runs", "far_depths"):
        W, H, n = 64, 48, {"long_lists_ties": 12_000, "mid_lists": 3_300, "two_runs": 1_500, "far_depths": 3_300}[case]
        us = rng.uniform(-10, 74, (n, 2)).astype(np.float32)
        areas = rng.integers(30, 60, (n, 2)).astype(np.int32)
        depths = (1.0 + 0.5 * rng.integers(0, 5, n)).astype(np.float32)
        if case == "far_depths":
            # quantised depths just below 0x7FEFFFFF (0x7FEFFF80), just above (0x7FF00380) and up to 4e9
            far = rng.choice(n, 12, replace=False)
            depths[far[:2]] = np.float32(2146435.0)
            depths[far[2:4]] = np.float32(2146436.0)
            depths[far[4:]] = rng.uniform(2.2e6, 4.0e6, 8).astype(np.float32)
    elif case.startswith("huge"):
        W, H = (1920, 1080) if case == "huge_1080p" else (3840, 2160)
        n = 20_000
        us = np.stack([rng.uniform(-50, W + 50, n), rng.uniform(-50, H + 50, n)], 1).astype(np.float32)
        areas = rng.integers(1, 40, (n, 2)).astype(np.int32)
        big = rng.choice(n, 40, replace=False)
        areas[big] = rng.integers(600, 1600, (40, 2))
        depths = rng.uniform(0.1, 80.0, n).astype(np.float32)
    else:
        W, H = (3840, 2160) if case == "4k" else (7680, 4320)
        n = 100_000
        us = np.stack([rng.uniform(-50, W + 50, n), rng.uniform(-50, H + 50, n)], 1).astype(np.float32)
        areas = rng.integers(1, 48, (n, 2)).astype(np.int32)
        depths = rng.uniform(0.1, 80.0, n).astype(np.float32)
    a_o, d_o = areas.copy(), depths.copy()
    rects, counts = O.get_rects(us, a_o, d_o, W, H)
    keys, ids, ranges = O.bin_sort(d_o, rects, counts, W, H)
    gk, gi, gr, m = gpu_bin(us, areas, depths, W, H, device, int(counts.sum()) + 16)
    assert m == len(ids)
    same(gk, keys, "keys")
    same(gi, ids, "ids")
    same(gr, ranges, "ranges")
    lens = ranges[:, 1] - ranges[:, 0]
    if case == "long_lists_ties":
        assert lens.max() > 2 * 4096
    if case == "mid_lists":
        assert 2048 < lens.max() <= 4096
    if case == "two_runs":
        assert lens.min() <= 1024 < lens.max() <= 2048
    if case == "far_depths":
        dq = keys & np.uint64(0xFFFFFFFF)  # (tile << 32) | quantised depth
        assert (dq == 0x7FEFFF80).sum() >= 4 and (dq == 0x7FF00380).sum() >= 4
        assert (dq > 0x7FF00380).sum() >= 10 and (dq < 0x7FEFFFFF).sum() > 30_000
    if case.startswith("huge"):
        st = 4 if case == "huge_1080p" else 8
        sw = (rects[:, 2] - 1) // st - rects[:, 0] // st + 1
        sh = (rects[:, 3] - 1) // st - rects[:, 1] // st + 1
        assert ((counts > 0) & (sw * sh > 32)).sum() >= 10


@pytest.mark.parametrize("sh_dim", [16, 9, 4, 1])
def test_calc_j_jacobians_bitexact(device, sh_dim):
    """calc_J=True outputs of project, computeCov3D, computeCov2D,
    inverseCov2D, sh2Weight and sh2Color against the oracle, bit for bit, on a scene
    with Gaussians behind the camera and outside the clamp (their entries
    keep the zero fill); the oracle's Jacobians are pinned by float64 finite
    differences."""
    print(f"case: calc_J sh_dim={sh_dim}", flush=True)
    sc = make_scene(3000, seed=23, extent=4.0, K=5)  # ~370 behind the camera, ~1600 clamped
    cam = orbit_cameras(3, 3.0, 320, 240)[1]
    tc = cam.to_torch(device)
    K = sc.K
    shs = np.ascontiguousarray(sc.shs[:, : sh_dim * K])
    pws = T(sc.pws, device)
    us, pcs, depths, du = gsc.project(pws, tc.Rcw, tc.tcw, cam.fx, cam.fy, cam.cx, cam.cy, True)
    c3, dr, ds = gsc.computeCov3D(T(sc.rots, device), T(sc.scales, device), depths, True)
    c2, d3, dp = gsc.computeCov2D(c3, pcs, tc.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height, True)
    w, basis, dwdp = gsc.sh2Weight(T(shs, device), pws, tc.twc, True, K)
    cinv, areas, dci = gsc.inverseCov2D(c2, depths, True)
    o_us, o_pcs, o_dep = O.project(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    assert (o_dep < 0.2).sum() > 100
    same(N(du), O.project_jac(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy), "du_dpcs")
    o_c3 = O.cov3d(sc.rots, sc.scales, o_dep)
    o_dr, o_ds = O.cov3d_jac(sc.rots, sc.scales, o_dep)
    same(N(dr), o_dr, "dcov3d_drots")
    same(N(ds), o_ds, "dcov3d_dscales")
    o_c2 = O.cov2d(o_c3, o_pcs, cam.Rcw, o_dep, cam.fx, cam.fy, cam.width, cam.height)
    o_d3, o_dp = O.cov2d_jac(o_c3, o_pcs, cam.Rcw, o_dep, cam.fx, cam.fy, cam.width, cam.height)
    same(N(d3), o_d3, "dcov2d_dcov3ds")
    same(N(dp), o_dp, "dcov2d_dpcs")
    O.inverse_cov2d(o_c2, o_dep)
    same(N(depths), o_dep, "depths")
    same(N(dci), O.inverse_jac(o_c2, o_dep), "dcinv_dcov2ds")
    same(N(dwdp), O.sh2weight_jac(shs, sc.pws, cam.twc, K), "dweight_dpws")
    same(N(basis).reshape(sc.n, sh_dim), O.sh2weight(shs, sc.pws, cam.twc, K)[1], "basis")
    # sh2Color with calc_J (colour coefficients: the first 3 sh_dim values of each row)
    cshs = np.ascontiguousarray(sc.shs[:, : sh_dim * 3])
    col, dcs, dcp = gsc.sh2Color(T(cshs, device), pws, tc.twc, True)
    o_dcs, o_dcp = O.sh2color_jac(cshs, sc.pws, cam.twc)
    same(N(col), O.sh2color(cshs, sc.pws, cam.twc), "colors")
    same(N(dcs), o_dcs, "dcolor_dshs")
    same(N(dcp), o_dcp, "dcolor_dpws")
