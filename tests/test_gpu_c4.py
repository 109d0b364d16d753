"""The large finetune configuration: 6,000,000 Gaussians, K=7 (5 chromatic
plus black and white), rank-2 factors P=7 x Q=15, 3840x2160, 4 orbit views
batched per launch (WeightGSFunctionFastViews: one preprocess launch for the
four cameras, one forward and one backward compositing launch over the four
views' tiles), forward + backward with the replay backward (compositing state
recomputed, not stored).

Checked at full size against the FP32 oracle, for every one of the four views:
* every tile's range and the whole sorted id list, bit for bit, and the depths;
* layer, contrib and final_tau on 16 seeded tiles, bit for bit (the oracle
  composites only those tiles; each tile is independent of the others);
and for the batch:
* dL/d(dc, A, B) of 16 seeded Gaussians for L = sum_v sum(G * layer_v),
  G ~ N(0,1): per view the oracle's replay backward over the tiles those
  Gaussians cover (which holds all of their terms) and the SH chain with that
  view's basis, summed over the views in float64; the GPU's sums must lie
  within 1e-5 * sum|terms|;
* the batch runs without pair-buffer overflow and gives finite gradients.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel
from rgp.scene import make_scene, orbit_cameras

pytestmark = pytest.mark.gpu
N, K, RANK, P, Q = 6_000_000, 7, 2, 7, 15
W, H = 3840, 2160
VIEWS = 4


def test_c4_large_finetune(device):
    print(f"case: C4 6M K=7 rank-2 3840x2160 x{VIEWS} views per launch: scene", flush=True)
    sc = make_scene(N, seed=4, extent=3.0, K=K, lowrank=True, rank=RANK, P=P, Q=Q)
    cams = orbit_cameras(VIEWS, 6.0, W, H)
    Tn = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    pws, alphas, scales, rots = Tn(sc.pws), Tn(sc.alphas), Tn(sc.scales), Tn(sc.rots)
    dc, A, B = (Tn(x).requires_grad_(True) for x in (sc.dc, sc.A, sc.B))
    rng = np.random.default_rng(9)
    G0 = rng.standard_normal((K, H, W)).astype(np.float32)
    layers, _ = gsmodel.WeightGSFunctionFastViews.apply(pws, dc, A, B, alphas, scales, rots, cams, K, Q)
    (layers * Tn(G0)).sum().backward()
    torch.cuda.synchronize()
    print("batched forward + backward done", flush=True)
    gdc, gA, gB = (t.grad.detach().cpu().numpy().astype(np.float64) for t in (dc, A, B))
    for g in (gdc, gA, gB):
        assert np.isfinite(g).all()
    with torch.no_grad():
        layers, depths, st = gsmodel.render_weights_views(pws, alphas, scales, rots, cams, K, dc=dc, A=A, B=B, Q=Q,
                                                          return_state=True)
    gx = (W + 15) // 16

    per_view = []
    for v, cam in enumerate(cams):
        m = st["m"][v]
        print(f"view {v}: oracle geometry and sort ({m} pairs)", flush=True)
        us, pcs, dep = O.project(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
        c3 = O.cov3d(sc.rots, sc.scales, dep)
        c2 = O.cov2d(c3, pcs, cam.Rcw, dep, cam.fx, cam.fy, W, H)
        cinv, areas = O.inverse_cov2d(c2, dep)
        rects, counts = O.get_rects(us, areas, dep, W, H)
        _, ids, ranges = O.bin_sort(dep, rects, counts, W, H)
        assert m == len(ids) and m > 10_000_000
        assert np.array_equal(st["ranges"][v].cpu().numpy(), ranges)
        assert np.array_equal(st["ids"][v][:m].cpu().numpy(), ids)
        assert np.array_equal(depths[v].cpu().numpy(), dep)

        w32, _ = O.sh2weight_fast(sc.dc, sc.A, sc.B, sc.pws, cam.twc, K, Q)
        nonempty = np.flatnonzero(ranges[:, 1] > ranges[:, 0])
        tiles = rng.choice(nonempty, 16, replace=False)
        sub = np.zeros_like(ranges)
        sub[tiles] = ranges[tiles]
        o_layer, o_contrib, o_ft, _, _ = O.draw(W, H, sub, ids, us, cinv, sc.alphas, w32, 0.0)
        g_layer = layers[v].cpu().numpy()
        g_contrib, g_ft = st["contrib"][v].cpu().numpy(), st["final_tau"][v].cpu().numpy()
        for t in tiles:
            ty, tx = divmod(int(t), gx)
            ys, xs = slice(16 * ty, min(16 * ty + 16, H)), slice(16 * tx, 16 * tx + 16)
            assert np.array_equal(g_layer[:, ys, xs], o_layer[:, ys, xs]), (v, int(t))
            assert np.array_equal(g_contrib[ys, xs], o_contrib[ys, xs]), (v, int(t))
            assert np.array_equal(g_ft[ys, xs], o_ft[ys, xs]), (v, int(t))
        per_view.append(dict(us=us, cinv=cinv, rects=rects, counts=counts, ids=ids, ranges=ranges,
                             nonempty=nonempty, cam=cam))

    print("gradients of 16 Gaussians, summed over the views", flush=True)
    # Gaussians at the head of a tile list of view 0: in a scene this dense the
    # pixels saturate after a few dozen splats, so most of the rest are hidden
    v0 = per_view[0]
    heads = np.unique(v0["ids"][v0["ranges"][v0["nonempty"], 0]])
    heads = heads[v0["counts"][heads] <= 16]
    pick = rng.choice(heads, 16, replace=False)
    ref = {k: 0.0 for k in ("ddc", "dA", "dB", "sdc", "sA", "sB")}
    C0 = np.float64(np.float32(0.28209479177387814))
    zeros_i, zeros_f = np.zeros((H, W), np.int32), np.zeros((H, W), np.float32)
    for pv in per_view:
        sub = np.zeros_like(pv["ranges"])
        for i in pick:
            if pv["counts"][i] == 0:
                continue
            x0, y0, x1, y1 = (int(c) for c in pv["rects"][i])
            for ty in range(y0, y1):
                t = np.arange(ty * gx + x0, ty * gx + x1)
                sub[t] = pv["ranges"][t]
        g, gabs = O.draw_backward(W, H, K, sub, pv["ids"], pv["us"], pv["cinv"], sc.alphas, zeros_i, zeros_f, G0,
                                  np.float32, mode=2)
        _, ylm = O.sh2weight_fast(sc.dc[pick], sc.A[pick], sc.B[pick], sc.pws[pick], pv["cam"].twc, K, Q)
        rA, rB = O.sh2weight_fast_backward(sc.A[pick], sc.B[pick], ylm, g[pick], K, Q, np.float64)
        aA, aB = O.sh2weight_fast_backward(np.abs(sc.A[pick]), np.abs(sc.B[pick]), np.abs(ylm), gabs[pick], K, Q,
                                           np.float64)
        for k, val in (("ddc", g[pick] * C0), ("dA", rA), ("dB", rB), ("sdc", gabs[pick] * C0), ("sA", aA),
                       ("sB", aB)):
            ref[k] = ref[k] + val
    nz = 0
    for name, got, r, s_ in (("ddc", gdc[pick], ref["ddc"], ref["sdc"]), ("dA", gA[pick], ref["dA"], ref["sA"]),
                             ("dB", gB[pick], ref["dB"], ref["sB"])):
        err = np.abs(got - r)
        lim = 1e-5 * (s_ * 1.0001) + 1e-12
        assert (err <= lim).all(), (name, float(np.max(err / lim)))
        nz += int((np.abs(r) > 0).sum(axis=tuple(range(1, r.ndim))).astype(bool).sum())
    assert nz >= 24, nz  # most of the 3 x 16 gradients are non-zero
