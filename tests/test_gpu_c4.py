"""The large finetune configuration: 6,000,000 Gaussians, K=7 (5 chromatic
plus black and white), rank-2 factors P=7 x Q=15, 3840x2160, 4 orbit views,
forward + backward with the replay backward (compositing state recomputed,
not stored).

Checked at full size against the FP32 oracle:
* every tile's range and the whole sorted id list, bit for bit;
* layer, contrib and final_tau on 64 seeded tiles, bit for bit (the oracle
  composites only those tiles; each tile is independent of the others);
* dL/d(dc, A, B) of 16 seeded visible Gaussians for L = sum(G * layer),
  G ~ N(0,1): the oracle's replay backward over the tiles those Gaussians
  cover (which holds all of their terms) forms the same FP32 terms and sums
  them in float64; the GPU's atomic sums must lie within 1e-5 * sum|terms|;
* the four views run without pair-buffer overflow and give finite gradients.
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


def test_c4_large_finetune(device):
    print("case: C4 6M K=7 rank-2 3840x2160 x4 views: scene", flush=True)
    sc = make_scene(N, seed=4, extent=3.0, K=K, lowrank=True, rank=RANK, P=P, Q=Q)
    cams = orbit_cameras(4, 6.0, W, H)
    Tn = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    pws, alphas, scales, rots = Tn(sc.pws), Tn(sc.alphas), Tn(sc.scales), Tn(sc.rots)
    dc, A, B = (Tn(x).requires_grad_(True) for x in (sc.dc, sc.A, sc.B))
    rng = np.random.default_rng(9)
    G0 = rng.standard_normal((K, H, W)).astype(np.float32)
    for v, cam in enumerate(cams):
        tc = cam.to_torch(device)
        layer, _ = gsmodel.WeightGSFunctionFast.apply(pws, dc, A, B, alphas, scales, rots, tc, K, Q)
        (layer * Tn(G0)).sum().backward()
        torch.cuda.synchronize()
        print(f"view {v}: forward + backward done", flush=True)
    for t in (dc, A, B):
        assert torch.isfinite(t.grad).all()

    # view 0 again, with its state, against the oracle
    cam = cams[0]
    tc = cam.to_torch(device)
    for t in (dc, A, B):
        t.grad = None
    with torch.no_grad():
        layer, depths, st = gsmodel.render_weights(pws, alphas, scales, rots, tc, K, dc=dc, A=A, B=B, Q=Q,
                                                   return_state=True)
    m = st["m"]
    g_ranges, g_ids = st["ranges"].cpu().numpy(), st["ids"][:m].cpu().numpy()
    print(f"oracle geometry and sort ({m} pairs)", flush=True)
    us, pcs, dep = O.project(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    c3 = O.cov3d(sc.rots, sc.scales, dep)
    c2 = O.cov2d(c3, pcs, cam.Rcw, dep, cam.fx, cam.fy, W, H)
    cinv, areas = O.inverse_cov2d(c2, dep)
    rects, counts = O.get_rects(us, areas, dep, W, H)
    _, ids, ranges = O.bin_sort(dep, rects, counts, W, H)
    assert m == len(ids) and m > 10_000_000
    assert np.array_equal(g_ranges, ranges)
    assert np.array_equal(g_ids, ids)
    assert np.array_equal(depths.cpu().numpy(), dep)

    print("oracle compositing of 64 tiles", flush=True)
    w32, _ = O.sh2weight_fast(sc.dc, sc.A, sc.B, sc.pws, cam.twc, K, Q)
    gx = (W + 15) // 16
    nonempty = np.flatnonzero(ranges[:, 1] > ranges[:, 0])
    tiles = rng.choice(nonempty, 64, replace=False)
    sub = np.zeros_like(ranges)
    sub[tiles] = ranges[tiles]
    o_layer, o_contrib, o_ft, _, _ = O.draw(W, H, sub, ids, us, cinv, sc.alphas, w32, 0.0)
    g_layer, g_contrib, g_ft = layer.cpu().numpy(), st["contrib"].cpu().numpy(), st["final_tau"].cpu().numpy()
    for t in tiles:
        ty, tx = divmod(int(t), gx)
        ys, xs = slice(16 * ty, min(16 * ty + 16, H)), slice(16 * tx, 16 * tx + 16)
        assert np.array_equal(g_layer[:, ys, xs], o_layer[:, ys, xs]), int(t)
        assert np.array_equal(g_contrib[ys, xs], o_contrib[ys, xs]), int(t)
        assert np.array_equal(g_ft[ys, xs], o_ft[ys, xs]), int(t)

    print("gradients of 16 Gaussians", flush=True)
    layer, _ = gsmodel.WeightGSFunctionFast.apply(pws, dc, A, B, alphas, scales, rots, tc, K, Q)
    (layer * Tn(G0)).sum().backward()
    gdc, gA, gB = (t.grad.detach().cpu().numpy().astype(np.float64) for t in (dc, A, B))
    # Gaussians at the head of a tile list: in a scene this dense the pixels
    # saturate after a few dozen splats, so most of the rest are hidden
    heads = np.unique(ids[ranges[nonempty, 0]])
    heads = heads[counts[heads] <= 16]
    pick = rng.choice(heads, 16, replace=False)
    sub = np.zeros_like(ranges)
    for i in pick:
        x0, y0, x1, y1 = (int(v) for v in rects[i])
        for ty in range(y0, y1):
            t = np.arange(ty * gx + x0, ty * gx + x1)
            sub[t] = ranges[t]
    zeros_i, zeros_f = np.zeros((H, W), np.int32), np.zeros((H, W), np.float32)
    g, gabs = O.draw_backward(W, H, K, sub, ids, us, cinv, sc.alphas, zeros_i, zeros_f, G0, np.float32, mode=2)
    _, ylm = O.sh2weight_fast(sc.dc[pick], sc.A[pick], sc.B[pick], sc.pws[pick], cam.twc, K, Q)
    rA, rB = O.sh2weight_fast_backward(sc.A[pick], sc.B[pick], ylm, g[pick], K, Q, np.float64)
    aA, aB = O.sh2weight_fast_backward(np.abs(sc.A[pick]), np.abs(sc.B[pick]), np.abs(ylm), gabs[pick], K, Q,
                                       np.float64)
    C0 = np.float64(np.float32(0.28209479177387814))
    nz = 0
    for name, got, ref, s_ in (("ddc", gdc[pick], g[pick] * C0, gabs[pick] * C0), ("dA", gA[pick], rA, aA),
                               ("dB", gB[pick], rB, aB)):
        err = np.abs(got - ref)
        assert (err <= 1e-5 * s_ + 1e-12).all(), (name, float(np.max(err / (1e-5 * s_ + 1e-12))))
        nz += int((np.abs(ref) > 0).sum(axis=tuple(range(1, ref.ndim))).astype(bool).sum())
    assert nz >= 24, nz  # most of the 3 x 16 gradients are non-zero
