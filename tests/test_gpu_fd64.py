"""Float64 central finite differences of the whole low-rank chain
(sh2Weight_fast -> splat_weight -> splat_weight_B -> sh2Weight_fast_B,
gsplat/gsmodel.py:151-200) against the GPU gradient.

Workload: 1,000,000 Gaussians, K=5, rank-2 factors A (N,5,2), B (N,2,15),
one 1920x1080 orbit view; loss L = sum(G * layer) with G ~ N(0,1) (seed 1);
64 seeded visible Gaussians, eps = 1e-4 on every one of their 45 weight
parameters (dc 5, A 10, B 30).

The geometry (u, conic, alpha, tile lists) is frozen at its FP32 values,
which the GPU reproduces bit for bit (test_gpu_parity): the finetune trains
the palette weights only.  L(theta +- eps) is evaluated in float64 by the
oracle on the crop of tiles that Gaussian covers, with the full-frame tile
lists and u shifted by the crop origin (a shift by whole tiles is exact in
float64), which equals the full-frame difference: no other pixel depends on
the Gaussian's weights.  L is linear in dc, in A and in B separately, so the
central difference has no truncation error.

Bar: |GPU - FD| <= 1e-4 * S + 1e-9, S = sum of |terms| of that gradient
(the GPU sums FP32 terms with atomics, in varying order; 1e-9 is above the
float64 rounding noise of the difference quotient, ~1e-16 |L| / eps)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel
from rgp.scene import workload_c2

pytestmark = pytest.mark.gpu
EPS = 1e-4
TOL = 1e-4
N_PICK = 64


def crop_problem(g, fw, i, Wf, Hf):
    """Local arrays of the tiles covered by Gaussian i (full-frame rect)."""
    x0, y0, x1, y1 = (int(v) for v in fw["rects"][i])
    gx = (Wf + 15) // 16
    tiles = [(tx, ty) for ty in range(y0, y1) for tx in range(x0, x1)]
    lists = [fw["ids"][fw["ranges"][ty * gx + tx, 0]:fw["ranges"][ty * gx + tx, 1]] for tx, ty in tiles]
    glob = np.unique(np.concatenate(lists))
    loc = {int(v): j for j, v in enumerate(glob)}
    ranges, ids, o = np.zeros((len(tiles), 2), np.int32), [], 0
    for t, lst in enumerate(lists):
        ids.extend(loc[int(v)] for v in lst)
        ranges[t] = (o, o + len(lst))
        o += len(lst)
    Wc = 16 * (x1 - x0)
    Hc = min(16 * y1, Hf) - 16 * y0
    us = fw["us"][glob].astype(np.float64) - np.array([16.0 * x0, 16.0 * y0])
    G = g[:, 16 * y0:16 * y0 + Hc, 16 * x0:16 * x0 + Wc]
    if Wc > G.shape[2]:  # right image border: the missing columns do not exist
        Gp = np.zeros((G.shape[0], Hc, Wc))
        Gp[:, :, :G.shape[2]] = G
        G = Gp
    return dict(glob=glob, me=loc[i], ranges=ranges, ids=np.asarray(ids, np.int32), us=us,
                cinv=fw["cinv"][glob].astype(np.float64), alphas=fw["alphas"][glob].astype(np.float64),
                G=G.astype(np.float64), Wc=Wc, Hc=Hc)


def test_fd64_lowrank_c2(device):
    print("case: FD64 rank-2 1M 1080p, 64 Gaussians", flush=True)
    scene, cams = workload_c2(seed=0, n=1_000_000, rank=2, n_views=8)
    cam = cams[0]
    K, Q = scene.K, scene.Q
    W, H = cam.width, cam.height
    G = np.random.default_rng(1).standard_normal((K, H, W)).astype(np.float32)

    # GPU: forward + backward through the autograd wrapper
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    dc, A, B = (T(x).requires_grad_(True) for x in (scene.dc, scene.A, scene.B))
    layer, _ = gsmodel.WeightGSFunctionFast.apply(T(scene.pws), dc, A, B, T(scene.alphas), T(scene.scales),
                                                  T(scene.rots), cam.to_torch(device), K, Q)
    (layer * T(G)).sum().backward()
    gdc, gA, gB = (x.grad.detach().cpu().numpy().astype(np.float64) for x in (dc, A, B))
    print("gpu gradient done", flush=True)

    # FP32 geometry and tile lists (the oracle; bit-identical to the GPU)
    us, pcs, depths = O.project(scene.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    c3 = O.cov3d(scene.rots, scene.scales, depths)
    c2 = O.cov2d(c3, pcs, cam.Rcw, depths, cam.fx, cam.fy, W, H)
    cinv, areas = O.inverse_cov2d(c2, depths)
    rects, counts = O.get_rects(us, areas, depths, W, H)
    _, ids, ranges = O.bin_sort(depths, rects, counts, W, H)
    fw = dict(us=us, cinv=cinv, alphas=scene.alphas, rects=rects, ids=ids, ranges=ranges)
    # near, small splats: mostly unoccluded, so most gradients are non-zero
    vis = np.flatnonzero((counts > 0) & (counts <= 12))
    vis = vis[depths[vis] < np.percentile(depths[vis], 30)]
    pick = np.random.default_rng(7).choice(vis, N_PICK, replace=False)
    C0 = 0.28209479177387814

    nonzero = 0
    for n_done, i in enumerate(pick):
        cp = crop_problem(G, fw, int(i), W, H)
        glob = cp["glob"]
        w_loc, _ = O.sh2weight_fast(scene.dc[glob], scene.A[glob], scene.B[glob], scene.pws[glob], cam.twc, K, Q,
                                    np.float64)
        me = cp["me"]

        def draw(w):
            return O.draw(cp["Wc"], cp["Hc"], cp["ranges"], cp["ids"], cp["us"], cp["cinv"], cp["alphas"], w, 0.0,
                          np.float64)

        def loss(dci, Ai, Bi):
            wi, _ = O.sh2weight_fast(dci[None], Ai[None], Bi[None], scene.pws[i:i + 1], cam.twc, K, Q, np.float64)
            w = w_loc.copy()
            w[me] = wi[0]
            return float(np.sum(cp["G"] * draw(w)[0]))

        base = [scene.dc[i].astype(np.float64), scene.A[i].astype(np.float64), scene.B[i].astype(np.float64)]
        fd = [np.zeros_like(b) for b in base]
        for t in range(3):
            for idx in np.ndindex(base[t].shape):
                pp = [b.copy() for b in base]
                mm = [b.copy() for b in base]
                pp[t][idx] += EPS
                mm[t][idx] -= EPS
                fd[t][idx] = (loss(*pp) - loss(*mm)) / (2 * EPS)

        # |terms| bound: the float64 backward of the crop with absolute values
        _, contrib, ftau, _, _ = draw(w_loc)
        _, gabs = O.draw_backward(cp["Wc"], cp["Hc"], K, cp["ranges"], cp["ids"], cp["us"], cp["cinv"], cp["alphas"],
                                  contrib, ftau, cp["G"], np.float64, mode=0)
        _, ylm = O.sh2weight_fast(scene.dc[i:i + 1], scene.A[i:i + 1], scene.B[i:i + 1], scene.pws[i:i + 1], cam.twc,
                                  K, Q, np.float64)
        aA, aB = O.sh2weight_fast_backward(np.abs(scene.A[i:i + 1]), np.abs(scene.B[i:i + 1]), np.abs(ylm),
                                           gabs[me:me + 1], K, Q, np.float64)
        S = [gabs[me] * C0, aA[0], aB[0]]
        got = [gdc[i], gA[i], gB[i]]
        for name, g_, f_, s_ in zip(("ddc", "dA", "dB"), got, fd, S):
            err = np.abs(g_ - f_)
            lim = TOL * s_ + 1e-9
            assert (err <= lim).all(), (int(i), name, float(np.max(err / lim)), g_.ravel()[:3], f_.ravel()[:3])
        nonzero += bool(np.max(np.abs(fd[0])) > 1e-6)
        if n_done % 16 == 15:
            print(f"{n_done + 1} Gaussians checked", flush=True)
    assert nonzero >= N_PICK // 2, nonzero
