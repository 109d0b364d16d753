"""Serving at the stated workload: the six resident scenes of default_specs()
(200k, 500k, 1M, 1M, 2M and 3M Gaussians; K 5 or 7; full rank or rank 2)
admitted under the 200 GiB budget, a request stream at 1280x720, 1920x1080
and 2560x1440 with 40% poses and 60% palette-only edits, served by the
continuously batched loop (multi-view render graphs, batched mixes).

The reference has no serving loop; its palette edit is a bake + re-render
(finetune.py:472-478, gsmodel.py:488-509), which by linearity equals the
palette mix of the weight image.  A seeded sample of the served frames
(every resolution, poses and edits) is checked against the oracle: the
scene's oracle geometry, sort and sh2Weight(_fast) for the frame's view, the
oracle's compositing of 8 seeded tiles and the oracle's mix with the frame's
palette (the scene palette for a pose, the edited one for an edit), bit for
bit on those tiles.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp.scene import make_scene
from rgp.serving import RESOLUTIONS, Engine, default_specs, edited_palette, make_requests

pytestmark = pytest.mark.gpu


def oracle_view(sc, cam):
    us, pcs, dep = O.project(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    c3 = O.cov3d(sc.rots, sc.scales, dep)
    c2 = O.cov2d(c3, pcs, cam.Rcw, dep, cam.fx, cam.fy, cam.width, cam.height)
    cinv, areas = O.inverse_cov2d(c2, dep)
    rects, counts = O.get_rects(us, areas, dep, cam.width, cam.height)
    _, ids, ranges = O.bin_sort(dep, rects, counts, cam.width, cam.height)
    if sc.lowrank:
        w, _ = O.sh2weight_fast(sc.dc, sc.A, sc.B, sc.pws, cam.twc, sc.K, sc.Q)
    else:
        w, _ = O.sh2weight(sc.shs, sc.pws, cam.twc, sc.K)
    return us, cinv, ids, ranges, w


def test_c3_served_frames_match_oracle(device):
    specs = default_specs()
    print("case: C3 serving, 6 scenes (200k-3M): admit and capture", flush=True)
    eng = Engine(specs, device)
    reqs = make_requests(specs, 96, rate=0.0, seed=5)
    eng.prepare(reqs)
    rep = eng.run(reqs, "batched", realtime=False, keep_outputs=True)
    torch.cuda.synchronize()
    assert rep["requests"] == len(reqs) and all(r.out is not None for r in reqs)
    rng = np.random.default_rng(12)
    src_res = lambda r: r.res if r.kind == "pose" else r.dep_view[1]  # an edit is served at its view's size
    sample = []
    for res in RESOLUTIONS:
        for kind in ("pose", "edit"):
            cand = [r for r in reqs if r.kind == kind and src_res(r) == res]
            assert cand, (res, kind)
            sample.extend(rng.choice(cand, min(2, len(cand)), replace=False).tolist())
    spec_of = {s.name: s for s in specs}
    hosts = {}
    for r in sample:
        sp = spec_of[r.scene]
        if sp.name not in hosts:
            hosts[sp.name] = make_scene(sp.n, seed=sp.seed, extent=3.0, K=sp.K, lowrank=sp.lowrank, rank=sp.rank,
                                        P=(sp.P or sp.K) if sp.lowrank else None, Q=sp.Q)
        sc = hosts[sp.name]
        pose, (w_, h_) = (r.pose, r.res) if r.kind == "pose" else r.dep_view
        pal = sc.palette.astype(np.float32) if r.kind == "pose" else edited_palette(sc.palette, r.dpal, r.lum)
        print(f"request {r.rid}: {r.kind} scene {sp.name} ({sp.n}, K={sp.K}) pose {pose} {w_}x{h_}", flush=True)
        cam = eng.pose_camera(pose, w_, h_)
        us, cinv, ids, ranges, w = oracle_view(sc, cam)
        nonempty = np.flatnonzero(ranges[:, 1] > ranges[:, 0])
        tiles = rng.choice(nonempty, 8, replace=False)
        sub = np.zeros_like(ranges)
        sub[tiles] = ranges[tiles]
        layer, _, _, _, _ = O.draw(w_, h_, sub, ids, us, cinv, sc.alphas, w, 0.0)
        ref = O.mix(pal, layer)
        got = r.out.cpu().numpy()
        assert got.shape == (3, h_, w_)
        assert np.isfinite(got).all()
        gx = (w_ + 15) // 16
        for t in tiles:
            ty, tx = divmod(int(t), gx)
            ys, xs = slice(16 * ty, min(16 * ty + 16, h_)), slice(16 * tx, min(16 * tx + 16, w_))
            assert np.array_equal(got[:, ys, xs], ref[:, ys, xs]), (r.rid, r.kind, int(t))
    kinds = {(r.kind, src_res(r)) for r in sample}
    assert len(kinds) == 2 * len(RESOLUTIONS)
