"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
the oracle (CPU) and the kernels (GPU) must reproduce them bit for bit."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from rgp.scene import Camera, Scene

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "*.npz")))
KEYS = ("us", "depths", "cinv", "areas", "weights", "ids", "ranges", "layer", "contrib", "final_tau")


def load(path):
    d = dict(np.load(path))
    w, h, fx, fy, cx, cy = d["cam"]
    cam = Camera(int(w), int(h), fx, fy, cx, cy, d["Rcw"], d["tcw"])
    cam.twc = d["twc"]
    K = d["palette"].shape[0]
    sc = dict(pws=d["pws"], rots=d["rots"], scales=d["scales"], alphas=d["alphas"], palette=d["palette"], K=K,
              lowrank="A" in d)
    for k in ("shs", "dc", "A", "B"):
        if k in d:
            sc[k] = d[k]
    if "A" in d:
        sc["P"], sc["Q"], sc["rank"] = d["A"].shape[1], d["B"].shape[2], d["A"].shape[2]
    return Scene(**sc), cam, d


def test_fixtures_present():
    assert len(FILES) >= 3


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_oracle_reproduces_fixture(path):
    sc, cam, d = load(path)
    o = O.forward_weight(sc, cam)
    for k in KEYS:
        assert np.array_equal(o[k], d["out_" + k]), k


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_kernels_reproduce_fixture(path, device):
    import torch
    from rgp import gsmodel

    sc, cam, d = load(path)
    print(f"case: golden {os.path.basename(path)}", flush=True)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
    kw = dict(shs=T(sc.shs)) if not sc.lowrank else dict(dc=T(sc.dc), A=T(sc.A), B=T(sc.B), Q=sc.Q)
    layer, depths, st = gsmodel.render_weights(T(sc.pws), T(sc.alphas), T(sc.scales), T(sc.rots),
                                               cam.to_torch(device), sc.K, return_state=True, **kw)
    assert np.array_equal(layer.cpu().numpy(), d["out_layer"])
    assert np.array_equal(depths.cpu().numpy(), d["out_depths"])
    assert np.array_equal(st["contrib"].cpu().numpy(), d["out_contrib"])
    assert np.array_equal(st["final_tau"].cpu().numpy(), d["out_final_tau"])
    assert np.array_equal(st["ranges"].cpu().numpy(), d["out_ranges"])
    assert np.array_equal(st["ids"][: st["m"]].cpu().numpy(), d["out_ids"])
