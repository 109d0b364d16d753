"""Writes the golden fixtures of tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`).

* cameras_insect.json: three camera poses (img_name, width, height, fx, fy,
  rotation, tcw) copied as data from the reference's camera paths
  (vanilla_3dgs/cameras/data/supersplat/insect.json, entries r_0000, r_0010,
  r_0020); the reference builds Rcw = rotation^T
  (vanilla_3dgs/render_images.py:58).
* *.npz: inputs and FP32-oracle outputs of the palette-weight chain on tiny
  seeded scenes (full rank K=5; rank-2 low rank K=5; the colour path), so
  that the kernels and later versions of the oracle can be compared with
  fixed vectors.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "reparamgs-palette_amd")]

from oracle import oracle as O  # noqa: E402
from rgp.scene import Camera, make_scene, orbit_cameras  # noqa: E402

CAM_SRC = "/root/reference/vanilla_3dgs/cameras/data/supersplat/insect.json"


def cameras():
    path = os.path.join(HERE, "cameras_insect.json")
    if not os.path.exists(path):
        data = json.load(open(CAM_SRC))
        keep = [c for c in data if c["img_name"] in ("r_0000", "r_0010", "r_0020")]
        json.dump(keep, open(path, "w"), indent=1)
    return json.load(open(path))


def camera_from_json(c, scale=0.15):
    """Camera of render_images.py:48-58 at a reduced resolution."""
    w, h = int(c["width"] * scale), int(c["height"] * scale)
    R = np.array(c["rotation"], np.float64).T
    return Camera(w, h, c["fx"] * scale, c["fy"] * scale, w / 2.0, h / 2.0, R, np.array(c["tcw"]))


def save(name, scene, cam, fwd, extra=None):
    d = dict(pws=scene.pws, rots=scene.rots, scales=scene.scales, alphas=scene.alphas, palette=scene.palette,
             Rcw=cam.Rcw, tcw=cam.tcw, twc=cam.twc,
             cam=np.array([cam.width, cam.height, cam.fx, cam.fy, cam.cx, cam.cy], np.float64))
    for k in ("us", "depths", "cinv", "areas", "weights", "ids", "ranges", "layer", "contrib", "final_tau"):
        d["out_" + k] = fwd[k]
    for k in ("shs", "dc", "A", "B"):
        if hasattr(scene, k):
            d[k] = getattr(scene, k)
    if extra:
        d.update(extra)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)


def main():
    cams = cameras()
    sc = make_scene(400, seed=101, extent=1.0, K=5)
    cam = camera_from_json(cams[0])
    save("full_rank_k5", sc, cam, O.forward_weight(sc, cam))
    sc = make_scene(400, seed=102, extent=1.0, K=5, lowrank=True, rank=2, P=5, Q=15)
    cam = camera_from_json(cams[1])
    f = O.forward_weight(sc, cam)
    save("low_rank2_k5", sc, cam, f, dict(out_ylm=f["ylm"]))
    sc = make_scene(300, seed=103, extent=1.0, K=7)
    cam = orbit_cameras(3, 3.0, 72, 56)[2]
    save("full_rank_k7", sc, cam, O.forward_weight(sc, cam))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
