"""Colour backward with geometry gradients (gsplatcu.splatB / splat_color_B on
rgp_draw_color_backward) against the FP32 oracle (oracle draw_color_backward,
pinned by float64 finite differences in tests/test_oracle_vanilla.py).  The
forward inputs are the GPU's own geometry and tile lists (bit-exact with the
oracle, tests/test_gpu_parity.py).  Each gradient sums per-(pixel, entry)
terms with atomics: it must lie within 1e-5 * sum|terms| of the oracle's
float64 sum of the same FP32 terms."""
import numpy as np
import pytest
import torch

import gsplatcu as gsc
from oracle import oracle as O
from rgp.scene import make_scene, orbit_cameras, workload_c0

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def N(t):
    return t.detach().cpu().numpy()


def colour_forward(scene, cam, dev, seed):
    tc = cam.to_torch(dev)
    pws = T(scene.pws, dev)
    us, pcs, depths = gsc.project(pws, tc.Rcw, tc.tcw, cam.fx, cam.fy, cam.cx, cam.cy, False)
    c3, = gsc.computeCov3D(T(scene.rots, dev), T(scene.scales, dev), depths, False)
    c2, = gsc.computeCov2D(c3, pcs, tc.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height, False)
    cinv, areas = gsc.inverseCov2D(c2, depths, False)
    colors = T(np.random.default_rng(seed).uniform(0, 1, (scene.n, 3)).astype(np.float32), dev)
    alphas = T(scene.alphas, dev)
    image, contrib, final_tau, ranges, ids = gsc.splat(cam.height, cam.width, us, cinv, alphas, depths, colors, areas)
    return dict(us=us, cinv=cinv, alphas=alphas, depths=depths, colors=colors, image=image, contrib=contrib,
                final_tau=final_tau, ranges=ranges, ids=ids)


def check(f, cam, dimg, dtau, with_tau, dev):
    args = (cam.height, cam.width, f["us"], f["cinv"], f["alphas"], f["depths"], f["colors"], f["contrib"],
            f["final_tau"], f["ranges"], f["ids"], T(dimg, dev))
    if with_tau:
        out = gsc.splatB(*args, T(dtau, dev))
    else:
        out = gsc.splat_color_B(*args)
    torch.cuda.synchronize()
    n = f["us"].shape[0]
    assert [tuple(o.shape) for o in out] == [(n, 1, 2), (n, 1, 3), (n, 1, 1), (n, 1, 3)]
    got = np.concatenate([N(o).reshape(n, -1) for o in out], axis=1).astype(np.float64)
    ref, rabs = O.draw_color_backward(cam.width, cam.height, N(f["ranges"]), N(f["ids"]), N(f["us"]), N(f["cinv"]),
                                      N(f["alphas"]), N(f["colors"]), N(f["contrib"]), N(f["final_tau"]), dimg,
                                      dtau if with_tau else None)
    err = np.abs(got - ref)
    lim = 1e-5 * rabs + 1e-30
    assert (err <= lim).all(), float(np.max(err / lim))
    assert np.count_nonzero(ref[:, 0]) > n // 20 and np.count_nonzero(ref[:, 5]) > n // 20


@pytest.mark.parametrize("with_tau", [False, True])
def test_colour_backward_c0(device, with_tau):
    print(f"case: colour backward C0 with_tau={with_tau}", flush=True)
    scene, cam = workload_c0()
    f = colour_forward(scene, cam, device, 1)
    rng = np.random.default_rng(2)
    dimg = rng.normal(size=(3, cam.height, cam.width)).astype(np.float32)
    dtau = rng.normal(size=(cam.height, cam.width)).astype(np.float32)
    check(f, cam, dimg, dtau, with_tau, device)


def test_colour_backward_200k_640x480(device):
    print("case: colour backward 200k 640x480", flush=True)
    sc = make_scene(200_000, seed=5, extent=2.0, K=5)
    cam = orbit_cameras(4, 5.0, 640, 480)[1]
    f = colour_forward(sc, cam, device, 3)
    rng = np.random.default_rng(4)
    dimg = rng.normal(size=(3, 480, 640)).astype(np.float32)
    dtau = rng.normal(size=(480, 640)).astype(np.float32)
    check(f, cam, dimg, dtau, True, device)


def test_gsfunction_backward_chain(device):
    """rgp.vanilla.GSFunction (the vanilla training step's autograd function)
    on C0 with colour coefficients: the forward image and final_tau equal the
    oracle's bit for bit; dL/d(pws, shs, alphas, scales, rots) for
    L = sum(G * image) + sum(G_tau * final_tau) lie within 1e-5 of the bound
    |J| |terms| of the same chain assembled in float64 from the FP32 oracle
    (its Jacobians bit-exact with the kernels, tests/test_gpu_parity.py)."""
    from rgp.vanilla import GSFunction
    from test_oracle_vanilla import oracle_chain, oracle_forward
    print("case: vanilla GSFunction backward chain (C0)", flush=True)
    scene, cam = workload_c0()
    rng = np.random.default_rng(12)
    shs = rng.normal(0, 0.3, (scene.n, 48)).astype(np.float32)
    sc = dict(pws=scene.pws, rots=scene.rots, scales=scene.scales, alphas=scene.alphas)
    f = oracle_forward(sc, cam, shs, np.float32)
    dimg = rng.normal(size=(3, cam.height, cam.width)).astype(np.float32)
    dtau = rng.normal(size=(cam.height, cam.width)).astype(np.float32)
    ref, bound = oracle_chain(sc, cam, shs, f, dimg, dtau, np.float32)
    leaves = {k: T(v, device).requires_grad_(True) for k, v in
              (("pws", scene.pws), ("shs", shs), ("alphas", scene.alphas.reshape(-1, 1)), ("scales", scene.scales),
               ("rots", scene.rots))}
    us0 = torch.zeros((scene.n, 2), device=device)
    img, ft, mask = GSFunction.apply(leaves["pws"], leaves["shs"], leaves["alphas"], leaves["scales"],
                                     leaves["rots"], us0, cam.to_torch(device))
    assert np.array_equal(N(img), f["img"]) and np.array_equal(N(ft), f["ft"])
    ((img * T(dimg, device)).sum() + (ft * T(dtau, device)).sum()).backward()
    for k in ("pws", "shs", "alphas", "scales", "rots"):
        got = N(leaves[k].grad).astype(np.float64).reshape(ref[k].shape)
        err = np.abs(got - ref[k])
        lim = 1e-5 * bound[k] + 1e-30
        assert (err <= lim).all(), (k, float(np.max(err / lim)))
        assert np.count_nonzero(ref[k]) > 0
