"""The oracle's FP32 stages of the finetune step that only the CPU timing
sample uses (bench.py's cpu_baseline): L1 + SSIM against the float64
reference of tests/ssim_ref.py (loss and every gradient element within 1e-5
of their sums of |terms|), the full-rank SH backward against numpy, Adam
against torch.optim.Adam, and the staged step timer's stages."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp.scene import workload_c0
from ssim_ref import l1_ssim_ref


@pytest.mark.parametrize("shape,seed", [((3, 53, 77), 0), ((3, 128, 96), 1)])
def test_l1_ssim_against_reference(shape, seed):
    rng = np.random.default_rng(seed)
    tgt = rng.uniform(0, 1, shape).astype(np.float32)
    img = np.clip(tgt + rng.normal(0, 0.1, shape), 0, 1).astype(np.float32)
    img[:, ::6, ::5] = tgt[:, ::6, ::5]
    L, ref, mag, Lmag = l1_ssim_ref(img, tgt)
    loss, g = O.l1_ssim(img, tgt)
    assert abs(loss - L) <= 1e-5 * Lmag
    assert (np.abs(g.astype(np.float64) - ref) <= 1e-5 * mag).all()


def test_sh_backward_and_adam():
    rng = np.random.default_rng(2)
    dw = rng.normal(size=(1000, 5)).astype(np.float32)
    basis = rng.normal(size=(1000, 16)).astype(np.float32)
    got = O.sh_backward_full(dw, basis)
    assert np.array_equal(got, (dw[:, None, :] * basis[:, :, None]).reshape(1000, 80))
    p = rng.normal(0, 0.1, 4000).astype(np.float32)
    m, v = np.zeros_like(p), np.zeros_like(p)
    pt = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([pt], lr=3e-3, foreach=False)
    for step in range(1, 4):
        g = rng.normal(0, 1e-2, p.size).astype(np.float32)
        O.adam(p, g, m, v, step, 3e-3)
        pt.grad = torch.tensor(g)
        opt.step()
    ref = pt.detach().numpy().astype(np.float64)
    assert (np.abs(p - ref) <= 1e-6 * (np.abs(ref) + np.sqrt(np.mean(ref * ref)))).all()


def test_timed_step_stages():
    sc, cam = workload_c0()
    o = O.forward_weight(sc, cam)
    tgt = O.mix(np.clip(sc.palette + 0.05, 0, 1).astype(np.float32), o["layer"])
    t = O.timed_step(sc, cam, tgt)
    for k in ("preprocess", "scan", "key_duplication", "sort", "ranges", "forward_compositing", "mix", "loss",
              "mix_backward", "backward_compositing", "sh_backward", "adam"):
        assert t[k] >= 0.0, k
    assert t["entry_share"] == 1.0
    tiles = np.arange(0, 256, 4)
    assert 0.1 < O.timed_step(sc, cam, tgt, tiles=tiles)["entry_share"] < 0.5
