"""rgp_bary_loss_noalloc (the barycentric loss with its float64 accumulators in
the library's per-device scratch array, no workspace argument) against
rgp_bary_loss with a caller workspace, through the C ABI: two calls in a row
on one stream, the second on a smaller image and a smaller palette, each
compared with the workspace form on the same inputs.  W_bary and dL/dW_tilde
bit for bit; dL/dpalette and the four loss values within 1e-6 relative (the
float64 accumulators are summed with atomics in a run-dependent order before
the final rounding to float32)."""
import numpy as np
import pytest
import torch

from rgp._lib import call, lib, ptr, stream

pytestmark = pytest.mark.gpu


def inputs(Kc, H, W, seed):
    rng = np.random.default_rng(seed)
    h = np.sort(rng.uniform(0, 2 * np.pi, Kc))
    u, v = np.array([1, -1, 0]) / np.sqrt(2), np.array([1, 1, -2]) / np.sqrt(6)
    Pal = 0.5 + 0.35 * (np.cos(h)[:, None] * u + np.sin(h)[:, None] * v)
    C = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    Wt = rng.normal(0.15, 0.35, (Kc + 2, H, W)).astype(np.float32)
    a = rng.uniform(0, 1, (H, W)).astype(np.float32)
    return np.clip(Pal, 0.02, 0.98).astype(np.float32), C, Wt, a


def run(dev, fn_name, Kc, H, W, seed, clamp, with_alpha):
    Pal, C, Wt, a = inputs(Kc, H, W, seed)
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    Ct, Wtt, Pt = T(C), T(Wt), T(Pal)
    at = T(a) if with_alpha else None
    wb, dW = torch.full_like(Wtt, float("nan")), torch.full_like(Wtt, float("nan"))
    dP, out = torch.full_like(Pt, float("nan")), torch.full((4,), float("nan"), dtype=torch.float32, device=dev)
    args = [W, H, Kc, ptr(Ct), ptr(Wtt), ptr(at), ptr(Pt), 1e-9, int(clamp), 0.5, 1.0, ptr(wb), ptr(dW), ptr(dP),
            ptr(out)]
    if fn_name == "noalloc":
        call("rgp_bary_loss_noalloc", *args, stream(dev))
    else:
        ws = torch.empty(int(lib().rgp_bary_workspace_bytes(Kc)), dtype=torch.uint8, device=dev)
        call("rgp_bary_loss", *args, ptr(ws), stream(dev))
    return [t.cpu().numpy() for t in (wb, dW, dP, out)]


def test_noalloc_equals_workspace_form(device):
    # the second call: smaller image, fewer palette colours (fewer accumulators than the first call left)
    for Kc, H, W, clamp, alpha in ((6, 270, 480, False, True), (3, 61, 97, True, False)):
        print(f"case: bary loss noalloc Kc={Kc} {W}x{H} clamp={clamp} alpha={alpha}", flush=True)
        got = run(device, "noalloc", Kc, H, W, 40 + Kc, clamp, alpha)
        ref = run(device, "workspace", Kc, H, W, 40 + Kc, clamp, alpha)
        assert np.array_equal(got[0], ref[0]), "w_bary"
        assert np.array_equal(got[1], ref[1]), "dw_tilde"
        for name, g, r in (("dpalette", got[2], ref[2]), ("loss", got[3], ref[3])):
            assert np.isfinite(g).all(), name
            assert (np.abs(g.astype(np.float64) - r) <= 1e-6 * np.abs(r) + 1e-30).all(), (name, g, r)


def test_noalloc_refuses_oversize_palette(device):
    z = torch.zeros(64, dtype=torch.float32, device=device)
    rc = lib().rgp_bary_loss_noalloc(8, 8, 7, ptr(z), ptr(z), None, ptr(z), 1e-9, 0, 0.5, 1.0, None, ptr(z), ptr(z), ptr(z),
                 stream(device))
    assert rc != 0 and "Kc" in lib().rgp_last_error().decode()
