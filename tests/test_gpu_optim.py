"""The full-rank SH backward fused with Adam (rgp_sh_backward_adam_full,
rgp.optim.PaletteAdam) against torch.optim.Adam applied to the oracle's
gradient dL/dshs[n, lm*K+k] = dL/dw[n,k] * Y_lm(n) (gsmodel.py:135-144: one
float32 product per element, Y from the oracle's sh2Weight basis): same
dL/dw in, parameters and both moments out.

torch's elementwise kernels may contract a multiply-add into one rounding
where the fused kernel rounds twice: a few float32 ulp of the operands, which
for a value near 0 (a parameter crossing zero, p - update cancelling) is
large relative to the value itself.  The bar is 1e-6 * (|value| + the RMS of
the tensor)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rgp import gsmodel
from rgp._lib import I64, call, ptr, stream
from rgp.optim import PaletteAdam
from rgp.scene import workload_c0

pytestmark = pytest.mark.gpu


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def oracle_dshs(pws, twc, dw, K, sh_dim, dev):
    """dL/dshs from dL/dw (N,K) with the oracle's SH basis, in float32."""
    dw = np.asarray(dw, np.float32)
    n = dw.shape[0]
    _, basis = O.sh2weight(np.zeros((n, sh_dim * K), np.float32), np.asarray(pws, np.float32), twc, K)
    g = (dw[:, None, :] * basis.reshape(n, sh_dim)[:, :, None].astype(np.float32)).reshape(n, sh_dim * K)
    return T(g.astype(np.float32), dev)


def close(a, b, what, rtol=1e-6):
    a, b = a.detach().cpu().numpy().astype(np.float64), b.detach().cpu().numpy().astype(np.float64)
    err = np.abs(a - b)
    lim = rtol * (np.abs(b) + np.sqrt(np.mean(b * b))) + 1e-30
    assert (err <= lim).all(), f"{what}: max err/lim {float(np.max(err / lim)):.3g}"


@pytest.mark.parametrize("sh_dim,K,wd", [(16, 5, 0.0), (16, 7, 1e-2), (9, 3, 0.0)])
def test_fused_kernel_matches_torch_adam(device, sh_dim, K, wd):
    print(f"case: fused adam sh_dim={sh_dim} K={K} wd={wd}", flush=True)
    from rgp import ops
    scene, cam = workload_c0()
    n = 5000 + 37  # a partial last block
    rng = np.random.default_rng(3)
    pws = T(scene.pws[:n], device)
    camd = ops.camera_struct(cam.to_torch(device), device)
    p_f = T(rng.normal(0, 0.1, (n, sh_dim * K)).astype(np.float32), device)
    m_f, v_f = torch.zeros_like(p_f), torch.zeros_like(p_f)
    p_t = p_f.clone().requires_grad_(True)
    ref = torch.optim.Adam([p_t], lr=3e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=wd, foreach=False)
    for step in range(1, 4):
        dw = T(rng.normal(0, 1e-3, (n, K)).astype(np.float32), device)
        call("rgp_sh_backward_adam_full", n, K, sh_dim, ptr(camd), ptr(pws), ptr(dw), ptr(p_f), ptr(m_f), ptr(v_f),
             I64(step), 3e-3, 0.9, 0.999, 1e-8, wd, stream(device))
        p_t.grad = oracle_dshs(scene.pws[:n], cam.twc, dw.cpu().numpy(), K, sh_dim, device)
        ref.step()
    st = ref.state[p_t]
    # with weight decay g = dw*Y + wd*p can cancel to near 0, where Adam's
    # normalised step m / sqrt(v) turns an ulp of g into a visible change
    rtol = 1e-6 if wd == 0 else 1e-5
    close(p_f, p_t, "param", rtol)
    close(m_f, st["exp_avg"], "exp_avg", rtol)
    close(v_f, st["exp_avg_sq"], "exp_avg_sq", rtol)


def test_palette_adam_in_the_finetune_step(device, monkeypatch):
    """WeightGSFunctionTest hands dL/dw to PaletteAdam (no dL/dshs is
    stored, shs.grad stays None); two steps equal torch.optim.Adam fed the
    materialised gradient of the same dL/dw."""
    from rgp import ops
    scene, cam = workload_c0()
    tc = cam.to_torch(device)
    seen = []
    take = PaletteAdam.take

    def spy(self, p, dw, camd, pws, K, sh_dim):
        seen.append((dw.clone(), camd, pws))
        return take(self, p, dw, camd, pws, K, sh_dim)

    monkeypatch.setattr(PaletteAdam, "take", spy)
    args = [T(x, device) for x in (scene.alphas, scene.scales, scene.rots)]
    pws = T(scene.pws, device)
    shs = T(scene.shs, device).requires_grad_(True)
    opt = PaletteAdam([shs], lr=1e-2, fused_sh=[shs])
    shs_ref = shs.detach().clone().requires_grad_(True)
    ref = torch.optim.Adam([shs_ref], lr=1e-2, foreach=False)
    G = T(np.random.default_rng(1).normal(size=(scene.K, cam.height, cam.width)).astype(np.float32), device)
    for step in range(2):
        layer, _ = gsmodel.WeightGSFunctionTest.apply(pws, shs, args[0], args[1], args[2], tc, scene.K)
        (layer * G).sum().backward()
        assert shs.grad is None and len(seen) == step + 1
        opt.step()
        opt.zero_grad()
        dw = seen[-1][0]
        shs_ref.grad = oracle_dshs(scene.pws, cam.twc, dw.cpu().numpy(), scene.K, 16, device)
        ref.step()
    close(shs, shs_ref, "shs")
    close(opt.state[shs]["exp_avg"], ref.state[shs_ref]["exp_avg"], "exp_avg")
    close(opt.state[shs]["exp_avg_sq"], ref.state[shs_ref]["exp_avg_sq"], "exp_avg_sq")
    # a second backward before the step falls back to the materialised gradient
    for _ in range(2):
        layer, _ = gsmodel.WeightGSFunctionTest.apply(pws, shs, args[0], args[1], args[2], tc, scene.K)
        (layer * G).sum().backward()
    assert shs.grad is not None
    opt.step()
    assert torch.isfinite(shs).all()


@pytest.mark.parametrize("n,offset,wd", [(1 << 20, 0, 0.0), (1000003, 0, 1e-2), (4099, 1, 0.0)])
def test_flat_adam_matches_torch_adam(device, n, offset, wd):
    """rgp_adam (one pass; float4 when aligned, the scalar path at a 4-byte
    offset and for the tail) against torch.optim.Adam over three steps."""
    print(f"case: rgp_adam n={n} offset={offset} wd={wd}", flush=True)
    rng = np.random.default_rng(7)
    base = torch.empty(n + offset, dtype=torch.float32, device=device)
    p_f = base[offset:]
    p_f.copy_(T(rng.normal(0, 0.1, n).astype(np.float32), device))
    m_f, v_f = torch.zeros_like(p_f), torch.zeros_like(p_f)
    p_t = p_f.clone().requires_grad_(True)
    ref = torch.optim.Adam([p_t], lr=3e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=wd, foreach=False)
    for step in range(1, 4):
        g = T(rng.normal(0, 1e-2, n).astype(np.float32), device)
        call("rgp_adam", I64(n), ptr(p_f), ptr(g), ptr(m_f), ptr(v_f), I64(step), 3e-3, 0.9, 0.999, 1e-8, wd,
             stream(device))
        p_t.grad = g.clone()
        ref.step()
    torch.cuda.synchronize()
    st = ref.state[p_t]
    # torch rounds g + wd * p once (a contracted multiply-add), the kernel twice
    tol = 1e-6 if wd == 0 else 1e-5
    close(p_f, p_t, "param", tol)
    close(m_f, st["exp_avg"], "exp_avg", tol)
    close(v_f, st["exp_avg_sq"], "exp_avg_sq", tol)
