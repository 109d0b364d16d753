"""Autograd wrappers and model glue of the palette-weight finetune, on the
fused gfx950 kernels.

Same classes, signatures and semantics as the reference gsplat/gsmodel.py:
`WeightGSFunctionTest.apply(pws, shs, alphas, scales, rots, cam, K)` and
`WeightGSFunctionFast.apply(pws, dc, A, B, alphas, scales, rots, cam, K, Q)`
return `(layer_W (K,H,W), depths > 0.2)` and their backward returns gradients
for the SH coefficients only (gsmodel.py:104-200).  Instead of the reference's
chain of six extension calls, the forward runs one fused preprocess kernel
(project + covariances + SH weights + inverse + tile rects, writing packed
splat records), the binning kernels and the compositing kernel; the backward
runs the reverse compositing kernel and one SH-chain kernel that recomputes
the basis from the positions rather than storing it.  The values are the ones
of the reference chain.
"""
import ctypes
import math

import torch
import torch.nn.functional as F

from . import ops
from ._lib import call, f32c, I64, ptr, require_cuda, stream
from .profiling import SPANS


def _rec(n, K, dev):
    return torch.empty((n, ops.record_floats(K)), dtype=torch.float32, device=dev)


# How the wrappers' backward forms the per-Gaussian weight gradient:
#  "replay"    : front-to-back replay of the forward (rgp_draw_backward_replay);
#                the terms use the forward's own transmittance, the exact
#                derivative of the computed forward.
#  "reference" : the reference's reverse traversal with the transmittance
#                recovered by division from final_tau (kernel.cu:1342-1378).
BACKWARD_MODE = "replay"


def _weight_grad(W, H, K, n, records, contrib, final_tau, ranges, ids, dlayer, strips=None):
    # the mode in force at the forward decides: a replay-mode forward keeps
    # strip masks and no contrib image, which the reference form would need
    if BACKWARD_MODE == "reference" and contrib is not None:
        return ops.draw_backward(W, H, K, ranges, ids, records, contrib, final_tau, dlayer, n)
    return ops.draw_backward_replay(W, H, K, ranges, ids, records, dlayer, n, strips)


def render_weights(pws, alphas, scales, rots, cam, K, shs=None, dc=None, A=None, B=None, Q=None,
                   sh_dim=None, return_state=False, exact=False, low=None, high=None, strips=False, contrib=True):
    """Forward of the palette-weight rasterizer.  Full rank when `shs`
    (N, sh_dim*K) is given, or the reference's two leaves low (N,K) and high
    (N,(sh_dim-1)K) (the values of their concatenation, without forming it);
    low rank (dc, A (N,P,r), B (N,r,Q)) otherwise.  The pair buffers keep the
    cached capacity unless exact=True (see ops.splat_pipeline); the results
    do not depend on it.  strips=True also records the strip masks of the
    replay backward (state["strips"]); contrib=False skips the contrib image
    (state["contrib"] is None), which the replay backward and plain frames do
    not need (without return_state nothing could read it, so it is skipped
    then too)."""
    require_cuda(pws, alphas, scales, rots)
    dev = pws.device
    n = pws.shape[0]
    W, H = int(cam.width), int(cam.height)
    camd = ops.camera_struct(cam, dev)
    pws_c, rots_c, scales_c, alphas_c = f32c(pws), f32c(rots), f32c(scales), f32c(alphas).reshape(-1)
    depths = torch.empty((n,), dtype=torch.float32, device=dev)
    records = _rec(n, K, dev)
    mode, P_, Q_, r_, shs_c, dc_c, A_c, B_c = _coef_args(K, shs, dc, A, B, Q, sh_dim, low, high)

    def fill(ws, cap):
        call("rgp_preprocess_weight", n, K, mode, P_, Q_, r_, ptr(camd), ptr(pws_c), ptr(rots_c), ptr(scales_c),
             ptr(alphas_c), ptr(shs_c), ptr(dc_c), ptr(A_c), ptr(B_c), ptr(depths), ptr(records), ptr(ws),
             I64(int(cap)), W, H, None, None, None, None, stream(dev))

    layer, contrib, final_tau, ranges, ids, m, _, smask = ops.splat_pipeline(n, W, H, K, records, depths, fill, 0.0,
                                                                             dev, exact=exact, strips=strips,
                                                                             contrib=contrib and return_state)
    if return_state:
        return layer, depths, dict(records=records, contrib=contrib, final_tau=final_tau, ranges=ranges, ids=ids,
                                   camd=camd, m=m, mode=mode, P=P_, Q=Q_, rank=r_, strips=smask)
    return layer, depths


def _coef_args(K, shs, dc, A, B, Q, sh_dim, low=None, high=None):
    """(mode, P, Q, rank, shs, dc, A, B) of the preprocess entry points."""
    if low is not None:  # mode 2: the two leaves, high bands in the shs slot, DC band in the dc slot
        high_c = f32c(high) if high is not None and high.numel() > 0 else None
        sd = 1 + (high.shape[1] // K if high_c is not None else 0)
        return 2, sd, 0, 1, high_c, f32c(low), None, None
    if shs is not None:
        shs_c = f32c(shs)
        sd = shs_c.shape[1] // K if sh_dim is None else sh_dim
        return 0, sd, 0, 1, shs_c, None, None, None
    dc_c, A_c, B_c = f32c(dc), f32c(A), f32c(B)
    P, r = (A_c.shape[1], 1) if A_c.dim() == 2 else (A_c.shape[1], A_c.shape[2])
    return 1, P, int(Q), r, None, dc_c, A_c, B_c


def render_weights_views(pws, alphas, scales, rots, cams, K, shs=None, dc=None, A=None, B=None, Q=None,
                         sh_dim=None, return_state=False, exact=False, strips=False, contrib=True):
    """render_weights for several cameras of one resolution at once (no
    reference counterpart: the reference renders one camera per call).  One
    preprocess launch reads the Gaussians once for all views, and one
    compositing launch covers every view's tiles.  Returns layers (V,K,H,W)
    and depths (V,N); each view's values are those render_weights gives for
    its camera alone."""
    require_cuda(pws, alphas, scales, rots)
    dev = pws.device
    n, V = pws.shape[0], len(cams)
    W, H = int(cams[0].width), int(cams[0].height)
    if any(int(c.width) != W or int(c.height) != H for c in cams):
        raise ValueError("render_weights_views: all cameras must share one resolution")
    camsd = ops.camera_structs(cams, dev)
    pws_c, rots_c, scales_c, alphas_c = f32c(pws), f32c(rots), f32c(scales), f32c(alphas).reshape(-1)
    depths = torch.empty((V, n), dtype=torch.float32, device=dev)
    records = torch.empty((V, n, ops.record_floats(K)), dtype=torch.float32, device=dev)
    mode, P_, Q_, r_, shs_c, dc_c, A_c, B_c = _coef_args(K, shs, dc, A, B, Q, sh_dim)

    def fill(ws, ws_stride, cap):
        call("rgp_preprocess_weight_views", V, n, K, mode, P_, Q_, r_, ptr(camsd), ptr(pws_c), ptr(rots_c),
             ptr(scales_c), ptr(alphas_c), ptr(shs_c), ptr(dc_c), ptr(A_c), ptr(B_c), ptr(depths), ptr(records),
             ptr(ws), ctypes.c_size_t(ws_stride), I64(int(cap)), W, H, stream(dev))

    layer, contrib, final_tau, ranges, ids, ms, _, _, smask = ops.splat_pipeline_views(
        n, V, W, H, K, records, depths, fill, 0.0, dev, exact=exact, strips=strips,
        contrib=contrib and return_state)
    if return_state:
        return layer, depths, dict(records=records, contrib=contrib, final_tau=final_tau, ranges=ranges, ids=ids,
                                   camsd=camsd, m=ms, mode=mode, P=P_, Q=Q_, rank=r_, strips=smask)
    return layer, depths


class WeightGSFunctionTestViews(torch.autograd.Function):
    """WeightGSFunctionTest over a list of cameras of one resolution in one
    batch: returns (layers (V,K,H,W), masks (V,N)); the gradient of the
    full-rank coefficients is the sum over the views."""

    @staticmethod
    def forward(ctx, pws, shs, alphas, scales, rots, cams, num_palette):
        K = int(num_palette)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the unused mask
        layer, depths, st = render_weights_views(pws, alphas, scales, rots, cams, K, shs=shs, return_state=True,
                                                 strips=True, contrib=False)
        ctx.meta = (int(cams[0].width), int(cams[0].height), K, pws.shape[0], st["P"], len(cams))
        ctx.save_for_backward(f32c(pws), st["records"], st["ranges"], st["ids"], st["camsd"], st["strips"])
        return layer, depths > 0.2

    @staticmethod
    def backward(ctx, dloss_dlayer, _):
        if dloss_dlayer is None:
            return (None,) * 7
        W, H, K, n, sh_dim, V = ctx.meta
        pws, records, ranges, ids, camsd, strips = ctx.saved_tensors
        dw = ops.draw_backward_replay_views(V, W, H, K, ranges, ids, records, dloss_dlayer.contiguous(), n, strips)
        dshs = torch.empty((n, sh_dim * K), dtype=torch.float32, device=pws.device)
        with SPANS.span("sh_backward"):
            call("rgp_sh_backward_full_views", V, n, K, sh_dim, ptr(camsd), ptr(pws), ptr(dw), ptr(dshs),
                 stream(pws.device))
        return (None, dshs, None, None, None, None, None)


class WeightGSFunctionFastViews(torch.autograd.Function):
    """WeightGSFunctionFast over a list of cameras of one resolution in one
    batch: returns (layers (V,K,H,W), masks (V,N)); the gradients of dc, A, B
    are sums over the views."""

    @staticmethod
    def forward(ctx, pws, shs_dc, shs_A, shs_B, alphas, scales, rots, cams, num_palette, Q):
        K = int(num_palette)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the unused mask
        layer, depths, st = render_weights_views(pws, alphas, scales, rots, cams, K, dc=shs_dc, A=shs_A, B=shs_B,
                                                 Q=Q, return_state=True, strips=True, contrib=False)
        ctx.meta = (int(cams[0].width), int(cams[0].height), K, pws.shape[0], st["P"], st["Q"], st["rank"],
                    len(cams))
        fused = getattr(shs_dc, "_rgp_fused_adam", None)  # PaletteAdam with the triple in fused_lowrank
        ctx.fused = (fused, shs_dc, shs_A, shs_B) if fused is not None else None
        ctx.save_for_backward(f32c(pws), f32c(shs_A), f32c(shs_B), st["records"], st["ranges"], st["ids"],
                              st["camsd"], st["strips"])
        return layer, depths > 0.2

    @staticmethod
    def backward(ctx, dloss_dlayer, _):
        if dloss_dlayer is None:
            ctx.fused = None
            return (None,) * 10
        W, H, K, n, P, Q, r, V = ctx.meta
        pws, A, B, records, ranges, ids, camsd, strips = ctx.saved_tensors
        dw = ops.draw_backward_replay_views(V, W, H, K, ranges, ids, records, dloss_dlayer.contiguous(), n, strips)
        fused, ctx.fused = ctx.fused, None
        if fused is not None and fused[0].take_lowrank(fused[1], fused[2], fused[3], dw, camsd, pws, K, P, Q, r,
                                                       views=V):
            return (None,) * 10
        ddc = torch.empty((n, K), dtype=torch.float32, device=pws.device)
        dA = torch.empty(A.shape, dtype=torch.float32, device=pws.device)
        dB = torch.empty(B.shape, dtype=torch.float32, device=pws.device)
        with SPANS.span("sh_backward"):
            call("rgp_sh_backward_lowrank_views", V, n, K, P, Q, r, ptr(camsd), ptr(pws), ptr(A), ptr(B), ptr(dw),
                 ptr(ddc), ptr(dA), ptr(dB), stream(pws.device))
        return (None, ddc, dA, dB, None, None, None, None, None, None)


class WeightGSFunctionSplit(torch.autograd.Function):
    """WeightGSFunctionTest on the reference's two coefficient leaves
    (FinetuneModel's low_shs_w (N,K) and high_shs_w (N,15K), which
    gsplat/utils.py:157-158 concatenates before every render): the same
    (layer, mask) and the same gradients, split over the leaves, without the
    concatenation.  An optimizer that fuses the SH update (rgp.optim.PaletteAdam
    with the pair in fused_sh) takes dL/dw instead of the two gradients."""

    @staticmethod
    def forward(ctx, pws, low, high, alphas, scales, rots, cam, num_palette):
        K = int(num_palette)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the unused mask
        layer, depths, st = render_weights(pws, alphas, scales, rots, cam, K, low=low, high=high, return_state=True,
                                           strips=BACKWARD_MODE == "replay",
                                           contrib=BACKWARD_MODE != "replay")
        ctx.meta = (int(cam.width), int(cam.height), K, pws.shape[0], st["P"])
        fused = getattr(low, "_rgp_fused_adam", None)
        ctx.fused = (fused, low, high) if fused is not None else None
        ctx.save_for_backward(f32c(pws), st["records"], st["contrib"], st["final_tau"], st["ranges"], st["ids"],
                              st["camd"], st["strips"])
        return layer, depths > 0.2

    @staticmethod
    def backward(ctx, dloss_dlayer, _):
        if dloss_dlayer is None:
            ctx.fused = None
            return (None,) * 8
        W, H, K, n, sh_dim = ctx.meta
        pws, records, contrib, final_tau, ranges, ids, camd, strips = ctx.saved_tensors
        dw = _weight_grad(W, H, K, n, records, contrib, final_tau, ranges, ids, dloss_dlayer.contiguous(), strips)
        if ctx.fused is not None and ctx.fused[0].take_split(ctx.fused[1], ctx.fused[2], dw, camd, pws, K, sh_dim):
            ctx.fused = None
            return (None, None, None, None, None, None, None, None)
        ctx.fused = None
        dshs = torch.empty((n, sh_dim * K), dtype=torch.float32, device=pws.device)
        with SPANS.span("sh_backward"):
            call("rgp_sh_backward_full", n, K, sh_dim, ptr(camd), ptr(pws), ptr(dw), ptr(dshs), stream(pws.device))
        return (None, dshs[:, :K], dshs[:, K:], None, None, None, None, None)


class WeightGSFunctionTest(torch.autograd.Function):
    """Full-rank palette weights (gsmodel.py:104-144)."""

    @staticmethod
    def forward(ctx, pws, shs, alphas, scales, rots, cam, num_palette):
        K = int(num_palette)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the unused mask
        layer, depths, st = render_weights(pws, alphas, scales, rots, cam, K, shs=shs, return_state=True,
                                           strips=BACKWARD_MODE == "replay",
                                           contrib=BACKWARD_MODE != "replay")
        ctx.meta = (int(cam.width), int(cam.height), K, pws.shape[0], st["P"])
        # an optimizer that fuses the SH update (rgp.optim.PaletteAdam) takes dL/dw instead of dL/dshs
        fused = getattr(shs, "_rgp_fused_adam", None)
        ctx.fused = (fused, shs) if fused is not None else None
        ctx.save_for_backward(f32c(pws), st["records"], st["contrib"], st["final_tau"], st["ranges"], st["ids"],
                              st["camd"], st["strips"])
        return layer, depths > 0.2

    @staticmethod
    def backward(ctx, dloss_dlayer, _):
        if dloss_dlayer is None:
            ctx.fused = None
            return (None,) * 7
        W, H, K, n, sh_dim = ctx.meta
        pws, records, contrib, final_tau, ranges, ids, camd, strips = ctx.saved_tensors
        dw = _weight_grad(W, H, K, n, records, contrib, final_tau, ranges, ids, dloss_dlayer.contiguous(), strips)
        if ctx.fused is not None and ctx.fused[0].take(ctx.fused[1], dw, camd, pws, K, sh_dim):
            ctx.fused = None
            return (None, None, None, None, None, None, None)
        ctx.fused = None
        dshs = torch.empty((n, sh_dim * K), dtype=torch.float32, device=pws.device)
        with SPANS.span("sh_backward"):
            call("rgp_sh_backward_full", n, K, sh_dim, ptr(camd), ptr(pws), ptr(dw), ptr(dshs), stream(pws.device))
        return (None, dshs, None, None, None, None, None)


class WeightGSFunctionFast(torch.autograd.Function):
    """Low-rank palette weights (gsmodel.py:151-200), rank r factors
    A (N,P,r) and B (N,r,Q) or the reference's rank-1 (N,P) and (N,Q)."""

    @staticmethod
    def forward(ctx, pws, shs_dc, shs_A, shs_B, alphas, scales, rots, cam, num_palette, Q):
        K = int(num_palette)
        ctx.set_materialize_grads(False)  # no zero-filled gradient for the unused mask
        layer, depths, st = render_weights(pws, alphas, scales, rots, cam, K, dc=shs_dc, A=shs_A, B=shs_B, Q=Q,
                                           return_state=True, strips=BACKWARD_MODE == "replay",
                                           contrib=BACKWARD_MODE != "replay")
        ctx.meta = (int(cam.width), int(cam.height), K, pws.shape[0], st["P"], st["Q"], st["rank"])
        # an optimizer that fuses the factor update (rgp.optim.PaletteAdam with the triple in
        # fused_lowrank) takes dL/dw instead of the three gradients
        fused = getattr(shs_dc, "_rgp_fused_adam", None)
        ctx.fused = (fused, shs_dc, shs_A, shs_B) if fused is not None else None
        ctx.save_for_backward(f32c(pws), f32c(shs_A), f32c(shs_B), st["records"], st["contrib"], st["final_tau"],
                              st["ranges"], st["ids"], st["camd"], st["strips"])
        return layer, depths > 0.2

    @staticmethod
    def backward(ctx, dloss_dlayer, _):
        if dloss_dlayer is None:
            ctx.fused = None
            return (None,) * 10
        W, H, K, n, P, Q, r = ctx.meta
        pws, A, B, records, contrib, final_tau, ranges, ids, camd, strips = ctx.saved_tensors
        dw = _weight_grad(W, H, K, n, records, contrib, final_tau, ranges, ids, dloss_dlayer.contiguous(), strips)
        fused, ctx.fused = ctx.fused, None
        if fused is not None and fused[0].take_lowrank(fused[1], fused[2], fused[3], dw, camd, pws, K, P, Q, r):
            return (None,) * 10
        ddc = torch.empty((n, K), dtype=torch.float32, device=pws.device)
        dA = torch.empty(A.shape, dtype=torch.float32, device=pws.device)
        dB = torch.empty(B.shape, dtype=torch.float32, device=pws.device)
        with SPANS.span("sh_backward"):
            call("rgp_sh_backward_lowrank", n, K, P, Q, r, ptr(camd), ptr(pws), ptr(A), ptr(B), ptr(dw), ptr(ddc),
                 ptr(dA), ptr(dB), stream(pws.device))
        return (None, ddc, dA, dB, None, None, None, None, None, None)


# ---------------------------------------------------------------------------
# model glue, restated from gsplat/gsmodel.py (torch code, off the kernel path)
# ---------------------------------------------------------------------------
def hsl_to_rgb(h, s, l):
    """gsmodel.py:33-49."""
    h6 = h / (2 * math.pi) * 6.0
    C = (1.0 - torch.abs(2.0 * l - 1.0)) * s
    X = C * (1.0 - torch.abs(h6 % 2.0 - 1.0))
    m = l - C / 2.0
    z = torch.zeros_like(C)
    sectors = torch.stack([torch.stack([C, X, z], -1), torch.stack([X, C, z], -1), torch.stack([z, C, X], -1),
                           torch.stack([z, X, C], -1), torch.stack([X, z, C], -1), torch.stack([C, z, X], -1)], 1)
    idx = h6.long().clamp(0, 5)
    rgb1 = sectors[torch.arange(len(h), device=h.device), idx]
    return (rgb1 + m.unsqueeze(-1)).clamp(0.0, 1.0)


def get_palette_rgb(hue_base, gap_raw, sat_raw, lit_raw):
    """gsmodel.py:203-212."""
    two_pi = 2 * math.pi
    g = F.softplus(gap_raw)
    gaps = two_pi * g / g.sum()
    hues = ((hue_base % two_pi) + torch.cumsum(F.pad(gaps[:-1], (1, 0)), 0)) % two_pi
    return hsl_to_rgb(hues, torch.sigmoid(sat_raw), torch.sigmoid(lit_raw))


def reconstruct_high_shs(params):
    """gsmodel.py:90-97, rank r: C = A @ B per Gaussian, flattened."""
    if params["low_rank"]:
        A, B = params["high_shs_A"], params["high_shs_B"]
        n = A.shape[0]
        if A.dim() == 2:
            return (A.unsqueeze(2) * B.unsqueeze(1)).reshape(n, -1)
        return torch.bmm(A, B.contiguous()).reshape(n, -1)
    return params["high_shs_w"]


def get_shs(low_shs, high_shs):
    """gsplat/utils.py:157-158."""
    return torch.cat((low_shs, high_shs), dim=1)


class FinetuneModel(torch.nn.Module):
    """gsmodel.py:292-329.  split_leaves: full-rank coefficients held as two
    leaves render through WeightGSFunctionSplit instead of concatenating them
    (get_shs) every forward; the values are the same."""

    def __init__(self, sh_order=3, num_palette=4, use_fast=True, split_leaves=True):
        super().__init__()
        self.sh_order = sh_order
        self.num_palette = num_palette
        self.splat_dim = num_palette + 2
        self.use_fast = use_fast
        self.split_leaves = split_leaves
        self.mask = None

    def forward(self, pws, params, alphas, scales, rots, cam):
        if params["low_rank"] and self.use_fast:
            W_tilde, self.mask = WeightGSFunctionFast.apply(pws, params["low_shs_w"], params["high_shs_A"],
                                                            params["high_shs_B"], alphas, scales, rots, cam,
                                                            self.splat_dim, params["reshape_Q"])
        elif not params["low_rank"] and self.split_leaves:
            W_tilde, self.mask = WeightGSFunctionSplit.apply(pws, params["low_shs_w"], params["high_shs_w"], alphas,
                                                             scales, rots, cam, self.splat_dim)
        else:
            shs_w = get_shs(params["low_shs_w"], reconstruct_high_shs(params))
            W_tilde, self.mask = WeightGSFunctionTest.apply(pws, shs_w, alphas, scales, rots, cam, self.splat_dim)
        palette_rgb = get_palette_rgb(params["hue_base"], params["gap_raw"], params["sat_raw"], params["lit_raw"])
        return W_tilde, palette_rgb


def bake_palette_edit(params, palette_rgb_new=None):
    """gsmodel.py:488-509: per-Gaussian colour SH = weight SH @ P_full."""
    with torch.no_grad():
        P_chroma = palette_rgb_new if palette_rgb_new is not None else get_palette_rgb(
            params["hue_base"], params["gap_raw"], params["sat_raw"], params["lit_raw"])
        dev, dt = P_chroma.device, P_chroma.dtype
        P_full = torch.cat([torch.zeros(1, 3, device=dev, dtype=dt), torch.ones(1, 3, device=dev, dtype=dt),
                            P_chroma], 0)
        shs_w = get_shs(params["low_shs_w"], reconstruct_high_shs(params))
        n, K = shs_w.shape[0], P_full.shape[0]
        num_sh = shs_w.shape[1] // K
        baked = torch.einsum("nlk,kc->nlc", shs_w.reshape(n, num_sh, K), P_full)
        return baked.reshape(n, num_sh * 3)


@torch.no_grad()
def render_vanilla(pws, shs, alphas, scales, rots, cam):
    """gsmodel.py:13-26: the colour path through the gsplatcu-compatible ops."""
    import gsplatcu as gsc
    us, pcs, depths = gsc.project(pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, False)
    cov3ds, = gsc.computeCov3D(rots, scales, depths, False)
    cov2ds, = gsc.computeCov2D(cov3ds, pcs, cam.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height, False)
    colors, = gsc.sh2Color(shs, pws, cam.twc, False)
    cinv2ds, areas = gsc.inverseCov2D(cov2ds, depths, False)
    image, _, _, _, _ = gsc.splat_color(cam.height, cam.width, us, cinv2ds, alphas, depths, colors, areas)
    return image
