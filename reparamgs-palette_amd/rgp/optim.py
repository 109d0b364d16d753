"""Adam for the palette finetune with the full-rank SH update fused into the
backward.

The finetune step (gsplat/finetune.py:346-399) runs torch.optim.Adam over the
SH coefficients after WeightGSFunctionTest.backward has formed
dL/dshs[n, lm*K+k] = dL/dw[n,k] * Y_lm(n) (gsmodel.py:135-144), an (N, 16K)
tensor written once and read back by the optimizer.  PaletteAdam takes dL/dw
(N, K) from that backward instead and applies the update with
rgp_sh_backward_adam_full, which forms each gradient element where it is used.
The arithmetic of the update is torch.optim.Adam's (no amsgrad, no maximize),
element by element:

    g += weight_decay * p
    m += (1 - beta1) * (g - m)
    v  = v * beta2 + ((1 - beta2) * g) * g
    p += -(lr / (1 - beta1^t)) * (m / (sqrt(v) / sqrt(1 - beta2^t) + eps))

Parameters listed in `fused_sh` must be the tensor passed as `shs` to
WeightGSFunctionTest (a float32, contiguous (N, sh_dim*K) CUDA tensor), or a
pair (low, high) passed to WeightGSFunctionSplit ((N,K) and (N,15K), updated
together by rgp_sh_backward_adam_split with the same arithmetic).  Triples
in `fused_lowrank` are the (shs_dc, shs_A, shs_B) of WeightGSFunctionFast
with Q = 15, P = K and rank 1 or 2: rgp_sh_backward_adam_lowrank forms their
gradients from dL/dw (N,K) and updates all three in one kernel.  Every other
parameter, a fused one whose gradient also arrives the usual way
(several backward passes before one step), and the tensors of a fused group
whose step counts differ (one of them was updated alone earlier) are updated
by the same formula:
in one kernel pass (rgp_adam) for contiguous float32 GPU tensors, with torch
operations otherwise.
"""
import math

import torch

from ._lib import I64, call, ptr, stream
from .profiling import SPANS


class PaletteAdam(torch.optim.Optimizer):
    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0, fused_sh=(),
                 fused_lowrank=()):
        defaults = dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay)
        super().__init__(params, defaults)
        self._pending = {}
        self._pair_of = {}  # id(high) -> low, for the pairs of split leaves
        self._triple_of = {}  # id(dc) -> (A, B), for the low-rank triples
        known = {id(p) for g in self.param_groups for p in g["params"]}
        group_of = {id(p): gi for gi, g in enumerate(self.param_groups) for p in g["params"]}
        for triple in fused_lowrank:
            if len(triple) != 3:
                raise ValueError("PaletteAdam: fused_lowrank entries are (shs_dc, shs_A, shs_B)")
            if len({group_of.get(id(p)) for p in triple}) != 1:
                raise ValueError("PaletteAdam: a fused_lowrank triple must sit in one parameter group")
            for p in triple:
                if id(p) not in known:
                    raise ValueError("PaletteAdam: a fused_lowrank tensor is not among the optimised parameters")
                if not (p.is_cuda and p.dtype == torch.float32 and p.is_contiguous()):
                    raise ValueError("PaletteAdam: fused_lowrank tensors must be contiguous float32 on the GPU")
            dc, A, B = triple
            n, K = dc.shape[0], dc.shape[-1]
            if dc.dim() != 2 or A.shape[0] != n or B.shape[0] != n or A.numel() % (n * K) or \
                    B.numel() != (A.numel() // (n * K)) * 15 * n or A.numel() // (n * K) not in (1, 2):
                raise ValueError("PaletteAdam: a low-rank triple must be (N,K), (N,K[,r]) and ([N,r,]15), r = 1 or 2")
            self._triple_of[id(dc)] = (A, B)
            dc._rgp_fused_adam = self
        for entry in fused_sh:
            pair = tuple(entry) if isinstance(entry, (tuple, list)) else (entry,)
            for p in pair:
                if id(p) not in known:
                    raise ValueError("PaletteAdam: a fused_sh tensor is not among the optimised parameters")
                if not (p.is_cuda and p.dtype == torch.float32 and p.dim() == 2 and p.is_contiguous()):
                    raise ValueError("PaletteAdam: fused_sh tensors must be contiguous 2-D float32 on the GPU")
            if len(pair) == 2:
                low, high = pair
                if high.shape[0] != low.shape[0] or high.shape[1] != 15 * low.shape[1]:
                    raise ValueError("PaletteAdam: a split pair must be (N,K) and (N,15K)")
                self._pair_of[id(high)] = low
            pair[0]._rgp_fused_adam = self

    # -- called by WeightGSFunctionTest.backward ---------------------------
    def take(self, p, dw, camd, pws, K, sh_dim):
        """Keep dL/dw for p's update; False when p's gradient must be
        materialised instead (a second backward before the step)."""
        if id(p) in self._pending:
            return False
        self._pending[id(p)] = (dw, camd, pws, int(K), int(sh_dim))
        return True

    def take_split(self, low, high, dw, camd, pws, K, sh_dim):
        """The same for a pair of split leaves (updated together at low's turn)."""
        if id(low) in self._pending or sh_dim != 16 or self._pair_of.get(id(high)) is not low:
            return False
        self._pending[id(low)] = ("split", high, dw, camd, pws, int(K))
        self._pending[id(high)] = ("split_high",)
        return True

    def take_lowrank(self, dc, A, B, dw, camd, pws, K, P, Q, rank, views=1):
        """The same for a low-rank triple (updated together at dc's turn);
        views > 1: dw is (views, N, K) and camd the device array of cameras."""
        if id(dc) in self._pending or Q != 15 or P != K or rank not in (1, 2):
            return False
        tri = self._triple_of.get(id(dc))
        if tri is None or tri[0] is not A or tri[1] is not B:
            return False
        self._pending[id(dc)] = ("lowrank", A, B, dw, camd, pws, int(K), int(rank), int(views))
        self._pending[id(A)] = ("lowrank_part",)
        self._pending[id(B)] = ("lowrank_part",)
        return True

    def zero_grad(self, set_to_none=True):
        self._pending.clear()
        super().zero_grad(set_to_none=set_to_none)

    @staticmethod
    def _materialise(p, pend):
        dw, camd, pws, K, sh_dim = pend
        g = torch.empty_like(p)
        call("rgp_sh_backward_full", p.shape[0], K, sh_dim, ptr(camd), ptr(pws), ptr(dw), ptr(g), stream(p.device))
        return g

    def _same_step(self, *ts):
        """Whether the tensors have taken the same number of Adam steps (a
        fused group update uses one step count for all of them)."""
        return len({self.state[t].get("step", 0) for t in ts}) == 1

    def _state(self, p):
        st = self.state[p]
        if not st:
            st["step"] = 0
            st["exp_avg"] = torch.zeros_like(p, memory_format=torch.preserve_format)
            st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.preserve_format)
        return st

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        extra = {}  # id -> gradient of a split leaf whose pair also holds an ordinary gradient
        for group in self.param_groups:
            for p in group["params"]:
                pend = self._pending.get(id(p))
                if pend is None or pend[0] != "lowrank":
                    continue
                _, A, B, dw, camd, pws, K, r, V = pend
                if p.grad is None and A.grad is None and B.grad is None and self._same_step(p, A, B):
                    continue
                # a second gradient arrived the usual way, or the three tensors have taken
                # different numbers of steps (the fused kernel applies one bias correction):
                # materialise the fused gradients and update each tensor on its own
                n = p.shape[0]
                gdc = torch.empty((n, K), dtype=torch.float32, device=p.device)
                gA, gB = torch.empty_like(A), torch.empty_like(B)
                if V == 1:
                    call("rgp_sh_backward_lowrank", n, K, K, 15, r, ptr(camd), ptr(pws), ptr(A), ptr(B), ptr(dw),
                         ptr(gdc), ptr(gA), ptr(gB), stream(p.device))
                else:
                    call("rgp_sh_backward_lowrank_views", V, n, K, K, 15, r, ptr(camd), ptr(pws), ptr(A), ptr(B),
                         ptr(dw), ptr(gdc), ptr(gA), ptr(gB), stream(p.device))
                extra[id(p)], extra[id(A)], extra[id(B)] = gdc.view_as(p), gA, gB
                for t in (p, A, B):
                    self._pending.pop(id(t), None)
        for group in self.param_groups:
            for p in group["params"]:
                pend = self._pending.get(id(p))
                if pend is None or pend[0] != "split":
                    continue
                _, high, dw, camd, pws, K = pend
                if p.grad is None and high.grad is None and self._same_step(p, high):
                    continue
                full = self._materialise(torch.empty((p.shape[0], 16 * K), dtype=p.dtype, device=p.device),
                                         (dw, camd, pws, K, 16))
                extra[id(p)], extra[id(high)] = full[:, :K].contiguous(), full[:, K:].contiguous()
                self._pending.pop(id(p))
                self._pending.pop(id(high), None)
        for group in self.param_groups:
            lr, (b1, b2), eps, wd = group["lr"], group["betas"], group["eps"], group["weight_decay"]
            for p in group["params"]:
                pend = self._pending.pop(id(p), None)
                if pend is not None and pend[0] in ("split_high", "lowrank_part"):
                    continue  # updated with its low leaf / its dc
                if pend is not None and pend[0] == "lowrank":
                    _, A, B, dw, camd, pws, K, r, V = pend
                    sd, sa, sb = self._state(p), self._state(A), self._state(B)
                    for st_ in (sd, sa, sb):
                        st_["step"] += 1
                    head = ("rgp_sh_backward_adam_lowrank",) if V == 1 else ("rgp_sh_backward_adam_lowrank_views", V)
                    with SPANS.span("sh_backward_adam"):
                        call(*head, p.shape[0], K, K, 15, r, ptr(camd), ptr(pws), ptr(dw),
                             ptr(p), ptr(sd["exp_avg"]), ptr(sd["exp_avg_sq"]), ptr(A), ptr(sa["exp_avg"]),
                             ptr(sa["exp_avg_sq"]), ptr(B), ptr(sb["exp_avg"]), ptr(sb["exp_avg_sq"]),
                             I64(sd["step"]), float(lr), float(b1), float(b2), float(eps), float(wd),
                             stream(p.device))
                    continue
                if pend is not None and pend[0] == "split":
                    _, high, dw, camd, pws, K = pend
                    sl, sh = self._state(p), self._state(high)
                    sl["step"] += 1
                    sh["step"] += 1
                    with SPANS.span("sh_backward_adam"):
                        call("rgp_sh_backward_adam_split", p.shape[0], K, ptr(camd), ptr(pws), ptr(dw), ptr(p),
                             ptr(sl["exp_avg"]), ptr(sl["exp_avg_sq"]), ptr(high), ptr(sh["exp_avg"]),
                             ptr(sh["exp_avg_sq"]), I64(sl["step"]), float(lr), float(b1), float(b2), float(eps),
                             float(wd), stream(p.device))
                    continue
                if id(p) in extra:
                    g = extra.pop(id(p))
                    p.grad = g if p.grad is None else p.grad + g
                if pend is None and p.grad is None:
                    continue
                st = self._state(p)
                st["step"] += 1
                t = st["step"]
                m, v = st["exp_avg"], st["exp_avg_sq"]
                if pend is not None and p.grad is None:
                    dw, camd, pws, K, sh_dim = pend
                    with SPANS.span("sh_backward_adam"):
                        call("rgp_sh_backward_adam_full", p.shape[0], K, sh_dim, ptr(camd), ptr(pws), ptr(dw), ptr(p),
                             ptr(m), ptr(v), I64(t), float(lr), float(b1), float(b2), float(eps), float(wd),
                             stream(p.device))
                    continue
                g = p.grad if pend is None else p.grad + self._materialise(p, pend)
                if (p.is_cuda and p.dtype == torch.float32 and g.dtype == torch.float32 and p.is_contiguous()
                        and g.is_contiguous() and m.is_contiguous() and v.is_contiguous()):
                    # one pass (rgp_adam): the same element update as below
                    call("rgp_adam", I64(p.numel()), ptr(p), ptr(g), ptr(m), ptr(v), I64(t), float(lr), float(b1),
                         float(b2), float(eps), float(wd), stream(p.device))
                    continue
                if wd != 0:
                    g = g.add(p, alpha=wd)
                m.lerp_(g, 1 - b1)
                v.mul_(b2).addcmul_(g, g, value=1 - b2)
                denom = (v.sqrt() / math.sqrt(1 - b2 ** t)).add_(eps)
                p.addcdiv_(m, denom, value=-lr / (1 - b1 ** t))
        return loss
