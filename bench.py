#!/usr/bin/env python
"""Benchmark of the palette-weight rasterizer on one MI355X.

Workload (the finetune step the headline metric is quoted on): 1,000,000
Gaussians in [-3,3]^3, K = 5 palette channels at full rank (16 SH
coefficients per channel), 32 orbit cameras at radius 6, 1920x1080, 60-degree
horizontal field of view.  One step = forward rasterization of the weight image
+ palette mix + 0.8 L1 + 0.2 (1 - SSIM) against a target rendered from the same
Gaussians with a perturbed palette + backward + Adam on the SH coefficients and
the chromatic palette.  Inputs are synthetic (seeded) and resident in HBM
before the timed region.

Printed: one JSON line with fwd+bwd steps/s as `value`, forward frames/s
(render + mix) beside it, the roofline of the dominant kernel measured with
HIP events on its stream during the timed steps, and the CPU oracle timed on
this host on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "reparamgs-palette_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "frames/s @1080p 1M Gaussians K=5; p99 palette-edit latency; fwd+bwd steps/s"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--views", type=int, default=32)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--cpu-views", type=int, default=1, help="cameras in the CPU-oracle sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving measurement (profiling)")
    ap.add_argument("--no-lowrank", action="store_true", help="skip the rank-2 1M / 3M finetune steps")
    ap.add_argument("--no-bary", action="store_true", help="skip the barycentric-loss finetune step")
    ap.add_argument("--bary-only", action="store_true", help="run only the barycentric-loss finetune step (profiling)")
    ap.add_argument("--no-vanilla", action="store_true", help="skip the vanilla 3DGS training step")
    ap.add_argument("--no-c4", action="store_true", help="skip the 6M / 4K / 4-view batched finetune step")
    ap.add_argument("--c4-only", action="store_true", help="run only the 6M / 4K / 4-view step (profiling)")
    ap.add_argument("--vanilla-only", action="store_true", help="run only the vanilla 3DGS training step (profiling)")
    ap.add_argument("--serve-requests", type=int, default=1500)
    ap.add_argument("--serve-rate", type=float, default=500.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy (the known-size kernel that scales "
                         "the FETCH_SIZE / WRITE_SIZE counters in a --pmc profile)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measure_serving(args, dev):
    """Serving workload: six resident scenes (0.2M-3M Gaussians, K 5 or 7, full
    rank or rank 2) under a 200 GiB budget, a Poisson stream of pose and
    palette-only requests at 720p/1080p/1440p.  Latency at the given rate,
    then the saturated throughput of continuous batching and of one request
    at a time on the same render graphs."""
    from rgp.serving import Engine, default_specs, make_requests
    s0 = time.perf_counter()
    specs = default_specs()
    eng = Engine(specs, dev)
    reqs = make_requests(specs, args.serve_requests, rate=args.serve_rate, seed=args.seed + 11)
    eng.prepare(reqs)
    log(f"serving setup {time.perf_counter() - s0:.1f}s ({len(eng.slots)} render graphs)")
    # one untimed pass warms the allocator's frame pool and pinned buffers
    eng.run(reqs, "batched", realtime=False, keep_outputs=False)
    live = eng.run(reqs, "batched", realtime=True, keep_outputs=False)
    sat_b = eng.run(reqs, "batched", realtime=False, keep_outputs=False)
    sat_r = eng.run(reqs, "serial", realtime=False, rerender_edits=True, keep_outputs=False)
    sat_s = eng.run(reqs, "serial", realtime=False, keep_outputs=False)
    out = {"scenes": [f"{s.name}:{s.n}:K{s.K}:{'rank2' if s.lowrank else 'full'}" for s in specs],
           "rate_req_s": args.serve_rate, "requests": len(reqs), "edits": live["edits"],
           "p50_ms": round(live["p50_ms"], 3), "p99_ms": round(live["p99_ms"], 3),
           "edit_p50_ms": round(live["edit_p50_ms"], 3), "edit_p99_ms": round(live["edit_p99_ms"], 3),
           "sustained_frames_per_s": round(live["frames_per_s"], 1),
           "batched_max_frames_per_s": round(sat_b["frames_per_s"], 1),
           "one_at_a_time_frames_per_s": round(sat_r["frames_per_s"], 1),
           "batched_over_one_at_a_time": round(sat_b["frames_per_s"] / sat_r["frames_per_s"], 2),
           "one_at_a_time_cached_frames_per_s": round(sat_s["frames_per_s"], 1),
           "batched_over_one_at_a_time_cached": round(sat_b["frames_per_s"] / sat_s["frames_per_s"], 2),
           "baselines": "one_at_a_time: each request rendered alone and synchronised, a palette edit "
                        "re-rendering its view (the reference's bake + render); _cached: the same but an "
                        "edit re-mixes the cached weight image",
           "hbm_budget_gib": 200, "admitted_gib": round(eng.store.used / 2**30, 2)}
    log(f"serving: p99 {out['p99_ms']} ms, edit p99 {out['edit_p99_ms']} ms, batched "
        f"{out['batched_max_frames_per_s']} vs one-at-a-time {out['one_at_a_time_frames_per_s']} "
        f"(cached {out['one_at_a_time_cached_frames_per_s']}) frames/s")
    del eng
    return out


# bench.py's span names -> the kernel each one times in the default workload
SPAN_KERNEL = {"draw_backward": "rgp::k_draw_bwd_replay<5, true>", "draw_forward": "rgp::k_draw_fwd<5, false, true>",
               "sh_backward_adam": "rgp::k_sh_adam_full<5, 16>", "preprocess": "rgp::k_preprocess<5, 4>"}
PMC_PROFILE = os.path.join(ROOT, "profiles", "pmc_finetune_step_1M_1080p_k5.json")


def pmc_profile(args, span):
    """HBM bytes per launch (FETCH_SIZE / WRITE_SIZE, copy-calibrated) and
    VALU instructions of the kernel behind `span`, from the committed
    rocprofv3 counter profile of this workload; None for other workloads."""
    if (args.n, args.width, args.height) != (1_000_000, 1920, 1080) or span not in SPAN_KERNEL:
        return None
    try:
        with open(PMC_PROFILE) as f:
            prof = json.load(f)
    except OSError:
        return None
    k = prof["kernels"].get(SPAN_KERNEL[span])
    if k is None:
        return None
    return {"traffic_bytes": k["traffic_bytes"], "valu_insts": k["valu_insts"] * 1.0,
            "source": "profiles/pmc_finetune_step_1M_1080p_k5.json (FETCH_SIZE x %.3f + WRITE_SIZE x %.3f)"
                      % (prof["fetch_scale"], prof["write_scale"])}


def measure_lowrank(args, dev, sizes=(1_000_000, 3_000_000), steps=16, warmup=2):
    """Rank-2 low-rank finetune step (WeightGSFunctionFast, K=5, A (N,5,2),
    B (N,2,15)) at 1080p over 8 orbit views, for 1M and 3M Gaussians."""
    import numpy as np
    import torch

    from rgp import gsmodel, ops
    from rgp.losses import L1SSIMLoss
    from rgp.optim import PaletteAdam
    from rgp.scene import make_scene, orbit_cameras

    out = {}
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    for n in sizes:
        sc = make_scene(n, seed=args.seed, extent=3.0, K=5, lowrank=True, rank=2, P=5, Q=15)
        tcams = [c.to_torch(dev) for c in orbit_cameras(8, 6.0, args.width, args.height)]
        pws, rots, scales, alphas = T(sc.pws), T(sc.rots), T(sc.scales), T(sc.alphas)
        dc, A, B = (T(x).requires_grad_(True) for x in (sc.dc, sc.A, sc.B))
        pal = T(sc.palette)
        with torch.no_grad():
            targets = [ops.mix(gsmodel.render_weights(pws, alphas, scales, rots, tc, 5, dc=dc, A=A, B=B, Q=15)[0],
                               torch.clamp(pal + 0.05, 0, 1)) for tc in tcams]
        # torch.optim.Adam's update of the three factor tensors, fused with their backward
        opt = PaletteAdam([dc, A, B], lr=1e-3, fused_lowrank=[(dc, A, B)])

        def step(i):
            tc = tcams[i % len(tcams)]
            layer, _ = gsmodel.WeightGSFunctionFast.apply(pws, dc, A, B, alphas, scales, rots, tc, 5, 15)
            loss = L1SSIMLoss.apply(ops.Mix.apply(layer, pal), targets[i % len(tcams)], 0.2)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()

        for i in range(warmup):
            step(i)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            step(i)
        torch.cuda.synchronize()
        sps = steps / (time.perf_counter() - t0)
        out[f"{n // 1_000_000}M"] = round(sps, 2)
        log(f"rank-2 low-rank fwd+bwd, {n} Gaussians: {sps:.2f} steps/s")
        del sc, pws, rots, scales, alphas, dc, A, B, targets, opt
        torch.cuda.empty_cache()
    return {"steps_per_s": out, "workload": "rank-2 factors A (N,5,2), B (N,2,15), K=5, 1920x1080, 8 orbit views, "
                                            "L1+SSIM, Adam on dc/A/B fused with their backward"}


def measure_c4(args, dev, n=6_000_000, views=4, steps=4, warmup=1):
    """The large finetune configuration: 6M Gaussians, K=7, rank-2 factors
    P=7 x Q=15, 3840x2160, 4 views per step.  Batched: one
    WeightGSFunctionFastViews call (one preprocess launch for the four
    cameras, one compositing launch per direction over the four views' tiles)
    against the same step with one WeightGSFunctionFast call per view; both
    with the palette mix, L1+SSIM over the four images and Adam on dc/A/B."""
    import numpy as np
    import torch

    from rgp import gsmodel, ops
    from rgp.losses import L1SSIMLoss
    from rgp.optim import PaletteAdam
    from rgp.scene import make_scene, orbit_cameras

    K, Q = 7, 15
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sc = make_scene(n, seed=4, extent=3.0, K=K, lowrank=True, rank=2, P=7, Q=Q)
    cams = orbit_cameras(views, 6.0, 3840, 2160)
    tcams = [c.to_torch(dev) for c in cams]
    pws, rots, scales, alphas = T(sc.pws), T(sc.rots), T(sc.scales), T(sc.alphas)
    dc, A, B = (T(x).requires_grad_(True) for x in (sc.dc, sc.A, sc.B))
    pal = T(sc.palette)
    with torch.no_grad():
        lay, _ = gsmodel.render_weights_views(pws, alphas, scales, rots, cams, K, dc=dc, A=A, B=B, Q=Q)
        target = torch.cat([ops.mix(lay[v], torch.clamp(pal + 0.05, 0, 1)) for v in range(views)], 0)
        del lay
    opt = PaletteAdam([dc, A, B], lr=1e-3, fused_lowrank=[(dc, A, B)])  # update fused with the factor backward
    # the one-view-per-launch baseline: its own copies of the factors and an unfused Adam (the four views'
    # gradients are summed by autograd, then one rgp_adam pass per tensor)
    dc1, A1, B1 = (t.detach().clone().requires_grad_(True) for t in (dc, A, B))
    opt1 = PaletteAdam([dc1, A1, B1], lr=1e-3)

    def step_batched():
        layers, _ = gsmodel.WeightGSFunctionFastViews.apply(pws, dc, A, B, alphas, scales, rots, cams, K, Q)
        rgb = torch.cat([ops.Mix.apply(layers[v], pal) for v in range(views)], 0)
        loss = L1SSIMLoss.apply(rgb, target, 0.2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    def step_serial():
        rgb = []
        for tc in tcams:
            layer, _ = gsmodel.WeightGSFunctionFast.apply(pws, dc1, A1, B1, alphas, scales, rots, tc, K, Q)
            rgb.append(ops.Mix.apply(layer, pal))
        loss = L1SSIMLoss.apply(torch.cat(rgb, 0), target, 0.2)
        opt1.zero_grad(set_to_none=True)
        loss.backward()
        opt1.step()

    out = {}
    for name, fn in (("batched", step_batched), ("one_view_per_launch", step_serial)):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        out[name] = round(steps / (time.perf_counter() - t0), 3)
        log(f"C4 6M K=7 rank-2 4K x{views} views, {name}: {out[name]:.3f} steps/s")
    del sc, pws, rots, scales, alphas, dc, A, B, opt, dc1, A1, B1, opt1, target
    torch.cuda.empty_cache()
    return {"steps_per_s": out, "views_per_s": round(out["batched"] * views, 2),
            "workload": f"{n} Gaussians, K=7 rank-2 (A (N,7,2), B (N,2,15)), 3840x2160, {views} views per step; "
                        "mix, L1+SSIM over the views; batched: Adam on dc/A/B fused with their view-summed "
                        "backward; one_view_per_launch: one WeightGSFunctionFast call per view, gradients "
                        "materialised and summed, one Adam pass per tensor"}


def torch_bary_loss(Wt, P, C, tau=1e-9, eps=1e-12, lam_pos=0.5):
    """The finetune step's loss as the reference builds it, with torch
    operations on (HW, K_c, K_c+2) tensors (barycentric_coords_3d,
    gsplat/gsmodel.py:219-285, then finetune.py:368-385): the baseline the
    fused kernel is timed against."""
    import torch
    import torch.nn.functional as F
    Kc = P.shape[0]
    H, W = C.shape[1:]
    dev = P.device
    i, j = torch.arange(Kc, device=dev), (torch.arange(Kc, device=dev) + 1) % Kc
    A = torch.stack([torch.ones(Kc, 3, device=dev), P[i], P[j]], dim=2)
    det = torch.linalg.det(A)
    inv = torch.linalg.inv(A) * (det / (det + eps))[:, None, None]
    rhs = C.permute(1, 2, 0).reshape(-1, 3)
    bgd = torch.einsum("kij,nj->nki", inv, rhs)
    alpha = 1.0 - bgd.sum(-1)
    coords = torch.cat([alpha[..., None], bgd], -1)
    n = rhs.shape[0]
    rows = torch.arange(Kc, device=dev)
    raw = torch.zeros(n, Kc, Kc + 2, device=dev)
    raw[:, :, 0], raw[:, :, 1] = coords[..., 0], coords[..., 1]
    raw[:, rows, i + 2] = coords[..., 2]
    raw[:, rows, j + 2] = coords[..., 3]
    sp = F.softplus(coords)
    q = sp / (sp.sum(-1, keepdim=True) + eps)
    proj = torch.zeros_like(raw)
    proj[:, :, 0], proj[:, :, 1] = q[..., 0], q[..., 1]
    proj[:, rows, i + 2] = q[..., 2]
    proj[:, rows, j + 2] = q[..., 3]
    Paug = torch.cat([torch.zeros(1, 3, device=dev), torch.ones(1, 3, device=dev), P], 0)
    err = ((proj @ Paug - rhs[:, None, :]) ** 2).sum(-1)
    p = F.softmax(-err / tau, dim=1)
    Wb = (p[..., None] * raw).sum(1).T.reshape(Kc + 2, H, W)
    return F.mse_loss(Wt, Wb) + lam_pos * torch.mean(F.relu(-Wt))


def measure_bary(args, dev, n=1_000_000, views=8, steps=96, warmup=4):
    """The reference's own finetune step (finetune.py:346-399): palette-weight
    render, barycentric target of the frozen render, L_geo + L_pos +
    compactness + hue separation, backward, Adam on the SH weights and the
    palette parameters.  Also times the fused loss kernel alone (HIP events on
    the stream it runs on) against the same loss built from torch operations."""
    import numpy as np
    import torch

    from rgp import gsmodel, ops
    from rgp.losses import bary_loss, hue_separation_loss, palette_compactness_loss
    from rgp.optim import PaletteAdam
    from rgp.scene import make_scene, orbit_cameras

    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sc = make_scene(n, seed=args.seed, extent=3.0, K=5)
    tcams = [c.to_torch(dev) for c in orbit_cameras(views, 6.0, args.width, args.height)]
    pws, rots, scales, alphas = T(sc.pws), T(sc.rots), T(sc.scales), T(sc.alphas)
    params = {"low_shs_w": T(sc.shs[:, :5]).requires_grad_(True), "high_shs_w": T(sc.shs[:, 5:]).requires_grad_(True),
              "low_rank": False}
    for k, v in (("hue_base", [0.4]), ("gap_raw", [0.1, -0.2, 0.05]), ("sat_raw", [0.6, 0.3, 0.9]),
                 ("lit_raw", [0.1, -0.1, 0.2])):
        params[k] = torch.tensor(v, dtype=torch.float32, device=dev, requires_grad=True)
    with torch.no_grad():  # frozen renders: the scene's weights mixed with a colour palette
        shs, pal0 = T(sc.shs), T(sc.palette)
        frozen = [torch.clamp(ops.mix(gsmodel.render_weights(pws, alphas, scales, rots, tc, 5, shs=shs)[0], pal0), 0, 1)
                  for tc in tcams]
        del shs
    model = gsmodel.FinetuneModel(num_palette=3, use_fast=False)  # two leaves, rendered without concatenation
    # Adam on every parameter; the two SH leaves take the update fused with their backward
    opt = PaletteAdam([v for v in params.values() if torch.is_tensor(v)], lr=1e-3,
                      fused_sh=[(params["low_shs_w"], params["high_shs_w"])])

    def step(i):
        W_tilde, pal = model(pws, params, alphas, scales, rots, tcams[i % views])
        loss, _ = bary_loss(W_tilde, pal, frozen[i % views])
        loss = loss + 0.001 * palette_compactness_loss(pal) + 1.0 * hue_separation_loss(params["gap_raw"], 0.3)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step(i)
    torch.cuda.synchronize()
    sps = steps / (time.perf_counter() - t0)

    # the loss alone: fused kernel vs torch operations (forward + backward)
    with torch.no_grad():
        W_tilde, pal = model(pws, params, alphas, scales, rots, tcams[0])
    Wt = W_tilde.detach().requires_grad_(True)
    P = pal.detach().requires_grad_(True)
    C = frozen[0]

    def timed(fn, reps=20):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def fused():
        loss, _ = bary_loss(Wt, P, C)
        g = torch.autograd.grad(loss, (Wt, P))
        return g

    def reference():
        loss = torch_bary_loss(Wt, P, C)
        return torch.autograd.grad(loss, (Wt, P))

    ms_fused = timed(fused)
    ms_torch = timed(reference, reps=5)
    gf, gr = fused(), reference()
    hw = args.width * args.height
    nbytes = hw * (12 + 8 * 5)  # C read, W_tilde read, dL/dW_tilde written (K = 5 channels)
    out = {"steps_per_s": round(sps, 2), "loss_fwd_bwd_ms": round(ms_fused, 4),
           "torch_loss_fwd_bwd_ms": round(ms_torch, 3), "loss_speedup_vs_torch": round(ms_torch / ms_fused, 1),
           "loss_kernel_bytes": nbytes, "loss_kernel_GBps": round(nbytes / (ms_fused * 1e-3) / 1e9, 1),
           "max_abs_diff_dW_vs_torch": float((gf[0] - gr[0]).abs().max()),
           "workload": f"finetune.py step: 1M Gaussians, K=5 (3 chromatic), {args.width}x{args.height}, {views} "
                       "views, barycentric target (tau 1e-9) + L_pos + compactness + hue separation, Adam"}
    log(f"barycentric finetune step: {sps:.2f} steps/s; loss fwd+bwd {ms_fused:.3f} ms fused vs "
        f"{ms_torch:.2f} ms torch ({ms_torch / ms_fused:.0f}x)")
    del sc, pws, rots, scales, alphas, params, frozen, opt, Wt, P
    torch.cuda.empty_cache()
    return out


def measure_vanilla(args, dev, n=1_000_000, views=8, steps=16, warmup=2):
    """The vanilla 3DGS training step (vanilla_3dgs/gsplat/gsmodel.py:6-85 and
    train.py's loop): colour render with the calc_J Jacobians, L1+SSIM against a
    target plus an alpha term on final_tau, the colour backward with geometry
    gradients, the Jacobian chain, Adam on pws/shs/alphas/scales/rots."""
    import numpy as np
    import torch

    from rgp.losses import L1SSIMLoss
    from rgp.optim import PaletteAdam
    from rgp.scene import make_scene, orbit_cameras
    from rgp.vanilla import GSFunction

    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    sc = make_scene(n, seed=args.seed, extent=3.0, K=5)
    rng = np.random.default_rng(args.seed + 3)
    tcams = [c.to_torch(dev) for c in orbit_cameras(views, 6.0, args.width, args.height)]
    leaves = [T(sc.pws), T(rng.normal(0, 0.3, (n, 48)).astype(np.float32)), T(sc.alphas.reshape(-1, 1)),
              T(sc.scales), T(sc.rots)]
    for t in leaves:
        t.requires_grad_(True)
    us0 = torch.zeros((n, 2), device=dev)
    with torch.no_grad():
        targets = [GSFunction.apply(*leaves, us0, tc)[0] * 0.9 + 0.05 for tc in tcams]
    opt = PaletteAdam(leaves, lr=1e-3)

    def step(i):
        img, ft, _ = GSFunction.apply(*leaves, us0, tcams[i % views])
        loss = L1SSIMLoss.apply(img, targets[i % views], 0.2) + 0.1 * (1.0 - ft).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step(i)
    torch.cuda.synchronize()
    sps = steps / (time.perf_counter() - t0)
    log(f"vanilla training step, {n} Gaussians: {sps:.2f} steps/s")
    del sc, leaves, targets, opt
    torch.cuda.empty_cache()
    return {"steps_per_s": round(sps, 2),
            "workload": f"vanilla 3DGS step: 1M Gaussians, 48 colour SH coefficients, {args.width}x{args.height}, "
                        f"{views} views, L1+SSIM + final_tau term, geometry gradients, Adam on all five leaves"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measure_cpu(args, scene, cams, pal_target):
    """The CPU oracle (FP32 C++ restatement of the path, OpenMP) timed on this
    host with the GPU step's stages (oracle.timed_step): the finetune step of
    the headline workload on one view at all threads and at one thread, plus
    configs[0] (20k, 256x256; median of 5), the rank-2 low-rank step at
    1M and 3M Gaussians (1080p) and the 6M / K=7 / 3840x2160 / 4-view step.
    The one-thread samples of the large workloads and the 4-view step run the
    compositing stages on a seeded subset of the tiles (1/8; 1/32 at 4K) and
    scale them by the subset's share of the list entries (labelled
    'extrapolated'); every other stage runs in full."""
    import numpy as np

    from oracle import oracle as O
    from rgp.scene import make_scene, orbit_cameras, workload_c0

    threads = O.num_threads()
    comp = ("forward_compositing", "backward_compositing")

    def total(t):
        share = t["entry_share"]
        return sum(v / share if k in comp else v for k, v in t.items() if k != "entry_share")

    def rounded(t):
        return {k: round(v * 1e3, 3) if k != "entry_share" else round(v, 4) for k, v in t.items()}

    def target_of(sc, cam):
        o = O.forward_weight(sc, cam) if sc.n <= 20000 else None
        if o is None:
            return None
        return O.mix(np.clip(sc.palette + 0.05, 0, 1).astype(np.float32), o["layer"])

    def tile_subset(cam, every, seed):
        gx, gy = O.grid_dims(cam.width, cam.height)
        return np.random.default_rng(seed).permutation(gx * gy)[: (gx * gy) // every]

    c0 = time.perf_counter()
    log(f"cpu oracle: headline step on {threads} threads and on 1 thread")
    # the headline step: 1M full rank, 1080p, L1+SSIM against the perturbed-palette target of view 0
    o = O.forward_weight(scene, cams[0])
    target = O.mix(np.asarray(pal_target.cpu()), o["layer"])
    del o
    t_all = O.timed_step(scene, cams[0], target)
    O.set_threads(1)
    t_one = O.timed_step(scene, cams[0], target, tiles=tile_subset(cams[0], 8, 1))
    O.set_threads(threads)
    # median of 5 full steps at each thread count
    log("cpu oracle: 20k Gaussians 256x256, median of 5")
    sc0, cam0 = workload_c0()
    tgt0 = target_of(sc0, cam0)
    c0s = {}
    for th in (threads, 1):
        O.set_threads(th)
        runs = [O.timed_step(sc0, cam0, tgt0) for _ in range(5)]
        c0s[th] = sorted(runs, key=total)[2]
    O.set_threads(threads)
    # rank-2 low-rank, 1M and 3M, 1080p (one view; no SSIM target: a seeded dL/drgb)
    c2 = {}
    for n in (1_000_000, 3_000_000):
        log(f"cpu oracle: {n} Gaussians rank-2 1920x1080")
        sc = make_scene(n, seed=args.seed, extent=3.0, K=5, lowrank=True, rank=2, P=5, Q=15)
        cam = orbit_cameras(8, 6.0, args.width, args.height)[0]
        ta = O.timed_step(sc, cam)
        O.set_threads(1)
        t1 = O.timed_step(sc, cam, tiles=tile_subset(cam, 8, 1))
        O.set_threads(threads)
        c2[f"{n // 1_000_000}M"] = {"all_threads_steps_per_s": round(1.0 / total(ta), 4),
                                     "one_thread_steps_per_s_extrapolated": round(1.0 / total(t1), 5),
                                     "stages_ms_all_threads": rounded(ta), "stages_ms_one_thread_subset": rounded(t1)}
        del sc
    # 6M, K=7 rank-2, 3840x2160, 4 views; compositing on 1/32 of the tiles, extrapolated
    sc = make_scene(6_000_000, seed=4, extent=3.0, K=7, lowrank=True, rank=2, P=7, Q=15)
    log("cpu oracle: 6M K=7 rank-2 3840x2160, 4 views")
    c4_t = [O.timed_step(sc, cam, tiles=tile_subset(cam, 32, 7 + v))
            for v, cam in enumerate(orbit_cameras(4, 6.0, 3840, 2160))]
    del sc
    # one Adam update per 4-view step
    c4_total = sum(total(t) for t in c4_t) - sum(t["adam"] for t in c4_t[1:])
    wall = time.perf_counter() - c0
    return {"value": round(1.0 / total(t_all), 5), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (FP32 C++ restatement of the path, OpenMP over Gaussians, tiles and rows) "
                      "timed on one view of the headline workload: preprocess, scan, key duplication, sort, "
                      f"ranges, forward compositing, mix, L1+SSIM forward and backward, mix backward, backward "
                      f"compositing, SH backward, Adam; {wall:.1f}s for every sample below",
            "value_1_thread_extrapolated": round(1.0 / total(t_one), 5),
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "stages_ms": {"all_threads": rounded(t_all), "one_thread_subset": rounded(t_one)},
            "configs0_20k_256": {"all_threads_steps_per_s": round(1.0 / total(c0s[threads]), 3),
                                 "one_thread_steps_per_s": round(1.0 / total(c0s[1]), 3),
                                 "stages_ms_all_threads": rounded(c0s[threads]),
                                 "stages_ms_one_thread": rounded(c0s[1])},
            "configs2_rank2_1080p": c2,
            "configs4_6M_4K_4views": {"all_threads_steps_per_s_extrapolated": round(1.0 / c4_total, 5),
                                      "method": "compositing stages on a seeded 1/32 of the tiles per view, scaled "
                                                "by 1 / (their share of the list entries); every other stage in "
                                                "full, per view, with one Adam update per step; no SSIM target "
                                                "(seeded dL/drgb)",
                                      "stages_ms_view0_subset": rounded(c4_t[0])}}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from rgp import gsmodel, ops
    from rgp.losses import L1SSIMLoss
    from rgp.optim import PaletteAdam
    from rgp.profiling import SPANS
    from rgp.scene import make_scene, orbit_cameras

    dev = torch.device("cuda:0")
    K = 5
    if args.serve_only:
        print(json.dumps({"serving": measure_serving(args, dev)}))
        return
    if args.bary_only:
        print(json.dumps({"barycentric_step": measure_bary(args, dev)}))
        return
    if args.vanilla_only:
        print(json.dumps({"vanilla_step": measure_vanilla(args, dev)}))
        return
    if args.c4_only:
        print(json.dumps({"large_finetune_4views": measure_c4(args, dev)}))
        return
    if args.copy_calibration:
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        a.fill_(1.0)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    t_setup = time.time()
    scene = make_scene(args.n, seed=args.seed, extent=3.0, K=K)
    cams = orbit_cameras(args.views, 6.0, args.width, args.height)
    tcams = [c.to_torch(dev) for c in cams]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    pws, rots, scales, alphas = T(scene.pws), T(scene.rots), T(scene.scales), T(scene.alphas)
    shs = T(scene.shs).requires_grad_(True)
    pal = T(scene.palette)
    chroma = pal[2:].clone().requires_grad_(True)
    bw = pal[:2].clone()
    rng = np.random.default_rng(args.seed + 7)
    pal_target = np.clip(scene.palette + rng.normal(0, 0.1, scene.palette.shape), 0, 1).astype(np.float32)
    pal_target[:2] = scene.palette[:2]
    pal_target = T(pal_target)

    # targets and per-camera work counts (outside the timed region)
    targets, work = [], []
    with torch.no_grad():
        for tc in tcams:
            layer, depths, st = gsmodel.render_weights(pws, alphas, scales, rots, tc, K, shs=shs.detach(),
                                                       return_state=True)
            targets.append(ops.mix(layer, pal_target))
            _, _, _, e_f, c_f = ops.draw_forward_counts(tc.width, tc.height, K, st["ranges"], st["ids"],
                                                        st["records"], 0.0)
            e_b = int(st["contrib"].sum().item())
            work.append(dict(e_f=e_f, c=c_f, e_b=e_b, pairs=st["m"]))
    torch.cuda.synchronize()
    log(f"setup {time.time() - t_setup:.1f}s; pairs/view mean {np.mean([w['pairs'] for w in work]):.0f}")

    # Adam on the SH coefficients and the chromatic palette; the SH update runs
    # fused with the SH backward (no (N, 80) gradient is stored)
    opt = PaletteAdam([shs, chroma], lr=1e-3, fused_sh=[shs])

    def step(i):
        tc = tcams[i % len(tcams)]
        layer, _ = gsmodel.WeightGSFunctionTest.apply(pws, shs, alphas, scales, rots, tc, K)
        rgb = ops.Mix.apply(layer, torch.cat([bw, chroma], 0))
        loss = L1SSIMLoss.apply(rgb, targets[i % len(tcams)], 0.2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss

    # Every span event is a barrier packet on the stream, so the timed steps
    # record only the dominant kernel's span (picked from the warm-up steps,
    # which record all of them); the per-kernel breakdown comes from an
    # untimed pass after the timed region.
    SPANS.reset()
    SPANS.enabled = True
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    SPANS.enabled = False
    warm = SPANS.summary()
    name = max(warm, key=lambda k: warm[k][1]) if warm else "draw_backward"
    SPANS.reset()
    SPANS.only = {name}
    SPANS.enabled = True
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    SPANS.enabled = False
    SPANS.only = None
    spans = SPANS.summary()
    step_s = (t1 - t0) / args.steps
    SPANS.reset()
    SPANS.enabled = True
    n_break = max(4, min(args.steps, 16))
    b0 = time.perf_counter()
    for i in range(n_break):
        step(args.warmup + i)
    torch.cuda.synchronize()
    b1 = time.perf_counter()
    SPANS.enabled = False
    spans_all = SPANS.summary()
    log(f"fwd+bwd: {args.steps / (t1 - t0):.2f} steps/s ({1e3 * step_s:.2f} ms/step)")

    # forward frames/s: render + mix, no autograd
    with torch.no_grad():
        for i in range(3):
            gsmodel.render_weights(pws, alphas, scales, rots, tcams[i % len(tcams)], K, shs=shs.detach())
        torch.cuda.synchronize()
        f0 = time.perf_counter()
        for i in range(args.frames):
            layer, _ = gsmodel.render_weights(pws, alphas, scales, rots, tcams[i % len(tcams)], K,
                                              shs=shs.detach())
            ops.mix(layer, pal)
        torch.cuda.synchronize()
        f1 = time.perf_counter()
    fps = args.frames / (f1 - f0)
    log(f"forward: {fps:.1f} frames/s")

    # roofline of the dominant kernel over the timed steps
    views = [i % len(tcams) for i in range(args.steps)]
    launches, total_ms, mean_ms = spans[name]
    # algorithmic flops per launch: 12 per evaluated (pixel, entry) pair + (3 + 2K) per contributing one.
    # The replay backward runs the forward's traversal (same stop rule), so its E is the forward's.
    if name in ("draw_forward", "draw_backward"):
        flops = sum(12 * work[v]["e_f"] + (3 + 2 * K) * work[v]["c"] for v in views)
    else:
        flops = None
    breakdown = {k: {"launches": v[0], "mean_ms": round(v[2], 4), "share": round(v[1] / (1e3 * (b1 - b0)), 4)}
                 for k, v in spans_all.items()}
    prof = pmc_profile(args, name)
    if flops is not None:
        achieved = flops / (total_ms * 1e-3) / 1e12
        roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": prof["traffic_bytes"] if prof else None, "kernel": name,
                "mean_launch_ms": round(mean_ms, 4), "launches": launches,
                "flops_total": int(flops), "flops_per_launch": int(flops / max(1, len(views))),
                "E_timed": int(sum(work[v]["e_f"] for v in views)),
                "C_timed": int(sum(work[v]["c"] for v in views)),
                "note": "FP32 VALU-bound compositing (no MFMA; 157.3 TFLOP/s is the FP32 vector peak, packed; "
                        "78.6 without packed FP32). Algorithmic flops = 12 E + (3+2K) C over the timed "
                        "launches: E evaluated (pixel, entry) pairs and C contributing ones, counted on the GPU "
                        "per view (work_per_view). achieved = flops_total / (launches x mean_launch_ms)"}
        if prof:
            # wave64 VALU instructions issue over 4 cycles on a 16-lane SIMD: 1024 SIMDs at 2.4 GHz
            busy_ms = prof["valu_insts"] * 4 / (1024 * 2.4e9) * 1e3
            roof["valu_issue_frac"] = round(busy_ms / mean_ms, 3)
            roof["traffic_source"] = prof["source"]
    else:
        roof = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": name, "mean_launch_ms": round(mean_ms, 4)}

    serving = None if args.no_serve else measure_serving(args, dev)
    lowrank = None if args.no_lowrank else measure_lowrank(args, dev)
    bary = None if args.no_bary else measure_bary(args, dev)
    vanilla = None if args.no_vanilla else measure_vanilla(args, dev)
    c4 = None if args.no_c4 else measure_c4(args, dev)

    cpu = None if args.no_cpu_baseline else measure_cpu(args, scene, cams, pal_target)
    if cpu is not None:
        log(f"cpu oracle: {cpu['value']} steps/s on {cpu['cores']} threads, "
            f"{cpu['value_1_thread_extrapolated']} on 1 thread")

    out = {
        "metric": METRIC, "value": round(args.steps / (t1 - t0), 3), "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * step_s, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "palette finetune step: 1M Gaussians, K=5 full rank, 1920x1080, 32 orbit views, "
                               "L1+SSIM, Adam", "n_gaussians": args.n, "width": args.width,
                   "height": args.height, "views": args.views, "K": K},
        "frames_per_s": round(fps, 2), "ms_per_frame": round(1e3 / fps, 3),
        "pairs_per_view": int(np.mean([w["pairs"] for w in work])),
        "work_per_view": {"E": [int(w["e_f"]) for w in work], "C": [int(w["c"]) for w in work],
                          "M": [int(w["pairs"]) for w in work], "timed_views": views},
        "p99_palette_edit_ms": serving["edit_p99_ms"] if serving else None,
        "serving": serving, "lowrank_rank2": lowrank, "barycentric_step": bary, "vanilla_step": vanilla,
        "large_finetune_4views": c4,
        "kernels": breakdown, "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
