"""Tensor-level wrappers over the librgp C ABI.

Outputs are allocated with torch's caching allocator on the input's device and
the kernels run on torch's current stream.  Functions that must know the
number of (tile, Gaussian) pairs read it back once, as the reference does
(gsplatcu/gausplat.cu:201); the capacity-based variants used for graph capture
never synchronise.
"""
import ctypes

import numpy as np
import torch

from ._lib import I64, call, f32c, i32c, lib, ptr, require_cuda, stream
from .profiling import SPANS

TILE = 16


def grid_dims(width, height):
    return (width + TILE - 1) // TILE, (height + TILE - 1) // TILE


def record_floats(dim):
    return lib().rgp_record_floats(int(dim))


# --------------------------------------------------------------------------
# camera
# --------------------------------------------------------------------------
def camera_struct_host(Rcw, tcw, twc, fx, fy, cx, cy, width, height):
    """The rgp_camera struct of include/rgp.h as 32 float32 words."""
    f32 = np.float32
    buf = np.zeros(32, f32)
    buf[0:9] = np.asarray(Rcw, f32).reshape(9)
    buf[9:12] = np.asarray(tcw, f32).reshape(3)
    buf[12:15] = np.asarray(twc, f32).reshape(3)
    buf[15:19] = [f32(fx), f32(fy), f32(cx), f32(cy)]
    # gausplat.cu:357-358: width / (2 * focal_x) in float
    buf[19] = f32(f32(width) / (f32(2) * f32(fx)))
    buf[20] = f32(f32(height) / (f32(2) * f32(fy)))
    buf[21:23] = np.array([width, height], np.int32).view(f32)
    return buf


def camera_struct(cam, device):
    """Device rgp_camera for a reference-style camera object (Rcw, tcw, twc
    tensors or arrays; fx, fy, cx, cy floats; width, height ints).  Cached on
    the object and rebuilt when any field changes."""
    def val(x):
        if isinstance(x, torch.Tensor):
            return x.detach().float().cpu().numpy()
        return np.asarray(x, np.float32)

    key_parts = []
    for name in ("Rcw", "tcw", "twc"):
        x = getattr(cam, name)
        if isinstance(x, torch.Tensor):
            key_parts.append((id(x), x._version))
        else:
            key_parts.append(np.asarray(x, np.float32).tobytes())
    key = (tuple(key_parts), float(cam.fx), float(cam.fy), float(cam.cx), float(cam.cy),
           int(cam.width), int(cam.height), str(device))
    cached = getattr(cam, "_rgp_cam", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    host = camera_struct_host(val(cam.Rcw), val(cam.tcw), val(cam.twc), cam.fx, cam.fy, cam.cx,
                              cam.cy, cam.width, cam.height)
    dev = torch.from_numpy(host).to(device)
    try:
        cam._rgp_cam = (key, dev)
    except AttributeError:
        pass
    return dev


# --------------------------------------------------------------------------
# binning
# --------------------------------------------------------------------------
def workspace_bytes(n, capacity, width, height):
    return int(lib().rgp_bin_workspace_bytes(int(n), I64(int(capacity)), int(width), int(height)))


def new_workspace(n, capacity, width, height, device):
    return torch.empty(workspace_bytes(n, capacity, width, height), dtype=torch.uint8, device=device)


def header(ws, words=4):
    """Header words [M, overflow, dq_max, M_eff] as a device int32 view."""
    return ws[: 4 * 64].view(torch.int32)[:words]


def bin_sort(n, depths, width, height, ws, capacity, device, keys=False):
    """createKeys + sort + getRanges into fresh ids (capacity,) / ranges."""
    gx, gy = grid_dims(width, height)
    ids = torch.empty(max(int(capacity), 1), dtype=torch.int32, device=device)
    ranges = torch.empty((gx * gy, 2), dtype=torch.int32, device=device)
    keys_t = torch.empty(max(int(capacity), 1), dtype=torch.int64, device=device) if keys else None
    call("rgp_bin_sort", int(n), ptr(depths), int(width), int(height), ptr(ws), I64(int(capacity)),
         ptr(ids), ptr(keys_t), ptr(ranges), stream(device))
    return ids, ranges, keys_t


def pack_records(us, cinv2ds, alphas, feat):
    require_cuda(us, cinv2ds, alphas, feat)
    n, dim = feat.shape[0], feat.shape[-1]
    rec = torch.empty((n, record_floats(dim)), dtype=torch.float32, device=us.device)
    call("rgp_pack_records", int(n), int(dim), ptr(f32c(us)), ptr(f32c(cinv2ds)),
         ptr(f32c(alphas).reshape(-1)), ptr(f32c(feat).reshape(n, dim)), ptr(rec), stream(us.device))
    return rec


def draw_forward(width, height, dim, ranges, ids, records, empty_tau, strips=None, contrib=True):
    """Forward compositing; strips (uint8, one per list position) receives,
    when given, each composited entry's mask of row strips with a
    contribution, which draw_backward_replay then uses.  contrib=False: the
    contrib image is not formed and None is returned for it (the training
    forward with strips, and frames that need only the images)."""
    dev = records.device
    layer = torch.empty((dim, height, width), dtype=torch.float32, device=dev)
    contrib = torch.empty((height, width), dtype=torch.int32, device=dev) if contrib else None
    final_tau = torch.empty((height, width), dtype=torch.float32, device=dev)
    with SPANS.span("draw_forward"):
        call("rgp_draw_forward", int(width), int(height), int(dim), ptr(ranges), ptr(ids), ptr(records),
             float(empty_tau), ptr(layer), ptr(contrib), ptr(final_tau), ptr(strips), stream(dev))
    return layer, contrib, final_tau


def draw_forward_counts(width, height, dim, ranges, ids, records, empty_tau):
    """Forward compositing that also counts, over all pixels, the list entries
    evaluated before each pixel stops (E) and the contributing ones (C)."""
    dev = records.device
    layer = torch.empty((dim, height, width), dtype=torch.float32, device=dev)
    contrib = torch.empty((height, width), dtype=torch.int32, device=dev)
    final_tau = torch.empty((height, width), dtype=torch.float32, device=dev)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    call("rgp_draw_forward_counts", int(width), int(height), int(dim), ptr(ranges), ptr(ids), ptr(records),
         float(empty_tau), ptr(layer), ptr(contrib), ptr(final_tau), ptr(counts), stream(dev))
    e, c = counts.tolist()
    return layer, contrib, final_tau, e, c


def draw_backward(width, height, dim, ranges, ids, records, contrib, final_tau, dlayer, n):
    dev = records.device
    dfeat = torch.zeros((n, dim), dtype=torch.float32, device=dev)
    ranges, ids, contrib, final_tau, dlayer = i32c(ranges), i32c(ids), i32c(contrib), f32c(final_tau), f32c(dlayer)
    with SPANS.span("draw_backward"):
        call("rgp_draw_backward", int(width), int(height), int(dim), ptr(ranges), ptr(ids), ptr(records),
             ptr(contrib), ptr(final_tau), ptr(dlayer), ptr(dfeat), stream(dev))
    return dfeat


def draw_backward_replay(width, height, dim, ranges, ids, records, dlayer, n, strips=None):
    """Gradient of the features by a front-to-back replay of the forward
    (strips: the forward's strip masks, or None to evaluate every reachable
    strip)."""
    dev = records.device
    dfeat = torch.zeros((n, dim), dtype=torch.float32, device=dev)
    ranges, ids, dlayer = i32c(ranges), i32c(ids), f32c(dlayer)
    with SPANS.span("draw_backward"):
        call("rgp_draw_backward_replay", int(width), int(height), int(dim), ptr(ranges), ptr(ids), ptr(records),
             ptr(strips), ptr(dlayer), ptr(dfeat), stream(dev))
    return dfeat


def draw_color_backward(width, height, ranges, ids, records, contrib, final_tau, dimage, dfinal_tau, n):
    """Colour backward with geometry gradients (rgp_draw_color_backward):
    dL/dus (N,2), dL/dcinv2ds (N,3), dL/dalphas (N,), dL/dcolors (N,3)."""
    dev = records.device
    dus = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    dcinv = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    dalpha = torch.zeros((n,), dtype=torch.float32, device=dev)
    dcolor = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    ranges, ids, contrib, final_tau, dimage = i32c(ranges), i32c(ids), i32c(contrib), f32c(final_tau), f32c(dimage)
    dtau = None if dfinal_tau is None else f32c(dfinal_tau)
    with SPANS.span("draw_color_backward"):
        call("rgp_draw_color_backward", int(width), int(height), ptr(ranges), ptr(ids), ptr(records), ptr(contrib),
             ptr(final_tau), ptr(dimage), ptr(dtau), ptr(dus), ptr(dcinv), ptr(dalpha), ptr(dcolor), stream(dev))
    return dus, dcinv, dalpha, dcolor


class CapacityCache:
    """Pair capacity per (n, width, height): grows to 1.25x the largest pair
    count seen, so repeated renders of a scene reuse one workspace size."""

    def __init__(self):
        self.cap = {}

    def get(self, n, width, height):
        return self.cap.get((n, width, height), max(1024, 4 * n))

    def update(self, n, width, height, m):
        key = (n, width, height)
        if m > self.cap.get(key, 0) or key not in self.cap:
            self.cap[key] = max(int(m * 1.25) + 1024, self.cap.get(key, 0))


CAPACITY = CapacityCache()


def splat_pipeline(n, width, height, dim, records, depths, ws_fill, empty_tau, device,
                   exact=True, strips=False, contrib=True):
    """Scan -> sort -> ranges -> draw.  ws_fill(ws, cap) must write
    rects/counts into a workspace (the getRects stage or the fused
    preprocess).

    exact=True: the pair count M is read back right after the scan and the
    pair buffers are sized from it (one host sync, as gausplat.cu:201).
    exact=False: the pair buffers take the cached capacity, the sort and draw
    are launched at once, and only then is the overflow flag read (the host
    waits for the scan while the GPU runs the sort and draw); on overflow the
    pipeline runs again with the grown capacity.  Both give the same M,
    ranges, ids[:M] and images.  strips=True: the forward also writes the
    strip masks of the replay backward (uint8, one per pair slot); they are
    returned last (None otherwise).  contrib=False: no contrib image (None in
    its place)."""
    cap = CAPACITY.get(n, width, height)
    while True:
        ws = new_workspace(n, cap, width, height, device)
        with SPANS.span("preprocess"):
            ws_fill(ws, cap)
        with SPANS.span("scan"):
            call("rgp_bin_scan", int(n), int(width), int(height), ptr(ws), I64(int(cap)), stream(device))
        if exact:
            m = int(header(ws, 1).item())
            CAPACITY.update(n, width, height, m)
            if m > cap:
                cap = CAPACITY.get(n, width, height)
                continue
        else:
            hdr = torch.empty(2, dtype=torch.int32, pin_memory=True)
            hdr.copy_(header(ws, 2), non_blocking=True)
            scanned = torch.cuda.Event()
            scanned.record(torch.cuda.current_stream(device))
        with SPANS.span("sort"):
            ids, ranges, _ = bin_sort(n, depths, width, height, ws, cap, device)
        smask = torch.empty(ids.shape[0], dtype=torch.uint8, device=device) if strips else None
        layer, contrib_img, final_tau = draw_forward(width, height, dim, ranges, ids, records, empty_tau, smask,
                                                     contrib)
        if not exact:
            scanned.synchronize()
            m = int(hdr[0])
            CAPACITY.update(n, width, height, m)
            if int(hdr[1]) != 0:
                cap = CAPACITY.get(n, width, height)
                continue
        return layer, contrib_img, final_tau, ranges, ids, m, ws, smask


def camera_structs(cams, device):
    """Contiguous device array (V, 32) of rgp_camera structs, one per camera."""
    return torch.stack([camera_struct(c, device) for c in cams])


def _align256(x):
    return (int(x) + 255) & ~255


def splat_pipeline_views(n, views, width, height, dim, records, depths, ws_fill, empty_tau, device,
                         exact=False, strips=False, contrib=True):
    """splat_pipeline for `views` cameras of one resolution: one preprocess
    launch (ws_fill(ws, stride, cap)), the binning of every view, and one
    compositing launch over all views' tiles.  Returns layers (V,dim,H,W),
    contrib (V,H,W), final_tau (V,H,W), ranges (V,tiles,2), ids (V,cap),
    the pair counts per view, the workspace, its per-view stride and the
    strip masks (V,cap) uint8 when strips=True (else None); contrib=False
    returns None for contrib.  The overflow check follows splat_pipeline's
    (exact=False: after the draw)."""
    gx, gy = grid_dims(width, height)
    cap = CAPACITY.get(n, width, height)
    while True:
        stride = _align256(workspace_bytes(n, cap, width, height))
        ws = torch.empty(stride * views, dtype=torch.uint8, device=device)
        with SPANS.span("preprocess"):
            ws_fill(ws, stride, cap)
        with SPANS.span("scan"):
            call("rgp_bin_scan_views", int(views), int(n), int(width), int(height), ptr(ws), ctypes.c_size_t(stride),
                 I64(int(cap)), stream(device))
        hdrs = ws.view(views, stride)[:, :8].contiguous().view(torch.int32)  # (V, 2): M, overflow
        if exact:
            ms = hdrs[:, 0].tolist()
            for m in ms:
                CAPACITY.update(n, width, height, m)
            if max(ms) > cap:
                cap = CAPACITY.get(n, width, height)
                continue
        else:
            host = torch.empty((views, 2), dtype=torch.int32, pin_memory=True)
            host.copy_(hdrs, non_blocking=True)
            scanned = torch.cuda.Event()
            scanned.record(torch.cuda.current_stream(device))
        ids = torch.empty((views, max(int(cap), 1)), dtype=torch.int32, device=device)
        ranges = torch.empty((views, gx * gy, 2), dtype=torch.int32, device=device)
        with SPANS.span("sort"):
            call("rgp_bin_sort_views", int(views), int(n), ptr(depths), int(width), int(height), ptr(ws),
                 ctypes.c_size_t(stride), I64(int(cap)), ptr(ids), ptr(ranges), stream(device))
        layer = torch.empty((views, dim, height, width), dtype=torch.float32, device=device)
        contrib_img = torch.empty((views, height, width), dtype=torch.int32, device=device) if contrib else None
        final_tau = torch.empty((views, height, width), dtype=torch.float32, device=device)
        smask = torch.empty(ids.shape, dtype=torch.uint8, device=device) if strips else None
        with SPANS.span("draw_forward"):
            call("rgp_draw_forward_views", int(views), int(width), int(height), int(dim), ptr(ranges), ptr(ids),
                 I64(ids.shape[1]), ptr(records), I64(int(n)), float(empty_tau), ptr(layer), ptr(contrib_img),
                 ptr(final_tau), ptr(smask), stream(device))
        if not exact:
            scanned.synchronize()
            ms = host[:, 0].tolist()
            for m in ms:
                CAPACITY.update(n, width, height, m)
            if int(host[:, 1].max()) != 0:
                cap = CAPACITY.get(n, width, height)
                continue
        return layer, contrib_img, final_tau, ranges, ids, ms, ws, stride, smask


def draw_backward_replay_views(views, width, height, dim, ranges, ids, records, dlayer, n, strips=None):
    """Replay backward of a multi-view batch: dfeat (V, n, dim), one row set per
    view (strips: the forward's (V, id stride) strip masks, or None)."""
    dev = records.device
    dfeat = torch.zeros((views, n, dim), dtype=torch.float32, device=dev)
    dlayer = f32c(dlayer)
    with SPANS.span("draw_backward"):
        call("rgp_draw_backward_replay_views", int(views), int(width), int(height), int(dim), ptr(ranges), ptr(ids),
             I64(ids.shape[1]), ptr(records), I64(int(n)), ptr(strips), ptr(dlayer), ptr(dfeat), stream(dev))
    return dfeat


def mix(W, palette):
    """RGB (3,H,W) = sum_k palette[k] W[k]."""
    require_cuda(W, palette)
    K = W.shape[0]
    hw = W.shape[1] * W.shape[2]
    rgb = torch.empty((3,) + tuple(W.shape[1:]), dtype=torch.float32, device=W.device)
    call("rgp_mix", int(K), I64(hw), ptr(f32c(W)), ptr(f32c(palette)), ptr(rgb), stream(W.device))
    return rgb


def mix_backward(W, palette, drgb):
    K = W.shape[0]
    hw = W.shape[1] * W.shape[2]
    dW = torch.empty_like(W)
    dpal = torch.zeros((K, 3), dtype=torch.float32, device=W.device)
    call("rgp_mix_backward", int(K), I64(hw), ptr(f32c(W)), ptr(f32c(palette)), ptr(f32c(drgb)),
         ptr(dW), ptr(dpal), stream(W.device))
    return dW, dpal


class Mix(torch.autograd.Function):
    """Palette mix with its backward on the GPU kernels."""

    @staticmethod
    def forward(ctx, W, palette):
        ctx.save_for_backward(W, palette)
        return mix(W, palette)

    @staticmethod
    def backward(ctx, drgb):
        W, palette = ctx.saved_tensors
        dW, dpal = mix_backward(W, palette, drgb.contiguous())
        return dW, dpal
