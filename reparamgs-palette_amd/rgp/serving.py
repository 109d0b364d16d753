"""Serving loop for palette-weight scenes resident in HBM.

There is no serving code in the reference.  Its nearest operation is the
palette edit by bake + re-render (finetune.py:472-478, gsplat/gsmodel.py:488-509).
Here a palette-only edit re-mixes the cached K-channel weight image of the
scene's current view (rgp_mix_batched), which by linearity equals the bake.

* Scenes are admitted against an HBM budget: their parameters plus the render
  buffers of every resolution they will be served at.
* Each (scene, resolution) owns a render slot with static buffers and a HIP
  graph of the capacity-bounded pipeline: preprocess -> scan -> sort ->
  composite.  A replay takes a new pose through the slot's device camera
  buffer and never synchronises the host.  A pair-count overflow is detected
  from the workspace header after the fact and the request is re-rendered
  with a larger capacity.
* Continuous batching: each loop iteration admits every request that has
  arrived, and each scene's requests go out in arrival order on the scene's
  own render stream, so scenes never wait for each other and their work
  overlaps on the device.  A run of pose requests of one resolution renders
  as one multi-view batch (up to batch_views cameras: one preprocess launch
  reading the scene once, the binning of each view, one compositing launch
  over all views' tiles, captured as a HIP graph per batch size).  Every
  view of a batch keeps its own weight image, so a palette edit is served
  from the view of the last pose of its scene that arrived before it, and
  one rgp_mix_batched launch after the batch mixes its poses and its edits.
  With batch_views = 0 each pose renders alone and the pending edits go out
  in one mix launch per K on a shared edit stream.
* Latency is the host time from a request's arrival to the moment the loop
  sees its completion event; the report gives p50/p99 overall and for
  palette-only edits, and frames (requests) per second.

Luminance edit (the reference has none): the edited palette is
clamp((P + dP) * s, 0, 1) on the chromatic rows; black and white stay fixed.

Scenes are synthetic (rgp.scene.make_scene) or read from an exported
.pply + .gswp pair (SceneSpec.from_files, rgp.formats).
"""
import collections
import ctypes
import dataclasses
import math
import time
import types

import numpy as np
import torch

from . import formats, ops
from ._lib import I64, call, f32c, ptr, stream
from .scene import look_at_camera, make_scene

RESOLUTIONS = ((1280, 720), (1920, 1080), (2560, 1440))


class AdmissionError(RuntimeError):
    pass


@dataclasses.dataclass
class SceneSpec:
    name: str
    n: int
    K: int
    lowrank: bool
    rank: int = 2
    seed: int = 0
    files: tuple = None  # (pply, gswp): read the scene instead of generating it
    P: int = 0           # low-rank factor shape P x Q (0: P = K, Q = 15)
    Q: int = 15

    @classmethod
    def from_files(cls, name, pply_path, sidecar_path=None):
        """A scene exported as .pply + .gswp; sizes come from the sidecar header."""
        sidecar_path = sidecar_path or formats.sidecar_path_for(pply_path)
        h = formats.read_gswp(sidecar_path)
        if h["num_sh"] != 16:
            raise ValueError(f"{sidecar_path}: num_sh = {h['num_sh']}; the serving loop renders 16 coefficients")
        return cls(name, h["N"], h["splat_dim"], h["low_rank"], rank=1, files=(pply_path, sidecar_path),
                   P=h["P"], Q=h["Q"] or 15)

    def param_bytes(self):
        geo = self.n * (12 + 16 + 12 + 4)
        if self.lowrank:
            P = self.P or self.K  # P x Q = K x 15
            return geo + self.n * 4 * (self.K + self.rank * (P + self.Q))
        return geo + self.n * 4 * 16 * self.K

    def slot_bytes(self, width, height, cap_per_gaussian=6.0):
        """Render buffers of one resolution: records, depths, workspace at the
        planning capacity, ids, ranges, weight image, final_tau."""
        rec = self.n * 4 * (8 + 4 * math.ceil(self.K / 4))
        cap = int(cap_per_gaussian * self.n)
        tiles = ((width + 15) // 16) * ((height + 15) // 16)
        ws = 256 + self.n * 16 + cap * 24 + 256 * 4 * (cap // 4096 + 1)
        img = width * height * 4 * (self.K + 1)
        return rec + self.n * 4 + ws + cap * 4 + tiles * 8 + img

    def footprint(self, resolutions=RESOLUTIONS, views=0):
        """Parameters plus, per resolution, a one-view render slot and a batch
        slot of `views` views (views = 0: the one-view slots only)."""
        return self.param_bytes() + sum(self.slot_bytes(w, h) * (1 + views) for w, h in resolutions)


def default_specs():
    """The six resident scenes of the serving workload."""
    return [SceneSpec("s200k", 200_000, 5, False, seed=1), SceneSpec("s500k", 500_000, 7, True, seed=2),
            SceneSpec("s1m_a", 1_000_000, 5, False, seed=3), SceneSpec("s1m_b", 1_000_000, 7, True, seed=4),
            SceneSpec("s2m", 2_000_000, 5, True, seed=5), SceneSpec("s3m", 3_000_000, 7, True, seed=6)]


class SceneStore:
    """Admission of scenes against an HBM budget (bytes)."""

    def __init__(self, budget_bytes, resolutions=RESOLUTIONS, batch_views=0):
        self.budget = int(budget_bytes)
        self.resolutions = resolutions
        self.batch_views = batch_views
        self.used = 0
        self.specs = {}

    def admit(self, spec):
        need = spec.footprint(self.resolutions, self.batch_views)
        if self.used + need > self.budget:
            raise AdmissionError(f"scene {spec.name} needs {need / 2**30:.2f} GiB; "
                                 f"{(self.budget - self.used) / 2**30:.2f} GiB of the budget left")
        self.used += need
        self.specs[spec.name] = spec
        return need

    def release(self, name):
        spec = self.specs.pop(name)
        self.used -= spec.footprint(self.resolutions, self.batch_views)


class ResidentScene:
    """Device copy of one finetuned scene (synthetic, or read from files)."""

    def __init__(self, spec, device):
        self.spec = spec
        if spec.files:
            sc = types.SimpleNamespace(rank=1, **formats.load_scene_files(*spec.files))
        else:
            sc = make_scene(spec.n, seed=spec.seed, extent=3.0, K=spec.K, lowrank=spec.lowrank, rank=spec.rank,
                            P=(spec.P or spec.K) if spec.lowrank else None, Q=spec.Q)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.n, self.K = spec.n, spec.K
        self.pws, self.rots, self.scales, self.alphas = T(sc.pws), T(sc.rots), T(sc.scales), T(sc.alphas)
        self.palette = np.array(sc.palette, dtype=np.float32)
        if spec.lowrank:
            self.mode, self.P, self.Q, self.rank = 1, sc.P, sc.Q, sc.rank
            self.shs, self.dc, self.A, self.B = None, T(sc.dc), T(sc.A), T(sc.B)
        else:
            self.mode, self.P, self.Q, self.rank = 0, 16, 0, 1
            self.shs, self.dc, self.A, self.B = T(sc.shs), None, None, None


class RenderSlot:
    """Static buffers and the captured render graph of one (scene, resolution)."""

    def __init__(self, scene, width, height, capacity, device, use_graph=True):
        self.scene, self.W, self.H, self.cap = scene, width, height, int(capacity)
        n, K = scene.n, scene.K
        f = dict(dtype=torch.float32, device=device)
        self.camd = torch.zeros(32, **f)
        self.depths = torch.empty(n, **f)
        self.records = torch.empty((n, ops.record_floats(K)), **f)
        self.ws = ops.new_workspace(n, self.cap, width, height, device)
        gx, gy = ops.grid_dims(width, height)
        self.ids = torch.empty(max(self.cap, 1), dtype=torch.int32, device=device)
        self.ranges = torch.empty((gx * gy, 2), dtype=torch.int32, device=device)
        self.layer = torch.empty((K, height, width), **f)
        self.final_tau = torch.empty((height, width), **f)
        self.graph = None
        if use_graph:
            s = torch.cuda.Stream(device)
            s.wait_stream(torch.cuda.current_stream(device))
            with torch.cuda.stream(s):
                self.launch()
            torch.cuda.current_stream(device).wait_stream(s)
            torch.cuda.synchronize(device)
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.launch()
            torch.cuda.synchronize(device)

    def launch(self):
        sc, W, H = self.scene, self.W, self.H
        st = stream(self.camd.device)
        call("rgp_preprocess_weight", sc.n, sc.K, sc.mode, sc.P, sc.Q, sc.rank, ptr(self.camd), ptr(sc.pws),
             ptr(sc.rots), ptr(sc.scales), ptr(sc.alphas), ptr(sc.shs), ptr(sc.dc), ptr(sc.A), ptr(sc.B),
             ptr(self.depths), ptr(self.records), ptr(self.ws), I64(self.cap), W, H, None, None, None, None, st)
        call("rgp_bin_scan", sc.n, W, H, ptr(self.ws), I64(self.cap), st)
        call("rgp_bin_sort", sc.n, ptr(self.depths), W, H, ptr(self.ws), I64(self.cap), ptr(self.ids), None,
             ptr(self.ranges), st)
        call("rgp_draw_forward", W, H, sc.K, ptr(self.ranges), ptr(self.ids), ptr(self.records), 0.0,
             ptr(self.layer), None, ptr(self.final_tau), None, st)  # no contrib image: frames only

    def render(self, cam_host):
        """Enqueue a render on the current stream with the pose in cam_host
        (a pinned float32 tensor holding the rgp_camera struct)."""
        self.camd.copy_(cam_host, non_blocking=True)
        if self.graph is not None:
            self.graph.replay()
        else:
            self.launch()


class BatchSlot:
    """Static buffers of up to `vmax` views of one (scene, resolution) and the
    captured graphs of the multi-view render, one per batch size."""

    def __init__(self, scene, width, height, capacity, vmax, device, use_graph=True):
        self.scene, self.W, self.H, self.cap, self.vmax = scene, width, height, int(capacity), int(vmax)
        n, K = scene.n, scene.K
        f = dict(dtype=torch.float32, device=device)
        self.camd = torch.zeros((vmax, 32), **f)
        self.depths = torch.empty((vmax, n), **f)
        self.records = torch.empty((vmax, n, ops.record_floats(K)), **f)
        self.stride = (ops.workspace_bytes(n, self.cap, width, height) + 255) & ~255
        self.ws = torch.empty(self.stride * vmax, dtype=torch.uint8, device=device)
        gx, gy = ops.grid_dims(width, height)
        self.ids = torch.empty((vmax, max(self.cap, 1)), dtype=torch.int32, device=device)
        self.ranges = torch.empty((vmax, gx * gy, 2), dtype=torch.int32, device=device)
        self.layer = torch.empty((vmax, K, height, width), **f)
        self.final_tau = torch.empty((vmax, height, width), **f)
        self.hdr_dev = self.ws.view(vmax, self.stride)[:, :8]
        self.use_graph = use_graph
        self.graphs = {}

    def launch(self, v):
        sc, W, H = self.scene, self.W, self.H
        st = stream(self.camd.device)
        sz = ctypes.c_size_t(self.stride)
        call("rgp_preprocess_weight_views", v, sc.n, sc.K, sc.mode, sc.P, sc.Q, sc.rank, ptr(self.camd), ptr(sc.pws),
             ptr(sc.rots), ptr(sc.scales), ptr(sc.alphas), ptr(sc.shs), ptr(sc.dc), ptr(sc.A), ptr(sc.B),
             ptr(self.depths), ptr(self.records), ptr(self.ws), sz, I64(self.cap), W, H, st)
        call("rgp_bin_scan_views", v, sc.n, W, H, ptr(self.ws), sz, I64(self.cap), st)
        call("rgp_bin_sort_views", v, sc.n, ptr(self.depths), W, H, ptr(self.ws), sz, I64(self.cap), ptr(self.ids),
             ptr(self.ranges), st)
        call("rgp_draw_forward_views", v, W, H, sc.K, ptr(self.ranges), ptr(self.ids), I64(self.ids.shape[1]),
             ptr(self.records), I64(sc.n), 0.0, ptr(self.layer), None, ptr(self.final_tau), None, st)

    def capture(self, v):
        """The graph of the v-view render (captured once; the capture
        synchronises the device, so serving captures every batch size before
        it starts)."""
        g = self.graphs.get(v)
        if g is None:
            dev = self.camd.device
            s = torch.cuda.Stream(dev)
            s.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(s):
                self.launch(v)  # warm-up launch outside the capture
            torch.cuda.current_stream(dev).wait_stream(s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                self.launch(v)
            torch.cuda.current_stream(dev).wait_stream(s)
            self.graphs[v] = g
        return g

    def render(self, cam_hosts):
        """Enqueue the render of len(cam_hosts) <= vmax poses on the current
        stream (pinned rgp_camera structs); view i lands in layer[i]."""
        v = len(cam_hosts)
        for i, c in enumerate(cam_hosts):
            self.camd[i].copy_(c, non_blocking=True)
        if not self.use_graph:
            self.launch(v)
            return
        self.capture(v).replay()


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float
    scene: str
    kind: str  # "pose" | "edit"
    pose: int = 0
    res: tuple = (1920, 1080)
    dpal: np.ndarray = None
    lum: float = 1.0
    done: float = None
    out: torch.Tensor = None
    dep: object = None        # edit: completion event of the render its weight image came from
    dep_view: tuple = None    # edit: (pose, resolution) of that render
    view: tuple = None        # edit to redo by re-rendering this (pose, resolution)
    redo: bool = False        # pose to render again after its render overflowed


def make_requests(specs, n_requests, rate, seed=0, pose_frac=0.4, n_poses=64):
    """Poisson arrivals; every scene's first request is a pose request."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests)) if rate > 0 else np.zeros(n_requests)
    seen = set()
    reqs = []
    for i in range(n_requests):
        sp = specs[rng.integers(len(specs))]
        kind = "pose" if (sp.name not in seen or rng.random() < pose_frac) else "edit"
        seen.add(sp.name)
        res = RESOLUTIONS[rng.integers(len(RESOLUTIONS))]
        dpal = rng.normal(0.0, 0.1, (sp.K, 3)).astype(np.float32)
        reqs.append(Request(i, float(t[i]), sp.name, kind, int(rng.integers(n_poses)), res, dpal,
                            float(rng.uniform(0.7, 1.3))))
    return reqs


def edited_palette(base, dpal, lum):
    P = base.astype(np.float32).copy()
    P[2:] = np.clip((P[2:] + dpal[2:]) * np.float32(lum), 0.0, 1.0)
    return P


class Engine:
    """Continuous-batching server over resident scenes."""

    def __init__(self, specs, device, budget_bytes=200 * 2**30, n_streams=4, use_graph=True, cap_margin=1.5,
                 radius=6.0, n_poses=64, batch_views=8):
        self.device = device
        self.batch_views = int(batch_views)
        self.store = SceneStore(budget_bytes, batch_views=self.batch_views)
        self.scenes = {}
        for sp in specs:
            self.store.admit(sp)
            self.scenes[sp.name] = ResidentScene(sp, device)
        self.use_graph = use_graph
        self.cap_margin = cap_margin
        self.n_poses = n_poses
        self.radius = radius
        self.render_streams = [torch.cuda.Stream(device) for _ in range(n_streams)]
        self.edit_stream = torch.cuda.Stream(device)
        self.scene_stream = {name: self.render_streams[i % n_streams] for i, name in enumerate(self.scenes)}
        self.slots = {}
        self.bslots = {}
        self.session = {}  # scene -> (weight image (K,H,W), render event, pose, resolution)
        self.last_edit = {}  # scene -> event of the last edit batch reading its weight image
        self.cams = {}
        self.base_pal = {}

    # -- cameras -----------------------------------------------------------
    def pose_camera(self, pose, width, height):
        """The host camera of pose `pose` (of n_poses on the orbit) at a resolution."""
        th = 2.0 * math.pi * pose / self.n_poses
        el = 0.3 * math.sin(3.0 * th)
        pos = (self.radius * math.cos(el) * math.sin(th), self.radius * math.sin(el),
               -self.radius * math.cos(el) * math.cos(th))
        return look_at_camera(pos, width, height)

    def camera(self, pose, width, height):
        """The pose's rgp_camera struct in pinned host memory (cached)."""
        key = (pose, width, height)
        if key not in self.cams:
            c = self.pose_camera(pose, width, height)
            host = ops.camera_struct_host(c.Rcw, c.tcw, c.twc, c.fx, c.fy, c.cx, c.cy, width, height)
            self.cams[key] = torch.from_numpy(host).pin_memory()
        return self.cams[key]

    def slot(self, name, res, min_cap=None):
        key = (name, res)
        sl = self.slots.get(key)
        if sl is None or (min_cap is not None and sl.cap < min_cap):
            if sl is not None:
                torch.cuda.synchronize(self.device)  # the old buffers may still be in use on a render stream
            sc = self.scenes[name]
            cap = min_cap or self._plan_capacity(sc, res)
            sl = RenderSlot(sc, res[0], res[1], cap, self.device, self.use_graph)
            self.slots[key] = sl
        return sl

    def bslot(self, name, res, min_cap=None):
        key = (name, res)
        bs = self.bslots.get(key)
        if bs is None or (min_cap is not None and bs.cap < min_cap):
            if bs is not None:
                torch.cuda.synchronize(self.device)  # the old buffers may still be in use on a render stream
            sc = self.scenes[name]
            one = self.slots.get(key)
            cap = min_cap or (one.cap if one is not None else self._plan_capacity(sc, res))
            bs = BatchSlot(sc, res[0], res[1], cap, self.batch_views, self.device, self.use_graph)
            self.bslots[key] = bs
        return bs

    def _grown(self, m):
        """Capacity after an overflow with m pairs: always above m."""
        return int(m * max(self.cap_margin, 1.25)) + 4096

    def _plan_capacity(self, sc, res, samples=4):
        """Capacity from exact pair counts of a few poses, times a margin."""
        m = 0
        for p in range(0, self.n_poses, max(1, self.n_poses // samples)):
            probe = RenderSlot(sc, res[0], res[1], 1, self.device, use_graph=False)
            probe.camd.copy_(self.camera(p, res[0], res[1]))
            call("rgp_preprocess_weight", sc.n, sc.K, sc.mode, sc.P, sc.Q, sc.rank, ptr(probe.camd), ptr(sc.pws),
                 ptr(sc.rots), ptr(sc.scales), ptr(sc.alphas), ptr(sc.shs), ptr(sc.dc), ptr(sc.A), ptr(sc.B),
                 ptr(probe.depths), ptr(probe.records), ptr(probe.ws), I64(1), res[0], res[1], None, None, None,
                 None, stream(self.device))
            call("rgp_bin_scan", sc.n, res[0], res[1], ptr(probe.ws), I64(1), stream(self.device))
            m = max(m, int(ops.header(probe.ws, 1).item()))
            del probe
        return int(m * self.cap_margin) + 4096

    def prepare(self, requests):
        """Build every render slot the requests need and capture their graphs
        for every batch size (outside any timing)."""
        for r in requests:
            if r.kind == "pose":
                self.slot(r.scene, r.res)
                if self.batch_views > 0:
                    self.bslot(r.scene, r.res)
        if self.use_graph:
            for bs in self.bslots.values():
                for v in range(1, bs.vmax + 1):
                    bs.capture(v)
        torch.cuda.synchronize(self.device)

    # -- launching ---------------------------------------------------------
    def _set_session(self, name, layer, done, pose, res, rid):
        """The scene's current view: the latest pose by arrival (a re-render of
        an earlier pose after an overflow does not replace a later one)."""
        cur = self.session.get(name)
        if cur is None or rid >= cur[4]:
            self.session[name] = (layer, done, pose, res, rid)

    def _launch_pose(self, r, pose=None, res=None, palette=None, update_session=True):
        """Render the request's view (or the given one) and mix it with the
        scene palette (or the given edited palette)."""
        sc = self.scenes[r.scene]
        st = self.scene_stream[r.scene]
        pose = r.pose if pose is None else pose
        res = r.res if res is None else res
        sl = self.slot(r.scene, res)
        with torch.cuda.stream(st):
            ev = self.last_edit.get(r.scene)
            if ev is not None:
                st.wait_event(ev)  # edits still reading the old weight image
            sl.render(self.camera(pose, *res))
            if palette is None:
                if r.scene not in self.base_pal:
                    self.base_pal[r.scene] = torch.from_numpy(sc.palette.astype(np.float32)).to(self.device)
                pal = self.base_pal[r.scene]
            else:
                pal = torch.from_numpy(palette).pin_memory().to(self.device, non_blocking=True)
            r.out = ops.mix(sl.layer, pal)
            hdr = torch.empty(4, dtype=torch.int32).pin_memory()
            hdr.copy_(ops.header(sl.ws), non_blocking=True)
            done = torch.cuda.Event()
            done.record(st)
        if update_session:
            self._set_session(r.scene, sl.layer, done, pose, res, r.rid)
        return done, hdr

    def _base_palette(self, name):
        return self.scenes[name].palette.astype(np.float32)

    def _mix_many(self, K, items):
        """One rgp_mix_batched launch on the current stream over items of
        (weight image (K,H,W), output (3,H,W), palette (K,3) host array)."""
        n = len(items)
        ints = np.array([w.data_ptr() for w, _, _ in items] + [o.data_ptr() for _, o, _ in items] +
                        [w.shape[1] * w.shape[2] for w, _, _ in items], dtype=np.int64)
        palf = np.concatenate([p for _, _, p in items]).astype(np.float32).reshape(-1)
        pin = torch.from_numpy(np.concatenate([ints.view(np.uint8), palf.view(np.uint8)])).pin_memory()
        dev = pin.to(self.device, non_blocking=True)
        base = dev.data_ptr()
        call("rgp_mix_batched", n, K, base, base + 8 * n, base + 16 * n, base + 24 * n, I64(int(ints[2 * n:].max())),
             stream(self.device))

    def _launch_scene(self, name, rs):
        """One scene's requests (arrival order) on its render stream.  The
        poses are rendered as multi-view batches grouped by resolution, in
        rounds: round k holds, per resolution, the k-th run of at most
        batch_views of its poses.  Each view keeps its own weight image in its
        resolution's batch slot until the next round, so every palette edit is
        mixed from the view of the last pose before it (in arrival order),
        and one rgp_mix_batched launch per round mixes the round's poses
        (scene palette) and the edits on its views (edited palettes).  Edits
        on the view left by earlier work are mixed before the first round.
        Returns the inflight entries (event, requests, per-view headers or
        None)."""
        sc = self.scenes[name]
        st = self.scene_stream[name]
        pal0 = self._base_palette(name)
        vmax = self.batch_views
        count = collections.Counter()
        rounds = []  # per round: {res: [poses]}, and [(edit, res, view)]
        head, src, last = [], None, None
        for r in rs:
            if r.kind == "pose":
                k = count[r.res] // vmax
                count[r.res] += 1
                while len(rounds) <= k:
                    rounds.append(({}, []))
                g = rounds[k][0].setdefault(r.res, [])
                g.append(r)
                src = last = (k, r.res, len(g) - 1)
            elif src is None:
                head.append(r)
            else:
                rounds[src[0]][1].append((r, src[1], src[2]))
        out = []
        with torch.cuda.stream(st):
            ev = self.last_edit.pop(name, None)
            if ev is not None:
                st.wait_event(ev)  # edits of a one-view-per-launch run still reading the old image
            if head:
                layer, rev, pose, res, _ = self.session[name]
                items = []
                for r in head:
                    r.dep, r.dep_view = rev, (pose, res)
                    r.out = torch.empty((3, res[1], res[0]), dtype=torch.float32, device=self.device)
                    items.append((layer, r.out, edited_palette(sc.palette, r.dpal, r.lum)))
                self._mix_many(sc.K, items)
                done = torch.cuda.Event()
                done.record(st)
                out.append((done, head, None))
            for k, (groups, edits) in enumerate(rounds):
                items, hdrs = [], []
                for res, g in groups.items():
                    bs = self.bslot(name, res)
                    bs.render([self.camera(r.pose, *res) for r in g])
                    for i, r in enumerate(g):
                        r.out = torch.empty((3, res[1], res[0]), dtype=torch.float32, device=self.device)
                        items.append((bs.layer[i], r.out, pal0))
                    hdr = torch.empty((len(g), 2), dtype=torch.int32).pin_memory()
                    hdr.copy_(bs.hdr_dev[:len(g)].view(torch.int32), non_blocking=True)
                    hdrs.append((g, hdr))
                for r, res, i in edits:
                    r.out = torch.empty((3, res[1], res[0]), dtype=torch.float32, device=self.device)
                    items.append((self.bslots[(name, res)].layer[i], r.out, edited_palette(sc.palette, r.dpal, r.lum)))
                self._mix_many(sc.K, items)
                done = torch.cuda.Event()
                done.record(st)
                for g, hdr in hdrs:
                    out.append((done, g, hdr))
                for r, res, i in edits:
                    r.dep, r.dep_view = done, (groups[res][i].pose, res)
                if edits:
                    out.append((done, [r for r, _, _ in edits], None))
                if last is not None and last[0] == k:
                    res, i = last[1], last[2]
                    p = groups[res][i]
                    self._set_session(name, self.bslots[(name, res)].layer[i], done, p.pose, res, p.rid)
        return out

    def _stale(self, r):
        """An edit whose weight image came from a render that overflowed: it
        is redone by re-rendering that view (r.view) with the edited palette."""
        if r.kind == "edit" and r.dep is not None and id(r.dep) in self._failed:
            r.view, r.dep = r.dep_view, None
            return True
        return False

    def _launch_view_edit(self, r):
        """Re-render the view an edit was made on and mix it with the edited
        palette (after that view's render overflowed); the scene's current
        view is left as it is."""
        ev, hdr = self._launch_pose(r, r.view[0], r.view[1],
                                    edited_palette(self.scenes[r.scene].palette, r.dpal, r.lum),
                                    update_session=False)
        return ev, [r], hdr

    def _launch_edits(self, edits):
        """One batched mix launch per K for the given palette edits."""
        st = self.edit_stream
        events = []
        by_k = collections.defaultdict(list)
        for r in edits:
            by_k[self.scenes[r.scene].K].append(r)
        with torch.cuda.stream(st):
            for K, group in by_k.items():
                waits = set()
                Wp, Rp, hw, pals = [], [], [], []
                for r in group:
                    layer, rev, pose, res, _ = self.session[r.scene]
                    r.dep, r.dep_view = rev, (pose, res)  # redone from this view if its render overflowed
                    if id(rev) not in waits:
                        st.wait_event(rev)
                        waits.add(id(rev))
                    r.out = torch.empty((3, res[1], res[0]), dtype=torch.float32, device=self.device)
                    Wp.append(layer.data_ptr())
                    Rp.append(r.out.data_ptr())
                    hw.append(res[0] * res[1])
                    pals.append(edited_palette(self.scenes[r.scene].palette, r.dpal, r.lum))
                # one upload: [W pointers | output pointers | pixel counts | palettes]
                n = len(group)
                ints = np.array(Wp + Rp + hw, dtype=np.int64)
                palf = np.concatenate(pals).astype(np.float32).reshape(-1)
                pin = torch.from_numpy(np.concatenate([ints.view(np.uint8), palf.view(np.uint8)])).pin_memory()
                dev = pin.to(self.device, non_blocking=True)
                base = dev.data_ptr()
                call("rgp_mix_batched", n, K, base, base + 8 * n, base + 16 * n, base + 24 * n, I64(max(hw)),
                     stream(self.device))
            ev = torch.cuda.Event()
            ev.record(st)
        for r in edits:
            self.last_edit[r.scene] = ev
        return ev

    # -- running -----------------------------------------------------------
    def run(self, requests, mode="batched", realtime=True, max_batch=256, rerender_edits=False,
            keep_outputs=True):
        """Serve the requests.  mode "batched": continuous batching;
        "serial": one request at a time, each synchronised before the next.
        realtime: honour arrival times (else everything arrives at t=0).
        rerender_edits (serial only): a palette edit re-renders the scene's
        current view, the reference's way (bake + render, finetune.py:472-478),
        instead of re-mixing the cached weight image.
        keep_outputs=False hands each frame off when it completes (r.out is
        dropped), so frame memory is recycled as a server would."""
        for r in requests:
            r.done, r.out, r.dep, r.dep_view, r.view, r.redo = None, None, None, None, None, False
        self._failed = {}  # id -> event of renders that overflowed (kept alive so ids stay unique)
        self.session = {}  # a run is a fresh request stream: views come from its own poses
        pending = collections.deque(sorted(requests, key=lambda q: q.arrival))
        inflight = []  # (event, [requests], hdr or None)
        redo = []
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        arrive = (lambda r: r.arrival) if realtime else (lambda r: 0.0)
        while pending or inflight or redo:
            now = time.perf_counter() - t0
            batch = []
            while pending and arrive(pending[0]) <= now and len(batch) < max_batch:
                batch.append(pending.popleft())
            batch = redo + batch
            redo = []
            if mode == "serial":
                for r in batch:
                    if r.kind == "pose":
                        res = r.res
                        ev, hdr = self._launch_pose(r)
                    elif rerender_edits or r.view is not None:
                        pose, res = r.view if r.view is not None else self.session[r.scene][2:4]
                        ev, hdr = self._launch_pose(r, pose, res,
                                                    edited_palette(self.scenes[r.scene].palette, r.dpal, r.lum),
                                                    update_session=r.view is None)
                    else:
                        ev, hdr = self._launch_edits([r]), None
                    ev.synchronize()
                    if hdr is not None and int(hdr[1]) != 0:
                        self._failed[id(ev)] = ev
                        self.slot(r.scene, res, min_cap=self._grown(int(hdr[0])))
                        redo.append(r)
                        continue
                    if self._stale(r):
                        redo.append(r)
                        continue
                    r.done = time.perf_counter() - t0
                    if not keep_outputs:
                        r.out = None
            elif self.batch_views <= 0:
                edits = []
                for r in batch:
                    if r.kind == "edit" and r.view is not None:
                        inflight.append(self._launch_view_edit(r))
                    elif r.kind == "pose":
                        if any(e.scene == r.scene for e in edits):
                            # earlier edits of this scene use the view before this pose
                            inflight.append((self._launch_edits(edits), edits, None))
                            edits = []
                        ev, hdr = self._launch_pose(r)
                        inflight.append((ev, [r], hdr))
                    else:
                        edits.append(r)
                if edits:
                    inflight.append((self._launch_edits(edits), edits, None))
            else:
                # Each scene's requests of this iteration go out on the scene's
                # own stream in arrival order (see _launch_scene).
                by_scene = collections.OrderedDict()
                for r in batch:
                    if r.kind == "edit" and r.view is not None:
                        inflight.append(self._launch_view_edit(r))
                    elif r.redo:
                        # an overflowed pose renders again on its own; later
                        # requests of its scene do not take their view from it
                        ev, hdr = self._launch_pose(r)
                        inflight.append((ev, [r], hdr))
                    else:
                        by_scene.setdefault(r.scene, []).append(r)
                for name, rs in by_scene.items():
                    inflight.extend(self._launch_scene(name, rs))
            still = []
            for ev, rs, hdr in inflight:
                if ev.query():
                    t = time.perf_counter() - t0
                    if hdr is not None and hdr.dim() == 2:  # a multi-view batch: per-view (M, overflow)
                        over = [i for i in range(len(rs)) if int(hdr[i, 1]) != 0]
                        if over:
                            self._failed[id(ev)] = ev
                            m = max(int(hdr[i, 0]) for i in over)
                            self.bslot(rs[0].scene, rs[0].res, min_cap=self._grown(m))
                            for i in over:
                                rs[i].redo = True
                                redo.append(rs[i])
                        done_rs = [r for i, r in enumerate(rs) if i not in over]
                    elif hdr is not None and int(hdr[1]) != 0:
                        self._failed[id(ev)] = ev
                        r = rs[0]
                        res = r.view[1] if r.kind == "edit" else r.res
                        self.slot(r.scene, res, min_cap=self._grown(int(hdr[0])))
                        r.redo = True
                        redo.append(r)
                        continue
                    else:
                        done_rs = rs
                    for r in done_rs:
                        if self._stale(r):
                            redo.append(r)
                            continue
                        r.done = t
                        if not keep_outputs:
                            r.out = None
                else:
                    still.append((ev, rs, hdr))
            inflight = still
            if not batch and inflight and not redo:
                time.sleep(0)
            elif not batch and not inflight and pending and realtime:
                time.sleep(max(0.0, min(2e-4, arrive(pending[0]) - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        return report(requests, wall, realtime)


def report(requests, wall, realtime=True):
    lat = np.array([(r.done - (r.arrival if realtime else 0.0)) * 1e3 for r in requests])
    edit = np.array([(r.done - (r.arrival if realtime else 0.0)) * 1e3 for r in requests if r.kind == "edit"])
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None
    return {"requests": len(requests), "wall_s": wall, "frames_per_s": len(requests) / wall,
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "edit_p50_ms": pct(edit, 50),
            "edit_p99_ms": pct(edit, 99), "edits": int(len(edit))}
