"""Seeded synthetic scenes and cameras for the palette-weight rasterizer.

The reference ships no scenes or fixtures, so every workload is generated here
from the distributions its workloads name:

* means uniform in [-extent, extent]^3; log-scales N(-4.5, 0.5) per axis
  (scales are stored linear, as gsplatcu's computeCov3D expects them,
  gsplat/gau_io.py:126 applies the exp at load); quaternions (w, x, y, z)
  normalised N(0, I4); opacities sigmoid(N(0, 1)).
* K palette channels ordered [black, white, chromatic...] (gsmodel.py:498-500).
* DC palette weights: a Dirichlet(1) draw divided by SH_C0, so that the
  evaluated DC weight SH_C0 * dc (kernel.cu:714, :910) is the Dirichlet
  sample itself and sums to 1 over the K channels.
* view-dependent weights: 15 x K values N(0, 0.05), band-major
  (index lm * K + k), the layout of gsplat/utils.py:157-158 get_shs.
* rank-r factors: A (N, P, r) and B (N, r, Q), entries N(0, 0.1).

Cameras follow the reference's Camera contract (gsplat/gausplat_dataset.py:21-33):
Rcw (3,3), tcw (3,), twc = -Rcw^-1 tcw, fx, fy, cx, cy, width, height, OpenCV
axes (x right, y down, z forward).
"""
import math

import numpy as np

SH_C0 = 0.28209479177387814


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


class Scene:
    """Host (numpy, float32) copy of one synthetic scene."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    @property
    def n(self):
        return self.pws.shape[0]

    def to_torch(self, device):
        import torch
        out = {}
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray):
                out[k] = torch.from_numpy(np.ascontiguousarray(v)).to(device)
            else:
                out[k] = v
        return Scene(**out)


def make_scene(n, seed=0, extent=1.0, K=5, lowrank=False, rank=1, P=None, Q=15,
               sh_std=0.05, factor_std=0.1):
    """Generate a scene of n Gaussians with K palette channels.

    lowrank=False: full-rank coefficients `shs` (N, 16K) = cat(dc, high).
    lowrank=True:  `dc` (N, K), `A` (N, P, rank), `B` (N, rank, Q), P*Q = 15K.
    """
    rng = _rng(seed)
    f32 = np.float32
    pws = rng.uniform(-extent, extent, (n, 3)).astype(f32)
    scales = np.exp(rng.normal(-4.5, 0.5, (n, 3))).astype(f32)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    rots = q.astype(f32)
    alphas = (1.0 / (1.0 + np.exp(-rng.normal(size=n)))).astype(f32)
    palette = np.zeros((K, 3), f32)
    palette[1] = 1.0
    palette[2:] = rng.uniform(0.0, 1.0, (K - 2, 3)).astype(f32)
    dc = (rng.dirichlet(np.ones(K), size=n) / SH_C0).astype(f32)
    sc = dict(pws=pws, scales=scales, rots=rots, alphas=alphas, palette=palette, K=K,
              lowrank=lowrank, seed=seed)
    if not lowrank:
        high = rng.normal(0.0, sh_std, (n, 15 * K)).astype(f32)
        sc["shs"] = np.concatenate([dc, high], axis=1)
        sc["dc"] = dc
        sc["high"] = high
    else:
        if P is None:
            P = 15 * K // Q
        assert P * Q == 15 * K, "P*Q must equal 15*K"
        sc["dc"] = dc
        sc["A"] = rng.normal(0.0, factor_std, (n, P, rank)).astype(f32)
        sc["B"] = rng.normal(0.0, factor_std, (n, rank, Q)).astype(f32)
        sc["P"], sc["Q"], sc["rank"] = P, Q, rank
    return Scene(**sc)


def reconstruct_high(A, B):
    """C = A @ B per Gaussian, flattened row-major: the rank-r form of
    reconstruct_high_shs (gsmodel.py:90-97), numpy."""
    n, P, r = A.shape
    Q = B.shape[2]
    C = np.zeros((n, P, Q), np.float32)
    for j in range(r):
        C = C + A[:, :, j:j + 1] * B[:, j:j + 1, :]
    return C.reshape(n, P * Q)


class Camera:
    """Pinhole camera with the reference Camera fields (numpy float32)."""

    def __init__(self, width, height, fx, fy, cx, cy, Rcw, tcw):
        self.width, self.height = int(width), int(height)
        self.fx, self.fy, self.cx, self.cy = float(fx), float(fy), float(cx), float(cy)
        self.Rcw = np.asarray(Rcw, np.float32).reshape(3, 3)
        self.tcw = np.asarray(tcw, np.float32).reshape(3)
        self.twc = (-np.linalg.inv(self.Rcw.astype(np.float64)) @ self.tcw.astype(np.float64)).astype(np.float32)

    def to_torch(self, device):
        """A reference-style camera object whose Rcw/tcw/twc are device tensors."""
        import torch

        class TorchCamera:
            pass
        c = TorchCamera()
        c.width, c.height = self.width, self.height
        c.fx, c.fy, c.cx, c.cy = self.fx, self.fy, self.cx, self.cy
        c.Rcw = torch.from_numpy(self.Rcw.copy()).to(device)
        c.tcw = torch.from_numpy(self.tcw.copy()).to(device)
        c.twc = torch.from_numpy(self.twc.copy()).to(device)
        return c


def look_at_camera(pos, width, height, fov_x_deg=60.0, target=(0.0, 0.0, 0.0)):
    pos = np.asarray(pos, np.float64)
    f = np.asarray(target, np.float64) - pos
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0])
    if abs(np.dot(up, f)) > 0.99:
        up = np.array([0.0, 0.0, 1.0])
    x = np.cross(up, f)
    x /= np.linalg.norm(x)
    y = np.cross(f, x)
    R = np.stack([x, y, f])
    t = -R @ pos
    fx = 0.5 * width / math.tan(math.radians(fov_x_deg) / 2.0)
    return Camera(width, height, fx, fx, width / 2.0, height / 2.0, R, t)


def orbit_cameras(n_views, radius, width, height, fov_x_deg=60.0, elevation=0.3):
    cams = []
    for i in range(n_views):
        th = 2.0 * math.pi * i / n_views
        pos = (radius * math.cos(elevation) * math.sin(th), radius * math.sin(elevation),
               -radius * math.cos(elevation) * math.cos(th))
        cams.append(look_at_camera(pos, width, height, fov_x_deg))
    return cams


# Named workloads.  C0 is small enough for the oracle to finish in seconds.
def workload_c0(seed=0):
    """20,000 Gaussians in [-1,1]^3, K=5 full rank, one 256x256 60-degree
    camera at radius 4 looking at the origin."""
    scene = make_scene(20000, seed=seed, extent=1.0, K=5)
    cam = orbit_cameras(1, 4.0, 256, 256)[0]
    return scene, cam


def workload_c1(seed=0, n=1_000_000, n_views=32):
    """1,000,000 Gaussians in [-3,3]^3, K=5 full rank, 32 orbit cameras at
    radius 6, 1920x1080, 60-degree horizontal field of view."""
    scene = make_scene(n, seed=seed, extent=3.0, K=5)
    cams = orbit_cameras(n_views, 6.0, 1920, 1080)
    return scene, cams


def workload_c2(seed=0, n=1_000_000, rank=2, n_views=8):
    """Rank-2 factors P=5 x Q=15, K=5, 1080p."""
    scene = make_scene(n, seed=seed, extent=3.0, K=5, lowrank=True, rank=rank, P=5, Q=15)
    cams = orbit_cameras(n_views, 6.0, 1920, 1080)
    return scene, cams
