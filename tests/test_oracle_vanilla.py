"""Pin of the oracle's colour backward with geometry gradients (vanilla drawB,
vanilla_3dgs/gsplatcu/kernel.cu:811-971): float64 central finite differences
of L = sum(dL/dimage * image) + sum(dL/dtau * final_tau), the colour forward
(draw_color, final_tau starting at 1) with the tile lists held fixed.  The
scene keeps every alpha' below the 0.99 clamp (alpha <= 0.6) and away from
the 0.002 cut, and its transmittance far above the 1e-4 stop, where the
reference's derivative formulas are the exact ones."""
import numpy as np
import pytest

from oracle import oracle as O

f64 = np.float64


def scene2d(n, W, H, seed):
    rng = np.random.default_rng(seed)
    us = np.stack([rng.uniform(2, W - 2, n), rng.uniform(2, H - 2, n)], 1)
    th = rng.uniform(0, np.pi, n)
    s1, s2 = rng.uniform(1.5, 4.0, n), rng.uniform(1.5, 4.0, n)
    c, s = np.cos(th), np.sin(th)
    cov = np.stack([c * c * s1 ** 2 + s * s * s2 ** 2, c * s * (s1 ** 2 - s2 ** 2), s * s * s1 ** 2 + c * c * s2 ** 2], 1)
    depths = rng.uniform(1.0, 5.0, n)
    cinv, areas = O.inverse_cov2d(cov, depths.copy(), dt=f64)
    rects, counts = O.get_rects(us, areas, depths.copy(), W, H, dt=f64)
    _, ids, ranges = O.bin_sort(depths, rects, counts, W, H, dt=f64)
    return dict(us=us, cinv=cinv, alphas=rng.uniform(0.1, 0.6, n), colors=rng.uniform(0, 1, (n, 3)), ids=ids,
                ranges=ranges)


def alpha_primes(sc, W, H):
    py, px = np.mgrid[0:H, 0:W]
    dx = sc["us"][:, 0, None, None] - px
    dy = sc["us"][:, 1, None, None] - py
    cx, cy, cz = (sc["cinv"][:, k, None, None] for k in range(3))
    m = np.maximum(0, cx * dx * dx + 2 * cy * dx * dy + cz * dy * dy)
    return np.minimum(0.99, sc["alphas"][:, None, None] * np.exp(-0.5 * m))


def loss(sc, W, H, dimg, dtau):
    img, _, ft, _, _ = O.draw(W, H, sc["ranges"], sc["ids"], sc["us"], sc["cinv"], sc["alphas"], sc["colors"], 1.0,
                              dt=f64)
    return float(np.sum(dimg * img) + np.sum(dtau * ft))


@pytest.mark.parametrize("seed,with_tau", [(3, True), (4, False), (5, True)])
def test_draw_color_backward_fd64(seed, with_tau):
    W, H, n = 40, 28, 14
    sc = scene2d(n, W, H, seed)
    ap = alpha_primes(sc, W, H)
    # a step of 1e-6 moves an alpha' near the cut by ~1e-9: keep 100x that away from it
    assert not np.any(np.abs(ap - 0.002) < 1e-7), "scene has an alpha' at the 0.002 cut"
    rng = np.random.default_rng(seed + 100)
    dimg = rng.normal(size=(3, H, W))
    dtau = rng.normal(size=(H, W)) if with_tau else np.zeros((H, W))
    img, contrib, ft, _, _ = O.draw(W, H, sc["ranges"], sc["ids"], sc["us"], sc["cinv"], sc["alphas"], sc["colors"],
                                    1.0, dt=f64)
    assert ft.min() > 1e-2  # far from the 1e-4 stop
    grad, _ = O.draw_color_backward(W, H, sc["ranges"], sc["ids"], sc["us"], sc["cinv"], sc["alphas"], sc["colors"],
                                    contrib, ft, dimg, dtau if with_tau else None, dt=f64)
    fd = np.zeros_like(grad)
    h = 1e-6
    fields = [("us", 0), ("us", 1), ("cinv", 0), ("cinv", 1), ("cinv", 2), ("alphas", None), ("colors", 0),
              ("colors", 1), ("colors", 2)]
    for g in range(n):
        for e, (key, k) in enumerate(fields):
            vals = []
            for sgn in (1, -1):
                s2 = {kk: (v.copy() if isinstance(v, np.ndarray) else v) for kk, v in sc.items()}
                if k is None:
                    s2[key][g] += sgn * h
                else:
                    s2[key][g, k] += sgn * h
                vals.append(loss(s2, W, H, dimg, dtau))
            fd[g, e] = (vals[0] - vals[1]) / (2 * h)
    scale = np.abs(fd).max()
    np.testing.assert_allclose(grad, fd, rtol=1e-5, atol=1e-7 * scale)
    assert np.abs(grad[:, 0:2]).max() > 0 and np.abs(grad[:, 5]).max() > 0


# ---------------------------------------------------------------------------
# the whole vanilla step (vanilla_3dgs/gsplat/gsmodel.py:6-85) on the oracle
# ---------------------------------------------------------------------------
def oracle_forward(sc, cam, shs, dt, lists=None):
    """project -> cov3d -> cov2d -> sh2color -> inverse -> rects/sort -> colour
    draw (final_tau starts at 1).  lists = (ranges, ids) holds the tile lists
    fixed (finite differences)."""
    W, H = cam.width, cam.height
    us, pcs, dep = O.project(sc["pws"], cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, dt)
    c3 = O.cov3d(sc["rots"], sc["scales"], dep, dt)
    c2 = O.cov2d(c3, pcs, cam.Rcw, dep, cam.fx, cam.fy, W, H, dt)
    col = O.sh2color(shs, sc["pws"], cam.twc, dt)
    cinv, areas = O.inverse_cov2d(c2, dep, dt)
    if lists is None:
        rects, counts = O.get_rects(us, areas, dep, W, H, dt)
        _, ids, ranges = O.bin_sort(dep, rects, counts, W, H, dt)
    else:
        ranges, ids = lists
    img, contrib, ft, _, _ = O.draw(W, H, ranges, ids, us, cinv, sc["alphas"], col, 1.0, dt=dt)
    return dict(us=us, pcs=pcs, dep=dep, c3=c3, c2=c2, col=col, cinv=cinv, img=img, contrib=contrib, ft=ft,
                ranges=ranges, ids=ids)


def oracle_chain(sc, cam, shs, f, dimg, dtau, dt, jac_dt=None):
    """The reference's backward chain in float64: colour backward sums, then
    dL/d(pws, shs, alphas, scales, rots) through the oracle's Jacobians
    (computed in jac_dt) and, with absolute values, a bound on the sum of
    |terms|."""
    jd = jac_dt or dt
    W, H = cam.width, cam.height
    g, gabs = O.draw_color_backward(W, H, f["ranges"], f["ids"], f["us"], f["cinv"], sc["alphas"], f["col"],
                                    f["contrib"], f["ft"], dimg, dtau, dt=dt)
    J = dict(du=O.project_jac(sc["pws"], cam.Rcw, cam.tcw, cam.fx, cam.fy, jd),
             dc3=O.cov3d_jac(sc["rots"], sc["scales"], f["dep"], jd),
             dc2=O.cov2d_jac(f["c3"], f["pcs"], cam.Rcw, f["dep"], cam.fx, cam.fy, W, H, jd),
             dci=O.inverse_jac(f["c2"], f["dep"], jd), dcol=O.sh2color_jac(shs, sc["pws"], cam.twc, jd))
    R = np.asarray(cam.Rcw, np.float64)

    def chain(gg, ab):
        A = (lambda x: np.abs(x.astype(np.float64))) if ab else (lambda x: x.astype(np.float64))
        du, dci, dal, dcol = gg[:, 0:2], gg[:, 2:5], gg[:, 5], gg[:, 6:9]
        dc2 = np.einsum("ni,nij->nj", dci, A(J["dci"]))
        dc3 = np.einsum("ni,nij->nj", dc2, A(J["dc2"][0]))
        drot = np.einsum("ni,nij->nj", dc3, A(J["dc3"][0]))
        dsc = np.einsum("ni,nij->nj", dc3, A(J["dc3"][1]))
        dshs = (dcol[:, None, :] * A(J["dcol"][0]).reshape(len(gg), -1, 1)).reshape(len(gg), -1)
        Rm = np.abs(R) if ab else R
        dpw = (np.einsum("ni,nij->nj", np.einsum("ni,nij->nj", du, A(J["du"])), Rm[None]) +
               np.einsum("ni,nij->nj", dcol, A(J["dcol"][1])) +
               np.einsum("ni,nij->nj", np.einsum("ni,nij->nj", dc2, A(J["dc2"][1])), Rm[None]))
        return dict(pws=dpw, shs=dshs, alphas=dal, scales=dsc, rots=drot)

    return chain(g, False), chain(gabs, True)


def vanilla_scene(n, seed, cam):
    from rgp.scene import make_scene
    sc = make_scene(n, seed=seed, extent=0.8, K=5)
    rng = np.random.default_rng(seed)
    shs = rng.normal(0, 0.3, (n, 48))
    shs[:, 27:30] = 0.0  # no band-9 coefficient: the sh2Color dY_9/dy term is out of play
    return dict(pws=sc.pws.astype(np.float64), rots=sc.rots.astype(np.float64),
                scales=(sc.scales * 6.0).astype(np.float64), alphas=rng.uniform(0.1, 0.6, n)), shs


def test_vanilla_chain_fd64():
    """The oracle's chain (colour backward + the reference's Jacobian
    products) against float64 central differences of the whole forward, on
    Gaussians in front of the camera and inside the image (within the
    computeCov2D clamp), tile lists fixed, no alpha' near the 0.002 cut."""
    from rgp.scene import orbit_cameras
    cam = orbit_cameras(1, 3.0, 48, 36)[0]
    sc, shs = vanilla_scene(24, 7, cam)
    d = np.float64
    f = oracle_forward(sc, cam, shs, d)
    inside = (f["dep"] > 0.3) & (f["us"][:, 0] > 0) & (f["us"][:, 0] < 48) & (f["us"][:, 1] > 0) & (f["us"][:, 1] < 36)
    assert inside.sum() >= 8
    rng = np.random.default_rng(1)
    dimg, dtau = rng.normal(size=(3, 36, 48)), rng.normal(size=(36, 48))
    grads, _ = oracle_chain(sc, cam, shs, f, dimg, dtau, d)
    lists = (f["ranges"], f["ids"])

    def L(s2, sh2):
        f2 = oracle_forward(s2, cam, sh2, d, lists)
        return float(np.sum(dimg * f2["img"]) + np.sum(dtau * f2["ft"]))

    h = 1e-6
    for g in np.flatnonzero(inside)[:8]:
        for key, width in (("pws", 3), ("rots", 4), ("scales", 3), ("alphas", 1)):
            for k in range(width):
                vals = []
                for sgn in (1, -1):
                    s2 = {kk: v.copy() for kk, v in sc.items()}
                    if width == 1:
                        s2[key][g] += sgn * h
                    else:
                        s2[key][g, k] += sgn * h
                    vals.append(L(s2, shs))
                fd = (vals[0] - vals[1]) / (2 * h)
                an = grads[key][g] if width == 1 else grads[key][g, k]
                assert abs(an - fd) <= 1e-5 * max(1.0, abs(fd)), (g, key, k, an, fd)
        for k in (0, 5, 40):
            vals = []
            for sgn in (1, -1):
                sh2 = shs.copy()
                sh2[g, k] += sgn * h
                vals.append(L(sc, sh2))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(grads["shs"][g, k] - fd) <= 1e-5 * max(1.0, abs(fd)), (g, "shs", k)
