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
