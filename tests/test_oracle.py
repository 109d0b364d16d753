"""Pins of the CPU oracle (oracle/oracle.py): each stage is checked against
something other than its own formula: closed forms built with numpy in
float64, invariants, brute force on tiny inputs and float64 central finite
differences.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from rgp.scene import SH_C0, make_scene, orbit_cameras, reconstruct_high

f32, f64 = np.float32, np.float64


def ulps(a, b):
    a = np.asarray(a, f32).view(np.int32).astype(np.int64)
    b = np.asarray(b, f32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


# --------------------------------------------------------------------------
# exp
# --------------------------------------------------------------------------
def test_expf_against_libm():
    x = np.concatenate([np.linspace(-104, 88.7, 2_000_003, dtype=f32),
                        np.random.default_rng(0).uniform(-20, 0, 1_000_000).astype(f32)])
    y = O.expf(x)
    ref = np.exp(x.astype(f64)).astype(f32)  # correctly rounded up to double rounding
    normal = ref > np.finfo(f32).tiny
    assert ulps(y[normal], ref[normal]).max() <= 2
    sub = ~normal & (ref > 0)
    assert np.abs(y[sub].astype(f64) - np.exp(x[sub].astype(f64))).max() <= 2 * 2.0 ** -149


def test_expf_special_values():
    assert O.expf(np.array([0.0], f32))[0] == 1.0
    for n in range(-40, 41):  # near the reduction boundaries r = 0
        x = np.array([n * math.log(2)], f32)
        assert ulps(O.expf(x), np.exp(x.astype(f64)).astype(f32))[0] <= 1
    v = O.expf(np.array([-np.inf, -200.0, np.nan, 100.0], f32))
    assert v[0] == 0.0 and v[1] == 0.0 and np.isnan(v[2]) and np.isinf(v[3])
    x = np.linspace(-30, 0, 100001, dtype=f32)
    assert (np.diff(O.expf(x)) >= 0).all()  # monotone on the compositing range


def test_gauss_half_form_exhaustive():
    """The compositing kernels form exp(-m/2) from the distance m without
    expf's range handling.  On every float of [0, 1e7] (1.26e9 values: every
    distance that can pass the cutoff of a finite opacity, capped at 1e7)
    it must equal expf(-0.5 m) bit for bit, including the subnormal results
    above m = 175.  A plausible slip in the scaled coefficients or the
    reduction constants shows up as differing values."""
    n, d, first = O.gauss_half_check(0.0, 1.0e7)
    assert n > 1_200_000_000
    assert d == 0, f"{d} differing values, first at m = {first}"


def test_gauss_half_fast_form_exhaustive():
    """For entries whose cutoff is at most 170 the kernels scale the
    polynomial by adding n << 23 to its bit pattern instead of ldexp.  On
    every float of [0, 170] (1.12e9 values) that must equal the range-test-
    free form bit for bit; just above, where the result leaves the normal
    range, the exponent add must stop agreeing (a check that the interval
    really is where the shortcut holds, not a vacuous comparison)."""
    n, d, first = O.gauss_half_fast_check(0.0, 170.0)
    assert n > 1_100_000_000
    assert d == 0, f"{d} differing values, first at m = {first}"
    n, d, first = O.gauss_half_fast_check(170.0, 200.0)
    assert d > 0 and 174.0 < first < 176.0


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
@pytest.fixture(scope="module")
def small():
    scene = make_scene(3000, seed=5, extent=1.0, K=5)
    cam = orbit_cameras(5, 4.0, 320, 240)[2]
    return scene, cam


def test_project_closed_form(small):
    scene, cam = small
    us, pcs, depths = O.project(scene.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    pc = scene.pws.astype(f64) @ cam.Rcw.astype(f64).T + cam.tcw.astype(f64)
    assert np.allclose(pcs, pc, rtol=1e-6, atol=1e-6)
    u = np.stack([cam.fx * pc[:, 0] / pc[:, 2] + cam.cx, cam.fy * pc[:, 1] / pc[:, 2] + cam.cy], 1)
    assert np.allclose(us, u, rtol=1e-5, atol=1e-3)
    assert (depths == pcs[:, 2]).all()
    # behind the camera: depth -1, outputs stay zero
    back = (cam.twc - 3.0 * cam.Rcw[2]).reshape(1, 3)
    u2, p2, d2 = O.project(back, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    assert d2[0] == -1 and (u2 == 0).all() and (p2 == 0).all()


def quat_to_R(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_cov3d_closed_form(small):
    scene, cam = small
    depths = np.ones(scene.n, f32)
    c = O.cov3d(scene.rots, scene.scales, depths)
    for i in range(0, scene.n, 97):
        R = quat_to_R(scene.rots[i].astype(f64))
        S = R @ np.diag(scene.scales[i].astype(f64) ** 2) @ R.T
        got = np.array([[c[i, 0], c[i, 1], c[i, 2]], [c[i, 1], c[i, 3], c[i, 4]], [c[i, 2], c[i, 4], c[i, 5]]])
        assert np.allclose(got, S, rtol=1e-4, atol=1e-12)
        assert np.linalg.det(got.astype(f64)) == pytest.approx(np.prod(scene.scales[i].astype(f64)) ** 2, rel=5e-3)
    # culled Gaussians keep zeros
    assert (O.cov3d(scene.rots[:4], scene.scales[:4], np.full(4, -1, f32)) == 0).all()


def test_cov2d_closed_form(small):
    scene, cam = small
    us, pcs, depths = O.project(scene.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    c3 = O.cov3d(scene.rots, scene.scales, depths)
    c2 = O.cov2d(c3, pcs, cam.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height)
    tfx, tfy = cam.width / (2 * cam.fx), cam.height / (2 * cam.fy)
    W = cam.Rcw.astype(f64)
    for i in range(0, scene.n, 101):
        x, y, z = pcs[i].astype(f64)
        x = min(1.3 * tfx, max(-1.3 * tfx, x / z)) * z
        y = min(1.3 * tfy, max(-1.3 * tfy, y / z)) * z
        J = np.array([[cam.fx / z, 0, -cam.fx * x / z ** 2], [0, cam.fy / z, -cam.fy * y / z ** 2]])
        a, b, c, d, e, f = c3[i].astype(f64)
        S = np.array([[a, b, c], [b, d, e], [c, e, f]])
        ref = J @ W @ S @ W.T @ J.T + 0.3 * np.eye(2)
        assert np.allclose([c2[i, 0], c2[i, 1], c2[i, 2]], [ref[0, 0], ref[0, 1], ref[1, 1]], rtol=1e-4, atol=1e-5)
        assert c2[i, 0] > 0 and c2[i, 0] * c2[i, 2] - c2[i, 1] ** 2 > 0


def test_inverse_cov2d_properties():
    rng = np.random.default_rng(1)
    L = rng.normal(size=(500, 2, 2))
    C = L @ L.transpose(0, 2, 1) + 0.3 * np.eye(2)
    cov = np.stack([C[:, 0, 0], C[:, 0, 1], C[:, 1, 1]], 1).astype(f32)
    depths = np.ones(500, f32)
    cinv, areas = O.inverse_cov2d(cov, depths)
    inv = np.stack([np.stack([cinv[:, 0], cinv[:, 1]], 1), np.stack([cinv[:, 1], cinv[:, 2]], 1)], 1).astype(f64)
    prod = inv @ np.stack([np.stack([cov[:, 0], cov[:, 1]], 1), np.stack([cov[:, 1], cov[:, 2]], 1)], 1)
    assert np.allclose(prod, np.eye(2)[None], atol=1e-3)
    assert (areas[:, 0] == np.ceil(3 * np.sqrt(cov[:, 0].astype(f64)))).mean() > 0.99
    # NaN determinant culls; a zero determinant gives inf and does not
    d2 = np.ones(2, f32)
    ci, ar = O.inverse_cov2d(np.array([[np.nan, 0, 1], [1, 1, 1]], f32), d2)
    assert d2[0] == -1 and d2[1] == 1 and np.isinf(ci[1]).all()


# --------------------------------------------------------------------------
# SH weights
# --------------------------------------------------------------------------
def test_sh_basis_addition_theorem(small):
    scene, cam = small
    K = scene.K
    w, basis = O.sh2weight(scene.shs, scene.pws, cam.twc, K, f64)
    for l, (a, b) in enumerate(((0, 1), (1, 4), (4, 9), (9, 16))):
        s = (basis[:, a:b] ** 2).sum(1)
        assert np.allclose(s, (2 * l + 1) / (4 * math.pi), rtol=1e-9)
    ref = np.einsum("nl,nlk->nk", basis, scene.shs.astype(f64).reshape(scene.n, 16, K))
    assert np.allclose(w, ref, rtol=1e-12, atol=1e-12)
    w32, b32 = O.sh2weight(scene.shs, scene.pws, cam.twc, K)
    assert np.allclose(w32, w, rtol=1e-5, atol=1e-6)
    assert (b32[:, 0] == f32(SH_C0)).all()


@pytest.mark.parametrize("rank", [1, 2, 3])
def test_fast_equals_full_on_reconstruction(small, rank):
    _, cam = small
    sc = make_scene(2000, seed=7, K=5, lowrank=True, rank=rank, P=5, Q=15)
    wf, ylm = O.sh2weight_fast(sc.dc, sc.A, sc.B, sc.pws, cam.twc, 5, 15, f64)
    shs = np.concatenate([sc.dc, reconstruct_high(sc.A, sc.B)], 1)
    w, basis = O.sh2weight(shs.astype(f64), sc.pws, cam.twc, 5, f64)
    C = np.einsum("npr,nrq->npq", sc.A.astype(f64), sc.B.astype(f64)).reshape(sc.n, -1)
    w_exact = O.sh2weight(np.concatenate([sc.dc.astype(f64), C], 1), sc.pws, cam.twc, 5, f64)[0]
    assert np.allclose(wf, w_exact, rtol=1e-12, atol=1e-14)
    assert np.allclose(ylm, basis[:, 1:], rtol=1e-14)
    wf32, _ = O.sh2weight_fast(sc.dc, sc.A, sc.B, sc.pws, cam.twc, 5, 15)
    assert np.allclose(wf32, w_exact, rtol=1e-5, atol=1e-6)


def test_rank1_layouts_agree(small):
    _, cam = small
    sc = make_scene(1000, seed=8, K=5, lowrank=True, rank=1, P=5, Q=15)
    a, ya = O.sh2weight_fast(sc.dc, sc.A, sc.B, sc.pws, cam.twc, 5, 15)
    b, yb = O.sh2weight_fast(sc.dc, sc.A[:, :, 0], sc.B[:, 0, :], sc.pws, cam.twc, 5, 15)
    assert np.array_equal(a, b) and np.array_equal(ya, yb)


def test_fast_backward_adjoint_and_fd(small):
    _, cam = small
    sc = make_scene(64, seed=9, K=5, lowrank=True, rank=2, P=5, Q=15)
    A, B = sc.A.astype(f64), sc.B.astype(f64)
    rng = np.random.default_rng(2)
    dw = rng.normal(size=(64, 5))
    _, ylm = O.sh2weight_fast(sc.dc, A, B, sc.pws, cam.twc, 5, 15, f64)
    dA, dB = O.sh2weight_fast_backward(A, B, ylm, dw, 5, 15, f64)

    def L(A_, B_):
        return (O.sh2weight_fast(sc.dc, A_, B_, sc.pws, cam.twc, 5, 15, f64)[0] * dw).sum()
    eps = 1e-4
    for idx in [(0, 0, 0), (3, 4, 1), (63, 2, 0), (10, 1, 1)]:
        Ap, Am = A.copy(), A.copy()
        Ap[idx] += eps
        Am[idx] -= eps
        assert dA[idx] == pytest.approx((L(Ap, B) - L(Am, B)) / (2 * eps), rel=1e-7, abs=1e-10)
    for idx in [(0, 0, 0), (5, 1, 14), (40, 0, 7)]:
        Bp, Bm = B.copy(), B.copy()
        Bp[idx] += eps
        Bm[idx] -= eps
        assert dB[idx] == pytest.approx((L(A, Bp) - L(A, Bm)) / (2 * eps), rel=1e-7, abs=1e-10)


# --------------------------------------------------------------------------
# binning: brute force
# --------------------------------------------------------------------------
def test_rects_cover_the_box():
    """A rect holds every tile that meets [u - r, u + r - 1] and nothing more
    than one tile beyond it.  (The upper bound floor((u + r + 15) / 16) of
    kernel.cu:149-150 leaves out the tile whose first pixel column lies in
    (u + r - 1, u + r].)"""
    rng = np.random.default_rng(4)
    n, W, H = 400, 200, 136
    us = np.stack([rng.uniform(-40, W + 40, n), rng.uniform(-40, H + 40, n)], 1).astype(f32)
    areas = rng.integers(1, 30, (n, 2)).astype(np.int32)
    depths = np.ones(n, f32)
    rects, counts = O.get_rects(us, areas.copy(), depths, W, H)
    gx, gy = O.grid_dims(W, H)
    for i in range(n):
        lo = np.floor((us[i].astype(f64) - areas[i]) / 16)
        hi = np.floor((us[i].astype(f64) + areas[i] - 1.001) / 16)
        need_x = set(range(max(0, int(lo[0])), min(gx, int(hi[0]) + 1)))
        need_y = set(range(max(0, int(lo[1])), min(gy, int(hi[1]) + 1)))
        x0, y0, x1, y1 = rects[i]
        if counts[i] == 0:
            assert depths[i] == -1
            continue
        got_x, got_y = set(range(x0, x1)), set(range(y0, y1))
        assert need_x <= got_x and need_y <= got_y
        assert len(got_x - need_x) <= 1 and len(got_y - need_y) <= 1
        assert counts[i] == len(got_x) * len(got_y)


def test_bin_sort_brute_force():
    rng = np.random.default_rng(6)
    n, W, H = 300, 100, 70
    gx, gy = O.grid_dims(W, H)
    depths = rng.uniform(0.1, 9.0, n).astype(f32)
    depths[::17] = depths[3]  # ties in quantised depth
    x0 = rng.integers(0, gx, n)
    y0 = rng.integers(0, gy, n)
    rects = np.stack([x0, y0, np.minimum(gx, x0 + rng.integers(1, 4, n)),
                      np.minimum(gy, y0 + rng.integers(1, 3, n))], 1).astype(np.uint32)
    counts = ((rects[:, 2] - rects[:, 0]) * (rects[:, 3] - rects[:, 1])).astype(np.uint32)
    counts[depths < 0.2] = 0
    keys, ids, ranges = O.bin_sort(depths, rects, counts, W, H)
    pairs = []
    for i in range(n):
        if depths[i] < 0.2:
            continue
        for y in range(rects[i, 1], rects[i, 3]):
            for x in range(rects[i, 0], rects[i, 2]):
                pairs.append((y * gx + x, int(np.float32(depths[i]) * np.float32(1000)), i))
    pairs.sort()  # (tile, quantised depth, id): the stable order
    assert [p[2] for p in pairs] == ids.tolist()
    assert [(p[0] << 32) | p[1] for p in pairs] == keys.tolist()
    for t in range(gx * gy):
        sel = [j for j, p in enumerate(pairs) if p[0] == t]
        assert tuple(ranges[t]) == ((sel[0], sel[-1] + 1) if sel else (0, 0))


def test_bin_sort_m0_m1():
    d = np.array([1.0, -1.0], f32)
    r = np.array([[1, 1, 2, 2], [0, 0, 0, 0]], np.uint32)
    keys, ids, ranges = O.bin_sort(d, r, np.array([1, 0], np.uint32), 64, 64)
    assert ids.tolist() == [0] and (ranges == 0).all()  # M == 1: the reference leaves (0,0)
    keys, ids, ranges = O.bin_sort(d, r, np.array([0, 0], np.uint32), 64, 64)
    assert len(ids) == 0 and (ranges == 0).all()


# --------------------------------------------------------------------------
# compositing
# --------------------------------------------------------------------------
def tiny_scene():
    sc = make_scene(40, seed=12, extent=0.6, K=3)
    sc.scales *= 6.0
    cam = orbit_cameras(1, 3.0, 40, 36)[0]
    return sc, cam


def fma32(a, b, c):
    """a * b + c rounded once to float32 (nearest, ties to even), from the
    exact rational value: the single rounding of a fused multiply-add."""
    from fractions import Fraction
    q = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    best = np.float32(float(q))  # within one float32 ulp of q
    for y in (np.nextafter(best, np.float32(-np.inf)), np.nextafter(best, np.float32(np.inf))):
        dy, db = abs(Fraction(float(y)) - q), abs(Fraction(float(best)) - q)
        if dy < db or (dy == db and int(y.view(np.int32)) % 2 == 0):
            best = y
    return np.float32(best)


def test_fma32_is_one_rounding():
    """fma32 against cases where the two-rounding result differs."""
    a = np.float32(1 + 2 ** -12)
    # a*a = 1 + 2^-11 + 2^-24: separately rounded the 2^-24 is lost
    assert fma32(a, a, np.float32(-1)) == np.float32(2 ** -11 + 2 ** -24)
    assert a * a - np.float32(1) == np.float32(2 ** -11)
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = rng.normal(size=3).astype(np.float32)
        exact = float(x) * float(y) + float(z)
        r = fma32(x, y, z)
        assert abs(float(r) - exact) <= 0.5 * float(np.spacing(np.float32(abs(exact)))) * (1 + 1e-12)


def test_draw_pure_python_recurrence():
    """Per pixel over the whole sorted list of its tile, the recurrence of
    kernel.cu:326-362 written as a plain Python loop (float32 scalars, the
    oracle's exp; mahaSqDist's two multiply-adds and the weight update
    contracted into single roundings, as the oracle and the kernels do)."""
    sc, cam = tiny_scene()
    o = O.forward_weight(sc, cam)
    us, cinv, w, al = o["us"], o["cinv"], o["weights"], sc.alphas
    gx, _ = O.grid_dims(cam.width, cam.height)
    strips = np.zeros(len(o["ids"]), np.uint8)  # rows strips (py mod 16) // 4 each entry contributed in
    nz = 0
    for py in range(cam.height):
        for px in range(cam.width):
            t = (py // 16) * gx + px // 16
            a, b = o["ranges"][t]
            acc = [f32(0)] * 3
            tau, cont = f32(1), 0
            for j in range(a, b):
                g = o["ids"][j]
                dx, dy = us[g, 0] - f32(px), us[g, 1] - f32(py)
                m = fma32(f32(2) * cinv[g, 1] * dx, dy, fma32(cinv[g, 2] * dy, dy, cinv[g, 0] * dx * dx))
                m = max(f32(0), m)
                ap = min(f32(0.99), al[g] * O.expf(np.array([f32(-0.5) * m], f32))[0])
                if ap < f32(0.002):
                    continue
                acc = [fma32(tau * ap, w[g, k], acc[k]) for k in range(3)]
                strips[j] |= 1 << ((py % 16) // 4)
                cont = j - a + 1
                tau = tau * (f32(1) - ap)
                if tau < f32(0.0001):
                    break
            if b - a == 0:
                tau = f32(0)
            assert [o["layer"][k, py, px] for k in range(3)] == acc
            assert o["contrib"][py, px] == cont and o["final_tau"][py, px] == tau
            nz += cont > 0
    assert nz > 100
    got = O.draw(cam.width, cam.height, o["ranges"], o["ids"], us, cinv, al, w, strips=True)[5]
    assert np.array_equal(got, strips) and (strips != 0).sum() > 50 and len(np.unique(strips)) > 4


def test_draw_unit_weights_invariant():
    """With every weight 1, each channel equals 1 - final_tau (exact in R)."""
    sc, cam = tiny_scene()
    sc = make_scene(2000, seed=3, extent=1.0, K=3)
    cam = orbit_cameras(1, 4.0, 128, 96)[0]
    o = O.forward_weight(sc, cam)
    layer, contrib, tau, _, _ = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"],
                                      sc.alphas, np.ones((sc.n, 3), f32), 1.0)
    nonempty = contrib > 0
    assert nonempty.sum() > 1000
    assert np.allclose(layer[:, nonempty], 1 - tau[nonempty], atol=2e-5)


def test_draw_fp32_vs_fp64(small):
    scene, cam = small
    a = O.forward_weight(scene, cam)
    b = O.forward_weight(scene, cam, dt=f64)
    assert np.array_equal(a["ids"], b["ids"])
    same = a["contrib"] == b["contrib"]
    assert same.mean() > 0.999
    assert np.abs(a["layer"] - b["layer"])[:, same].max() < 1e-5


def test_draw_backward_fd64(small):
    """The layer is linear in the per-Gaussian weights, so central
    differences of L = sum(dL * layer) are exact up to rounding in float64."""
    scene, cam = small
    o = O.forward_weight(scene, cam, dt=f64)
    dl = np.random.default_rng(3).normal(size=o["layer"].shape)
    g, gabs = O.draw_backward(cam.width, cam.height, scene.K, o["ranges"], o["ids"], o["us"], o["cinv"],
                              scene.alphas, o["contrib"], o["final_tau"], dl, f64)
    w = o["weights"]
    busy = np.argsort(-np.abs(g).sum(1))[:8]
    eps = 1e-4
    for i in busy:
        for k in (0, 3):
            wp, wm = w.copy(), w.copy()
            wp[i, k] += eps
            wm[i, k] -= eps
            lp = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"], scene.alphas, wp, 0.0, f64)[0]
            lm = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"], scene.alphas, wm, 0.0, f64)[0]
            fd = ((lp - lm) * dl).sum() / (2 * eps)
            assert g[i, k] == pytest.approx(fd, rel=1e-7, abs=1e-9)


def test_mix_linear_identity(small):
    """mix(draw(w), P) == draw(w @ P): bake-then-render (gsmodel.py:488-509)."""
    scene, cam = small
    o = O.forward_weight(scene, cam, dt=f64)
    P = scene.palette.astype(f64)
    rgb = O.mix(P, o["layer"], f64)
    assert np.allclose(rgb, np.einsum("kc,khw->chw", P, o["layer"]), rtol=1e-12, atol=1e-14)
    baked = O.draw(cam.width, cam.height, o["ranges"], o["ids"], o["us"], o["cinv"], scene.alphas,
                   o["weights"] @ P, 0.0, f64)[0]
    assert np.allclose(rgb, baked, rtol=1e-10, atol=1e-12)


def test_backward_replay_vs_division(small):
    """The replayed transmittance (tau in front of each entry, as the forward
    used it) and the reference's division-recovered one give the same
    gradient up to FP32 rounding; both agree with float64."""
    scene, cam = small
    o32 = O.forward_weight(scene, cam)
    o64 = O.forward_weight(scene, cam, dt=f64)
    same = (o32["contrib"] == o64["contrib"]).all()
    dl = np.random.default_rng(8).normal(size=o32["layer"].shape).astype(f32)
    div = O.backward_weight(scene, o32, dl)
    rep = O.backward_weight(scene, o32, dl, replay=True)
    ref = O.backward_weight(scene, o64, dl.astype(f64))
    scale = div["dweights_abs"] + 1e-30
    assert (np.abs(div["dweights"] - rep["dweights"]) <= 1e-5 * scale).all()
    if same:
        # FP32 vs FP64 differ mostly through alpha' itself (~1e-5 of sum|terms|)
        assert (np.abs(rep["dweights"] - ref["dweights"]) <= 1e-4 * scale).all()
        assert (np.abs(div["dweights"] - ref["dweights"]) <= 1e-4 * scale).all()
        # both errors come mostly from the FP32 alpha' they share: the replay is
        # not less accurate than the division (within 1% of its total error)
        assert np.abs(rep["dweights"] - ref["dweights"]).sum() <= 1.01 * np.abs(div["dweights"] - ref["dweights"]).sum()


# --------------------------------------------------------------------------
# calc_J outputs: float64 central differences of the forward functions
# --------------------------------------------------------------------------
def _fd(fn, x, eps):
    """d fn(x) / dx by central differences: fn maps (n, d) -> (n, ...)."""
    cols = []
    for j in range(x.shape[1]):
        xp, xm = x.copy(), x.copy()
        xp[:, j] += eps
        xm[:, j] -= eps
        cols.append((fn(xp) - fn(xm)) / (2 * eps))
    return np.stack(cols, axis=-1)


def _rel(a, b):
    return float(np.abs(a - b).max() / (np.abs(b).max() + 1e-300))


def test_jacobians_fd64():
    """Every calc_J output of project, computeCov3D, computeCov2D,
    inverseCov2D and sh2Weight against central differences of the float64
    forward it differentiates (computeCov2D on Gaussians inside the +-1.3
    tan_fov clamp, whose derivative the reference leaves out)."""
    sc = make_scene(300, seed=17, extent=1.0, K=5)
    cam = orbit_cameras(1, 3.5, 160, 120)[0]
    n, d = sc.n, f64
    pws = sc.pws.astype(d)
    R = cam.Rcw.astype(d)
    us, pcs, dep = O.project(pws, R, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, d)
    ok = dep > 0.2
    assert ok.sum() > 200
    # project: d us / d pws = du/dpc R
    du = O.project_jac(pws, R, cam.tcw, cam.fx, cam.fy, d)
    fd = _fd(lambda p: O.project(p, R, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, d)[0], pws, 1e-6)
    assert _rel((du @ R)[ok], fd[ok]) < 1e-7
    assert (du[:, 0, 1] == 0).all() and (du[:, 1, 0] == 0).all()
    # computeCov3D
    rots, scales = sc.rots.astype(d), sc.scales.astype(d)
    dr, ds = O.cov3d_jac(rots, scales, dep, d)
    assert _rel(dr[ok], _fd(lambda q: O.cov3d(q, scales, dep, d), rots, 1e-6)[ok]) < 1e-7
    assert _rel(ds[ok], _fd(lambda s: O.cov3d(rots, s, dep, d), scales, 1e-8)[ok]) < 1e-6
    # computeCov2D
    c3 = O.cov3d(rots, scales, dep, d)
    W, H = cam.width, cam.height
    tfx, tfy = O.tan_fov(W, H, cam.fx, cam.fy, d)
    inside = ok & (np.abs(pcs[:, 0] / pcs[:, 2]) < 1.29 * tfx) & (np.abs(pcs[:, 1] / pcs[:, 2]) < 1.29 * tfy)
    assert inside.sum() > 100
    d3, dp = O.cov2d_jac(c3, pcs, R, dep, cam.fx, cam.fy, W, H, d)
    c2fn = lambda c, p: O.cov2d(c, p, R, dep, cam.fx, cam.fy, W, H, d)
    assert _rel(d3[inside], _fd(lambda c: c2fn(c, pcs), c3, 1e-9)[inside]) < 1e-6
    assert _rel(dp[inside], _fd(lambda p: c2fn(c3, p), pcs, 1e-7)[inside]) < 1e-6
    # inverseCov2D
    c2 = c2fn(c3, pcs)
    dep2 = dep.copy()
    O.inverse_cov2d(c2, dep2, d)
    di = O.inverse_jac(c2, dep2, d)
    ok2 = dep2 > 0.2

    def inv(c):
        dd = dep.copy()
        return O.inverse_cov2d(c, dd, d)[0]
    assert _rel(di[ok2], _fd(inv, c2, 1e-6)[ok2]) < 1e-6
    # sh2Weight: dW/dpw through the view direction.  Quirk kept from the
    # reference (kernel.cu:840): dY_9/dy, Y_9 = C30 y (3x^2 - y^2), is taken
    # as C30 (3x^2 - 2y^2) instead of C30 (3x^2 - 3y^2), so the result
    # exceeds the true derivative by C30 sh_9 y^2 ddir_y/dpw.
    shs = sc.shs.astype(d)
    K = sc.K
    dw = O.sh2weight_jac(shs, pws, cam.twc, K, d)
    fdw = _fd(lambda p: O.sh2weight(shs, p, cam.twc, K, d)[0], pws, 1e-6)
    dv = pws - cam.twc.astype(d)
    nrm = np.linalg.norm(dv, axis=1)
    dirv = dv / nrm[:, None]
    ddir_y = (np.eye(3)[1][None, :] - dirv[:, 1:2] * dirv) / nrm[:, None]  # d dir_y / d pw
    quirk = (-0.5900435899266435 * shs[:, 9 * K:10 * K] * (dirv[:, 1:2] ** 2))[:, :, None] * ddir_y[:, None, :]
    assert _rel(dw, fdw + quirk) < 1e-7
    assert np.abs(quirk).max() > 1e-3 * np.abs(fdw).max()  # the quirk is visible here
    for sh_dim in (1, 4, 9):
        sub = np.ascontiguousarray(shs[:, : sh_dim * sc.K])
        dws = O.sh2weight_jac(sub, pws, cam.twc, sc.K, d)
        fds = _fd(lambda p: O.sh2weight(sub, p, cam.twc, sc.K, d)[0], pws, 1e-6)
        assert np.abs(dws - fds).max() <= 1e-7 * (np.abs(fdw).max() + 1e-12), sh_dim


def test_sh2color_jacobians_fd64():
    """sh2Color's calc_J outputs against central differences of the float64
    colour: dcolor/dpw carries the reference's dY_9/dy = C30 (3x^2 - 2y^2)
    (kernel.cu:1152), exceeding the true derivative by C30 sh_9 y^2 ddir_y/dpw;
    dcolor/dshs is the basis (the colour is linear in the coefficients)."""
    d = np.float64
    sc = make_scene(300, seed=19, extent=1.0, K=5)
    cam = orbit_cameras(1, 3.5, 160, 120)[0]
    n = sc.n
    pws = sc.pws.astype(d)
    shs = np.random.default_rng(5).normal(0, 0.3, (n, 48)).astype(d)
    dshs, dpws = O.sh2color_jac(shs, pws, cam.twc, d)
    fd = _fd(lambda p: O.sh2color(shs, p, cam.twc, d), pws, 1e-6)
    dv = pws - cam.twc.astype(d)
    nrm = np.linalg.norm(dv, axis=1)
    dirv = dv / nrm[:, None]
    ddir_y = (np.eye(3)[1][None, :] - dirv[:, 1:2] * dirv) / nrm[:, None]
    quirk = (-0.5900435899266435 * shs[:, 27:30] * (dirv[:, 1:2] ** 2))[:, :, None] * ddir_y[:, None, :]
    assert _rel(dpws, fd + quirk) < 1e-7
    assert np.abs(quirk).max() > 1e-3 * np.abs(fd).max()
    fds = _fd(lambda h: O.sh2color(h, pws, cam.twc, d), shs, 1e-6)  # (n, 3, 48)
    for c in range(3):
        np.testing.assert_allclose(fds[:, c, c::3], dshs[:, 0, :], rtol=0, atol=1e-8)
    for sh_dim in (1, 4, 9):
        sub = np.ascontiguousarray(shs[:, : 3 * sh_dim])
        _, dp = O.sh2color_jac(sub, pws, cam.twc, d)
        fdp = _fd(lambda p: O.sh2color(sub, p, cam.twc, d), pws, 1e-6)
        assert np.abs(dp - fdp).max() <= 1e-7 * (np.abs(fd).max() + 1e-12), sh_dim


def test_jacobians_fp32_vs_fp64():
    """The FP32 Jacobians follow their float64 values within FP32 rounding
    (cancellation aside: the bar is relative to the largest entry)."""
    sc = make_scene(300, seed=18, extent=1.0, K=5)
    cam = orbit_cameras(1, 3.5, 160, 120)[0]
    out = {}
    for dt in (f32, f64):
        us, pcs, dep = O.project(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, dt)
        c3 = O.cov3d(sc.rots, sc.scales, dep, dt)
        c2 = O.cov2d(c3, pcs, cam.Rcw, dep, cam.fx, cam.fy, cam.width, cam.height, dt)
        out[dt] = [O.project_jac(sc.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, dt),
                   *O.cov3d_jac(sc.rots, sc.scales, dep, dt),
                   *O.cov2d_jac(c3, pcs, cam.Rcw, dep, cam.fx, cam.fy, cam.width, cam.height, dt),
                   O.inverse_jac(c2, dep, dt), O.sh2weight_jac(sc.shs, sc.pws, cam.twc, sc.K, dt)]
    for a, b in zip(out[f32], out[f64]):
        assert _rel(a.astype(f64), b) < 2e-5
