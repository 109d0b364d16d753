"""Pins of the barycentric palette loss oracle (oracle/rgp_oracle_bary.cpp,
restating gsplat/gsmodel.py:219-285 and the loss terms of
finetune.py:353-392) by means independent of its formulas: numpy.linalg.solve
for the facet systems, numpy.logaddexp for softplus, scipy.special.softmax,
the identities the construction implies for any tau, special pixels, and
float64 central finite differences for the backward."""
import numpy as np
import pytest
from scipy.special import softmax

from oracle import oracle as O

f32, f64 = np.float32, np.float64


def palette(Kc, seed):
    """Well-conditioned chromatic colours around the grey axis."""
    rng = np.random.default_rng(seed)
    h = np.sort(rng.uniform(0, 2 * np.pi, Kc)) if Kc > 1 else np.array([0.3])
    r = rng.uniform(0.25, 0.45, Kc)
    u, v = np.array([1, -1, 0]) / np.sqrt(2), np.array([1, 1, -2]) / np.sqrt(6)
    P = 0.5 + r[:, None] * (np.cos(h)[:, None] * u + np.sin(h)[:, None] * v) + rng.uniform(-0.1, 0.1, (Kc, 1))
    return np.clip(P, 0.02, 0.98)


def reference_numpy(C, P, tau):
    """barycentric_coords_3d in float64 by other means: each facet system by
    numpy.linalg.solve, softplus by logaddexp, softmax by scipy."""
    Kc = P.shape[0]
    H, W = C.shape[1:]
    rhs = C.reshape(3, -1).T
    Paug = np.vstack([np.zeros(3), np.ones(3), P])
    raws, errs = [], []
    for k in range(Kc):
        i, j = k, (k + 1) % Kc
        A = np.stack([np.ones(3), P[i], P[j]], axis=1)
        bgd = np.linalg.solve(A, rhs.T).T  # beta, gamma, delta
        coords = np.concatenate([1 - bgd.sum(1, keepdims=True), bgd], axis=1)
        raw = np.zeros((rhs.shape[0], Kc + 2))
        raw[:, 0], raw[:, 1], raw[:, i + 2] = coords[:, 0], coords[:, 1], coords[:, 2]
        raw[:, j + 2] = coords[:, 3]
        sp = np.logaddexp(0.0, coords)
        proj = np.zeros_like(raw)
        q = sp / (sp.sum(1, keepdims=True) + 1e-12)
        proj[:, 0], proj[:, 1], proj[:, i + 2] = q[:, 0], q[:, 1], q[:, 2]
        proj[:, j + 2] = q[:, 3]
        errs.append(((proj @ Paug - rhs) ** 2).sum(1))
        raws.append(raw)
    p = softmax(-np.stack(errs, 1) / tau, axis=1)
    Wb = np.einsum("nk,nkm->nm", p, np.stack(raws, 1))
    return Wb.T.reshape(Kc + 2, H, W)


def test_log1pf_accuracy():
    y = np.concatenate([np.geomspace(1e-30, 4.85e8, 200_000), np.linspace(0, 4, 100_001)]).astype(f32)
    got = O.log1pf(y).astype(f64)
    want = np.log1p(y.astype(f64))
    ulp = np.spacing(np.abs(want).astype(f32)).astype(f64)
    ulp[want == 0] = np.finfo(f32).tiny
    assert np.max(np.abs(got - want) / ulp) <= 4.0
    assert O.log1pf(np.array([0.0], f32))[0] == 0.0
    tiny = np.array([2.0 ** -25, 1e-10, 1e-30], f32)
    assert np.array_equal(O.log1pf(tiny), tiny)


@pytest.mark.parametrize("Kc,tau", [(3, 1e-9), (3, 0.05), (5, 1e-9), (5, 0.5), (2, 0.05)])
def test_forward_matches_numpy(Kc, tau):
    rng = np.random.default_rng(Kc)
    P = palette(Kc, Kc)
    C = rng.uniform(0, 1, (3, 16, 16))
    Wt = np.zeros((Kc + 2, 16, 16))
    r = O.bary_loss(C, Wt, P, tau=tau, dt=f64)
    want = reference_numpy(C, P, tau)
    if tau < 1e-6:  # hard argmin: pixels within 1e-6 of a tie are ill-posed, skip them
        errs_gap_ok = np.ones(C.shape[1:], bool)
        w2 = reference_numpy(C, P, 1e-7)
        errs_gap_ok &= np.all(np.isclose(w2, want, atol=1e-9), axis=0)
        assert errs_gap_ok.mean() > 0.95
        np.testing.assert_allclose(r["wbary"][:, errs_gap_ok], want[:, errs_gap_ok], rtol=1e-9, atol=1e-9)
    else:
        np.testing.assert_allclose(r["wbary"], want, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("tau", [1e-9, 0.05, 3.0])
def test_weights_sum_to_one_and_reconstruct(tau):
    """sum_m W = sum_k p_k (alpha+beta+gamma+delta) = 1 and W @ [black; white;
    P] = sum_k p_k C = C, for any tau, since every facet's coordinates solve
    its own system exactly."""
    rng = np.random.default_rng(7)
    P = palette(4, 11)
    C = rng.uniform(0, 1, (3, 12, 12))
    r = O.bary_loss(C, np.zeros((6, 12, 12)), P, tau=tau, dt=f64)
    Wb = r["wbary"]
    np.testing.assert_allclose(Wb.sum(0), 1.0, atol=1e-12)
    Paug = np.vstack([np.zeros(3), np.ones(3), P])
    np.testing.assert_allclose(np.einsum("mhw,mc->chw", Wb, Paug), C, atol=1e-11)
    r32 = O.bary_loss(C, np.zeros((6, 12, 12)), P, tau=tau, dt=f32)
    np.testing.assert_allclose(r32["wbary"].sum(0), 1.0, atol=2e-5)


def test_palette_colour_is_one_hot():
    P = palette(4, 3)
    C = np.repeat(P.T[:, :, None], 2, axis=2)  # (3, 4, 2): pixel (i, .) = P_i
    r = O.bary_loss(C, np.zeros((6, 4, 2)), P, dt=f64)
    for i in range(4):
        want = np.zeros(6)
        want[i + 2] = 1.0
        # the coordinates divide by det + 1e-12 (gsmodel.py:249): a 1e-12/det bias
        np.testing.assert_allclose(r["wbary"][:, i, 0], want, atol=1e-10)


def test_single_colour_gamma_slot():
    """K_c = 1: i = j = 0, so delta's assignment overwrites gamma's slot
    (gsmodel.py:264-267) and the facet is singular (det + eps = 1e-12)."""
    P = np.array([[0.8, 0.2, 0.1]])
    C = np.array([[0.5], [0.4], [0.3]])[:, :, None]
    r = O.bary_loss(C, np.zeros((3, 1, 1)), P, dt=f64)
    A = np.stack([np.ones(3), P[0], P[0]], axis=1)
    # adjugate from the 2x2 minors (cofactor transpose), det from numpy
    cof = np.array([[(-1) ** (r + c) * np.linalg.det(np.delete(np.delete(A, r, 0), c, 1)) for c in range(3)]
                    for r in range(3)])
    # two equal columns: the first-row expansion cancels to exactly 0 in floating point
    adj, det = cof.T, 0.0
    b, g, d = adj @ C[:, 0, 0] / (det + 1e-12)
    np.testing.assert_allclose(r["wbary"][:, 0, 0], [1 - b - g - d, b, d], rtol=1e-6)


def test_loss_assembly():
    rng = np.random.default_rng(2)
    P = palette(3, 5)
    C = rng.uniform(0, 1, (3, 10, 9))
    Wt = rng.normal(0.2, 0.4, (5, 10, 9))
    a = rng.uniform(0.0, 1.0, (10, 9))
    for clamp in (False, True):
        for alpha in (None, a):
            r = O.bary_loss(C, Wt, P, alpha=alpha, clamp=clamp, lam_pos=0.5, lam_alpha=1.0, dt=f64)
            Cs = C if alpha is None else np.clip(C / np.maximum(alpha, 1e-3), 0, 1)
            Wb = reference_numpy(Cs, P, 1e-9)
            if clamp:
                Wb = np.maximum(Wb, 0)
                Wb = Wb / np.maximum(Wb.sum(0, keepdims=True), 1e-8)
            if alpha is not None:
                Wb = Wb * alpha
            ok = np.all(np.isclose(reference_numpy(Cs, P, 1e-7), reference_numpy(Cs, P, 1e-9), atol=1e-9), axis=0)
            np.testing.assert_allclose(r["wbary"][:, ok], Wb[:, ok], atol=1e-9)
            wb = r["wbary"]
            want = np.mean((Wt - wb) ** 2) + 0.5 * np.mean(np.maximum(-Wt, 0))
            if alpha is not None:
                want += np.mean((Wt.sum(0) - alpha) ** 2)
            assert abs(r["loss"] - want) <= 1e-12 * max(1.0, want)


def fd_case(Kc, clamp, with_alpha, seed):
    rng = np.random.default_rng(seed)
    P = palette(Kc, seed)
    C = rng.uniform(0.05, 0.95, (3, 5, 6))
    Wt = rng.normal(0.2, 0.4, (Kc + 2, 5, 6))
    a = rng.uniform(0.2, 1.0, (5, 6)) if with_alpha else None
    return P, C, Wt, a


@pytest.mark.parametrize("Kc,clamp,with_alpha", [(3, False, False), (3, True, False), (3, False, True),
                                                 (5, True, True), (2, False, False)])
def test_gradients_fd64(Kc, clamp, with_alpha):
    """dL/dpalette and dL/dW_tilde against float64 central differences at
    tau = 0.05 (smooth softmax); the loss uses lam_pos = 0.5, lam_alpha = 1."""
    P, C, Wt, a = fd_case(Kc, clamp, with_alpha, 10 + Kc)
    kw = dict(alpha=a, tau=0.05, clamp=clamp, dt=f64)
    r = O.bary_loss(C, Wt, P, **kw)
    h = 1e-6
    fd = np.zeros_like(P)
    for idx in np.ndindex(*P.shape):
        Pp, Pm = P.copy(), P.copy()
        Pp[idx] += h
        Pm[idx] -= h
        fd[idx] = (O.bary_loss(C, Wt, Pp, **kw)["loss"] - O.bary_loss(C, Wt, Pm, **kw)["loss"]) / (2 * h)
    np.testing.assert_allclose(r["dP"], fd, rtol=2e-6, atol=1e-8 * np.abs(fd).max())
    rng = np.random.default_rng(1)
    for _ in range(12):
        idx = tuple(rng.integers(0, s) for s in Wt.shape)
        Wp, Wm = Wt.copy(), Wt.copy()
        Wp[idx] += h
        Wm[idx] -= h
        f = (O.bary_loss(C, Wp, P, **kw)["loss"] - O.bary_loss(C, Wm, P, **kw)["loss"]) / (2 * h)
        assert abs(r["dW"][idx] - f) <= 1e-7 * max(1.0, abs(f)), (idx, r["dW"][idx], f)


def test_fp32_against_fp64():
    rng = np.random.default_rng(3)
    P = palette(5, 9)
    C = rng.uniform(0, 1, (3, 32, 32))
    Wt = rng.normal(0.2, 0.3, (7, 32, 32))
    r32 = O.bary_loss(C, Wt, P, dt=f32)
    r64 = O.bary_loss(C, Wt, P, dt=f64)
    ok = np.all(np.isclose(reference_numpy(C, P, 1e-7), reference_numpy(C, P, 1e-9), atol=1e-9), axis=0)
    assert ok.mean() > 0.95
    np.testing.assert_allclose(r32["wbary"][:, ok], r64["wbary"][:, ok], atol=2e-5)
    np.testing.assert_allclose(r32["dW"][:, ok], r64["dW"][:, ok], atol=2e-5 * np.abs(r64["dW"]).max())
    assert abs(r32["loss"] - r64["loss"]) <= 1e-5 * r64["loss"]
