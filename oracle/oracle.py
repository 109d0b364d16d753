"""CPU oracle of the palette-weight rasterizer: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module.  It wraps oracle/rgp_oracle.cpp (a C++ restatement of the
reference gsplatcu arithmetic, FP32 and FP64) through ctypes and composes the
reference call chains of gsplat/gsmodel.py:104-200 on numpy arrays.

How each function is pinned (tests/test_oracle.py):
  rgp_expf            against libm exp in double: max error <= 2 ulp over a
                      dense sweep of [-104, 88.7]; exact at 0 and on 2^n.
  gauss_half_form     bit for bit equal to rgp_expf(-m/2) on every float of
                      [0, 1e7]; gauss_half_fast_form equal to it on every float
                      of [0, 170] and differing just above 174 (gauss_half_check,
                      gauss_half_fast_check).
  project             closed form: an unprojection of (u, depth) returns the
                      camera-space point; z < 0.2 culls; FP32 vs FP64.
  cov3d               det(Sigma) = (s0 s1 s2)^2 for unit quaternions, Sigma
                      equals R diag(s^2) R^T built by numpy in float64, PSD.
  cov2d               equals J W Sigma W^T J^T + 0.3 I built by numpy in
                      float64 (with the +-1.3 tan_fov clamp), symmetric PSD.
  inverse_cov2d       cinv . cov = I; areas = ceil(3 sqrt(diag)); a NaN
                      determinant culls (depth -1) but det = 0 does not.
  sh2weight           SH basis: addition theorem sum_m Y_lm^2 = (2l+1)/4pi per
                      band; weights equal basis @ coefficients in float64.
  sh2weight_fast      equals sh2weight on cat(dc, reconstruct(A, B)) (the
                      reference's fused/unfused pair, gsmodel.py:303-316) and,
                      at rank 1, A (N,P,1) / B (N,1,Q) reduce to (N,P) / (N,Q).
  sh2weight_fast_bwd  adjoint identity <dw, J dtheta> = <J^T dw, dtheta> and
                      float64 central finite differences.
  sh2color_jac        float64 central differences of sh2color in pws (plus
                      the reference's dY_9/dy term, kernel.cu:1152, pinned
                      exactly) and in the coefficients (linear).
  get_rects/bin_sort  brute force on tiny inputs: per-Gaussian tile sets from
                      an independent enumeration, a Python stable sort of
                      (tile, uint32(depth*1000)) keys, ranges by scanning,
                      including M = 0 and M = 1 (the reference leaves (0,0)).
  draw (forward)      pure-Python per-pixel recurrence on tiny scenes (bit
                      exact in float32); with all weights 1 every channel
                      equals 1 - final_tau up to rounding; FP32 vs FP64.
  draw_backward       float64 central finite differences of sum(dL * layer)
                      w.r.t. the per-Gaussian weights (exactly linear).
  draw_color_backward float64 central finite differences of
                      sum(dL/dimage * image) + sum(dL/dtau * final_tau)
                      w.r.t. every Gaussian's u, cinv, alpha and colour, the
                      tile lists held fixed, no alpha' near the 0.002 cut.
  mix                 linear identity mix(render(w)) = render(bake(w, P))
                      (gsmodel.py:488-509) and a numpy einsum.
  log1pf              against libm log1p in double: max error <= 4 ulp over
                      a dense sweep of [0, e^20]; exact at 0; log1p(y) = y for
                      y below 2^-24.
  bary_loss           (oracle/rgp_oracle_bary.cpp) barycentric coordinates
                      against numpy.linalg.solve of each facet system and the
                      identities alpha+beta+gamma+delta = 1 and
                      beta*white + gamma*P_i + delta*P_j = C; softplus against
                      numpy.logaddexp; the softmax and the weighted sum against
                      scipy.special.softmax in float64; a pixel equal to a
                      palette colour; the loss against numpy means; the
                      gradients (W_tilde and palette, plain, clamped and with
                      alpha) against float64 central finite differences at
                      tau = 0.05, where the softmax path is smooth.
No function of this module is left unpinned.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "librgp_oracle.so")
_lib = None

f32, f64, i32, u32, u64, i64 = np.float32, np.float64, np.int32, np.uint32, np.uint64, np.int64
BLOCK = 16


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        _lib = ctypes.CDLL(_SO)
        _lib.orc_expf.restype = ctypes.c_float
        _lib.orc_expf.argtypes = [ctypes.c_float]
        for s in ("f", "d"):
            getattr(_lib, "orc_bin_sort_" + s).restype = ctypes.c_int64
        _lib.orc_l1_ssim_f.restype = ctypes.c_double
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _sfx(dt):
    return "f" if dt == f32 else "d"


def _scal(dt, v):
    return ctypes.c_float(v) if dt == f32 else ctypes.c_double(v)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def set_threads(t):
    lib().orc_set_num_threads(int(t))


def num_threads():
    return lib().orc_num_threads()


def log1pf(y):
    y = _c(y, f32)
    out = np.empty_like(y)
    lib().orc_log1pf_array(ctypes.c_int64(y.size), _p(y), _p(out))
    return out


def bary_loss(C, W_tilde, palette, alpha=None, tau=1e-9, clamp=False, lam_pos=0.5, lam_alpha=1.0, dt=f32):
    """Barycentric palette loss of the finetune step (gsmodel.py:219-285,
    finetune.py:353-392) on one view.  C (3,H,W), W_tilde (Kc+2,H,W),
    palette (Kc,3) chromatic colours, alpha (H,W) or None.  Returns a dict:
    wbary (Kc+2,H,W), dW (Kc+2,H,W) = dL/dW_tilde, dP (Kc,3) = dL/dpalette
    and dP_abs (bound on the sum of |terms|), both summed in float64, sums
    (L_geo, L_pos, L_alpha numerators) and loss = L_geo + lam_pos L_pos +
    lam_alpha L_alpha."""
    C, Wt, P = _c(C, dt), _c(W_tilde, dt), _c(palette, dt)
    Kc = P.shape[0]
    Kt, H, W = Wt.shape
    assert Kt == Kc + 2 and C.shape == (3, H, W) and 1 <= Kc <= 6
    A = None if alpha is None else _c(alpha, dt).reshape(H, W)
    wb, dW = np.zeros_like(Wt), np.zeros_like(Wt)
    parts, dP, dPa = np.zeros(3, f64), np.zeros((Kc, 3), f64), np.zeros((Kc, 3), f64)
    acc = np.zeros((Kc, 16), f64)
    getattr(lib(), "orc_bary_loss_" + _sfx(dt))(H, W, Kc, _p(C), _p(Wt), _p(A), _p(P), _scal(dt, tau), int(clamp),
                                                _scal(dt, lam_pos), _scal(dt, lam_alpha), _p(wb), _p(dW),
                                                _p(parts), _p(dP), _p(dPa), _p(acc))
    nk, hw = Kt * H * W, H * W
    loss = parts[0] / nk + lam_pos * parts[1] / nk + (lam_alpha * parts[2] / hw if alpha is not None else 0.0)
    return dict(wbary=wb, dW=dW, dP=dP, dP_abs=dPa, sums=parts, loss=loss, facet_acc=acc)


def bary_coords_backward(C, palette, dW_bary, tau=1e-9, dt=f32):
    """dL/dpalette of barycentric_coords_3d for a given dL/dW_bary (Kc+2,H,W):
    (dP, dP_abs), summed in float64."""
    C, P, dWb = _c(C, dt), _c(palette, dt), _c(dW_bary, dt)
    Kc = P.shape[0]
    H, W = C.shape[1:]
    assert dWb.shape == (Kc + 2, H, W)
    dP, dPa = np.zeros((Kc, 3), f64), np.zeros((Kc, 3), f64)
    getattr(lib(), "orc_bary_coords_backward_" + _sfx(dt))(H, W, Kc, _p(C), _p(P), _scal(dt, tau), _p(dWb), _p(dP),
                                                           _p(dPa))
    return dP, dPa


def expf(x):
    x = _c(x, f32)
    y = np.empty_like(x)
    lib().orc_expf_array(ctypes.c_int64(x.size), _p(x), _p(y))
    return y


def gauss_half_check(lo, hi):
    """(checked, differing, first differing m) over every float m in [lo, hi]:
    exp(-m/2) in the compositing kernels' range-test-free form against
    expf(-0.5 m)."""
    n, d, f = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_float()
    lib().orc_gauss_half_check(ctypes.c_float(lo), ctypes.c_float(hi), ctypes.byref(n), ctypes.byref(d),
                               ctypes.byref(f))
    return n.value, d.value, f.value


def gauss_half_fast_check(lo, hi):
    """(checked, differing, first differing m) over every float m in [lo, hi]:
    the kernels' exponent-add form of exp(-m/2) (entries whose cutoff is at
    most 170) against the range-test-free form."""
    n, d, f = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_float()
    lib().orc_gauss_half_fast_check(ctypes.c_float(lo), ctypes.c_float(hi), ctypes.byref(n), ctypes.byref(d),
                                    ctypes.byref(f))
    return n.value, d.value, f.value


# --------------------------------------------------------------------------
# per-stage wrappers (shapes follow gsplatcu/ext.cpp:10-124)
# --------------------------------------------------------------------------
def project(pws, Rcw, tcw, fx, fy, cx, cy, dt=f32):
    pws, R, t = _c(pws, dt), _c(Rcw, dt), _c(tcw, dt)
    n = pws.shape[0]
    us, pcs, depths = np.zeros((n, 2), dt), np.zeros((n, 3), dt), np.zeros(n, dt)
    getattr(lib(), "orc_project_" + _sfx(dt))(n, _p(pws), _p(R), _p(t), _scal(dt, fx), _scal(dt, fy),
                                              _scal(dt, cx), _scal(dt, cy), _p(us), _p(pcs), _p(depths))
    return us, pcs, depths


def cov3d(rots, scales, depths, dt=f32):
    rots, scales, depths = _c(rots, dt), _c(scales, dt), _c(depths, dt)
    n = rots.shape[0]
    out = np.zeros((n, 6), dt)
    getattr(lib(), "orc_cov3d_" + _sfx(dt))(n, _p(rots), _p(scales), _p(depths), _p(out))
    return out


def tan_fov(width, height, fx, fy, dt=f32):
    """gausplat.cu:357-358: width / (2 * focal_x) in the arithmetic type."""
    return dt(dt(width) / (dt(2) * dt(fx))), dt(dt(height) / (dt(2) * dt(fy)))


def cov2d(cov3ds, pcs, Rcw, depths, fx, fy, width, height, dt=f32):
    cov3ds, pcs, R, depths = _c(cov3ds, dt), _c(pcs, dt), _c(Rcw, dt), _c(depths, dt)
    n = pcs.shape[0]
    out = np.zeros((n, 3), dt)
    tfx, tfy = tan_fov(width, height, fx, fy, dt)
    getattr(lib(), "orc_cov2d_" + _sfx(dt))(n, _p(cov3ds), _p(pcs), _p(R), _p(depths), _scal(dt, fx),
                                            _scal(dt, fy), _scal(dt, float(tfx)), _scal(dt, float(tfy)), _p(out))
    return out


def inverse_cov2d(cov2ds, depths, dt=f32):
    """Returns cinv2ds, areas; depths is updated in place (kernel.cu:397)."""
    cov2ds = _c(cov2ds, dt)
    assert depths.dtype == dt and depths.flags.c_contiguous
    n = cov2ds.shape[0]
    cinv, areas = np.zeros((n, 3), dt), np.zeros((n, 2), i32)
    getattr(lib(), "orc_inverse_cov2d_" + _sfx(dt))(n, _p(cov2ds), _p(depths), _p(cinv), _p(areas))
    return cinv, areas


def sh2weight(shs, pws, twc, K, dt=f32):
    shs, pws, twc = _c(shs, dt), _c(pws, dt), _c(twc, dt)
    n = pws.shape[0]
    sh_dim = shs.shape[1] // K
    w, basis = np.zeros((n, K), dt), np.zeros((n, sh_dim), dt)
    getattr(lib(), "orc_sh2weight_" + _sfx(dt))(n, K, sh_dim, _p(shs), _p(pws), _p(twc), _p(w), _p(basis))
    return w, basis


def _factor_shapes(A, B):
    if A.ndim == 2:
        A = A[:, :, None]
    if B.ndim == 2:
        B = B[:, None, :]
    return A, B


def sh2weight_fast(dc, A, B, pws, twc, K, Q, dt=f32):
    A, B = _factor_shapes(A, B)
    dc, A, B, pws, twc = _c(dc, dt), _c(A, dt), _c(B, dt), _c(pws, dt), _c(twc, dt)
    n, P, r = A.shape
    L = P * Q // K
    w, ylm = np.zeros((n, K), dt), np.zeros((n, L), dt)
    getattr(lib(), "orc_sh2weight_fast_" + _sfx(dt))(n, K, P, Q, r, _p(dc), _p(A), _p(B), _p(pws), _p(twc),
                                                     _p(w), _p(ylm))
    return w, ylm


def sh2weight_fast_backward(A, B, ylm, dw, K, Q, dt=f32):
    A2, B2 = _factor_shapes(A, B)
    A2, B2, ylm, dw = _c(A2, dt), _c(B2, dt), _c(ylm, dt), _c(dw, dt)
    n, P, r = A2.shape
    dA, dB = np.zeros_like(A2), np.zeros_like(B2)
    getattr(lib(), "orc_sh2weight_fast_backward_" + _sfx(dt))(n, K, P, Q, r, _p(A2), _p(B2), _p(ylm), _p(dw),
                                                              _p(dA), _p(dB))
    return dA.reshape(A.shape), dB.reshape(B.shape)


def sh2color(shs, pws, twc, dt=f32):
    shs, pws, twc = _c(shs, dt), _c(pws, dt), _c(twc, dt)
    n = pws.shape[0]
    sh_dim = shs.shape[1] // 3
    out = np.zeros((n, 3), dt)
    getattr(lib(), "orc_sh2color_" + _sfx(dt))(n, sh_dim, _p(shs), _p(pws), _p(twc), _p(out))
    return out


# --------------------------------------------------------------------------
# calc_J outputs (gausplat.cu:298-462): zero-filled like the reference's
# torch::full(..., 0.0) allocations, then the written entries
# --------------------------------------------------------------------------
def project_jac(pws, Rcw, tcw, fx, fy, dt=f32):
    """du/dpc (N,2,3)."""
    pws, R, t = _c(pws, dt), _c(Rcw, dt), _c(tcw, dt)
    n = pws.shape[0]
    du = np.zeros((n, 2, 3), dt)
    getattr(lib(), "orc_project_jac_" + _sfx(dt))(n, _p(pws), _p(R), _p(t), _scal(dt, fx), _scal(dt, fy), _p(du))
    return du


def cov3d_jac(rots, scales, depths, dt=f32):
    """dcov3d/drot (N,6,4), dcov3d/dscale (N,6,3)."""
    rots, scales, depths = _c(rots, dt), _c(scales, dt), _c(depths, dt)
    n = rots.shape[0]
    dr, ds = np.zeros((n, 6, 4), dt), np.zeros((n, 6, 3), dt)
    getattr(lib(), "orc_cov3d_jac_" + _sfx(dt))(n, _p(rots), _p(scales), _p(depths), _p(dr), _p(ds))
    return dr, ds


def cov2d_jac(cov3ds, pcs, Rcw, depths, fx, fy, width, height, dt=f32):
    """dcov2d/dcov3d (N,3,6), dcov2d/dpc (N,3,3)."""
    cov3ds, pcs, R, depths = _c(cov3ds, dt), _c(pcs, dt), _c(Rcw, dt), _c(depths, dt)
    n = pcs.shape[0]
    d3, dp = np.zeros((n, 3, 6), dt), np.zeros((n, 3, 3), dt)
    tfx, tfy = tan_fov(width, height, fx, fy, dt)
    getattr(lib(), "orc_cov2d_jac_" + _sfx(dt))(n, _p(cov3ds), _p(pcs), _p(R), _p(depths), _scal(dt, fx),
                                                _scal(dt, fy), _scal(dt, float(tfx)), _scal(dt, float(tfy)), _p(d3),
                                                _p(dp))
    return d3, dp


def inverse_jac(cov2ds, depths, dt=f32):
    """dcinv/dcov2d (N,3,3); depths as returned by inverse_cov2d."""
    cov2ds, depths = _c(cov2ds, dt), _c(depths, dt)
    n = cov2ds.shape[0]
    out = np.zeros((n, 3, 3), dt)
    getattr(lib(), "orc_inverse_jac_" + _sfx(dt))(n, _p(cov2ds), _p(depths), _p(out))
    return out


def sh2weight_jac(shs, pws, twc, K, dt=f32):
    """dW/dpw (N,K,3)."""
    shs, pws, twc = _c(shs, dt), _c(pws, dt), _c(twc, dt)
    n = pws.shape[0]
    sh_dim = shs.shape[1] // K
    out = np.zeros((n, K, 3), dt)
    getattr(lib(), "orc_sh2weight_jac_" + _sfx(dt))(n, K, sh_dim, _p(shs), _p(pws), _p(twc), _p(out))
    return out


def sh2color_jac(shs, pws, twc, dt=f32):
    """sh2Color's calc_J outputs: dcolor_dshs (N,1,sh_dim), dcolor_dpws (N,3,3)."""
    shs, pws, twc = _c(shs, dt), _c(pws, dt), _c(twc, dt)
    n = pws.shape[0]
    sh_dim = shs.shape[1] // 3
    dshs, dpws = np.zeros((n, 1, sh_dim), dt), np.zeros((n, 3, 3), dt)
    getattr(lib(), "orc_sh2color_jac_" + _sfx(dt))(n, sh_dim, _p(shs), _p(pws), _p(twc), _p(dshs), _p(dpws))
    return dshs, dpws


def grid_dims(width, height):
    return (width + BLOCK - 1) // BLOCK, (height + BLOCK - 1) // BLOCK


def get_rects(us, areas, depths, width, height, dt=f32):
    """rects (N,4) u32 and counts (N,) u32; areas/depths updated in place."""
    us = _c(us, dt)
    assert areas.dtype == i32 and depths.dtype == dt
    n = us.shape[0]
    gx, gy = grid_dims(width, height)
    rects, counts = np.zeros((n, 4), u32), np.zeros(n, u32)
    getattr(lib(), "orc_get_rects_" + _sfx(dt))(n, _p(us), _p(areas), _p(depths), gx, gy, _p(rects), _p(counts))
    return rects, counts


def bin_sort(depths, rects, counts, width, height, dt=f32, secs=None):
    """Returns (keys (M,) u64, ids (M,) i32, ranges (T,2) i32); secs (a
    float64 array of 4, optional) receives the seconds of the scan, the key
    duplication, the sort and the ranges."""
    depths = _c(depths, dt)
    n = depths.shape[0]
    gx, gy = grid_dims(width, height)
    M = int(counts.astype(np.uint64).sum())
    keys, ids = np.zeros(max(M, 1), u64), np.zeros(max(M, 1), i32)
    ranges = np.zeros((gx * gy, 2), i32)
    got = getattr(lib(), "orc_bin_sort_" + _sfx(dt))(n, _p(depths), _p(_c(rects, u32)), _p(_c(counts, u32)), gx, gy,
                                                     _p(keys), _p(ids), ctypes.c_int64(M), _p(ranges), _p(secs))
    assert got == M
    return keys[:M], ids[:M], ranges


def draw(width, height, ranges, ids, us, cinv, alphas, feat, final_tau_init=0.0, dt=f32, strips=False):
    """draw_weight (feat (N,K)) / draw_color (feat (N,3)).  Returns
    (layer (dim,H,W), contrib (H,W), final_tau (H,W), n_eval, n_contrib), and
    with strips=True also the per-list-position masks of row strips (uint8,
    bit (py mod 16) / 4) in which the entry contributed to some pixel."""
    feat = _c(feat, dt)
    dim = feat.shape[1]
    layer = np.zeros((dim, height, width), dt)
    contrib = np.zeros((height, width), i32)
    final_tau = np.full((height, width), final_tau_init, dt)
    ne, nc = ctypes.c_int64(0), ctypes.c_int64(0)
    smask = np.zeros(max(len(ids), 1), np.uint8) if strips else None
    getattr(lib(), "orc_draw_" + _sfx(dt))(width, height, dim, _p(_c(ranges, i32)), _p(_c(ids, i32)), _p(_c(us, dt)),
                                           _p(_c(cinv, dt)), _p(_c(alphas, dt).reshape(-1)), _p(feat), _p(layer),
                                           _p(contrib), _p(final_tau), ctypes.byref(ne), ctypes.byref(nc),
                                           _p(smask) if strips else None)
    if strips:
        return layer, contrib, final_tau, ne.value, nc.value, smask[:len(ids)]
    return layer, contrib, final_tau, ne.value, nc.value


def draw_backward(width, height, K, ranges, ids, us, cinv, alphas, contrib, final_tau, dlayer, dt=f32, mode=0):
    """Returns (grad (N,K) float64, gabs (N,K) float64) for mode 0 (the
    reference's division-recovered tau) and mode 2 (tau replayed front to
    back); for mode 1 the mode-0 gradient accumulated with atomic adds in the
    arithmetic type."""
    n = us.shape[0]
    grad, gabs = np.zeros((n, K), f64), np.zeros((n, K), f64)
    grad_t = np.zeros((n, K), dt)
    getattr(lib(), "orc_draw_backward_" + _sfx(dt))(
        width, height, K, _p(_c(ranges, i32)), _p(_c(ids, i32)), _p(_c(us, dt)), _p(_c(cinv, dt)),
        _p(_c(alphas, dt).reshape(-1)), _p(_c(contrib, i32)), _p(_c(final_tau, dt)), _p(_c(dlayer, dt)), mode,
        _p(grad), _p(gabs), _p(grad_t))
    return grad_t if mode == 1 else (grad, gabs)


def draw_color_backward(width, height, ranges, ids, us, cinv, alphas, colors, contrib, final_tau, dimg, dtau=None,
                        dt=f32):
    """vanilla drawB (vanilla_3dgs/gsplatcu/kernel.cu:811-971): per-Gaussian
    (N, 9) gradient sums [du (2), dcinv (3), dalpha, dcolor (3)] and the sums of
    |terms|, in float64."""
    us, cinv, alphas, colors = _c(us, dt), _c(cinv, dt), _c(alphas, dt), _c(colors, dt)
    ft, dimg = _c(final_tau, dt), _c(dimg, dt)
    dtau = None if dtau is None else _c(dtau, dt)
    n = us.shape[0]
    grad, gabs = np.zeros((n, 9), f64), np.zeros((n, 9), f64)
    getattr(lib(), "orc_draw_color_backward_" + _sfx(dt))(width, height, _p(_c(ranges, i32)), _p(_c(ids, i32)), _p(us),
                                                          _p(cinv), _p(alphas), _p(colors), _p(_c(contrib, i32)),
                                                          _p(ft), _p(dimg), _p(dtau), _p(grad), _p(gabs))
    return grad, gabs


def mix(palette, W, dt=f32):
    W = _c(W, dt)
    K = W.shape[0]
    HW = W.shape[1] * W.shape[2]
    rgb = np.zeros((3,) + W.shape[1:], dt)
    getattr(lib(), "orc_mix_" + _sfx(dt))(K, ctypes.c_int64(HW), _p(_c(palette, dt)), _p(W), _p(rgb))
    return rgb


def ssim_taps():
    """The normalised float32 taps of create_window (pytorch_ssim.py:17-29)."""
    import math
    g = np.array([math.exp(-(x - 5) ** 2 / float(2 * 1.5 ** 2)) for x in range(11)], np.float32)
    return (g / g.sum()).astype(np.float32)


def l1_ssim(img, tgt, lam=0.2, grad=True):
    """FP32 (1 - lam) L1 + lam (1 - SSIM) of (C,H,W) images and, with grad,
    dL/dimg (OpenMP over rows; the CPU timing sample's loss stage)."""
    img, tgt = _c(img, f32), _c(tgt, f32)
    C, H, W = img.shape
    d = np.zeros_like(img) if grad else None
    loss = lib().orc_l1_ssim_f(C, H, W, _p(img), _p(tgt), _p(ssim_taps()), ctypes.c_float(0.01 ** 2),
                               ctypes.c_float(0.03 ** 2), ctypes.c_float(lam), _p(d))
    return loss, d


def sh_backward_full(dw, basis):
    """dL/dshs (N, L*K) = dL/dw (N,K) x Y (N,L) per element, FP32."""
    dw, basis = _c(dw, f32), _c(basis, f32)
    n, K = dw.shape
    L = basis.shape[1]
    out = np.empty((n, L * K), f32)
    lib().orc_sh_backward_full_f(ctypes.c_int64(n), K, L, _p(dw), _p(basis), _p(out))
    return out


def adam(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """torch.optim.Adam's update of p, m, v in place (FP32, OpenMP)."""
    lib().orc_adam_f(ctypes.c_int64(p.size), _p(p), _p(_c(g, f32)), _p(m), _p(v), ctypes.c_int64(step),
                     ctypes.c_double(lr), ctypes.c_double(b1), ctypes.c_double(b2), ctypes.c_double(eps))


def timed_step(scene, cam, target=None, tiles=None, lr=1e-3):
    """One finetune step of the palette-weight rasterizer on the CPU with the
    stages of the GPU step, each timed in seconds: preprocess (project,
    covariances, SH weights, inverse, rects), scan, key duplication, sort,
    ranges, forward compositing, palette mix, loss (L1 + SSIM against target,
    forward and backward, when target is given), mix backward, backward
    compositing, SH / low-rank backward, Adam.  tiles: composite only these
    tiles (a seeded subset; the compositing stages then cover that share of
    the list entries, reported as 'entry_share')."""
    import time
    t = {}
    clock = time.perf_counter
    c0 = clock()
    K = scene.K
    us, pcs, depths = project(scene.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy)
    c3 = cov3d(scene.rots, scene.scales, depths)
    c2 = cov2d(c3, pcs, cam.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height)
    if scene.lowrank:
        w, ylm = sh2weight_fast(scene.dc, scene.A, scene.B, scene.pws, cam.twc, K, scene.Q)
    else:
        w, basis = sh2weight(scene.shs, scene.pws, cam.twc, K)
    cinv, areas = inverse_cov2d(c2, depths)
    rects, counts = get_rects(us, areas, depths, cam.width, cam.height)
    t["preprocess"] = clock() - c0
    secs = np.zeros(4, f64)
    _, ids, ranges = bin_sort(depths, rects, counts, cam.width, cam.height, secs=secs)
    t["scan"], t["key_duplication"], t["sort"], t["ranges"] = (float(x) for x in secs)
    share = 1.0
    if tiles is not None:
        sub = np.zeros_like(ranges)
        sub[tiles] = ranges[tiles]
        share = float((sub[:, 1] - sub[:, 0]).sum()) / max(1, len(ids))
        ranges = sub
    c0 = clock()
    layer, contrib, ftau, _, _ = draw(cam.width, cam.height, ranges, ids, us, cinv, scene.alphas, w, 0.0)
    t["forward_compositing"] = clock() - c0
    c0 = clock()
    rgb = mix(scene.palette, layer)
    t["mix"] = clock() - c0
    if target is not None:
        c0 = clock()
        _, drgb = l1_ssim(rgb, target)
        t["loss"] = clock() - c0
    else:
        drgb = np.random.default_rng(0).standard_normal(rgb.shape).astype(f32)
    c0 = clock()
    P = scene.palette.astype(f32)
    dl = ((P[:, 0, None, None] * drgb[0] + P[:, 1, None, None] * drgb[1]) + P[:, 2, None, None] * drgb[2])
    dl = dl.astype(f32)
    t["mix_backward"] = clock() - c0
    c0 = clock()
    g = draw_backward(cam.width, cam.height, K, ranges, ids, us, cinv, scene.alphas, contrib, ftau, dl, f32, mode=1)
    t["backward_compositing"] = clock() - c0
    c0 = clock()
    if scene.lowrank:
        dA, dB = sh2weight_fast_backward(scene.A, scene.B, ylm, g, K, scene.Q)
        grads = [(g * f32(0.28209479177387814)).astype(f32), dA, dB]
    else:
        grads = [sh_backward_full(g, basis)]
    t["sh_backward"] = clock() - c0
    params = [scene.dc, scene.A, scene.B] if scene.lowrank else [scene.shs]
    state = [(pa.astype(f32).copy(), np.zeros(pa.shape, f32), np.zeros(pa.shape, f32)) for pa in params]
    c0 = clock()
    for (p_, m_, v_), gr in zip(state, grads):
        adam(p_, gr, m_, v_, 1, lr)
    t["adam"] = clock() - c0
    t["entry_share"] = share
    return t


# --------------------------------------------------------------------------
# reference call chains (gsplat/gsmodel.py:104-200)
# --------------------------------------------------------------------------
def forward_weight(scene, cam, dt=f32, fast=None):
    """The forward of WeightGSFunctionTest (full rank) or WeightGSFunctionFast
    (low rank): project -> cov3d -> cov2d -> sh2Weight(_fast) -> inverseCov2D
    -> splat_weight.  Returns a dict with every intermediate."""
    K = scene.K
    if fast is None:
        fast = scene.lowrank
    us, pcs, depths = project(scene.pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, dt)
    c3 = cov3d(scene.rots, scene.scales, depths, dt)
    c2 = cov2d(c3, pcs, cam.Rcw, depths, cam.fx, cam.fy, cam.width, cam.height, dt)
    out = dict(K=K, width=cam.width, height=cam.height, dt=dt, fast=fast)
    if fast:
        w, ylm = sh2weight_fast(scene.dc, scene.A, scene.B, scene.pws, cam.twc, K, scene.Q, dt)
        out["ylm"] = ylm
    else:
        shs = scene.shs if not scene.lowrank else np.concatenate(
            [scene.dc, _recon(scene)], axis=1)
        w, basis = sh2weight(shs, scene.pws, cam.twc, K, dt)
        out["basis"] = basis
    cinv, areas = inverse_cov2d(c2, depths, dt)
    rects, counts = get_rects(us, areas, depths, cam.width, cam.height, dt)
    keys, ids, ranges = bin_sort(depths, rects, counts, cam.width, cam.height, dt)
    layer, contrib, final_tau, ne, nc = draw(cam.width, cam.height, ranges, ids, us, cinv, scene.alphas, w, 0.0, dt)
    out.update(us=us, pcs=pcs, depths=depths, cov3ds=c3, cov2ds=c2, weights=w, cinv=cinv, areas=areas,
               rects=rects, counts=counts, keys=keys, ids=ids, ranges=ranges, layer=layer, contrib=contrib,
               final_tau=final_tau, mask=depths > f32(0.2), n_eval=ne, n_contrib=nc)
    return out


def _recon(scene):
    A, B = _factor_shapes(scene.A, scene.B)
    n, P, r = A.shape
    C = np.zeros((n, P, B.shape[2]), f32)
    for j in range(r):
        C = C + A[:, :, j:j + 1] * B[:, j:j + 1, :]
    return C.reshape(n, -1)


def backward_weight(scene, fwd, dlayer, parallel=False, replay=False):
    """The backward of WeightGSFunctionTest / Fast: splat_weight_B, then the
    SH chain.  Returns a dict: dweights (N,K) float64 plus its abs-sum bound,
    and dshs (full rank) or ddc, dA, dB (low rank), in float64.
    parallel=True: the timing form (OpenMP over tiles, atomic adds in the
    arithmetic type, reference order of the SH chain in that type)."""
    dt = fwd["dt"]
    K = fwd["K"]
    if parallel:
        g = draw_backward(fwd["width"], fwd["height"], K, fwd["ranges"], fwd["ids"], fwd["us"], fwd["cinv"],
                          scene.alphas, fwd["contrib"], fwd["final_tau"], dlayer, dt, mode=1)
        if fwd["fast"]:
            dA, dB = sh2weight_fast_backward(scene.A, scene.B, fwd["ylm"], g, K, scene.Q, dt)
            return dict(dweights=g, ddc=g * dt(0.28209479177387814), dA=dA, dB=dB)
        basis = fwd["basis"]
        n, L = basis.shape
        return dict(dweights=g, dshs=(g[:, None, :] * basis[:, :, None]).reshape(n, L * K))
    g, gabs = draw_backward(fwd["width"], fwd["height"], K, fwd["ranges"], fwd["ids"], fwd["us"], fwd["cinv"],
                            scene.alphas, fwd["contrib"], fwd["final_tau"], dlayer, dt, mode=2 if replay else 0)
    out = dict(dweights=g, dweights_abs=gabs)
    if fwd["fast"]:
        C0 = f64(0.28209479177387814) if dt == f64 else f64(f32(0.28209479177387814))
        out["ddc"] = g * C0
        out["ddc_abs"] = gabs * C0
        dA, dB = sh2weight_fast_backward(scene.A, scene.B, fwd["ylm"], g, K, scene.Q, f64)
        aA, aB = sh2weight_fast_backward(np.abs(scene.A), np.abs(scene.B), np.abs(fwd["ylm"]), gabs, K, scene.Q, f64)
        out.update(dA=dA, dB=dB, dA_abs=aA, dB_abs=aB)
    else:
        basis = fwd["basis"].astype(f64)
        n, L = basis.shape
        out["dshs"] = (g[:, None, :] * basis[:, :, None]).reshape(n, L * K)
        out["dshs_abs"] = (gabs[:, None, :] * np.abs(basis)[:, :, None]).reshape(n, L * K)
    return out
