"""Float64 reference of the L1 + SSIM loss (gsplat/pytorch_ssim.py:17-66 with
zero padding, combined as (1 - lam) L1 + lam (1 - SSIM), vanilla_3dgs/gsplat/
pytorch_ssim.py:64-67) and of its gradient, with a per-element magnitude that
bounds the rounding error of any float32 evaluation of the same expression.

Per pixel q (windowed sums over the 11x11 window w, taps the float32 values
of create_window):
  m1 = w*x, m2 = w*y, e11 = w*x^2, e22 = w*y^2, e12 = w*(x y)
  s1 = e11 - m1^2, s2 = e22 - m2^2, s12 = e12 - m1 m2
  N1 = 2 m1 m2 + C1, N2 = 2 s12 + C2, D1 = m1^2 + m2^2 + C1, D2 = s1 + s2 + C2
  S = N1 N2 / (D1 D2)
  dS/dm1  = 2 m2 N2/(D1 D2) - 2 m1 N1 N2/(D1^2 D2) + 2 m1 N1 N2/(D1 D2^2) - 2 m2 N1/(D1 D2)
  dS/de11 = -N1 N2 / (D1 D2^2),  dS/de12 = 2 N1 / (D1 D2)
and dL/dx_p = (1 - lam)/n sign(x_p - y_p)
              - lam/n [w*dS/dm1 + 2 x_p (w*dS/de11) + y_p (w*dS/de12)]_p.

The magnitude ("sum of |terms|") follows each quantity through its terms:
a windowed sum of |inputs|, a difference s1 = e11 - m1^2 carries e11 + m1^2
(the cancellation of the variance), and a product or quotient of factors with
magnitudes a_i carries |t| sum_i a_i / |f_i| (numerator factors entering as
the product of the other factors times a_i, so a factor near zero is no
division by zero).  Any float32 evaluation in this operation order lies within
a small multiple of 2^-24 times that magnitude of the exact value.
"""
import math

import numpy as np
import torch
import torch.nn.functional as F


def window_taps(size=11, sigma=1.5):
    g = np.array([math.exp(-(x - size // 2) ** 2 / float(2 * sigma ** 2)) for x in range(size)], np.float32)
    return (g / g.sum()).astype(np.float32)


def _blur(a, taps):
    """Zero-padded separable 11x11 window over (C,H,W) float64 tensors."""
    c = a.shape[0]
    g = torch.tensor(taps.astype(np.float64))
    wx = g.view(1, 1, 1, -1).repeat(c, 1, 1, 1)
    wy = g.view(1, 1, -1, 1).repeat(c, 1, 1, 1)
    p = len(taps) // 2
    return F.conv2d(F.conv2d(a[None], wx, padding=(0, p), groups=c), wy, padding=(p, 0), groups=c)[0]


def l1_ssim_ref(img, tgt, lam=0.2, range_=1.0):
    """img, tgt: (C,H,W) float32 arrays.  Returns (loss, grad, grad_mag,
    loss_mag): the float64 loss, dL/dimg, its per-element magnitude bound and
    the magnitude bound of the loss itself."""
    taps = window_taps()
    x = torch.from_numpy(np.asarray(img, np.float32).astype(np.float64))
    y = torch.from_numpy(np.asarray(tgt, np.float32).astype(np.float64))
    n = x.numel()
    C1 = float(np.float32((0.01 * range_) ** 2))
    C2 = float(np.float32((0.03 * range_) ** 2))
    bl = lambda a: _blur(a, taps)
    m1, m2, e11, e22, e12 = bl(x), bl(y), bl(x * x), bl(y * y), bl(x * y)
    am1, am2, ae12 = bl(x.abs()), bl(y.abs()), bl((x * y).abs())
    s1, s2, s12 = e11 - m1 * m1, e22 - m2 * m2, e12 - m1 * m2
    as1, as2, as12 = e11 + am1 * am1, e22 + am2 * am2, ae12 + am1 * am2
    N1, N2 = 2 * m1 * m2 + C1, 2 * s12 + C2
    D1, D2 = m1 * m1 + m2 * m2 + C1, s1 + s2 + C2
    aN1, aN2 = 2 * am1 * am2 + C1, 2 * as12 + C2
    aD1, aD2 = am1 * am1 + am2 * am2 + C1, as1 + as2 + C2
    rD1, rD2 = aD1 / D1.abs(), aD2 / D2.abs()
    S = N1 * N2 / (D1 * D2)
    loss = (1 - lam) * (x - y).abs().mean() + lam * (1 - S.mean())
    mS = (aN1 * N2.abs() + N1.abs() * aN2) / (D1 * D2).abs() + S.abs() * (rD1 + rD2)
    loss_mag = (1 - lam) * ((x.abs() + y.abs()).mean()) + lam * (1 + mS.mean())

    DD = D1 * D2
    t1 = 2 * m2 * N2 / DD
    t2 = -2 * m1 * N1 * N2 / (D1 * DD)
    t3 = 2 * m1 * N1 * N2 / (D2 * DD)
    t4 = -2 * m2 * N1 / DD
    mt1 = 2 * (am2 * N2.abs() + m2.abs() * aN2) / DD.abs() + t1.abs() * (rD1 + rD2)
    num3 = am1 * N1.abs() * N2.abs() + m1.abs() * aN1 * N2.abs() + m1.abs() * N1.abs() * aN2
    mt2 = 2 * num3 / (D1 * DD).abs() + t2.abs() * (2 * rD1 + rD2)
    mt3 = 2 * num3 / (D2 * DD).abs() + t3.abs() * (rD1 + 2 * rD2)
    mt4 = 2 * (am2 * N1.abs() + m2.abs() * aN1) / DD.abs() + t4.abs() * (rD1 + rD2)
    dA, mA = t1 + t2 + t3 + t4, mt1 + mt2 + mt3 + mt4
    dB = -N1 * N2 / (D2 * DD)
    mB = (aN1 * N2.abs() + N1.abs() * aN2) / (D2 * DD).abs() + dB.abs() * (rD1 + 2 * rD2)
    dC = 2 * N1 / DD
    mC = 2 * aN1 / DD.abs() + dC.abs() * (rD1 + rD2)

    grad = (1 - lam) / n * torch.sign(x - y) - lam / n * (bl(dA) + 2 * x * bl(dB) + y * bl(dC))
    mag = (1 - lam) / n + lam / n * (bl(mA) + 2 * x.abs() * bl(mB) + y.abs() * bl(mC))
    return float(loss), grad.numpy(), mag.numpy(), float(loss_mag)
