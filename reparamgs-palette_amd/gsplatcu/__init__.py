"""gsplatcu-compatible module on the MI355X kernels of librgp.

Drop-in for the `gsplatcu` extension of the reference (gsplatcu/ext.cpp:126-139):
the same function names, argument order, return lists, dtypes and shapes,
including the in-place updates of `depths` and `areas` (kernel.cu:156-157,
397 through gausplat.cu:191-193).  Put `reparamgs-palette_amd/` on sys.path and
`import gsplatcu as gsc` resolves here.

Differences from the reference, all deliberate:
* errors raise (the reference prints CUDA errors and carries on,
  common.cuh:17-25); K is checked against MAX_SPLAT_DIM = 8;
* kernels run on torch's current stream instead of the legacy default stream,
  with no device synchronisation after each launch;
* sh2Weight_fast accepts rank-r factors A (N,P,r), B (N,r,Q) besides the
  reference's rank-1 (N,P), (N,Q), and calc_J=False returns the weights
  instead of dereferencing null (kernel.cu:911, 961);
* the calc_J Jacobians of project / computeCov3D / computeCov2D /
  inverseCov2D / sh2Weight / sh2Color are computed by kernels that follow
  the reference's formulas (kernel.cu:402-414, 463-500, 557-607, 656-660,
  769-869, 1112-1167), including dY_9/dy = C30 (3x^2 - 2y^2) in sh2Weight
  and sh2Color (kernel.cu:840, 1152);
* splat_color_B and the vanilla module's splat / splatB (with the gradient
  of a loss on final_tau) are provided.
"""
import torch

from rgp import ops
from rgp._lib import I64, RgpError, call, f32c, i32c, ptr, require_cuda, stream

__all__ = ["project", "computeCov3D", "computeCov2D", "inverseCov2D", "sh2Weight", "sh2Weight_fast",
           "sh2Weight_fast_B", "sh2Color", "splat_weight", "splat_weight_B", "splat_color",
           "splat_color_B", "splat", "splatB", "lab2rgb"]


def _inplace_f32(t):
    """The tensor the reference writes through `.contiguous().data_ptr()`:
    the caller's own storage when it is contiguous float32, else a copy."""
    if t.dtype == torch.float32 and t.is_contiguous():
        return t
    return t.float().contiguous()


def _inplace_i32(t):
    if t.dtype == torch.int32 and t.is_contiguous():
        return t
    return t.int().contiguous()


def project(pws, Rcw, tcw, focal_x, focal_y, center_x, center_y, calc_J):
    """gausplat.cu:381-419; calc_J adds du/dpc (N,2,3)."""
    require_cuda(pws, Rcw, tcw)
    n = pws.shape[0]
    dev = pws.device
    us = torch.zeros((n, 2), dtype=torch.float32, device=dev)
    pcs = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    depths = torch.zeros((n,), dtype=torch.float32, device=dev)
    du = torch.zeros((n, 2, 3), dtype=torch.float32, device=dev) if calc_J else None
    call("rgp_project", n, ptr(f32c(pws)), ptr(f32c(Rcw)), ptr(f32c(tcw)), float(focal_x), float(focal_y),
         float(center_x), float(center_y), ptr(us), ptr(pcs), ptr(depths), ptr(du), stream(dev))
    return [us, pcs, depths, du] if calc_J else [us, pcs, depths]


def computeCov3D(rots, scales, depths, calc_J):
    """gausplat.cu:298-331; calc_J adds dcov3d/drot (N,6,4) and
    dcov3d/dscale (N,6,3)."""
    require_cuda(rots, scales, depths)
    n = rots.shape[0]
    dev = rots.device
    cov3ds = torch.zeros((n, 6), dtype=torch.float32, device=dev)
    dr = torch.zeros((n, 6, 4), dtype=torch.float32, device=dev) if calc_J else None
    ds = torch.zeros((n, 6, 3), dtype=torch.float32, device=dev) if calc_J else None
    call("rgp_compute_cov3d", n, ptr(f32c(rots)), ptr(f32c(scales)), ptr(f32c(depths)), ptr(cov3ds), ptr(dr),
         ptr(ds), stream(dev))
    return [cov3ds, dr, ds] if calc_J else [cov3ds]


def computeCov2D(cov3ds, pcs, Rcw, depths, focal_x, focal_y, width, height, calc_J):
    """gausplat.cu:333-379 (width/height arrive as floats); calc_J adds
    dcov2d/dcov3d (N,3,6) and dcov2d/dpc (N,3,3)."""
    require_cuda(cov3ds, pcs, Rcw, depths)
    n = pcs.shape[0]
    dev = pcs.device
    cov2ds = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    d3 = torch.zeros((n, 3, 6), dtype=torch.float32, device=dev) if calc_J else None
    dp = torch.zeros((n, 3, 3), dtype=torch.float32, device=dev) if calc_J else None
    call("rgp_compute_cov2d", n, ptr(f32c(cov3ds)), ptr(f32c(pcs)), ptr(f32c(Rcw)), ptr(f32c(depths)),
         float(focal_x), float(focal_y), float(width), float(height), ptr(cov2ds), ptr(d3), ptr(dp), stream(dev))
    return [cov2ds, d3, dp] if calc_J else [cov2ds]


def inverseCov2D(cov2ds, depths, calc_J):
    """gausplat.cu:590-617; depths updated in place; calc_J adds
    dcinv/dcov2d (N,3,3)."""
    require_cuda(cov2ds, depths)
    n = cov2ds.shape[0]
    dev = cov2ds.device
    cinv2ds = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    areas = torch.zeros((n, 2), dtype=torch.int32, device=dev)
    dci = torch.zeros((n, 3, 3), dtype=torch.float32, device=dev) if calc_J else None
    d = _inplace_f32(depths)
    call("rgp_inverse_cov2d", n, ptr(f32c(cov2ds)), ptr(d), ptr(cinv2ds), ptr(areas), ptr(dci), stream(dev))
    return [cinv2ds, areas, dci] if calc_J else [cinv2ds, areas]


def sh2Weight(shs, pws, twc, calc_J, num_palette):
    """gausplat.cu:425-462.  With calc_J the outputs are [weights,
    dweight_dshs (N,1,sh_dim), dweight_dpws (N,K,3)]."""
    require_cuda(shs, pws, twc)
    K = int(num_palette)
    n = pws.shape[0]
    sh_dim = shs.shape[1] // K
    dev = pws.device
    weights = torch.zeros((n, K), dtype=torch.float32, device=dev)
    basis = torch.zeros((n, 1, sh_dim), dtype=torch.float32, device=dev) if calc_J else None
    dwdp = torch.zeros((n, K, 3), dtype=torch.float32, device=dev) if calc_J else None
    call("rgp_sh2weight", n, K, sh_dim, ptr(f32c(shs)), ptr(f32c(pws)), ptr(f32c(twc)), ptr(weights),
         ptr(basis), ptr(dwdp), stream(dev))
    return [weights, basis, dwdp] if calc_J else [weights]


def _factors(shs_A, shs_B, K, Q):
    if shs_A.dim() == 3:
        n, P, r = shs_A.shape
        if shs_B.dim() != 3 or shs_B.shape[1] != r or shs_B.shape[2] != Q:
            raise RgpError("sh2Weight_fast: B must be (N, r, Q) when A is (N, P, r)")
    else:
        n, P = shs_A.shape
        r = 1
        if shs_B.shape[-1] != Q:
            raise RgpError("sh2Weight_fast: B must be (N, Q)")
    return P, r


def sh2Weight_fast(shs_dc, shs_A, shs_B, pws, twc, calc_J, num_palette, Q):
    """gausplat.cu:469-509, rank r.  Returns [weights(, dweight_ddc, Ylm)]."""
    require_cuda(shs_dc, shs_A, shs_B, pws, twc)
    K, Q = int(num_palette), int(Q)
    P, r = _factors(shs_A, shs_B, K, Q)
    n = pws.shape[0]
    L = P * Q // K
    dev = pws.device
    weights = torch.zeros((n, K), dtype=torch.float32, device=dev)
    dwdc = torch.zeros((n, K), dtype=torch.float32, device=dev) if calc_J else None
    ylm = torch.zeros((n, L), dtype=torch.float32, device=dev) if calc_J else None
    call("rgp_sh2weight_fast", n, K, P, Q, r, ptr(f32c(shs_dc)), ptr(f32c(shs_A)), ptr(f32c(shs_B)),
         ptr(f32c(pws)), ptr(f32c(twc)), ptr(weights), ptr(dwdc), ptr(ylm), stream(dev))
    return [weights, dwdc, ylm] if calc_J else [weights]


def sh2Weight_fast_B(shs_A, shs_B, Ylm_stored, dloss_dw, num_palette, Q):
    """gausplat.cu:512-541, rank r.  Returns [dA, dB] shaped like A and B."""
    require_cuda(shs_A, shs_B, Ylm_stored, dloss_dw)
    K, Q = int(num_palette), int(Q)
    P, r = _factors(shs_A, shs_B, K, Q)
    n = shs_A.shape[0]
    dev = shs_A.device
    dA = torch.empty(shs_A.shape, dtype=torch.float32, device=dev)
    dB = torch.empty(shs_B.shape, dtype=torch.float32, device=dev)
    call("rgp_sh2weight_fast_backward", n, K, P, Q, r, ptr(f32c(shs_A)), ptr(f32c(shs_B)),
         ptr(f32c(Ylm_stored)), ptr(f32c(dloss_dw).reshape(n, K)), ptr(dA), ptr(dB), stream(dev))
    return [dA, dB]


def sh2Color(shs, pws, twc, calc_J):
    """gausplat.cu:547-584: [colors (N,3)] or, with calc_J, [colors,
    dcolor_dshs (N,1,sh_dim), dcolor_dpws (N,3,3)]."""
    require_cuda(shs, pws, twc)
    n = pws.shape[0]
    sh_dim = shs.shape[1] // 3
    dev = pws.device
    colors = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    dshs = torch.zeros((n, 1, sh_dim), dtype=torch.float32, device=dev) if calc_J else None
    dpws = torch.zeros((n, 3, 3), dtype=torch.float32, device=dev) if calc_J else None
    call("rgp_sh2color", n, sh_dim, ptr(f32c(shs)), ptr(f32c(pws)), ptr(f32c(twc)), ptr(colors), ptr(dshs),
         ptr(dpws), stream(dev))
    return [colors, dshs, dpws] if calc_J else [colors]


def _splat(height, width, dim, us, cinv2ds, alphas, depths, feat, areas, empty_tau):
    require_cuda(us, cinv2ds, alphas, depths, feat, areas)
    n = us.shape[0]
    dev = us.device
    d = _inplace_f32(depths)
    a = _inplace_i32(areas)
    us_c = f32c(us)
    records = ops.pack_records(us_c, cinv2ds, alphas, feat.reshape(n, dim))

    def fill(ws, cap):
        call("rgp_bin_rects", n, ptr(us_c), ptr(a), ptr(d), int(width), int(height), ptr(ws), I64(int(cap)),
             stream(dev))

    layer, contrib, final_tau, ranges, ids, m, _, _ = ops.splat_pipeline(
        n, int(width), int(height), dim, records, d, fill, empty_tau, dev, exact=True)
    return [layer, contrib, final_tau, ranges, ids[:m].clone()]


def splat_weight(height, width, num_palette, us, cinv2ds, alphas, depths, weights, areas):
    """gausplat.cu:161-249: [layer (K,H,W), contrib (H,W) i32, final_tau (H,W),
    ranges (T,2) i32, ids (M,) i32].  depths/areas are culled in place."""
    return _splat(int(height), int(width), int(num_palette), us, cinv2ds, alphas, depths, weights, areas, 0.0)


def splat_weight_B(height, width, num_palette, us, cinv2ds, alphas, depths, weights, contrib, final_tau,
                   patch_range_per_tile, gsid_per_patch, dloss_dgammas):
    """gausplat.cu:252-292: [dL/dweights (N,1,K)]."""
    require_cuda(us, cinv2ds, alphas, weights, contrib, final_tau, patch_range_per_tile, dloss_dgammas)
    K = int(num_palette)
    n = us.shape[0]
    records = ops.pack_records(us, cinv2ds, alphas, weights.reshape(n, K))
    ids = gsid_per_patch if gsid_per_patch.numel() > 0 else torch.zeros(1, dtype=torch.int32, device=us.device)
    d = ops.draw_backward(int(width), int(height), K, patch_range_per_tile, ids, records, contrib, final_tau,
                          dloss_dgammas, n)
    return [d.reshape(n, 1, K)]


def splat_color(height, width, us, cinv2ds, alphas, depths, colors, areas):
    """gausplat.cu:25-109 (final_tau of empty tiles is 1, gausplat.cu:39)."""
    return _splat(int(height), int(width), 3, us, cinv2ds, alphas, depths, colors, areas, 1.0)


def _color_B(height, width, us, cinv2ds, alphas, colors, contrib, final_tau, patch_range_per_tile, gsid_per_patch,
             dloss_dgammas, dloss_dfinal_tau):
    require_cuda(us, cinv2ds, alphas, colors, contrib, final_tau, patch_range_per_tile, dloss_dgammas,
                 dloss_dfinal_tau)
    n = us.shape[0]
    records = ops.pack_records(us, cinv2ds, alphas, colors.reshape(n, 3))
    ids = gsid_per_patch if gsid_per_patch.numel() > 0 else torch.zeros(1, dtype=torch.int32, device=us.device)
    dus, dcinv, dalpha, dcolor = ops.draw_color_backward(int(width), int(height), patch_range_per_tile, ids, records,
                                                         contrib, final_tau, dloss_dgammas, dloss_dfinal_tau, n)
    return [dus.reshape(n, 1, 2), dcinv.reshape(n, 1, 3), dalpha.reshape(n, 1, 1), dcolor.reshape(n, 1, 3)]


def splat_color_B(height, width, us, cinv2ds, alphas, depths, colors, contrib, final_tau, patch_range_per_tile,
                  gsid_per_patch, dloss_dgammas):
    """gausplat.cu:111-155 (draw_color_B, kernel.cu:1176-1289): [dL/dus (N,1,2),
    dL/dcinv2ds (N,1,3), dL/dalphas (N,1,1), dL/dcolors (N,1,3)]."""
    return _color_B(height, width, us, cinv2ds, alphas, colors, contrib, final_tau, patch_range_per_tile,
                    gsid_per_patch, dloss_dgammas, None)


def splat(height, width, us, cinv2ds, alphas, depths, colors, areas):
    """The vanilla module's name for the colour forward
    (vanilla_3dgs/gsplatcu/gausplat.cu:24-112): splat_color."""
    return splat_color(height, width, us, cinv2ds, alphas, depths, colors, areas)


def splatB(height, width, us, cinv2ds, alphas, depths, colors, contrib, final_tau, patch_range_per_tile,
           gsid_per_patch, dloss_dgammas, dloss_dfinal_tau):
    """vanilla_3dgs/gsplatcu/gausplat.cu:114-159 (drawB, kernel.cu:811-971):
    splat_color_B plus the gradient of a loss on final_tau, dL/dfinal_tau (H,W)."""
    return _color_B(height, width, us, cinv2ds, alphas, colors, contrib, final_tau, patch_range_per_tile,
                    gsid_per_patch, dloss_dgammas, dloss_dfinal_tau)


def lab2rgb(lab_image):
    raise NotImplementedError("gsplatcu.lab2rgb is unused by the palette-weight path and is not provided")
