"""The vanilla 3DGS training step's autograd function on the gfx950 kernels.

GSFunction restates vanilla_3dgs/gsplat/gsmodel.py:6-85: the forward runs the
per-Gaussian stages with their calc_J Jacobians and the colour rasterizer;
the backward runs the colour backward with geometry gradients
(gsplatcu.splatB, rgp_draw_color_backward, including the gradient of a loss on
final_tau) and pulls dL/d(us, cinv2d, alpha, colour) back through the
Jacobians to dL/d(pws, shs, alphas, scales, rots) with the reference's
products (its batched float32 matrix products, here one kernel,
rgp_vanilla_chain, one Gaussian per thread).  The derivatives the
reference leaves out stay out: the +-1.3 tan_fov clamp of computeCov2D, the
0.99 clamp of alpha', max(0, .) of the Mahalanobis distance, and the
sh2Color term dY_9/dy = C30 (3x^2 - 2y^2) (kernel.cu:1152).
"""
import torch

import gsplatcu as gsc

from ._lib import call, f32c, ptr, stream


class GSFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pws, shs, alphas, scales, rots, us, cam):
        ctx.set_materialize_grads(False)  # an unused final_tau or mask gets no zero-filled gradient
        us, pcs, depths, du_dpcs = gsc.project(pws, cam.Rcw, cam.tcw, cam.fx, cam.fy, cam.cx, cam.cy, True)
        cov3ds, dcov3d_drots, dcov3d_dscales = gsc.computeCov3D(rots, scales, depths, True)
        cov2ds, dcov2d_dcov3ds, dcov2d_dpcs = gsc.computeCov2D(cov3ds, pcs, cam.Rcw, depths, cam.fx, cam.fy,
                                                               cam.width, cam.height, True)
        colors, dcolor_dshs, dcolor_dpws = gsc.sh2Color(shs, pws, cam.twc, True)
        cinv2ds, areas, dcinv2d_dcov2ds = gsc.inverseCov2D(cov2ds, depths, True)
        image, contrib, final_tau, ranges, ids = gsc.splat(cam.height, cam.width, us, cinv2ds, alphas, depths, colors,
                                                           areas)
        ctx.cam = cam
        ctx.save_for_backward(us, cinv2ds, alphas, depths, colors, contrib, final_tau, ranges, ids, dcinv2d_dcov2ds,
                              dcov2d_dcov3ds, dcov3d_drots, dcov3d_dscales, dcolor_dshs, du_dpcs, dcov2d_dpcs,
                              dcolor_dpws)
        return image, final_tau, depths > 0.2

    @staticmethod
    def backward(ctx, dloss_dgammas, dloss_dfinal_tau, _):
        cam = ctx.cam
        (us, cinv2ds, alphas, depths, colors, contrib, final_tau, ranges, ids, dcinv2d_dcov2ds, dcov2d_dcov3ds,
         dcov3d_drots, dcov3d_dscales, dcolor_dshs, du_dpcs, dcov2d_dpcs, dcolor_dpws) = ctx.saved_tensors
        if dloss_dgammas is None and dloss_dfinal_tau is None:
            return (None,) * 7
        if dloss_dgammas is None:
            dloss_dgammas = torch.zeros((3, cam.height, cam.width), dtype=torch.float32, device=us.device)
        dtau = None if dloss_dfinal_tau is None else dloss_dfinal_tau.contiguous()
        dus, dcinv, dalphas, dcolors = gsc.splatB(cam.height, cam.width, us, cinv2ds, alphas, depths, colors, contrib,
                                                  final_tau, ranges, ids, dloss_dgammas.contiguous(), dtau)
        n = us.shape[0]
        sh_dim = dcolor_dshs.shape[-1]
        dev = us.device
        dpws = torch.empty((n, 3), dtype=torch.float32, device=dev)
        dshs = torch.empty((n, 3 * sh_dim), dtype=torch.float32, device=dev)
        drots = torch.empty((n, 4), dtype=torch.float32, device=dev)
        dscales = torch.empty((n, 3), dtype=torch.float32, device=dev)
        # the chain of gsmodel.py:59-73 (dcov2d = dcinv J_inv, ..., dpw) in one kernel
        R = f32c(cam.Rcw).to(dev)
        call("rgp_vanilla_chain", n, sh_dim, ptr(R), ptr(dus), ptr(dcinv), ptr(dcolors), ptr(du_dpcs),
             ptr(dcinv2d_dcov2ds), ptr(dcov2d_dcov3ds), ptr(dcov2d_dpcs), ptr(dcov3d_drots), ptr(dcov3d_dscales),
             ptr(dcolor_dshs), ptr(dcolor_dpws), ptr(dpws), ptr(dshs), ptr(drots), ptr(dscales), stream(dev))
        return dpws, dshs, dalphas.reshape(n, 1), dscales, drots, dus.reshape(n, 2), None
