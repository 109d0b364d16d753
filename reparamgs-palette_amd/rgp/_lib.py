"""Loader of librgp.so, the gfx950 kernels behind the C ABI of include/rgp.h.

The library is built in-tree (reparamgs-palette_amd/lib/librgp.so) by
csrc/build.py.  There is no fallback: if the library is missing or the device
is not a GPU, every operator raises.
"""
import ctypes
import os

import torch

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "librgp.so")
_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
F = ctypes.c_float
D = ctypes.c_double
U32 = ctypes.c_uint32
SZ = ctypes.c_size_t

_SIGS = {
    "rgp_last_error": (ctypes.c_char_p, []),
    "rgp_version": (I, []),
    "rgp_project": (I, [I, P, P, P, F, F, F, F, P, P, P, P, P]),
    "rgp_compute_cov3d": (I, [I, P, P, P, P, P, P, P]),
    "rgp_compute_cov2d": (I, [I, P, P, P, P, F, F, F, F, P, P, P, P]),
    "rgp_inverse_cov2d": (I, [I, P, P, P, P, P, P]),
    "rgp_sh2weight": (I, [I, I, I, P, P, P, P, P, P, P]),
    "rgp_sh2weight_fast": (I, [I, I, I, I, I, P, P, P, P, P, P, P, P, P]),
    "rgp_sh2weight_fast_backward": (I, [I, I, I, I, I, P, P, P, P, P, P, P]),
    "rgp_sh2color": (I, [I, I, P, P, P, P, P, P, P]),
    "rgp_bin_workspace_bytes": (SZ, [I, I64, I, I]),
    "rgp_bin_rects": (I, [I, P, P, P, I, I, P, I64, P]),
    "rgp_bin_scan": (I, [I, I, I, P, I64, P]),
    "rgp_preprocess_weight_views": (I, [I, I, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, SZ, I64, I, I, P]),
    "rgp_bin_scan_views": (I, [I, I, I, I, P, SZ, I64, P]),
    "rgp_bin_sort_views": (I, [I, I, P, I, I, P, SZ, I64, P, P, P]),
    "rgp_draw_forward_views": (I, [I, I, I, I, P, P, I64, P, I64, F, P, P, P, P, P]),
    "rgp_draw_backward_replay_views": (I, [I, I, I, I, P, P, I64, P, I64, P, P, P, P]),
    "rgp_sh_backward_full_views": (I, [I, I, I, I, P, P, P, P, P]),
    "rgp_sh_backward_adam_split": (I, [I, I, P, P, P, P, P, P, P, P, P, I64, D, D, D, D, D, P]),
    "rgp_sh_backward_adam_lowrank": (I, [I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, D, D, D, D,
                                         P]),
    "rgp_sh_backward_adam_lowrank_views": (I, [I, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, I64, D, D, D,
                                               D, D, P]),
    "rgp_sh_backward_lowrank_views": (I, [I, I, I, I, I, I, P, P, P, P, P, P, P, P, P]),
    "rgp_bin_sort": (I, [I, P, I, I, P, I64, P, P, P, P]),
    "rgp_bin_header": (P, [P]),
    "rgp_record_floats": (I, [I]),
    "rgp_pack_records": (I, [I, I, P, P, P, P, P, P]),
    "rgp_draw_forward": (I, [I, I, I, P, P, P, F, P, P, P, P, P]),
    "rgp_draw_forward_counts": (I, [I, I, I, P, P, P, F, P, P, P, P, P]),
    "rgp_draw_backward": (I, [I, I, I, P, P, P, P, P, P, P, P]),
    "rgp_draw_backward_replay": (I, [I, I, I, P, P, P, P, P, P, P]),
    "rgp_preprocess_weight": (I, [I, I, I, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, I64, I, I,
                                  P, P, P, P, P]),
    "rgp_sh_backward_full": (I, [I, I, I, P, P, P, P, P]),
    "rgp_sh_backward_adam_full": (I, [I, I, I, P, P, P, P, P, P, I64, D, D, D, D, D, P]),
    "rgp_adam": (I, [I64, P, P, P, P, I64, D, D, D, D, D, P]),
    "rgp_draw_color_backward": (I, [I, I, P, P, P, P, P, P, P, P, P, P, P, P]),
    "rgp_vanilla_chain": (I, [I, I] + [P] * 16 + [P]),
    "rgp_sh_backward_lowrank": (I, [I, I, I, I, I, P, P, P, P, P, P, P, P, P]),
    "rgp_mix_batched": (I, [I, I, P, P, P, P, I64, P]),
    "rgp_mix": (I, [I, I64, P, P, P, P]),
    "rgp_mix_backward": (I, [I, I64, P, P, P, P, P, P]),
    "rgp_fill_u32": (I, [P, U32, I64, P]),
    "rgp_l1_ssim_sums": (I64, [I, I, I]),
    "rgp_l1_ssim_forward": (I, [I, I, I, P, P, P, F, F, F, P, P, P, P]),
    "rgp_l1_ssim_backward": (I, [I, I, I, P, P, P, P, F, P, P, P]),
    "rgp_bary_workspace_bytes": (ctypes.c_size_t, [I]),
    "rgp_bary_loss": (I, [I, I, I, P, P, P, P, F, I, F, F, P, P, P, P, P, P]),
    "rgp_bary_loss_noalloc": (I, [I, I, I, P, P, P, P, F, I, F, F, P, P, P, P, P]),
    "rgp_bary_coords": (I, [I, I, I, P, P, F, P, P]),
    "rgp_bary_coords_backward": (I, [I, I, I, P, P, F, P, P, P, P]),
}


class RgpError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python __graft_entry__.py build` "
                "(or reparamgs-palette_amd/csrc/build.py) first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise RgpError(f"{name} failed ({rc}): {lib().rgp_last_error().decode()}")
    return rc


def ptr(t):
    return None if t is None else t.data_ptr()


def stream(device=None):
    return torch.cuda.current_stream(device).cuda_stream


def require_cuda(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RgpError("rgp operators need GPU tensors (got a tensor on %s)" % t.device)


def f32c(t):
    """float32, contiguous (a copy only when needed, as .contiguous() in the
    reference host code)."""
    if t.dtype != torch.float32:
        t = t.float()
    return t.contiguous()


def i32c(t):
    if t.dtype != torch.int32:
        t = t.int()
    return t.contiguous()
