"""The C-ABI library loads on a machine without a GPU and exports every entry
point that include/rgp.h declares (no compute calls here)."""
import ctypes
import os
import re

from conftest import ROOT

LIB = os.path.join(ROOT, "reparamgs-palette_amd", "lib", "librgp.so")
HDR = os.path.join(ROOT, "include", "rgp.h")


def declared():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rgp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "rgp_draw_forward" in names and "rgp_bin_sort" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_pure_host_entry_points():
    lib = ctypes.CDLL(LIB)
    lib.rgp_record_floats.restype = ctypes.c_int
    assert lib.rgp_record_floats(5) == 16
    assert lib.rgp_record_floats(3) == 12
    assert lib.rgp_record_floats(9) == 0
    lib.rgp_bin_workspace_bytes.restype = ctypes.c_size_t
    lib.rgp_bin_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
    # tile-bucket grids: per pair of capacity one u64 key and one 16-byte coarse-list entry
    small = lib.rgp_bin_workspace_bytes(1000, 0, 64, 64)
    big = lib.rgp_bin_workspace_bytes(1000, 100000, 64, 64)
    assert 100000 * 24 <= big - small < 100000 * 25
    # grids above 32768 tiles (the radix path): two key and two id buffers
    small = lib.rgp_bin_workspace_bytes(1000, 0, 7680, 4320)
    big = lib.rgp_bin_workspace_bytes(1000, 100000, 7680, 4320)
    assert big - small >= 100000 * 24
    assert lib.rgp_version() >= 1
