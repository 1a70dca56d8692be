"""The C-ABI library: it loads, exports every function include/woxel_hip.h declares,
and reports errors through status codes and wx_last_error (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(wx_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


def test_header_declares_functions():
    names = declared_functions()
    assert "wx_trace_camera" in names and "wx_grid_create" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(wx):
    lib = C.CDLL(wx.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_hit_record_layout(wx, oracle_mod):
    assert wx.HIT_DTYPE == oracle_mod.HIT_DTYPE and wx.HIT_DTYPE.itemsize == 32
    assert C.sizeof(wx.Camera) == 64


def test_null_arguments_return_invalid(wx):
    assert wx.lib.wx_vdb_create(None) == wx.WX_ERR_INVALID
    assert b"NULL" in wx.lib.wx_last_error()
    assert wx.lib.wx_vdb_layout(None, None) == wx.WX_ERR_INVALID
    assert wx.lib.wx_trace_camera(None, None, 4, 4, 0.0, 0, None, None) == wx.WX_ERR_INVALID
    assert wx.lib.wx_trace_rays(None, None, 4, 0, None, None) == wx.WX_ERR_INVALID
    assert wx.lib.wx_grid_create(None, None) == wx.WX_ERR_INVALID
    assert wx.lib.wx_grid_destroy(None) == wx.WX_OK


def test_grid_bound_out_of_range_is_invalid(wx):
    """wx_grid_create rejects a bound outside (0, 2^100] before touching a device (the
    traversal's fast quotient relies on |p| <= bound <= 2^100)."""
    import ctypes as C

    for bound in (0.0, -1.0, float("nan"), float("inf"), 2.0 ** 100 * 1.0001, 1e31):
        d = wx.GridDesc()
        d.bound = bound
        h = C.c_void_p()
        assert wx.lib.wx_grid_create(C.byref(d), C.byref(h)) == wx.WX_ERR_INVALID, bound
        assert b"bound" in wx.lib.wx_last_error()


def test_camera_basis_matches_oracle(wx, oracle_mod):
    import numpy as np

    for eye, tgt, asp, fov, w in [((400, 300, -200), (128, 128, 128), 4 / 3, 60.0, 640.0),
                                  ((8048.0, 3000.0, 2048.0), (2048, 2048, 2048), 16 / 9, 45.0, 1920.0)]:
        p = wx.camera_basis(eye, tgt, (0, 1, 0), asp, fov, w).as_array()
        o = oracle_mod.camera_basis(eye, tgt, (0, 1, 0), asp, fov, w)
        assert np.array_equal(p.view(np.uint32), o.view(np.uint32))


def test_camera_basis_rejects_degenerate(wx):
    with pytest.raises(wx.WoxelError):
        wx.camera_basis((0, 0, 0), (0, 0, 0))
    with pytest.raises(wx.WoxelError):
        wx.camera_basis((0, 0, 0), (0, 5, 0), up=(0, 1, 0))


def _header_prototypes():
    protos = {}
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", h)).read(), flags=re.S)
        for m in re.finditer(r"(wx_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
            args = m.group(2).strip()
            protos[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return protos


def test_rust_binding_matches_header():
    """Every function of the extern "C" block of the integration guide is declared
    in the header with the same number of parameters."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = text[text.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    protos = _header_prototypes()
    seen = 0
    for m in re.finditer(r"pub fn (wx_\w+)\s*\(([^)]*)\)", block, flags=re.S):
        name, args = m.group(1), m.group(2).strip()
        n = 0 if not args else args.count(",") + 1 - args.rstrip().endswith(",")
        assert name in protos, name
        assert protos[name] == n, (name, protos[name], n)
        seen += 1
    assert seen >= 15
