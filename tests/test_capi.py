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


def _desc_from(wx, arrs, bound=4096.0):
    """A wx_grid_desc over numpy copies of a layout (kept alive by the caller)."""
    d = wx.GridDesc()
    d.n5, d.n4, d.n3 = (int(x) for x in arrs["n"])
    d.bound = bound
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    d.root = P(arrs["root"], C.c_int32)
    d.n5_records = P(arrs["n5"], C.c_uint32)
    d.n4_records = P(arrs["n4"], C.c_uint32)
    d.leaves = P(arrs["leaves"], C.c_uint32)
    return d


def test_grid_create_rejects_corrupted_descriptors(wx):
    """wx_grid_create checks every index a kernel derives from the descriptor before
    anything reaches a device: a root entry naming a node-5 outside the pool, a
    record whose walk bit has no child bit, a first-child index whose children run
    past the child pool. Each is WX_ERR_INVALID (no GPU needed); the intact layout
    passes the check (WX_OK on a GPU host, WX_ERR_DEVICE on a CPU-only host)."""
    import numpy as np
    from woxel_amd import scenes

    v = scenes.build_sphere_vdb(np.array([[100.0, 100.0, 100.0, 30.0], [-4000.0, 50.0, 50.0, 20.0]]))
    base = v.layout_arrays()
    n5, n4, n3 = base["n"]
    assert n5 == 2 and n4 > 0 and n3 > 0
    h = C.c_void_p()
    rc = wx.lib.wx_grid_create(C.byref(_desc_from(wx, base)), C.byref(h))
    assert rc in (wx.WX_OK, wx.WX_ERR_DEVICE)
    if rc == wx.WX_OK:
        wx.lib.wx_grid_destroy(h)
    # node-5 record / node-4 record that are walked into (walk bits set)
    r5 = int(np.nonzero((base["n5"][:, 0] >> 8) & 0xFF)[0][0])
    r4 = int(np.nonzero(base["n4"][:, 0] >> 16)[0][0])
    cases = []
    a = {k: (x.copy() if isinstance(x, np.ndarray) else x) for k, x in base.items()}
    a["root"][1, 3] = n5  # root entry names node-5 #n5
    cases.append(("root", a))
    a = {k: (x.copy() if isinstance(x, np.ndarray) else x) for k, x in base.items()}
    a["root"][0, 3] = -1
    cases.append(("root", a))
    a = {k: (x.copy() if isinstance(x, np.ndarray) else x) for k, x in base.items()}
    a["n5"][r5, 1] = n4  # first child beyond the node-4 pool
    cases.append(("node-5", a))
    a = {k: (x.copy() if isinstance(x, np.ndarray) else x) for k, x in base.items()}
    kids = int(a["n5"][r5, 0]) & 0xFF
    free = [b for b in range(8) if not (kids >> b) & 1]
    if free:  # a walk bit without its child bit
        a["n5"][r5, 0] |= np.uint32(1 << (8 + free[0]))
        cases.append(("node-5", a))
    a = {k: (x.copy() if isinstance(x, np.ndarray) else x) for k, x in base.items()}
    a["n4"][r4, 1] = n3 - bin(int(a["n4"][r4, 0]) & 0xFFFF).count("1") + 1  # last child = leaf #n3
    cases.append(("node-4", a))
    a = {k: (x.copy() if isinstance(x, np.ndarray) else x) for k, x in base.items()}
    a["n4"][r4, 0] = np.uint32(0xFFFF0000) | (a["n4"][r4, 0] & np.uint32(0xFFFF))  # every walk bit
    if (int(a["n4"][r4, 0]) & 0xFFFF) != 0xFFFF:
        cases.append(("node-4", a))
    assert len(cases) >= 5
    for what, arrs in cases:
        rc = wx.lib.wx_grid_create(C.byref(_desc_from(wx, arrs)), C.byref(h))
        msg = wx.lib.wx_last_error().decode()
        assert rc == wx.WX_ERR_INVALID, (what, rc, msg)
        assert what in msg, (what, msg)


def test_static_batch_limits_without_a_device(wx):
    """wx_trace_batch_static rejects batches beyond its static scratch (64 grids,
    1024 requests) before anything reaches a device."""
    grids = (C.c_void_p * 65)()
    reqs = (wx.Request * 1025)()
    assert wx.lib.wx_trace_batch_static(grids, 65, reqs, 1, None, None) == wx.WX_ERR_INVALID
    assert wx.lib.wx_trace_batch_static(grids, 1, reqs, 1025, None, None) == wx.WX_ERR_INVALID
    assert b"64 grids" in wx.lib.wx_last_error()
    assert wx.lib.wx_trace_batch_static(None, 0, None, 0, None, None) == wx.WX_OK


_C_TO_RUST = {"uint32_t": "u32", "int32_t": "i32", "uint64_t": "u64", "int64_t": "i64", "uint16_t": "u16",
              "int16_t": "i16", "uint8_t": "u8", "int8_t": "i8", "float": "f32", "double": "f64", "char": "c_char",
              "int": "i32", "size_t": "usize"}


def _c_structs():
    """typedef structs of the header: name -> [(field, rust type)] with arrays as [T; N]
    and pointers as *const / *mut T."""
    out = {}
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", h)).read(), flags=re.S)
        for m in re.finditer(r"typedef struct (\w+) \{(.*?)\} \1;", text, flags=re.S):
            fields = []
            for decl in m.group(2).split(";"):
                decl = " ".join(decl.split())
                if not decl:
                    continue
                tm = re.match(r"(const )?(\w+)\s*(.*)$", decl)
                assert tm, (m.group(1), decl)
                const, ctype, rest = tm.groups()
                for one in rest.split(","):  # "uint32_t width, height"
                    fm = re.match(r"\s*(\*?)\s*(\w+)((?:\[\d+\])?)\s*$", one)
                    assert fm, (m.group(1), decl)
                    ptr, name, arr = fm.groups()
                    rt = _C_TO_RUST.get(ctype, ctype)
                    if ptr:
                        rt = ("*const " if const else "*mut ") + rt
                    if arr:
                        rt = f"[{rt}; {arr[1:-1]}]"
                    fields.append((name, rt))
            out[m.group(1)] = fields
    return out


def _rust_structs():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    out = {}
    for m in re.finditer(r"pub struct (\w+) \{(.*?)\}", text, flags=re.S):
        body = re.sub(r"//[^\n]*", "", m.group(2))
        fields = [(f.group(1), " ".join(f.group(2).split())) for f in
                  re.finditer(r"pub (\w+): ((?:\[[^\]]*\])|(?:\*(?:const|mut) \w+)|\w+)", body)]
        out[m.group(1)] = fields
    return out


def test_rust_structs_match_header():
    """Every #[repr(C)] struct of the integration guide's Rust binding has the header's
    fields in the same order with the same types and array lengths (so the layouts
    agree), and every header struct that crosses the ABI by value or pointer is bound."""
    cs, rs = _c_structs(), _rust_structs()
    assert len(rs) >= 12
    for name, fields in rs.items():
        assert name in cs, name
        assert fields == cs[name], (name, fields, cs[name])
    missing = sorted(set(cs) - set(rs))
    assert not missing, missing


def test_build_rs_uses_the_library_build_flags():
    """The integration guide's build.rs calls hipcc / g++ itself with the same sources
    and flags as woxel_amd/build.py (the flags the bit-exact results depend on)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("wxb", os.path.join(ROOT, "woxel_amd", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    rs = text[text.index("## build.rs"):text.index("## src/render/woxel_hip.rs")]
    assert "python" not in rs
    lists = {m.group(1): re.findall(r'"([^"]+)"', m.group(2))
             for m in re.finditer(r"const (\w+): &\[&str\] = &\[(.*?)\];", rs, flags=re.S)}
    assert lists["HIP"] == b.HIP_SOURCES and lists["CXX"] == b.CXX_SOURCES
    assert lists["COMMON"] == b.COMMON and lists["HIP_FLAGS"] == b.HIP_FLAGS
