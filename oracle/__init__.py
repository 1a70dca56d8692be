"""ctypes binding of the CPU oracle (oracle/woxel_oracle.c).

TEST INFRASTRUCTURE ONLY: the oracle is the checker. Only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product package ``woxel_amd`` never does.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# 32-byte hit record shared with the product (woxel_hip.h: wx_hit).
HIT_DTYPE = np.dtype(
    [
        ("voxel", "<i4", (3,)),
        ("t", "<f4"),
        ("steps", "<u2", (4,)),
        ("state", "u1"),
        ("normal", "i1", (3,)),
        ("lookups", "<u4"),
    ]
)
assert HIT_DTYPE.itemsize == 32
DIAG_DTYPE = np.dtype([("p", "<f4", (3,)), ("flags", "<u4")])

HIT, OOB, MAXSTEPS = 0, 1, 2
F_TIE, F_NONFINITE, F_ZERODIR, F_CLIPPED, F_CLIPMISS = 1, 2, 4, 8, 16
EP_OFFS, EP_LEAF, EP_INNR, EP_ROOT, EP_BKGR = 0, 1, 2, 3, 4


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else [])
        subprocess.run(cmd, check=True)
    return LIB_PATH


class _Layout(C.Structure):
    _fields_ = [
        ("n5", C.c_uint32),
        ("n4", C.c_uint32),
        ("n3", C.c_uint32),
        ("origins", C.POINTER(C.c_int32)),
        ("kids5", C.POINTER(C.c_uint32)),
        ("vals5", C.POINTER(C.c_uint32)),
        ("kids4", C.POINTER(C.c_uint32)),
        ("vals4", C.POINTER(C.c_uint32)),
        ("vals3", C.POINTER(C.c_uint32)),
        ("atlas5", C.POINTER(C.c_uint32)),
        ("atlas4", C.POINTER(C.c_uint32)),
        ("atlas3", C.POINTER(C.c_uint32)),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i32, u32, f32 = C.c_void_p, C.c_int32, C.c_uint32, C.c_float
        L.or_vdb_new.restype = vp
        L.or_vdb_free.argtypes = [vp]
        L.or_vdb_set_voxel.argtypes = [vp, i32, i32, i32, u32]
        L.or_vdb_get_voxel.argtypes = [vp, i32, i32, i32, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        L.or_vdb_get_voxel.restype = C.c_int
        L.or_vdb_compute_sdf.argtypes = [vp]
        L.or_vdb_count_nodes.argtypes = [vp, C.POINTER(C.c_uint64)]
        L.or_vdb_add_sphere.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.POINTER(i32), C.POINTER(i32), C.c_int, u32]
        L.or_vdb_add_leaf_masks.argtypes = [vp, C.POINTER(i32), C.POINTER(C.c_uint64), C.c_size_t, u32]
        L.or_vdb_add_node.argtypes = [vp, C.c_int, C.POINTER(i32), C.POINTER(C.c_uint64)]
        L.or_vdb_add_node.restype = None
        L.or_vdb_leaf.argtypes = [vp, i32, i32, i32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.or_vdb_leaf.restype = C.c_int
        L.or_layout_build.argtypes = [vp]
        L.or_layout_build.restype = C.POINTER(_Layout)
        L.or_layout_free.argtypes = [C.POINTER(_Layout)]
        L.or_layout_from_arrays.argtypes = [C.c_uint32] * 3 + [C.POINTER(C.c_int32)] + [C.POINTER(C.c_uint32)] * 8
        L.or_layout_from_arrays.restype = C.POINTER(_Layout)
        L.or_hdda_ray.argtypes = [C.POINTER(_Layout), C.POINTER(f32), C.POINTER(f32), f32, C.c_int, vp, vp]
        L.or_camera_ray.argtypes = [C.POINTER(f32), u32, u32, f32, C.POINTER(f32)]
        L.or_camera_basis.argtypes = [C.POINTER(f32)] * 3 + [f32, f32, f32, C.POINTER(f32)]
        L.or_trace_camera.argtypes = [C.POINTER(_Layout), C.POINTER(f32), u32, u32, f32, f32, C.c_int,
                                      vp, vp, C.c_int]
        L.or_trace_rays.argtypes = [C.POINTER(_Layout), C.POINTER(f32), C.c_size_t, f32, C.c_int, vp, vp,
                                    C.c_int]
        L.or_render_camera.argtypes = [C.POINTER(_Layout), C.POINTER(f32), u32, u32, f32, f32, C.c_int,
                                       C.POINTER(Shade), vp, vp, C.c_int]
        L.or_render_camera.restype = None
        L.or_shadow_ray.argtypes = [C.POINTER(f32), C.POINTER(f32), C.POINTER(C.c_int8), C.POINTER(f32),
                                    C.POINTER(f32), C.POINTER(f32)]
        L.or_global_to_node.argtypes = [C.POINTER(i32), C.c_int, C.POINTER(i32)]
        L.or_global_to_offset.argtypes = [C.POINTER(i32), C.c_int, C.c_int]
        L.or_global_to_offset.restype = u32
        L.or_offset_to_child.argtypes = [u32, C.c_int, C.POINTER(u32)]
        L.or_child_to_offset.argtypes = [C.POINTER(u32), C.c_int]
        L.or_child_to_offset.restype = u32
        _lib = L
    return _lib


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Vdb:
    """Reference-style VDB345 tree (vdb345.rs) held by the oracle."""

    def __init__(self):
        self.h = lib().or_vdb_new()
        self._layout = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def close(self):
        if getattr(self, "_layout", None) is not None:
            self._layout.close()
            self._layout = None
        if getattr(self, "h", None):
            lib().or_vdb_free(self.h)
            self.h = None

    def set_voxel(self, x, y, z, value_bits=0x3F800000):
        lib().or_vdb_set_voxel(self.h, int(x), int(y), int(z), int(value_bits) & 0xFFFFFFFF)

    def add_sphere(self, c, r, lo=(-(1 << 22),) * 3, hi=((1 << 22),) * 3, value_mode=0, value_bits=0x3F800000):
        lo = np.asarray(lo, np.int32)
        hi = np.asarray(hi, np.int32)
        lib().or_vdb_add_sphere(self.h, float(c[0]), float(c[1]), float(c[2]), float(r), _i32p(lo), _i32p(hi),
                                int(value_mode), int(value_bits))

    def add_leaf_masks(self, origins, masks, value_bits=0x3F800000):
        org = np.ascontiguousarray(origins, np.int32).reshape(-1, 3)
        m = np.ascontiguousarray(masks, np.uint64).reshape(-1, 8)
        lib().or_vdb_add_leaf_masks(self.h, _i32p(org), m.ctypes.data_as(C.POINTER(C.c_uint64)), org.shape[0],
                                    int(value_bits))

    def add_node(self, level, origin, mask):
        """or_vdb_add_node: a level 5/4/3 node with its value mask (512/64/8 u64)."""
        o = np.ascontiguousarray(origin, np.int32).reshape(3)
        m = np.ascontiguousarray(mask, np.uint64).reshape({5: 512, 4: 64, 3: 8}[level])
        lib().or_vdb_add_node(self.h, int(level), _i32p(o), m.ctypes.data_as(C.POINTER(C.c_uint64)))

    @classmethod
    def from_file_topology(cls, g):
        """A tree with the topology oracle.vdb_file.read_grid returned, in file order."""
        v = cls()
        for o, _, vm in g["n5"]:
            v.add_node(5, o, vm)
        for o, _, vm in g["n4"]:
            v.add_node(4, o, vm)
        for o, vm in g["leaves"]:
            v.add_node(3, o, vm)
        return v

    def get_voxel(self, x, y, z):
        v = C.c_uint64()
        lvl = C.c_int()
        kind = lib().or_vdb_get_voxel(self.h, int(x), int(y), int(z), C.byref(v), C.byref(lvl))
        return kind, v.value, lvl.value

    def compute_sdf(self):
        lib().or_vdb_compute_sdf(self.h)

    def count_nodes(self):
        a = (C.c_uint64 * 3)()
        lib().or_vdb_count_nodes(self.h, a)
        return [int(a[0]), int(a[1]), int(a[2])]

    def leaf(self, x, y, z):
        d = np.zeros(512, np.uint64)
        iv = np.zeros(8, np.uint64)
        ok = lib().or_vdb_leaf(self.h, int(x), int(y), int(z), d.ctypes.data_as(C.POINTER(C.c_uint64)),
                               iv.ctypes.data_as(C.POINTER(C.c_uint64)))
        return (d, iv) if ok else None

    def layout(self):
        """The reference layout (or_layout_build), cached."""
        if self._layout is None:
            self._layout = Layout(lib().or_layout_build(self.h), owner=self)
        return self._layout

    def layout_arrays(self):
        return self.layout().arrays()

    def hdda_ray(self, *a, **k):
        return self.layout().hdda_ray(*a, **k)

    def trace_camera(self, *a, **k):
        return self.layout().trace_camera(*a, **k)

    def trace_rays(self, *a, **k):
        return self.layout().trace_rays(*a, **k)


class Shade(C.Structure):
    """or_shade: mode, show_345[3], sun_dir[4], sun_color[4] (rgb, intensity)."""
    _fields_ = [("mode", C.c_uint32), ("show_345", C.c_uint32 * 3), ("sun_dir", C.c_float * 4),
                ("sun_color", C.c_float * 4)]


def shade_struct(d):
    """dict(mode, show_345, sun_dir, sun_color, intensity) -> Shade."""
    s = Shade()
    s.mode = int(d.get("mode", 3))
    for i, v in enumerate(d.get("show_345", (0, 0, 0))):
        s.show_345[i] = int(v)
    for i, v in enumerate(d["sun_dir"]):
        s.sun_dir[i] = float(v)
    for i, v in enumerate(d["sun_color"]):
        s.sun_color[i] = float(v)
    s.sun_color[3] = float(d.get("intensity", 1.0))
    return s


class Layout:
    """The reference GPU layout (origins, five mask arrays, three atlases) that
    the oracle's traversal reads."""

    def __init__(self, ptr, owner=None):
        self.ptr = ptr
        self._owner = owner

    @classmethod
    def from_arrays(cls, origins, kids5, vals5, kids4, vals4, vals3, atlas5, atlas4, atlas3):
        """Copy reference-format arrays (e.g. exported by another builder)."""
        u32 = lambda a: np.ascontiguousarray(a, np.uint32)
        org = np.zeros((len(origins), 4), np.int32)
        org[:, :3] = np.asarray(origins, np.int32).reshape(-1, 3)
        arrs = [u32(a) for a in (kids5, vals5, kids4, vals4, vals3, atlas5, atlas4, atlas3)]
        n5, n4, n3 = org.shape[0], arrs[2].size // 128, arrs[4].size // 16
        up = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
        ptr = lib().or_layout_from_arrays(n5, n4, n3, _i32p(org), *[up(a) for a in arrs])
        return cls(ptr)

    def close(self):
        if getattr(self, "ptr", None) is not None:
            lib().or_layout_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        """numpy copies of the layout arrays."""
        L = self.ptr.contents

        def view(ptr, n, dt):
            if n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(n,)).view(dt).copy()

        return {
            "n": (L.n5, L.n4, L.n3),
            "origins": view(L.origins, L.n5 * 4, np.int32).reshape(-1, 4),
            "kids5": view(L.kids5, L.n5 * 1024, np.uint32).reshape(-1, 1024),
            "vals5": view(L.vals5, L.n5 * 1024, np.uint32).reshape(-1, 1024),
            "kids4": view(L.kids4, L.n4 * 128, np.uint32).reshape(-1, 128),
            "vals4": view(L.vals4, L.n4 * 128, np.uint32).reshape(-1, 128),
            "vals3": view(L.vals3, L.n3 * 16, np.uint32).reshape(-1, 16),
            "atlas5": view(L.atlas5, L.n5 * 32768, np.uint32).reshape(-1, 32768),
            "atlas4": view(L.atlas4, L.n4 * 4096, np.uint32).reshape(-1, 4096),
            "atlas3": view(L.atlas3, L.n3 * 512, np.uint32).reshape(-1, 512),
        }

    def hdda_ray(self, src, d, bound=4096.0, clip=False):
        s = np.asarray(src, np.float32)
        dd = np.asarray(d, np.float32)
        hit = np.zeros(1, HIT_DTYPE)
        diag = np.zeros(1, DIAG_DTYPE)
        lib().or_hdda_ray(self.ptr, _f32p(s), _f32p(dd), bound, int(clip), hit.ctypes.data, diag.ctypes.data)
        return hit[0], diag[0]

    def trace_camera(self, cam, w, h, offset=0.001, bound=4096.0, clip=False, threads=None):
        cam = np.ascontiguousarray(cam, np.float32).reshape(16)
        out = np.zeros(w * h, HIT_DTYPE)
        diag = np.zeros(w * h, DIAG_DTYPE)
        lib().or_trace_camera(self.ptr, _f32p(cam), w, h, offset, bound, int(clip), out.ctypes.data,
                              diag.ctypes.data, threads or os.cpu_count() or 1)
        return out, diag

    def render_camera(self, cam, w, h, shade, offset=0.001, bound=4096.0, clip=False, threads=None):
        """or_render_camera: (rgba float32 (h*w, 4), rgba8 uint32 (h*w,))."""
        cam = np.ascontiguousarray(cam, np.float32).reshape(16)
        rgba = np.zeros((w * h, 4), np.float32)
        rgba8 = np.zeros(w * h, np.uint32)
        lib().or_render_camera(self.ptr, _f32p(cam), w, h, offset, bound, int(clip), C.byref(shade_struct(shade)),
                               rgba.ctypes.data, rgba8.ctypes.data, threads or os.cpu_count() or 1)
        return rgba, rgba8

    def trace_rays(self, rays, bound=4096.0, clip=False, threads=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        out = np.zeros(n, HIT_DTYPE)
        diag = np.zeros(n, DIAG_DTYPE)
        lib().or_trace_rays(self.ptr, _f32p(rays), n, bound, int(clip), out.ctypes.data, diag.ctypes.data,
                            threads or os.cpu_count() or 1)
        return out, diag


def camera_ray(cam, px, py, offset=0.001):
    cam = np.ascontiguousarray(cam, np.float32).reshape(16)
    d = np.zeros(3, np.float32)
    lib().or_camera_ray(_f32p(cam), px, py, offset, _f32p(d))
    return d


def camera_basis(eye, target, up, aspect, fovy_deg, width):
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    out = np.zeros(16, np.float32)
    lib().or_camera_basis(_f32p(e), _f32p(t), _f32p(u), aspect, fovy_deg, width, _f32p(out))
    return out


def shadow_ray(p, d, normal, light):
    p = np.asarray(p, np.float32)
    d = np.asarray(d, np.float32)
    n = np.asarray(normal, np.int8)
    l = np.asarray(light, np.float32)
    o = np.zeros(3, np.float32)
    dd = np.zeros(3, np.float32)
    lib().or_shadow_ray(_f32p(p), _f32p(d), n.ctypes.data_as(C.POINTER(C.c_int8)), _f32p(l), _f32p(o), _f32p(dd))
    return o, dd


def global_to_node(g, total_log2):
    g = np.asarray(g, np.int32)
    out = np.zeros(3, np.int32)
    lib().or_global_to_node(_i32p(g), total_log2, _i32p(out))
    return out.tolist()


def global_to_offset(g, log2_d, total_log2):
    g = np.asarray(g, np.int32)
    return int(lib().or_global_to_offset(_i32p(g), log2_d, total_log2))


def offset_to_child(off, log2_d):
    out = np.zeros(3, np.uint32)
    lib().or_offset_to_child(off, log2_d, out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out.tolist()


def child_to_offset(c, log2_d):
    c = np.asarray(c, np.uint32)
    return int(lib().or_child_to_offset(c.ctypes.data_as(C.POINTER(C.c_uint32)), log2_d))


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]
