"""woxel_amd — MI355X-native HDDA + grid-SDF ray traversal over VDB345.

Python binding of libwoxel_hip.so (C ABI in include/woxel_hip.h). The classes
mirror the reference's host interface:

  Vdb345        VDB345<T> (src/vdb/vdb345.rs): set_voxel, get_voxel, compute_sdf,
                count_nodes, origins, masks, atlas
  camera_basis  ComputeState::build (src/render/gpu_types/compute_state.rs:87-131)
  Grid          the device-resident grid (the atlas textures + mask buffers that
                WgpuContext uploads, src/render/wgpu_context.rs:101-159), with
                trace_camera (cp_main) and trace_rays (hdda_ray over a ray list)

There is no CPU fallback: importing this package without the built library
raises ImportError.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwoxel_hip.so")

WX_OK, WX_ERR_INVALID, WX_ERR_DEVICE, WX_ERR_UNSUPPORTED, WX_ERR_STATE, WX_ERR_BUDGET = 0, 1, 2, 3, 4, 5
WX_ERR_BUSY = 6
WX_ERR_NOT_FOUND = 7
HIT, OOB, MAXSTEPS, NO_RAY = 0, 1, 2, 3
TRACE_CLIP, TRACE_SHADOW, TRACE_SIMPLE = 1, 2, 4
REQ_CAMERA, REQ_RAYS = 0, 1
EP_OFFS, EP_LEAF, EP_INNR, EP_ROOT, EP_BKGR = 0, 1, 2, 3, 4

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


class WoxelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[wx {code}] {msg}")
        self.code = code


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 4), ("u", C.c_float * 4), ("mv", C.c_float * 4), ("wp", C.c_float * 4)]

    def as_array(self):
        return np.array(list(self.eye) + list(self.u) + list(self.mv) + list(self.wp), np.float32)

    @classmethod
    def from_array(cls, a):
        a = np.asarray(a, np.float32).reshape(16)
        c = cls()
        for i in range(4):
            c.eye[i], c.u[i], c.mv[i], c.wp[i] = a[i], a[4 + i], a[8 + i], a[12 + i]
        return c


class Request(C.Structure):
    """wx_request: one camera or ray set of a batched launch."""
    _fields_ = [
        ("kind", C.c_uint32),
        ("grid", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("flags", C.c_uint32),
        ("pixel_offset", C.c_float),
        ("camera", Camera),
        ("rays", C.c_void_p),
        ("light", C.c_float * 4),
        ("out", C.c_void_p),
        ("shadow_out", C.c_void_p),
    ]


assert C.sizeof(Request) == 128


class Tune(C.Structure):
    """wx_tune: launch shape of the persistent traversal."""
    _fields_ = [("persistent", C.c_uint32), ("refill", C.c_uint32), ("chunk", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("xcd_split", C.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("streams", C.c_uint32), ("max_batch_requests", C.c_uint32),
                ("max_batch_rays", C.c_uint64), ("flags", C.c_uint32), ("max_queued", C.c_uint32),
                ("wait_limit_us", C.c_uint32), ("idle_stop_us", C.c_uint32)]


class ServeRequest(C.Structure):
    _fields_ = [("grid", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("flags", C.c_uint32),
                ("pixel_offset", C.c_float), ("light", C.c_float * 3), ("camera", Camera)]


class ServeResult(C.Structure):
    _fields_ = [("primary", C.c_void_p), ("shadow", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32),
                ("submit_s", C.c_double), ("launch_s", C.c_double), ("done_s", C.c_double),
                ("batch_requests", C.c_uint32), ("status", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("batches", C.c_uint64), ("primary_rays", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("mean_batch_requests", C.c_double), ("budget_in_use", C.c_uint64), ("budget_peak", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SERVER_DRYRUN = 1
SERVER_RESIDENT = 2


class VdbFileInfo(C.Structure):
    _fields_ = [("file_version", C.c_uint32), ("library_major", C.c_uint32), ("library_minor", C.c_uint32),
                ("compression", C.c_uint32), ("half_float", C.c_uint32), ("n5", C.c_uint32), ("n4", C.c_uint32),
                ("n3", C.c_uint32), ("root_tiles", C.c_uint32), ("active_tiles", C.c_uint32),
                ("file_voxel_count", C.c_int64), ("active_voxels", C.c_uint64), ("voxel_size", C.c_double * 3),
                ("grid_name", C.c_char * 64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["voxel_size"] = tuple(self.voxel_size)
        d["grid_name"] = self.grid_name.decode()
        return d


class Shade(C.Structure):
    """wx_shade: render mode, grid overlays, sun (ComputeState, compute_state.rs:9-29)."""
    _fields_ = [("mode", C.c_uint32), ("show_345", C.c_uint32 * 3), ("sun_dir", C.c_float * 4),
                ("sun_color", C.c_float * 4)]


RENDER_F32 = 8
GRAY, RGB, RAY_LENGTH, DIFFUSE, GLOSSY = 0, 1, 2, 3, 4


def default_sun():
    """SunSettings::default (egui_dev.rs:355-368): dir normalize(1, -1, 0.5) as
    v * (1 / |v|) in f32, colour (255, 210, 160) / 255, intensity 1."""
    v = np.array([1.0, -1.0, 0.5], np.float32)
    inv = np.float32(1.0) / np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]), dtype=np.float32)
    d = (v * inv).astype(np.float32)
    return dict(sun_dir=tuple(float(x) for x in d),
                sun_color=tuple(float(np.float32(c) / np.float32(255.0)) for c in (255.0, 210.0, 160.0)),
                intensity=1.0)


def shade(mode: int = DIFFUSE, show_345=(0, 0, 0), sun_dir=None, sun_color=None, intensity=None) -> dict:
    """Shading settings as a dict (defaults: the reference's sun)."""
    d = default_sun()
    d["mode"], d["show_345"] = int(mode), tuple(int(x) for x in show_345)
    if sun_dir is not None:
        d["sun_dir"] = tuple(float(x) for x in sun_dir)
    if sun_color is not None:
        d["sun_color"] = tuple(float(x) for x in sun_color)
    if intensity is not None:
        d["intensity"] = float(intensity)
    return d


def shade_struct(d: dict) -> Shade:
    s = Shade()
    s.mode = int(d["mode"])
    for i, v in enumerate(d.get("show_345", (0, 0, 0))):
        s.show_345[i] = int(v)
    for i, v in enumerate(d["sun_dir"]):
        s.sun_dir[i] = float(v)
    for i, v in enumerate(d["sun_color"]):
        s.sun_color[i] = float(v)
    s.sun_color[3] = float(d.get("intensity", 1.0))
    return s


class BuildInfo(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("before_sdf_s", C.c_double), ("sdf_device_s", C.c_double),
                ("total_s", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GridDesc(C.Structure):
    _fields_ = [
        ("n5", C.c_uint32),
        ("n4", C.c_uint32),
        ("n3", C.c_uint32),
        ("bound", C.c_float),
        ("root", C.POINTER(C.c_int32)),
        ("n5_records", C.POINTER(C.c_uint32)),
        ("n4_records", C.POINTER(C.c_uint32)),
        ("leaves", C.POINTER(C.c_uint32)),
    ]


def _load():
    # One HIP runtime per process: PyTorch's libc10_hip binds "libamdhip64.so" from
    # torch/lib, our library "libamdhip64.so.7". Loading torch first makes the
    # loader resolve our dependency to the runtime torch already mapped (same
    # SONAME), so device pointers and streams are shared; loading ours first would
    # map a second runtime from /opt/rocm.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"woxel_amd: {LIB_PATH} is not built; run `python woxel_amd/build.py` "
            "(there is no CPU fallback for the traversal)"
        )
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, f32, sz = C.c_void_p, C.c_int32, C.c_uint32, C.c_float, C.c_size_t
    P = C.POINTER
    sig = {
        "wx_last_error": ([], C.c_char_p),
        "wx_version": ([], C.c_int),
        "wx_device_count": ([P(C.c_int)], C.c_int),
        "wx_vdb_create": ([P(vp)], C.c_int),
        "wx_vdb_destroy": ([vp], C.c_int),
        "wx_vdb_set_voxel": ([vp, i32, i32, i32, u32], C.c_int),
        "wx_vdb_set_voxels": ([vp, P(i32), sz], C.c_int),
        "wx_vdb_add_spheres": ([vp, P(C.c_double), sz, P(i32), P(i32), C.c_int], C.c_int),
        "wx_vdb_add_leaf_masks": ([vp, P(i32), P(C.c_uint64), sz], C.c_int),
        "wx_vdb_add_noise": ([vp, C.c_uint64, P(i32), P(i32), f32, f32, C.c_int, P(f32)], C.c_int),
        "wx_vdb_add_noise_device": ([vp, C.c_uint64, P(i32), P(i32), f32, f32, vp, P(f32)], C.c_int),
        "wx_vdb_finalize": ([vp, C.c_int], C.c_int),
        "wx_vdb_finalize_device": ([vp, C.c_int, vp, P(BuildInfo)], C.c_int),
        "wx_vdb_add_nodes": ([vp, C.c_int, P(i32), P(C.c_uint64), sz], C.c_int),
        "wx_vdb_read_file": ([vp, C.c_char_p, C.c_char_p, P(VdbFileInfo)], C.c_int),
        "wx_vdb_get_voxel": ([vp, i32, i32, i32, P(C.c_int), P(C.c_uint64), P(C.c_int)], C.c_int),
        "wx_vdb_count_nodes": ([vp, P(C.c_uint64)], C.c_int),
        "wx_vdb_active_voxels": ([vp, P(C.c_uint64)], C.c_int),
        "wx_vdb_ref_masks": ([vp] + [P(u32)] * 5, C.c_int),
        "wx_vdb_ref_atlas": ([vp] + [P(u32)] * 3, C.c_int),
        "wx_vdb_layout": ([vp, P(GridDesc)], C.c_int),
        "wx_vdb_leaf_origins": ([vp, P(i32)], C.c_int),
        "wx_camera_basis": ([P(f32), P(f32), P(f32), f32, f32, f32, P(Camera)], C.c_int),
        "wx_grid_create": ([P(GridDesc), P(vp)], C.c_int),
        "wx_vdb_grid_create": ([vp, f32, P(vp)], C.c_int),
        "wx_layout_save": ([P(GridDesc), C.c_char_p, C.c_char_p], C.c_int),
        "wx_grid_save": ([vp, C.c_char_p, C.c_char_p], C.c_int),
        "wx_layout_probe": ([C.c_char_p, C.c_char_p, P(u32)], C.c_int),
        "wx_grid_load": ([C.c_char_p, C.c_char_p, P(vp)], C.c_int),
        "wx_grid_destroy": ([vp], C.c_int),
        "wx_grid_bytes": ([vp, P(C.c_uint64)], C.c_int),
        "wx_trace_camera": ([vp, P(Camera), u32, u32, f32, u32, vp, vp], C.c_int),
        "wx_trace_rays": ([vp, vp, u32, u32, vp, vp], C.c_int),
        "wx_render_camera": ([vp, P(Camera), u32, u32, f32, P(Shade), u32, vp, vp], C.c_int),
        "wx_tune_get": ([P(Tune)], C.c_int),
        "wx_tune_set": ([P(Tune)], C.c_int),
        "wx_trace_batch": ([P(vp), u32, P(Request), u32, P(Tune), vp], C.c_int),
        "wx_trace_batch_static": ([P(vp), u32, P(Request), u32, P(Tune), vp], C.c_int),
        "wx_autotune": ([P(vp), u32, P(Request), u32, vp, P(Tune), P(f32)], C.c_int),
        "wx_server_create": ([P(ServerConfig), P(vp)], C.c_int),
        "wx_server_destroy": ([vp], C.c_int),
        "wx_server_add_grid": ([vp, vp, P(u32)], C.c_int),
        "wx_server_submit": ([vp, P(ServeRequest), P(C.c_uint64)], C.c_int),
        "wx_server_wait": ([vp, C.c_uint64, C.c_double, P(ServeResult)], C.c_int),
        "wx_server_release": ([vp, C.c_uint64], C.c_int),
        "wx_server_cancel": ([vp, C.c_uint64], C.c_int),
        "wx_server_get_stats": ([vp, P(ServerStats)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    return L


lib = _load()


def _check(rc):
    if rc != WX_OK:
        raise WoxelError(rc, lib.wx_last_error().decode(errors="replace"))


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def device_count() -> int:
    n = C.c_int()
    _check(lib.wx_device_count(C.byref(n)))
    return n.value


def camera_basis(eye, target, up=(0.0, 1.0, 0.0), aspect=4.0 / 3.0, fovy_deg=60.0, width=640.0) -> Camera:
    """ComputeState::build (compute_state.rs:87-131): the eye/u/mv/wp words."""
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)
    cam = Camera()
    _check(lib.wx_camera_basis(_p(e, C.c_float), _p(t, C.c_float), _p(u, C.c_float), float(aspect),
                               float(fovy_deg), float(width), C.byref(cam)))
    return cam


class Vdb345:
    """Host VDB345 tree (vdb345.rs). Build with set_voxel / add_spheres, then
    compute_sdf() (topology in DFS order + grid SDF + device layout)."""

    def __init__(self):
        h = C.c_void_p()
        _check(lib.wx_vdb_create(C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.wx_vdb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_voxel(self, x, y, z, value_bits=0):
        _check(lib.wx_vdb_set_voxel(self.h, int(x), int(y), int(z), int(value_bits) & 0xFFFFFFFF))

    def set_voxels(self, xyz):
        a = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        _check(lib.wx_vdb_set_voxels(self.h, _p(a, C.c_int32), a.shape[0]))

    def add_spheres(self, spheres, lo=(-(1 << 22),) * 3, hi=((1 << 22),) * 3, threads=None):
        s = np.ascontiguousarray(spheres, np.float64).reshape(-1, 4)
        lo = np.asarray(lo, np.int32)
        hi = np.asarray(hi, np.int32)
        _check(lib.wx_vdb_add_spheres(self.h, _p(s, C.c_double), s.shape[0], _p(lo, C.c_int32), _p(hi, C.c_int32),
                                      int(threads or os.cpu_count() or 1)))

    def add_noise(self, seed, lo, hi, base_freq, occupancy, threads=None, device=False, stream=None):
        """Isosurface of 3-octave value noise; returns the threshold used. With
        device=True the voxels are decided on the current GPU (identical leaves)."""
        lo = np.asarray(lo, np.int32)
        hi = np.asarray(hi, np.int32)
        thr = C.c_float()
        if device:
            _check(lib.wx_vdb_add_noise_device(self.h, int(seed), _p(lo, C.c_int32), _p(hi, C.c_int32),
                                               float(base_freq), float(occupancy),
                                               None if stream is None else C.c_void_p(stream.cuda_stream),
                                               C.byref(thr)))
        else:
            _check(lib.wx_vdb_add_noise(self.h, int(seed), _p(lo, C.c_int32), _p(hi, C.c_int32), float(base_freq),
                                        float(occupancy), int(threads or os.cpu_count() or 1), C.byref(thr)))
        return thr.value

    def leaf_origins(self):
        """Origins (n x 3 voxel coordinates) of every leaf, in pool (DFS) order."""
        org = np.zeros((self.count_nodes()[2], 3), np.int32)
        _check(lib.wx_vdb_leaf_origins(self.h, _p(org, C.c_int32)))
        return org

    def leaf_masks(self):
        """(origins n x 3, value masks n x 8 u64) of every leaf, in pool (DFS) order."""
        n3 = self.count_nodes()[2]
        org = self.leaf_origins()
        v3 = self.masks()[4]
        return org, np.ascontiguousarray(v3).view(np.uint64).reshape(n3, 8)

    def add_leaf_masks(self, keys, masks):
        k = np.ascontiguousarray(keys, np.int32).reshape(-1, 3)
        m = np.ascontiguousarray(masks, np.uint64).reshape(-1, 8)
        assert k.shape[0] == m.shape[0]
        _check(lib.wx_vdb_add_leaf_masks(self.h, _p(k, C.c_int32), _p(m, C.c_uint64), k.shape[0]))

    def add_nodes(self, level: int, keys, masks):
        """Explicit topology (read.rs:273-347): level 5/4/3 node keys (origin >> 12 / 7 / 3)
        with their value masks (512 / 64 / 8 u64 per node); empty leaves are kept."""
        words = {5: 512, 4: 64, 3: 8}[level]
        k = np.ascontiguousarray(keys, np.int32).reshape(-1, 3)
        m = np.ascontiguousarray(masks, np.uint64).reshape(-1, words)
        assert k.shape[0] == m.shape[0]
        _check(lib.wx_vdb_add_nodes(self.h, int(level), _p(k, C.c_int32), _p(m, C.c_uint64), k.shape[0]))

    def read_file(self, path: str, grid: str | None = None) -> dict:
        """Topology of one grid of an OpenVDB file (VdbReader, read.rs); returns what was read."""
        info = VdbFileInfo()
        _check(lib.wx_vdb_read_file(self.h, os.fsencode(path), grid.encode() if grid else None, C.byref(info)))
        return info.as_dict()

    def compute_sdf(self, threads=None, device: bool = False):
        """Finalise: DFS topology, compute_sdf (vdb345.rs:290-628), device layout. With
        device=True the SDF passes run on the GPU (same result); returns the launch count."""
        t = int(threads or os.cpu_count() or 1)
        if device:
            info = BuildInfo()
            _check(lib.wx_vdb_finalize_device(self.h, t, None, C.byref(info)))
            self.build_info = info.as_dict()
            return info.launches
        _check(lib.wx_vdb_finalize(self.h, t))

    def get_voxel(self, x, y, z):
        kind, val, lvl = C.c_int(), C.c_uint64(), C.c_int()
        _check(lib.wx_vdb_get_voxel(self.h, int(x), int(y), int(z), C.byref(kind), C.byref(val), C.byref(lvl)))
        return kind.value, val.value, lvl.value

    def count_nodes(self):
        a = (C.c_uint64 * 3)()
        _check(lib.wx_vdb_count_nodes(self.h, a))
        return [int(a[0]), int(a[1]), int(a[2])]

    def active_voxels(self):
        a = C.c_uint64()
        _check(lib.wx_vdb_active_voxels(self.h, C.byref(a)))
        return int(a.value)

    def desc(self) -> GridDesc:
        d = GridDesc()
        _check(lib.wx_vdb_layout(self.h, C.byref(d)))
        return d

    def origins(self):
        """VDB345::origins (vdb345.rs:108-117): sorted N5 origins."""
        d = self.desc()
        if d.n5 == 0:
            return np.zeros((0, 3), np.int32)
        r = np.ctypeslib.as_array(d.root, shape=(d.n5 * 4,)).reshape(-1, 4).copy()
        return (r[:, :3].astype(np.int64) << 12).astype(np.int32)

    def masks(self):
        """VDB345::masks (vdb345.rs:119-161) as u32 words."""
        n5, n4, n3 = self.count_nodes()
        k5 = np.zeros((n5, 1024), np.uint32)
        v5 = np.zeros((n5, 1024), np.uint32)
        k4 = np.zeros((n4, 128), np.uint32)
        v4 = np.zeros((n4, 128), np.uint32)
        v3 = np.zeros((n3, 16), np.uint32)
        _check(lib.wx_vdb_ref_masks(self.h, *[_p(a, C.c_uint32) for a in (k5, v5, k4, v4, v3)]))
        return k5, v5, k4, v4, v3

    def atlas(self):
        """VDB345::atlas (vdb345.rs:163-264) slot values per node (offset order)."""
        n5, n4, n3 = self.count_nodes()
        a5 = np.zeros((n5, 32768), np.uint32)
        a4 = np.zeros((n4, 4096), np.uint32)
        a3 = np.zeros((n3, 512), np.uint32)
        _check(lib.wx_vdb_ref_atlas(self.h, *[_p(a, C.c_uint32) for a in (a5, a4, a3)]))
        return a5, a4, a3

    def layout_arrays(self):
        """Copies of the device layout (root, n5/n4 records, leaf slots)."""
        d = self.desc()

        def view(ptr, n):
            return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, np.uint32)

        return {
            "n": (d.n5, d.n4, d.n3),
            "root": view(d.root, d.n5 * 4).reshape(-1, 4),
            "n5": view(d.n5_records, d.n5 * 4096 * 4).reshape(-1, 4),
            "n4": view(d.n4_records, d.n4 * 256 * 4).reshape(-1, 4),
            "leaves": view(d.leaves, d.n3 * 64).reshape(-1, 64),
        }


def save_layout(vdb: Vdb345, path: str, key: str, bound: float = 4096.0) -> None:
    """Writes a finalized tree's host layout to `path` under `key` (wx_layout_save)."""
    d = vdb.desc()
    d.bound = float(bound)
    _check(lib.wx_layout_save(C.byref(d), os.fsencode(path), key.encode()))


def probe_layout(path: str, key: str):
    """(n5, n4, n3) of the layout file at `path` written under `key`, or None when
    there is no such file or it was written under another key (wx_layout_probe);
    raises WoxelError for a damaged or foreign file."""
    a = (C.c_uint32 * 3)()
    rc = lib.wx_layout_probe(os.fsencode(path), key.encode(), a)
    if rc == WX_ERR_NOT_FOUND:
        return None
    _check(rc)
    return (int(a[0]), int(a[1]), int(a[2]))


class Grid:
    """A device-resident grid. Trace outputs are device buffers (torch tensors)."""

    def __init__(self, vdb: Vdb345 | None, bound: float = 4096.0):
        """The grid of a finalized tree (wx_vdb_grid_create: after a device build the
        leaf words are copied on the device, not through the host)."""
        if vdb is None:  # filled by from_desc / load
            self.h = None
            return
        h = C.c_void_p()
        _check(lib.wx_vdb_grid_create(vdb.h, float(bound), C.byref(h)))
        self.h = h
        self.counts = tuple(vdb.count_nodes())

    @classmethod
    def from_desc(cls, d: GridDesc) -> "Grid":
        """wx_grid_create over a host layout descriptor."""
        g = cls(None)
        h = C.c_void_p()
        _check(lib.wx_grid_create(C.byref(d), C.byref(h)))
        g.h, g.counts = h, (d.n5, d.n4, d.n3)
        return g

    @classmethod
    def load(cls, path: str, key: str) -> "Grid":
        """A grid from a layout file written under `key` (wx_grid_load); raises
        WoxelError with code WX_ERR_NOT_FOUND when there is none."""
        g = cls(None)
        h = C.c_void_p()
        _check(lib.wx_grid_load(os.fsencode(path), key.encode(), C.byref(h)))
        g.h, g.counts = h, tuple(probe_layout(path, key))
        return g

    def save(self, path: str, key: str) -> None:
        """Writes the grid's layout to `path` under `key` (wx_grid_save)."""
        _check(lib.wx_grid_save(self.h, os.fsencode(path), key.encode()))

    def close(self):
        if getattr(self, "h", None):
            lib.wx_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nbytes(self) -> int:
        a = C.c_uint64()
        _check(lib.wx_grid_bytes(self.h, C.byref(a)))
        return int(a.value)

    @staticmethod
    def _stream(stream):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def trace_camera(self, cam: Camera, w: int, h: int, pixel_offset: float = 0.001, clip: bool = False,
                     out=None, stream=None, simple: bool = False):
        """One primary ray per pixel (cp_main). Returns a (h*w, 32) uint8 cuda tensor."""
        import torch

        if out is None:
            out = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        assert out.is_cuda and out.numel() >= w * h * 32
        flags = (TRACE_CLIP if clip else 0) | (TRACE_SIMPLE if simple else 0)
        _check(lib.wx_trace_camera(self.h, C.byref(cam), w, h, float(pixel_offset), flags,
                                   C.c_void_p(out.data_ptr()), self._stream(stream)))
        return out

    def render_camera(self, cam: Camera, w: int, h: int, shade: dict, pixel_offset: float = 0.001,
                      clip: bool = False, f32: bool = False, out=None, stream=None):
        """Pixel colours (ray_trace, raycast.comp.wgsl:144-342): a (h*w,) int32 cuda tensor of
        rgba8 words, or (h*w, 4) float32 with f32=True."""
        import torch

        if out is None:
            out = (torch.empty((w * h, 4), dtype=torch.float32, device="cuda") if f32
                   else torch.empty(w * h, dtype=torch.int32, device="cuda"))
        assert out.is_cuda and out.numel() * out.element_size() >= w * h * (16 if f32 else 4)
        flags = (TRACE_CLIP if clip else 0) | (RENDER_F32 if f32 else 0)
        s = shade_struct(shade)
        _check(lib.wx_render_camera(self.h, C.byref(cam), w, h, float(pixel_offset), C.byref(s), flags,
                                    C.c_void_p(out.data_ptr()), self._stream(stream)))
        return out

    def trace_rays(self, rays, clip: bool = False, out=None, stream=None, simple: bool = False):
        """rays: cuda float32 tensor (n, 6) = origin, direction."""
        import torch

        assert rays.is_cuda and rays.dtype == torch.float32 and rays.is_contiguous()
        n = rays.numel() // 6
        if out is None:
            out = torch.empty((n, 32), dtype=torch.uint8, device="cuda")
        flags = (TRACE_CLIP if clip else 0) | (TRACE_SIMPLE if simple else 0)
        _check(lib.wx_trace_rays(self.h, C.c_void_p(rays.data_ptr()), n, flags,
                                 C.c_void_p(out.data_ptr()), self._stream(stream)))
        return out


def camera_request(grid_index: int, cam: Camera, w: int, h: int, out, pixel_offset: float = 0.5,
                   clip: bool = True, shadow_out=None, light=(8000.0, 8000.0, 8000.0)) -> Request:
    """A camera request; out / shadow_out are cuda tensors of >= w*h*32 bytes."""
    r = Request()
    r.kind = REQ_CAMERA
    r.grid = grid_index
    r.width, r.height = w, h
    r.flags = (TRACE_CLIP if clip else 0) | (TRACE_SHADOW if shadow_out is not None else 0)
    r.pixel_offset = pixel_offset
    r.camera = cam
    r.out = out.data_ptr()
    r.shadow_out = shadow_out.data_ptr() if shadow_out is not None else None
    for i in range(3):
        r.light[i] = light[i]
    return r


def rays_request(grid_index: int, rays, out, clip: bool = True, shadow_out=None,
                 light=(8000.0, 8000.0, 8000.0)) -> Request:
    r = Request()
    r.kind = REQ_RAYS
    r.grid = grid_index
    r.width, r.height = rays.numel() // 6, 1
    r.flags = (TRACE_CLIP if clip else 0) | (TRACE_SHADOW if shadow_out is not None else 0)
    r.rays = rays.data_ptr()
    r.out = out.data_ptr()
    r.shadow_out = shadow_out.data_ptr() if shadow_out is not None else None
    for i in range(3):
        r.light[i] = light[i]
    return r


def _batch_args(grids, requests):
    gh = (C.c_void_p * len(grids))(*[g.h.value for g in grids])
    rq = (Request * len(requests))(*requests)
    return gh, rq


def trace_batch(grids, requests, tune: Tune | None = None, stream=None):
    """All requests in one persistent launch (wx_trace_batch)."""
    gh, rq = _batch_args(grids, requests)
    _check(lib.wx_trace_batch(gh, len(grids), rq, len(requests), C.byref(tune) if tune is not None else None,
                              Grid._stream(stream)))


def trace_batch_static(grids, requests, tune: Tune | None = None, stream=None):
    """wx_trace_batch_static: trace_batch with its tables in the GPU's static scratch
    (no allocation; at most 64 grids and 1024 requests; overlapping calls race)."""
    gh, rq = _batch_args(grids, requests)
    _check(lib.wx_trace_batch_static(gh, len(grids), rq, len(requests),
                                     C.byref(tune) if tune is not None else None, Grid._stream(stream)))


def tune_get() -> Tune:
    t = Tune()
    _check(lib.wx_tune_get(C.byref(t)))
    return t


def tune_set(t: Tune):
    _check(lib.wx_tune_set(C.byref(t)))


def autotune(grids, requests, stream=None):
    """Times launch shapes on this batch, stores the best as the device default."""
    gh, rq = _batch_args(grids, requests)
    t = Tune()
    ms = C.c_float()
    _check(lib.wx_autotune(gh, len(grids), rq, len(requests), Grid._stream(stream), C.byref(t), C.byref(ms)))
    return t, ms.value


class Server:
    """Continuous-batching serving loop over resident grids (wx_server_*)."""

    def __init__(self, hbm_budget: int, streams: int = 2, max_batch_requests: int = 64, max_batch_rays: int = 0,
                 dryrun: bool = False, max_queued: int = 0, resident: bool = False, wait_limit_us: int = 0,
                 idle_stop_us: int = 0):
        """resident=True admits requests into one running kernel (WX_SERVER_RESIDENT);
        wait_limit_us bounds each of its waits, idle_stop_us ends it when idle."""
        flags = (SERVER_DRYRUN if dryrun else 0) | (SERVER_RESIDENT if resident else 0)
        cfg = ServerConfig(int(hbm_budget), int(streams), int(max_batch_requests), int(max_batch_rays), flags,
                           int(max_queued), int(wait_limit_us), int(idle_stop_us))
        h = C.c_void_p()
        _check(lib.wx_server_create(C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.wx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_grid(self, grid) -> int:
        gid = C.c_uint32()
        _check(lib.wx_server_add_grid(self.h, grid.h if grid is not None else None, C.byref(gid)))
        return gid.value

    def submit(self, grid_id: int, cam: Camera, w: int, h: int, pixel_offset: float = 0.5, clip: bool = True,
               shadow: bool = False, light=(8000.0, 8000.0, 8000.0)) -> int:
        r = ServeRequest()
        r.grid, r.width, r.height = grid_id, w, h
        r.flags = (TRACE_CLIP if clip else 0) | (TRACE_SHADOW if shadow else 0)
        r.pixel_offset = pixel_offset
        for i in range(3):
            r.light[i] = light[i]
        r.camera = cam
        t = C.c_uint64()
        _check(lib.wx_server_submit(self.h, C.byref(r), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_s: float = -1.0) -> ServeResult:
        res = ServeResult()
        _check(lib.wx_server_wait(self.h, ticket, float(timeout_s), C.byref(res)))
        return res

    def release(self, ticket: int):
        _check(lib.wx_server_release(self.h, ticket))

    def cancel(self, ticket: int):
        _check(lib.wx_server_cancel(self.h, ticket))

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib.wx_server_get_stats(self.h, C.byref(st)))
        return st


def device_copy(ptr: int, n: int):
    """Copy n hit records from a device pointer (e.g. a served result) into a new
    (n, 32) uint8 cuda tensor."""
    import torch

    buf = torch.empty((n, 32), dtype=torch.uint8, device="cuda")
    lib_hip = _hip_runtime()
    rc = lib_hip.hipMemcpy(C.c_void_p(buf.data_ptr()), C.c_void_p(ptr), C.c_size_t(n * 32), 3)
    if rc != 0:
        raise WoxelError(WX_ERR_DEVICE, f"hipMemcpy device->device failed ({rc})")
    return buf


def device_hits(ptr: int, n: int):
    """Copy n hit records from a device pointer into a numpy array (HIT_DTYPE)."""
    return hits_to_numpy(device_copy(ptr, n))


_hip = None


def _hip_runtime():
    global _hip
    if _hip is None:
        import torch  # noqa: F401  (torch has mapped its HIP runtime; reuse it)

        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    return _hip


def hits_to_numpy(t) -> np.ndarray:
    """Device hit buffer -> structured numpy array (HIT_DTYPE)."""
    a = t.detach().to("cpu").numpy().reshape(-1)
    return a.view(HIT_DTYPE)
