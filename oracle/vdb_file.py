"""OpenVDB file reader, topology part -- TEST INFRASTRUCTURE ONLY (the checker).

A plain restatement of the reference reader (src/vdb/read.rs) in Python: the
archive header (read.rs:62-120), metadata (:213-271), grid descriptors
(:169-211), the transform (:141-167) and the tree topology (:273-347) with the
internal-node headers (:349-372) and the framing of compressed value blocks
(:374-519), which are skipped by their byte counts. Root tiles are read and
ignored, as the reference's traversal ignores them.

Pinned by the reference's own known answers (read.rs:746-802): the number of
active leaf voxels equals the grid's `file_voxel_count` metadata, 1,452,218 for
assets/cube.vdb and 364,741 for assets/icosahedron.vdb.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 0x56444220
V_BOOST_UUID = 218          # minimum supported (read.rs:24-30)
V_SELECTIVE_COMPRESSION = 220
V_NODE_MASK_COMPRESSION = 222
V_PER_GRID_COMPRESSION = 223

C_NONE, C_ZIP, C_ACTIVE_MASK, C_BLOSC = 0, 1, 2, 4


class Reader:
    def __init__(self, data: bytes):
        self.b = data
        self.p = 0

    def take(self, n):
        if self.p + n > len(self.b):
            raise ValueError("vdb: unexpected end of file")
        s = self.b[self.p:self.p + n]
        self.p += n
        return s

    def u8(self):
        return self.take(1)[0]

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def i64(self):
        return struct.unpack("<q", self.take(8))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def f32(self):
        return struct.unpack("<f", self.take(4))[0]

    def vec3i(self):
        return struct.unpack("<3i", self.take(12))

    def vec3d(self):
        return struct.unpack("<3d", self.take(24))

    def string(self, n):
        return self.take(n).decode("utf-8")

    def lstring(self):
        return self.string(self.u32())

    def words(self, n):
        return np.frombuffer(self.take(8 * n), dtype="<u8").copy()


def read_metadata(r: Reader) -> dict:
    """read.rs:213-271."""
    out = {}
    for _ in range(r.u32()):
        name = r.lstring()
        typ = r.lstring()
        n = r.u32()
        if typ == "string":
            out[name] = r.string(n)
        elif typ == "bool":
            out[name] = r.u8() == 1
        elif typ == "int32":
            out[name] = r.i32()
        elif typ == "int64":
            out[name] = r.i64()
        elif typ == "float":
            out[name] = r.f32()
        elif typ == "vec3i":
            out[name] = r.vec3i()
        else:
            out[name] = (typ, r.take(n))
    return out


def read_header(r: Reader) -> dict:
    """read.rs:62-120 and the grid descriptors (read.rs:169-211)."""
    if r.u64() != MAGIC:
        raise ValueError("vdb: magic mismatch")
    ver = r.u32()
    if ver < V_BOOST_UUID:
        raise ValueError(f"vdb: unsupported file version {ver}")
    major, minor = r.u32(), r.u32()
    has_offsets = r.u8() != 0
    comp = (C_ZIP | C_ACTIVE_MASK) if ver < V_PER_GRID_COMPRESSION else (C_BLOSC | C_ACTIVE_MASK)
    if V_SELECTIVE_COMPRESSION <= ver < V_NODE_MASK_COMPRESSION:
        comp = C_ZIP if r.u8() == 1 else C_NONE
    uuid = r.string(36)
    meta = read_metadata(r)
    ngrids = r.u32()
    if not has_offsets:
        raise ValueError("vdb: grid offsets required")
    grids = {}
    for _ in range(ngrids):
        name, gtype, parent = r.lstring(), r.lstring(), r.lstring()
        gpos, bpos, epos = r.u64(), r.u64(), r.u64()
        gcomp = comp
        if ver >= V_NODE_MASK_COMPRESSION:
            gcomp = r.u32()
        gmeta = read_metadata(r)
        grids[name] = dict(name=name, type=gtype, parent=parent, grid_pos=gpos, block_pos=bpos, end_pos=epos,
                           compression=gcomp, meta=gmeta)
        r.p = epos
    return dict(version=ver, library=(major, minor), uuid=uuid, compression=comp, meta=meta, grids=grids)


def skip_compressed(r: Reader, ver: int, gd: dict, size: int, value_mask: np.ndarray, elem: int):
    """Framing of read_compressed / read_compressed_data (read.rs:374-519)."""
    meta = r.u8() if ver >= V_NODE_MASK_COMPRESSION else 6
    if meta > 6:
        raise ValueError(f"vdb: invalid node metadata {meta}")
    if meta in (2, 4):
        r.take(elem)
    elif meta == 5:
        r.take(2 * elem)
    if meta in (3, 4, 5):
        r.take(8 * ((size + 63) // 64))
    comp = gd["compression"]
    if (comp & C_ACTIVE_MASK) and meta != 6 and ver >= V_NODE_MASK_COMPRESSION:
        count = int(np.unpackbits(value_mask.view(np.uint8)).sum())
    else:
        count = size
    if comp & (C_BLOSC | C_ZIP):
        n = r.i64()
        r.take(-n if n <= 0 else n)
    else:
        r.take(count * elem)


def read_grid(path_or_bytes, name: str | None = None) -> dict:
    """Header, descriptor and topology of one grid (read.rs:122-139, 273-347).

    Returns n5 = [(origin, child_mask u64[512], value_mask u64[512])],
    n4 = [(origin, child_mask u64[64], value_mask u64[64])] and
    leaves = [(origin, value_mask u64[8])], each in file order."""
    data = open(path_or_bytes, "rb").read() if isinstance(path_or_bytes, str) else path_or_bytes
    r = Reader(data)
    h = read_header(r)
    if name is None:
        name = next(iter(h["grids"]))
    gd = h["grids"][name]
    ver = h["version"]
    r.p = gd["grid_pos"]
    if ver >= V_NODE_MASK_COMPRESSION:
        r.u32()
    read_metadata(r)
    tname = r.lstring()
    if tname == "UniformScaleMap":
        tr = [r.vec3d() for _ in range(5)]
    elif tname in ("UniformScaleTranslateMap", "ScaleTranslateMap"):
        tr = [r.vec3d() for _ in range(6)]
    else:
        raise ValueError(f"vdb: unsupported transform {tname}")
    if r.u32() != 1:
        raise ValueError("vdb: multi-buffer trees are not supported")
    background = r.u32()
    ntiles, nn5 = r.u32(), r.u32()
    for _ in range(ntiles):
        r.vec3i(), r.u32(), r.u8()
    elem = 2 if gd["meta"].get("is_saved_as_half_float", False) else 4
    n5, n4, leaves = [], [], []
    for _ in range(nn5):
        o = r.vec3i()
        # the node is keyed by its origin floored to 4096 (root_key_from_coords)
        o5 = tuple((c >> 12) << 12 for c in o)
        c5, v5 = r.words(512), r.words(512)
        size = 32768 if ver >= V_NODE_MASK_COMPRESSION else 32768 - int(np.unpackbits(c5.view(np.uint8)).sum())
        skip_compressed(r, ver, gd, size, v5, elem)
        n5.append((o5, c5, v5))
        bits5 = np.flatnonzero(np.unpackbits(c5.view(np.uint8), bitorder="little"))
        for i5 in bits5:
            c4, v4 = r.words(64), r.words(64)
            size = 4096 if ver >= V_NODE_MASK_COMPRESSION else 4096 - int(np.unpackbits(c4.view(np.uint8)).sum())
            skip_compressed(r, ver, gd, size, v4, elem)
            o4 = (o5[0] + int(i5 >> 10) * 128, o5[1] + int((i5 >> 5) & 31) * 128, o5[2] + int(i5 & 31) * 128)
            n4.append((o4, c4, v4))
            for i4 in np.flatnonzero(np.unpackbits(c4.view(np.uint8), bitorder="little")):
                o3 = (o4[0] + int(i4 >> 8) * 8, o4[1] + int((i4 >> 4) & 15) * 8, o4[2] + int(i4 & 15) * 8)
                leaves.append((o3, r.words(8)))
    return dict(header=h, grid=gd, transform=(tname, tr), background=background, root_tiles=ntiles,
                n5=n5, n4=n4, leaves=leaves)


def active_voxels(g: dict) -> int:
    """The reference test's count (read.rs:770-795): active voxels of the leaves."""
    return int(sum(int(np.unpackbits(m.view(np.uint8)).sum()) for _, m in g["leaves"]))
