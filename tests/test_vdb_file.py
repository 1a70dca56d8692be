"""The .vdb reader (src/vdb/read.rs) and the explicit topology it produces.

* The oracle's reader (oracle/vdb_file.py) is pinned by the reference's own
  known answers: active leaf voxels == file_voxel_count (read.rs:770-802).
* The product reader (wx_vdb_read_file) is compared with it node by node, and
  the finalised layout (DFS numbering, grid SDF) slot by slot.
* Explicit topology the synthetic grids never produce -- nodes without
  children, active tiles, leaves without active voxels -- is checked between
  product and oracle, and the oracle on it by closed forms.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from test_layout import compare

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# read.rs:770-802 asserts these against each grid's file_voxel_count metadata
KAT = {"cube": 1_452_218, "icosahedron": 364_741}


def fixture(name):
    path = os.path.join(GOLDEN, f"{name}.vdb")
    meta = json.load(open(os.path.join(GOLDEN, "vdb_fixtures.json")))[f"{name}.vdb"]
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == meta["sha256"]
    return path


@pytest.fixture(scope="module")
def vdb_file():
    from oracle import vdb_file

    return vdb_file


@pytest.mark.parametrize("name", ["cube", "icosahedron"])
def test_oracle_reader_known_answers(vdb_file, name):
    g = vdb_file.read_grid(fixture(name))
    assert g["grid"]["name"] == f"ls_{name}"
    assert g["grid"]["meta"]["file_voxel_count"] == KAT[name]
    assert vdb_file.active_voxels(g) == KAT[name]
    assert g["header"]["version"] == 222 and g["grid"]["meta"]["is_saved_as_half_float"]


def leaf_sets(origins, masks):
    return {tuple(int(c) for c in o): bytes(np.asarray(m, np.uint64).tobytes()) for o, m in zip(origins, masks)}


@pytest.mark.parametrize("name", ["cube", "icosahedron"])
def test_product_reader_matches_oracle(wx, oracle_mod, vdb_file, name):
    g = vdb_file.read_grid(fixture(name))
    p = wx.Vdb345()
    info = p.read_file(fixture(name))
    assert info["grid_name"] == f"ls_{name}"
    assert info["file_voxel_count"] == KAT[name] and info["active_voxels"] == KAT[name]
    assert (info["n5"], info["n4"], info["n3"]) == (len(g["n5"]), len(g["n4"]), len(g["leaves"]))
    assert info["half_float"] == 1 and info["file_version"] == 222 and info["compression"] == 2
    assert abs(info["voxel_size"][0] - 0.05) < 1e-7
    p.compute_sdf(threads=4)
    assert p.active_voxels() == KAT[name]
    org, masks = p.leaf_masks()
    assert leaf_sets(org, masks) == leaf_sets([o for o, _ in g["leaves"]], [m for _, m in g["leaves"]])
    o = oracle_mod.Vdb.from_file_topology(g)
    o.compute_sdf()
    compare(p, o)


def test_reader_selects_grid_by_name(wx):
    p = wx.Vdb345()
    assert p.read_file(fixture("icosahedron"), "ls_icosahedron")["n3"] == 1943
    with pytest.raises(wx.WoxelError) as e:
        wx.Vdb345().read_file(fixture("icosahedron"), "no_such_grid")
    assert e.value.code == wx.WX_ERR_INVALID


def test_reader_rejects_bad_files(wx, tmp_path):
    data = open(fixture("icosahedron"), "rb").read()
    bad = tmp_path / "bad.vdb"
    bad.write_bytes(b"\0" * 64)
    with pytest.raises(wx.WoxelError) as e:
        wx.Vdb345().read_file(str(bad))
    assert e.value.code == wx.WX_ERR_INVALID and "magic" in str(e.value)
    for cut in (12, 300, len(data) // 2, len(data) - 1):
        short = tmp_path / f"short{cut}.vdb"
        short.write_bytes(data[:cut])
        with pytest.raises(wx.WoxelError) as e:
            wx.Vdb345().read_file(str(short))
        assert e.value.code == wx.WX_ERR_INVALID
    old = tmp_path / "old.vdb"
    old.write_bytes(data[:8] + (217).to_bytes(4, "little") + data[12:])
    with pytest.raises(wx.WoxelError) as e:
        wx.Vdb345().read_file(str(old))
    assert e.value.code == wx.WX_ERR_UNSUPPORTED
    with pytest.raises(wx.WoxelError):
        wx.Vdb345().read_file(str(tmp_path / "missing.vdb"))
    v = wx.Vdb345()
    v.compute_sdf()
    with pytest.raises(wx.WoxelError) as e:
        v.read_file(fixture("icosahedron"))
    assert e.value.code == wx.WX_ERR_STATE


def explicit_tree(rng):
    """Nodes without children, active tiles, empty leaves beside ordinary leaves."""
    n5 = [((0, 0, 0), 40), ((4096, 0, 0), 0), ((-4096, 4096, 0), 7)]  # (origin, active tiles)
    n4 = [((128, 256, 0), 30), ((384, 0, 128), 0), ((4096 + 640, 128, 0), 11)]
    leaves = []
    for o in ((8, 16, 24), (136, 264, 8), (-4096 + 40, 4096 + 8, 64), (4096 + 640 + 8, 128 + 16, 0)):
        m = rng.integers(0, 2**63, size=8, dtype=np.uint64) & rng.integers(0, 2**63, size=8, dtype=np.uint64)
        leaves.append((o, m))
    leaves.append(((256, 256, 256), np.zeros(8, np.uint64)))           # an empty leaf
    leaves.append(((4096 + 640, 128 + 8, 8), np.zeros(8, np.uint64)))   # an empty leaf beside a full one

    def tiles(count, bits):
        m = np.zeros(bits // 64, np.uint64)
        for b in rng.choice(bits, size=count, replace=False):
            m[b >> 6] |= np.uint64(1) << np.uint64(b & 63)
        return m

    return [(o, tiles(c, 32768)) for o, c in n5], [(o, tiles(c, 4096)) for o, c in n4], leaves


def test_explicit_topology_matches_oracle(wx, oracle_mod):
    rng = np.random.default_rng(5)
    n5, n4, leaves = explicit_tree(rng)
    p = wx.Vdb345()
    p.add_nodes(5, [[c >> 12 for c in o] for o, _ in n5], np.stack([m for _, m in n5]))
    p.add_nodes(4, [[c >> 7 for c in o] for o, _ in n4], np.stack([m for _, m in n4]))
    p.add_nodes(3, [[c >> 3 for c in o] for o, _ in leaves], np.stack([m for _, m in leaves]))
    p.compute_sdf(threads=2)
    o = oracle_mod.Vdb()
    for org, m in n5:
        o.add_node(5, org, m)
    for org, m in n4:
        o.add_node(4, org, m)
    for org, m in leaves:
        o.add_node(3, org, m)
    o.compute_sdf()
    # every declared node exists: 3 N5 (one is only a parent of an N4), N4s = declared + leaf parents
    assert tuple(p.count_nodes()) == tuple(o.layout_arrays()["n"])
    assert p.count_nodes()[2] == len(leaves)
    compare(p, o)
    k5, v5, k4, v4, v3 = p.masks()
    assert int(np.unpackbits(v5.view(np.uint8)).sum()) == 47
    assert int(np.unpackbits(v4.view(np.uint8)).sum()) == 41


@pytest.mark.parametrize("level,dim", [(3, 8), (4, 16), (5, 32)])
def test_lone_node_sdf_closed_form(oracle_mod, level, dim):
    """A node with no children or active voxels, alone in the tree: every slot is a
    tile whose neighbours across the node border are background or another level
    (=1), so the two-pass sweep gives 1 + the chessboard distance to the border."""
    o = oracle_mod.Vdb()
    origin = {3: 4096 * 3 + 128 * 5 + 8 * 2, 4: 4096 * 3 + 128 * 5, 5: 4096 * 3}[level]
    o.add_node(level, (origin,) * 3, np.zeros({3: 8, 4: 64, 5: 512}[level], np.uint64))
    o.compute_sdf()
    la = o.layout_arrays()
    c = np.arange(dim)
    edge = np.minimum(c, dim - 1 - c)
    want = 1 + np.minimum(np.minimum(edge[:, None, None], edge[None, :, None]), edge[None, None, :])
    got = la[f"atlas{level}"][0].reshape(dim, dim, dim)  # the only node of its level
    assert np.array_equal(got.astype(np.int64), want)
