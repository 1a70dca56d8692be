"""GPU: traversal of grids read from .vdb files (the reference's level-set
assets) and of trees with active tiles, nodes without children and empty
leaves, bit for bit against the oracle."""
import sys

import numpy as np
import pytest

from test_vdb_file import explicit_tree, fixture

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("name,eye", [("cube", (260.0, 180.0, -300.0)), ("icosahedron", (-150.0, 90.0, 170.0))])
def test_vdb_file_grid_camera_bit_exact(wx, cuda, oracle_mod, name, eye):
    from oracle import vdb_file

    say(f"{name}.vdb camera from {eye}")
    p = wx.Vdb345()
    p.read_file(fixture(name))
    p.compute_sdf(threads=4)
    g = wx.Grid(p)
    o = oracle_mod.Vdb.from_file_topology(vdb_file.read_grid(fixture(name)))
    o.compute_sdf()
    W, H = 320, 240
    cam = wx.camera_basis(eye, (4.0, 4.0, 4.0), (0, 1, 0), W / H, 50.0, float(W))
    got = wx.hits_to_numpy(g.trace_camera(cam, W, H, 0.5))
    want, _ = o.trace_camera(cam.as_array(), W, H, 0.5)
    assert same(got, want)
    assert int((got["state"] == wx.HIT).sum()) > 5000


def test_active_tiles_and_empty_leaves_bit_exact(wx, cuda, oracle_mod):
    say("explicit topology with active tiles")
    rng = np.random.default_rng(5)
    n5, n4, leaves = explicit_tree(rng)
    p = wx.Vdb345()
    p.add_nodes(5, [[c >> 12 for c in o] for o, _ in n5], np.stack([m for _, m in n5]))
    p.add_nodes(4, [[c >> 7 for c in o] for o, _ in n4], np.stack([m for _, m in n4]))
    p.add_nodes(3, [[c >> 3 for c in o] for o, _ in leaves], np.stack([m for _, m in leaves]))
    p.compute_sdf(threads=2)
    g = wx.Grid(p)
    o = oracle_mod.Vdb()
    for org, m in n5:
        o.add_node(5, org, m)
    for org, m in n4:
        o.add_node(4, org, m)
    for org, m in leaves:
        o.add_node(3, org, m)
    o.compute_sdf()

    def bit(mask, b):
        return (int(mask[b >> 6]) >> (b & 63)) & 1

    act5 = {tuple(c >> 12 for c in org): m for org, m in n5}
    act4 = {tuple(c >> 7 for c in org): m for org, m in n4}

    def tile_level(v):  # 5 / 4 when the voxel lies in an active tile of that level
        m = act5.get(tuple(int(c) >> 12 for c in v))
        if m is not None and bit(m, int(((v[0] >> 7) & 31) << 10 | ((v[1] >> 7) & 31) << 5 | ((v[2] >> 7) & 31))):
            return 5
        m = act4.get(tuple(int(c) >> 7 for c in v))
        if m is not None and bit(m, int(((v[0] >> 3) & 15) << 8 | ((v[1] >> 3) & 15) << 4 | ((v[2] >> 3) & 15))):
            return 4
        return 3

    W, H = 256, 192
    tiles_hit = {5: 0, 4: 0, 3: 0}
    for eye, target in (((-2500.0, 1900.0, -3100.0), (2048.0, 2048.0, 2048.0)),
                        ((300.0, 200.0, -200.0), (200.0, 200.0, 100.0)),
                        ((4000.0, 300.0, 3000.0), (4700.0, 150.0, 10.0))):
        cam = wx.camera_basis(eye, target, (0, 1, 0), W / H, 60.0, float(W))
        got = wx.hits_to_numpy(g.trace_camera(cam, W, H, 0.5))
        want, _ = o.trace_camera(cam.as_array(), W, H, 0.5)
        assert same(got, want)
        for v in got["voxel"][got["state"] == wx.HIT]:
            tiles_hit[tile_level(v)] += 1
    say(f"hits in active node-5 tiles / node-4 tiles / leaves: {tiles_hit}")
    assert tiles_hit[5] > 0 and tiles_hit[4] > 0
