"""GPU: grids built without a host round trip of the leaf words, and the
grid-layout cache.

* After wx_vdb_finalize_device the leaf words stay on the device and
  wx_vdb_grid_create copies them there: the grid traces exactly like the grid of
  the host-built tree's layout (wx_grid_create), and the tree's host queries,
  which fetch the words on first use, equal the host build's.
* wx_grid_save streams a device grid to a file whose bytes equal wx_layout_save's
  file of the same layout; wx_grid_load gives back a grid that traces identically;
  a damaged leaf section or another key is refused and leaves no grid."""
import os
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu

W, H = 320, 240


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noise_tree(wx, device, side=512, occ=0.15, seed=2):
    v = wx.Vdb345()
    v.add_noise(seed, (0, 0, 0), (side, side, side), 8.0 / side, occ, threads=4, device=device)
    v.compute_sdf(threads=4, device=device)
    return v


def frame(wx, g, side=512):
    cam = wx.camera_basis((-0.6 * side, 1.4 * side, -side), (side / 2, side / 2, side / 2), (0, 1, 0), W / H, 50.0,
                          float(W))
    return g.trace_camera(cam, W, H, 0.5, clip=True).cpu().numpy()


def test_device_words_grid_equals_host_layout_grid(wx, cuda):
    say("device-built tree -> wx_vdb_grid_create vs host-built tree -> wx_grid_create")
    d, h = noise_tree(wx, True), noise_tree(wx, False)
    gd = wx.Grid(d)
    hd = h.desc()
    hd.bound = 4096.0
    gh = wx.Grid.from_desc(hd)
    a, b = frame(wx, gd), frame(wx, gh)
    assert np.array_equal(a, b)
    assert int((a.reshape(-1).view(wx.HIT_DTYPE)["state"] == wx.HIT).sum()) > 5000
    # host queries of the device-built tree fetch the words on first use
    la, lb = d.layout_arrays(), h.layout_arrays()
    for k in ("root", "n5", "n4", "leaves"):
        assert np.array_equal(la[k], lb[k]), k
    rng = np.random.default_rng(3)
    for x, y, z in rng.integers(0, 512, (500, 3)):
        assert d.get_voxel(x, y, z) == h.get_voxel(x, y, z)
    # a second grid from the same tree (the tree keeps its device copy)
    assert np.array_equal(frame(wx, wx.Grid(d)), a)


def test_save_load_round_trip(wx, cuda, tmp_path):
    say("wx_grid_save / wx_layout_save / wx_grid_load")
    v = noise_tree(wx, True, side=896, occ=0.2, seed=7)
    assert v.count_nodes()[2] * 256 > (64 << 20)  # more than one staging round
    g = wx.Grid(v, bound=4096.0)
    ref = frame(wx, g, 896)
    p_dev, p_host = str(tmp_path / "dev.wxgrid"), str(tmp_path / "host.wxgrid")
    key = scenes.layout_key(dict(kind="noise", side=896, occupancy=0.2, seed=7))
    g.save(p_dev, key)
    wx.save_layout(v, p_host, key, bound=4096.0)
    assert open(p_dev, "rb").read() == open(p_host, "rb").read()
    assert wx.probe_layout(p_dev, key) == tuple(v.count_nodes())
    g2 = wx.Grid.load(p_dev, key)
    assert g2.nbytes() == g.nbytes()
    assert np.array_equal(frame(wx, g2, 896), ref)
    # another key: not found; a flipped bit in the leaf words: refused, no grid
    with pytest.raises(wx.WoxelError) as e:
        wx.Grid.load(p_dev, key + "x")
    assert e.value.code == wx.WX_ERR_NOT_FOUND
    raw = bytearray(open(p_dev, "rb").read())
    import struct

    off3, size3 = struct.unpack_from("<Q", raw, 40 + 24)[0], struct.unpack_from("<Q", raw, 72 + 24)[0]
    raw[off3 + size3 - 5] ^= 0x10
    bad = str(tmp_path / "bad.wxgrid")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(wx.WoxelError) as e:
        wx.Grid.load(bad, key)
    assert e.value.code == wx.WX_ERR_INVALID and b"leaf section" in wx.lib.wx_last_error()


def test_cached_grid_builds_then_loads(wx, cuda, tmp_path):
    say("scenes.cached_grid: build + save, then load")
    spec = dict(kind="noise", side=384, occupancy=0.1, seed=12)
    g1, how1 = scenes.cached_grid(spec, str(tmp_path), threads=4)
    g2, how2 = scenes.cached_grid(spec, str(tmp_path), threads=4)
    assert (how1, how2) == ("built", "loaded")
    assert len(os.listdir(tmp_path)) == 1
    assert np.array_equal(frame(wx, g1, 384), frame(wx, g2, 384))
