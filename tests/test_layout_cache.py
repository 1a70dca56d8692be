"""Grid-layout cache files on the host (no GPU): the file format against an
independent reading of it (header fields, section offsets, pool bytes equal to the
tree's layout, section digests recomputed with the xxhash package's XXH64), and
wx_layout_probe's verdicts on other keys, damaged headers, truncated files and
missing files."""
import os
import struct

import numpy as np
import pytest

xxhash = pytest.importorskip("xxhash")

HEADER = 4096
CHUNK = 16 << 20


def _tree(wx):
    v = wx.Vdb345()
    rng = np.random.default_rng(11)
    # a domain across the origin: several root entries, negative node keys
    sp = np.concatenate([rng.uniform(-200, 500, (40, 3)), rng.uniform(3, 30, (40, 1))], axis=1)
    v.add_spheres(sp, (-256, -256, -256), (600, 600, 600), threads=4)
    v.compute_sdf(threads=4)
    return v


def _section_digest(data: bytes) -> int:
    pieces = [xxhash.xxh64(data[i:i + CHUNK], seed=0).intdigest() for i in range(0, len(data), CHUNK)]
    return xxhash.xxh64(struct.pack(f"<{len(pieces)}Q", *pieces), seed=0).intdigest()


@pytest.fixture(scope="module")
def saved(wx, tmp_path_factory):
    v = _tree(wx)
    path = str(tmp_path_factory.mktemp("layout") / "grid.wxgrid")
    key = "spheres seed 11, domain [-256, 600)"
    wx.save_layout(v, path, key, bound=2048.0)
    return v, path, key


def test_xxh64_reference_values():
    # the package's XXH64 is the published one: its value for the empty input and
    # for a short string are the algorithm's documented test values
    assert xxhash.xxh64(b"", seed=0).intdigest() == 0xEF46DB3751D8E999
    assert xxhash.xxh64(b"a", seed=0).intdigest() == 0xD24EC4F1A98C6E5B


def test_file_layout_and_digests(wx, saved):
    v, path, key = saved
    raw = open(path, "rb").read()
    assert raw[:8] == b"WXLAYOUT"
    fmt, hdr, n5, n4, n3, _, klen, chunk = struct.unpack_from("<8I", raw, 8)
    assert (fmt, hdr, chunk) == (1, HEADER, CHUNK)
    assert (n5, n4, n3) == tuple(v.count_nodes()) and n5 > 1
    assert struct.unpack_from("<f", raw, 28)[0] == 2048.0
    assert raw[256:256 + klen].decode() == key
    assert struct.unpack_from("<Q", raw, HEADER - 8)[0] == xxhash.xxh64(raw[:HEADER - 8], seed=0).intdigest()
    off = struct.unpack_from("<4Q", raw, 40)
    size = struct.unpack_from("<4Q", raw, 72)
    dig = struct.unpack_from("<4Q", raw, 104)
    lay = v.layout_arrays()
    pools = [lay["root"], lay["n5"], lay["n4"], lay["leaves"]]
    expect_off = HEADER
    for k, pool in enumerate(pools):
        data = np.ascontiguousarray(pool).tobytes()
        assert off[k] == expect_off and off[k] % 4096 == 0
        assert size[k] == len(data)
        assert raw[off[k]:off[k] + size[k]] == data, k
        assert dig[k] == _section_digest(data), k
        expect_off = (off[k] + size[k] + 4095) // 4096 * 4096
    assert len(raw) == expect_off


def test_probe_verdicts(wx, saved, tmp_path):
    v, path, key = saved
    assert wx.probe_layout(path, key) == tuple(v.count_nodes())
    assert wx.probe_layout(path, key + " ") is None  # another key
    assert wx.probe_layout(str(tmp_path / "absent.wxgrid"), key) is None
    raw = bytearray(open(path, "rb").read())
    damaged = tmp_path / "damaged.wxgrid"
    bad = bytearray(raw)
    bad[20] ^= 1  # n4
    damaged.write_bytes(bytes(bad))
    with pytest.raises(wx.WoxelError) as e:
        wx.probe_layout(str(damaged), key)
    assert e.value.code == wx.WX_ERR_INVALID and b"damaged header" in wx.lib.wx_last_error()
    short = tmp_path / "short.wxgrid"
    short.write_bytes(bytes(raw[:-4096]))
    with pytest.raises(wx.WoxelError) as e:
        wx.probe_layout(str(short), key)
    assert b"truncated" in wx.lib.wx_last_error()
    other = tmp_path / "other.bin"
    other.write_bytes(b"\0" * 8192)
    with pytest.raises(wx.WoxelError):
        wx.probe_layout(str(other), key)


def test_save_is_atomic_and_rejects_bad_keys(wx, saved, tmp_path):
    v, _, _ = saved
    p = str(tmp_path / "g.wxgrid")
    with pytest.raises(wx.WoxelError):
        wx.save_layout(v, p, "k" * 2049)
    assert not os.path.exists(p) and not os.path.exists(p + ".tmp")
    with pytest.raises(wx.WoxelError):  # directory does not exist: nothing is left behind
        wx.save_layout(v, str(tmp_path / "no" / "g.wxgrid"), "k")
    wx.save_layout(v, p, "k")
    assert os.path.exists(p) and not os.path.exists(p + ".tmp")
    assert wx.probe_layout(p, "k") == tuple(v.count_nodes())


def test_cache_key_names_spec_and_seed(wx):
    from woxel_amd import scenes

    a = dict(kind="noise", side=512, occupancy=0.1, seed=3)
    assert scenes.layout_key(a) == scenes.layout_key(dict(reversed(list(a.items()))))
    assert scenes.layout_key(a) != scenes.layout_key(dict(a, seed=4))


def test_multi_piece_sections(wx, tmp_path):
    """A leaf section of several 16 MiB pieces and two 64 MiB staging rounds: the
    digest is the XXH64 of the pieces' digests, in order."""
    v = wx.Vdb345()
    v.add_noise(5, (0, 0, 0), (800, 800, 800), 8.0 / 800, 0.2, threads=8)
    v.compute_sdf(threads=8)
    n3 = v.count_nodes()[2]
    assert n3 * 256 > (64 << 20)
    p = str(tmp_path / "noise.wxgrid")
    wx.save_layout(v, p, "noise 800")
    raw = open(p, "rb").read()
    off, size, dig = (struct.unpack_from("<4Q", raw, o) for o in (40, 72, 104))
    leaves = np.ascontiguousarray(v.layout_arrays()["leaves"]).tobytes()
    assert size[3] == len(leaves) and raw[off[3]:off[3] + size[3]] == leaves
    assert dig[3] == _section_digest(leaves)


def test_leaf_section_ending_on_a_4k_boundary(wx, tmp_path):
    """16 leaves: the leaf section (4096 bytes) ends exactly on a 4 KiB boundary, so
    the file ends with it and no padding may touch it."""
    v = wx.Vdb345()
    keys = np.array([[2 * i, 0, 0] for i in range(16)], np.int32)
    masks = np.full((16, 8), 0x8000000000000001, np.uint64)
    v.add_leaf_masks(keys, masks)
    v.compute_sdf(threads=2)
    assert v.count_nodes()[2] == 16
    p = str(tmp_path / "aligned.wxgrid")
    wx.save_layout(v, p, "aligned")
    raw = open(p, "rb").read()
    off, size, dig = (struct.unpack_from("<4Q", raw, o) for o in (40, 72, 104))
    leaves = np.ascontiguousarray(v.layout_arrays()["leaves"]).tobytes()
    assert size[3] == 4096 and off[3] + size[3] == len(raw)
    assert raw[off[3]:] == leaves and dig[3] == _section_digest(leaves)
