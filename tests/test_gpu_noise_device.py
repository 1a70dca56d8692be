"""GPU: the device generator of value-noise isosurfaces (wx_vdb_add_noise_device)
gives the host generator's leaves bit for bit, and the 64-bit leaf addressing that
grids of more than 2^24 leaves take traces the same rays as the 32-bit form.

The host generator is the checker here: its voxels are a closed-form test of the
shared noise definition (noise > threshold), and the traversal results are
checked against the oracle elsewhere. On the full-size grid (a 4096^3 isosurface
whose leaf pool exceeds 4 GiB) no oracle run fits, so the check there is a
property: every hit lands on an active voxel of the host tree, and hits reach
leaves past index 2^24."""
import os
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_layout(a, b):
    la, lb = a.layout_arrays(), b.layout_arrays()
    assert tuple(la["n"]) == tuple(lb["n"])
    for k in ("root", "n5", "n4", "leaves"):
        assert np.array_equal(la[k], lb[k]), k


BOXES = [
    ((0, 0, 0), (256, 256, 256), 8.0 / 256, 0.12),
    ((-37, 5, 100), (300, 260, 411), 0.02, 0.05),   # box edges inside leaves, negative keys
    ((0, 0, 0), (512, 512, 512), 8.0 / 512, 0.2),   # full leaves decided from the centre value
    # rows shorter than a classify lane's eight candidates: 7 leaves along z, 3 along y (a lane's
    # candidates wrap into the next row and the next plane), and a single leaf along z
    ((0, 0, 0), (2000, 24, 56), 0.03, 0.3),
    ((5, -800, 16), (405, 0, 24), 0.03, 0.3),
]


@pytest.mark.parametrize("lo,hi,f0,occ", BOXES)
def test_device_noise_matches_host(wx, cuda, lo, hi, f0, occ):
    say(f"device noise {lo} {hi} occupancy {occ}")
    h, d = wx.Vdb345(), wx.Vdb345()
    th = h.add_noise(9, lo, hi, f0, occ, threads=4)
    td = d.add_noise(9, lo, hi, f0, occ, device=True)
    assert np.float32(th).view(np.uint32) == np.float32(td).view(np.uint32)
    h.compute_sdf(threads=4)
    d.compute_sdf(threads=4)
    assert h.count_nodes()[2] > 500
    same_layout(h, d)


def test_device_noise_merges_into_existing_leaves(wx, cuda):
    say("device noise over a sphere set")
    sp = scenes.sphere_set(3, count=60, domain=256, rmin=4, rmax=30)

    def make(device):
        v = wx.Vdb345()
        v.add_spheres(sp, (0, 0, 0), (256, 256, 256), threads=4)
        v.add_noise(5, (64, 0, 0), (320, 256, 256), 8.0 / 256, 0.1, threads=4, device=device)
        v.compute_sdf(threads=4, device=device)
        return v

    same_layout(make(False), make(True))


def test_device_masks_resident_and_dropped(wx, cuda, capfd, monkeypatch):
    """A tree whose leaves all came from the device generator keeps their masks on the
    device, and the device replay reads them there instead of uploading the host copy
    (its phase line says so), with the host build's layout; after a voxel of an existing
    leaf is set, or with leaves from before the generator, the copy is gone, the replay
    uploads the host masks, and the layout follows the change."""
    monkeypatch.setenv("WX_TRACE_BUILD", "1")
    lo, hi, f0, occ = (0, 0, 0), (384, 384, 384), 8.0 / 384, 0.15

    def host(edit=None):
        v = wx.Vdb345()
        v.add_noise(3, lo, hi, f0, occ, threads=4)
        if edit:
            v.set_voxel(*edit)
        v.compute_sdf(threads=4)
        return v

    def device(edit=None, before=None):
        v = wx.Vdb345()
        if before:
            v.set_voxel(*before)
        v.add_noise(3, lo, hi, f0, occ, device=True)
        if edit:
            v.set_voxel(*edit)
        capfd.readouterr()
        v.compute_sdf(threads=4, device=True)
        return v, capfd.readouterr().err

    say("device masks: resident")
    h = host()
    d, err = device()
    assert "masks resident" in err, err
    same_layout(h, d)

    # an inactive voxel inside an existing leaf: setting it changes that leaf's mask
    edit = None
    for o in h.leaf_origins()[:: 97]:
        for k in range(512):
            p = (int(o[0]) + (k >> 6), int(o[1]) + ((k >> 3) & 7), int(o[2]) + (k & 7))
            if h.get_voxel(*p)[0] != 1:  # not an active leaf voxel
                edit = p
                break
        if edit:
            break
    assert edit is not None
    say(f"device masks: dropped after set_voxel{edit}")
    h2 = host(edit)
    d2, err = device(edit)
    assert "masks resident" not in err and "masks up" in err, err
    same_layout(h2, d2)
    assert not np.array_equal(h.layout_arrays()["leaves"], h2.layout_arrays()["leaves"])

    say("device masks: not kept over leaves from before the generator")
    far = (1000, 1000, 1000)
    h3 = wx.Vdb345()
    h3.set_voxel(*far)
    h3.add_noise(3, lo, hi, f0, occ, threads=4)
    h3.compute_sdf(threads=4)
    d3, err = device(before=far)
    assert "masks resident" not in err and "masks up" in err, err
    same_layout(h3, d3)


def test_wide_leaf_addressing_bit_exact(wx, cuda):
    """The same grid with the 64-bit leaf addressing forced: identical hit records
    from every launch shape, alone and batched with other grids (shadow rays
    included), and identical shading."""
    say("64-bit leaf addressing on the config-1 sphere and a noise grid")
    sp = scenes.build_sphere_vdb(scenes.single_sphere())
    nz = wx.Vdb345()
    nz.add_noise(2, (0, 0, 0), (512, 512, 512), 8.0 / 512, 0.15, device=True)
    nz.compute_sdf(threads=4, device=True)
    narrow = [wx.Grid(sp), wx.Grid(nz)]
    os.environ["WX_WIDE_LEAVES"] = "1"
    try:
        wide = [wx.Grid(sp), wx.Grid(nz)]
    finally:
        del os.environ["WX_WIDE_LEAVES"]
    cams = [scenes.basis(scenes.config1_camera()),
            wx.camera_basis((-300.0, 700.0, -500.0), (256.0, 256.0, 256.0), (0, 1, 0), 4 / 3, 50.0, 320.0)]
    for gn, gw, cam in zip(narrow, wide, cams):
        for simple in (False, True):
            a = gn.trace_camera(cam, 320, 240, 0.5, clip=True, simple=simple).cpu().numpy()
            b = gw.trace_camera(cam, 320, 240, 0.5, clip=True, simple=simple).cpu().numpy()
            assert np.array_equal(a, b)
            assert int((a.reshape(-1).view(wx.HIT_DTYPE)["state"] == wx.HIT).sum()) > 1000
        for shape in (2, 3):  # compact and LDS shapes on the wide grid
            o = cuda.empty((320 * 240, 32), dtype=cuda.uint8, device="cuda")
            wx.trace_batch([gw], [wx.camera_request(0, cam, 320, 240, o, 0.5, True)],
                           tune=wx.Tune(persistent=shape, refill=8, chunk=256, waves_per_cu=16, xcd_split=1))
            assert np.array_equal(o.cpu().numpy(), a), shape
        ra = gn.render_camera(cam, 320, 240, wx.shade(4), 0.5, clip=True).cpu().numpy()
        rb = gw.render_camera(cam, 320, 240, wx.shade(4), 0.5, clip=True).cpu().numpy()
        assert np.array_equal(ra, rb)
    # batched launches over a narrow and a wide grid at once, every launch shape, shadow rays on
    light = (8000.0, 8000.0, 8000.0)
    ref = []
    for gn, cam in zip(narrow, cams):
        o = cuda.empty((320 * 240, 32), dtype=cuda.uint8, device="cuda")
        s = cuda.empty_like(o)
        wx.trace_batch([gn], [wx.camera_request(0, cam, 320, 240, o, 0.5, True, shadow_out=s, light=light)],
                       tune=wx.Tune(persistent=0, refill=16, chunk=256, waves_per_cu=16, xcd_split=1))
        ref.append((o, s))
    for persistent in (0, 1):
        outs = [(cuda.empty_like(o), cuda.empty_like(s)) for o, s in ref]
        reqs = [wx.camera_request(k, cam, 320, 240, o, 0.5, True, shadow_out=s, light=light)
                for k, (cam, (o, s)) in enumerate(zip(cams, outs))]
        wx.trace_batch([narrow[0], wide[1]], reqs,
                       tune=wx.Tune(persistent=persistent, refill=16, chunk=128, waves_per_cu=16, xcd_split=1))
        outs2 = [(cuda.empty_like(o), cuda.empty_like(s)) for o, s in ref]
        reqs2 = [wx.camera_request(1 - k, cam, 320, 240, o, 0.5, True, shadow_out=s, light=light)
                 for k, (cam, (o, s)) in enumerate(zip(cams, outs2))]
        wx.trace_batch([wide[1], wide[0]], reqs2,
                       tune=wx.Tune(persistent=persistent, refill=16, chunk=128, waves_per_cu=16, xcd_split=1))
        cuda.cuda.synchronize()
        for (a, b), (c, d), (e, f) in zip(ref, outs, outs2):
            assert cuda.equal(a, c) and cuda.equal(b, d), persistent
            assert cuda.equal(a, e) and cuda.equal(b, f), persistent


@pytest.mark.timeout(600)
def test_leaf_pool_beyond_4gib(wx, cuda):
    """A 4096^3 isosurface at 20 % occupancy: more than 2^24 leaves, built on the
    device; hits fall on active voxels, including leaves past index 2^24."""
    say("4096^3 noise isosurface, 20 % occupancy: device build")
    v = wx.Vdb345()
    v.add_noise(4, (0, 0, 0), (4096, 4096, 4096), 8.0 / 4096, 0.2, device=True)
    v.compute_sdf(threads=16, device=True)
    n5, n4, n3 = v.count_nodes()
    say(f"nodes {n5} {n4} {n3}")
    assert n3 > (1 << 24)
    g = wx.Grid(v)
    W, H = 320, 240
    cam = wx.camera_basis((7500.0, 5000.0, 6000.0), (2048.0, 2048.0, 2048.0), (0, 1, 0), W / H, 40.0, float(W))
    hits = wx.hits_to_numpy(g.trace_camera(cam, W, H, 0.5, clip=True))
    hit = hits["state"] == wx.HIT
    assert hit.sum() > 20000
    vox = hits["voxel"][hit]
    for x, y, z in vox[:: max(1, len(vox) // 4000)]:
        kind, _, _ = v.get_voxel(int(x), int(y), int(z))
        assert kind == 1, (x, y, z, kind)  # EP_LEAF: an active voxel
    # DFS position of each hit's leaf: one root node, so the order is (off4, off3)
    org = v.leaf_origins()
    key = lambda o: ((o[:, 0] >> 7) << 10 | (o[:, 1] >> 7) << 5 | (o[:, 2] >> 7)) << 12 | \
        (((o[:, 0] >> 3) & 15) << 8 | ((o[:, 1] >> 3) & 15) << 4 | ((o[:, 2] >> 3) & 15))
    keys = key(org.astype(np.int64))
    assert np.all(np.diff(keys) > 0)
    idx = np.searchsorted(keys, key((vox.astype(np.int64) >> 3) << 3))
    say(f"hit leaves: index range {idx.min()}..{idx.max()}, past 2^24: {(idx >= (1 << 24)).sum()}")
    assert (idx >= (1 << 24)).sum() > 1000
