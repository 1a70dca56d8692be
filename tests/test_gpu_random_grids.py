"""GPU: seeded random trees -- sphere unions over several root nodes, value-noise
isosurfaces, explicit nodes with active tiles, empty leaves -- traced and shaded
from random cameras and ray sets, bit for bit against the oracle (which computes
its own SDF and layout from the same topology); and random trees built on the
device (noise voxels and SDF) against the same trees built on the host."""
import sys

import numpy as np
import pytest
from random_trees import random_tree

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.mark.parametrize("seed", range(24))
def test_random_tree_bit_exact(wx, oracle_mod, cuda, seed):
    rng = np.random.default_rng(1000 + seed)
    p, o, centre = random_tree(wx, oracle_mod, rng)
    say(f"seed {seed}: nodes {p.count_nodes()}")
    g = wx.Grid(p)
    lay = o.layout()
    for k in range(3):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        eye = centre + v * rng.uniform(50, 2500)
        w, h = int(rng.integers(24, 160)), int(rng.integers(16, 120))
        clip = bool(np.any(np.abs(eye) > 4096.0) or rng.uniform() < 0.5)
        cam = wx.camera_basis(eye, centre + rng.uniform(-200, 200, size=3), (0, 1, 0), w / h,
                              float(rng.uniform(20, 90)), float(w))
        off = float(rng.choice([0.5, 0.001]))
        got = wx.hits_to_numpy(g.trace_camera(cam, w, h, off, clip=clip))
        want, _ = lay.trace_camera(cam.as_array(), w, h, off, clip=clip)
        assert same(got, want), (seed, k)
        shade = wx.shade(int(rng.integers(0, 5)), tuple(int(x) for x in rng.integers(0, 2, size=3)),
                         sun_dir=tuple(float(x) for x in rng.standard_normal(3)), intensity=float(rng.uniform(0.3, 1.2)))
        gf = g.render_camera(cam, w, h, shade, off, clip=clip, f32=True).cpu().numpy()
        rf, _ = lay.render_camera(cam.as_array(), w, h, shade, off, clip=clip)
        assert same(gf, rf), (seed, k, shade["mode"])
    o3 = centre + rng.uniform(-600, 600, size=(3000, 3))
    d = rng.standard_normal((3000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o3, d], 1).astype(np.float32)
    for clip in (False, True):
        got = wx.hits_to_numpy(g.trace_rays(cuda.from_numpy(rays).cuda(), clip=clip))
        want, _ = lay.trace_rays(rays, clip=clip)
        assert same(got, want), (seed, clip)


def build_both_ways(wx, rng_seed, device):
    """A random tree (spheres over one or several roots, a value-noise box of any
    shape, maybe an explicit node-4 with active tiles) built with the noise voxels
    and the SDF on the host or on the device."""
    rng = np.random.default_rng(rng_seed)
    p = wx.Vdb345()
    centre = rng.uniform(-3000, 3000, size=3)
    n = int(rng.integers(1, 7))
    sp = np.column_stack([centre + rng.uniform(-400, 400, size=(n, 3)), rng.uniform(2, 60, size=n)])
    p.add_spheres(sp, threads=2)
    if rng.uniform() < 0.8:
        lo = (centre + rng.uniform(-300, 100, size=3)).astype(np.int32)
        hi = lo + rng.integers(8, 300, size=3).astype(np.int32)
        p.add_noise(int(rng.integers(1 << 30)), lo, hi, float(rng.uniform(0.01, 0.05)), float(rng.uniform(0.05, 0.4)),
                    threads=2, device=device)
    extra = None
    if rng.uniform() < 0.5:
        k4 = (np.floor(centre / 128).astype(np.int64) + rng.integers(-3, 4, size=3)).astype(np.int32)
        m = np.zeros(64, np.uint64)
        for b in rng.choice(4096, size=int(rng.integers(1, 40)), replace=False):
            m[b >> 6] |= np.uint64(1) << np.uint64(b & 63)
        p.add_nodes(4, k4[None, :], m[None, :])
        extra = (k4, m)
    p.compute_sdf(threads=2, device=device)
    return p, centre, rng, extra


@pytest.mark.parametrize("seed", range(16))
def test_random_tree_device_build(wx, oracle_mod, cuda, seed):
    """The same random tree built on the host and on the device (noise voxels, SDF,
    kept leaf words): identical layouts, and the device-built grid traces a random
    camera bit for bit like the oracle."""
    ph, centre, _, extra = build_both_ways(wx, 5000 + seed, False)
    pd, _, rng, _ = build_both_ways(wx, 5000 + seed, True)
    say(f"device build seed {seed}: nodes {ph.count_nodes()}")
    la, lb = ph.layout_arrays(), pd.layout_arrays()
    assert tuple(la["n"]) == tuple(lb["n"])
    for k in ("root", "n5", "n4", "leaves"):
        assert np.array_equal(la[k], lb[k]), (seed, k)
    o = oracle_mod.Vdb()
    org, masks = ph.leaf_masks()
    for i in range(org.shape[0]):
        o.add_node(3, org[i], masks[i])
    if extra is not None:
        o.add_node(4, (extra[0].astype(np.int64) << 7).astype(np.int32), extra[1])
    o.compute_sdf()
    lay = o.layout()
    g = wx.Grid(pd)
    v = rng.standard_normal(3)
    v /= np.linalg.norm(v)
    eye = centre + v * rng.uniform(50, 2500)
    w, h = int(rng.integers(24, 160)), int(rng.integers(16, 120))
    cam = wx.camera_basis(eye, centre + rng.uniform(-200, 200, size=3), (0, 1, 0), w / h, 60.0, float(w))
    got = wx.hits_to_numpy(g.trace_camera(cam, w, h, 0.5, clip=True))
    want, _ = lay.trace_camera(cam.as_array(), w, h, 0.5, clip=True)
    assert same(got, want), seed
