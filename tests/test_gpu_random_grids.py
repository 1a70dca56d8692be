"""GPU: seeded random trees -- sphere unions over several root nodes, value-noise
isosurfaces, explicit nodes with active tiles, empty leaves -- traced and shaded
from random cameras and ray sets, bit for bit against the oracle (which computes
its own SDF and layout from the same topology)."""
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
