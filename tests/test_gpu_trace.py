"""GPU parity: the HIP traversal against the CPU oracle, bit for bit.

Every field of the 32-byte hit record (state, voxel, t, normal, per-level step
counts, lookup counts) must be bit-identical to the oracle's. The oracle
traverses its own reference-format layout (masks + atlases); the GPU traverses
the product's packed device layout.
"""
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_hits_equal(gpu, ora, what=""):
    g = gpu.view(np.uint8).reshape(-1, 32)
    o = ora.view(np.uint8).reshape(-1, 32)
    bad = np.nonzero((g != o).any(axis=1))[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size}/{g.shape[0]} records differ; first #{i}: gpu={gpu[i]} oracle={ora[i]}")


@pytest.fixture(scope="module")
def cfg1(wx, oracle_mod, cuda):
    sp = scenes.single_sphere()
    p = scenes.build_sphere_vdb(sp)
    o = oracle_mod.Vdb()
    o.add_sphere(sp[0, :3], sp[0, 3])
    o.compute_sdf()
    return p, wx.Grid(p), o


@pytest.mark.parametrize("offset", [0.5, 0.001])
def test_config1_camera_bit_exact(wx, oracle_mod, cuda, cfg1, offset):
    say(f"config1 camera 640x480 offset={offset}")
    p, g, o = cfg1
    spec = scenes.config1_camera()
    cam = scenes.basis(spec)
    out = g.trace_camera(cam, 640, 480, offset)
    cuda.cuda.synchronize()
    gpu = wx.hits_to_numpy(out)
    ora, diag = o.trace_camera(cam.as_array(), 640, 480, offset)
    assert_hits_equal(gpu, ora, f"config1 offset {offset}")
    st = np.bincount(gpu["state"], minlength=3)
    assert st[0] > 20000 and st[2] == 0


@pytest.mark.parametrize("w,h", [(1, 1), (37, 29), (8, 4), (65, 3)])
def test_odd_image_sizes(wx, oracle_mod, cuda, cfg1, w, h):
    say(f"odd size {w}x{h}")
    p, g, o = cfg1
    cam = wx.camera_basis((400, 300, -200), (128, 128, 128), (0, 1, 0), w / h, 60.0, float(w))
    gpu = wx.hits_to_numpy(g.trace_camera(cam, w, h, 0.5))
    ora, _ = o.trace_camera(cam.as_array(), w, h, 0.5)
    assert_hits_equal(gpu, ora, f"{w}x{h}")


def random_rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.concatenate([o, d], 1).astype(np.float32)
    return r


def test_trace_rays_random(wx, oracle_mod, cuda, cfg1):
    say("trace_rays random 100k")
    p, g, o = cfg1
    rng = np.random.default_rng(3)
    rays = random_rays(rng, 100_000, -60.0, 320.0)
    # axis-aligned and exactly-diagonal directions exercise tMax ties and +-inf
    rays[:1000, 3:] = 0.0
    rays[np.arange(1000), 3 + np.arange(1000) % 3] = np.where(np.arange(1000) % 2, 1.0, -1.0)
    rays[1000:2000, 3:] = np.float32(1 / np.sqrt(3))
    rays[1000:2000, :3] = np.floor(rays[1000:2000, :3]) + 0.5  # voxel centres: every step is an exact tie
    t = cuda.from_numpy(rays).cuda()
    gpu = wx.hits_to_numpy(g.trace_rays(t))
    ora, diag = o.trace_rays(rays)
    assert_hits_equal(gpu, ora, "random rays")
    assert (diag["flags"] & oracle_mod.F_TIE).any()


def test_entry_clip_rays(wx, oracle_mod, cuda, cfg1):
    say("entry clip rays")
    p, g, o = cfg1
    rng = np.random.default_rng(4)
    n = 20_000
    eye = rng.standard_normal((n, 3))
    eye = 128 + 9000 * eye / np.linalg.norm(eye, axis=1, keepdims=True)
    tgt = 128 + rng.uniform(-120, 120, size=(n, 3))
    d = tgt - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([eye, d], 1).astype(np.float32)
    rays[:500, 3:] *= -1  # pointing away: the clip misses
    t = cuda.from_numpy(rays).cuda()
    gpu = wx.hits_to_numpy(g.trace_rays(t, clip=True))
    ora, diag = o.trace_rays(rays, clip=True)
    assert_hits_equal(gpu, ora, "clipped rays")
    assert (gpu["state"] == 0).sum() > 1000
    assert (diag["flags"][:500] & oracle_mod.F_CLIPMISS).all()


def test_deterministic_repeat(wx, cuda, cfg1):
    say("determinism")
    p, g, o = cfg1
    cam = scenes.basis(scenes.config1_camera())
    a = g.trace_camera(cam, 640, 480, 0.5).clone()
    b = g.trace_camera(cam, 640, 480, 0.5)
    assert cuda.equal(a, b)


def test_empty_grid_all_oob(wx, oracle_mod, cuda):
    say("empty grid")
    v = wx.Vdb345()
    v.compute_sdf()
    g = wx.Grid(v)
    cam = scenes.basis(scenes.config1_camera())
    gpu = wx.hits_to_numpy(g.trace_camera(cam, 64, 48, 0.5))
    assert (gpu["state"] == wx.OOB).all()
    assert (gpu["lookups"] == 0).all()


def test_multi_root_negative_grid(wx, oracle_mod, cuda):
    say("multi-root negative grid")
    sp = np.array([[0.5, -3.0, 2.0, 22.0], [-4096.0, 10.0, -9.0, 17.5], [4100.0, 4090.0, -4090.0, 12.0],
                   [-20.0, 4096.0, 0.0, 9.0], [-30.0, -30.0, -30.0, 25.0]])
    p = scenes.build_sphere_vdb(sp)
    o = oracle_mod.Vdb()
    for s in sp:
        o.add_sphere(s[:3], s[3])
    o.compute_sdf()
    g = wx.Grid(p)
    rng = np.random.default_rng(8)
    n = 50_000
    centres = sp[rng.integers(0, len(sp), n), :3]
    eye = centres + rng.uniform(-200, 200, size=(n, 3))
    d = centres + rng.uniform(-15, 15, size=(n, 3)) - eye
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([eye, d], 1).astype(np.float32)
    gpu = wx.hits_to_numpy(g.trace_rays(cuda.from_numpy(rays).cuda(), clip=True))
    ora, _ = o.trace_rays(rays, clip=True)
    assert_hits_equal(gpu, ora, "multi-root")
    assert (gpu["state"] == 0).sum() > n // 4


def test_many_roots_global_table(wx, oracle_mod, cuda):
    """27 node-5 roots (more than the 16 a workgroup stages in LDS, so the root
    table is scanned in global memory) and 2 roots (staged in LDS): long rays that
    cross root cells, cameras and shadow rays, bit for bit against the oracle."""
    lat = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], np.float64)
    for sp in (np.column_stack([lat * 4096.0 + 1900.0, np.full(len(lat), 30.0)]),
               np.array([[100.0, 100.0, 100.0, 40.0], [-4000.0, 90.0, 60.0, 35.0]])):
        say(f"{len(sp)} spheres over several roots")
        p = scenes.build_sphere_vdb(sp)
        o = oracle_mod.Vdb()
        for s in sp:
            o.add_sphere(s[:3], s[3])
        o.compute_sdf()
        assert p.count_nodes()[0] == len(sp)
        g = wx.Grid(p)
        rng = np.random.default_rng(len(sp))
        n = 40_000
        tgt = sp[rng.integers(0, len(sp), n), :3] + rng.uniform(-20, 20, size=(n, 3))
        eye = rng.uniform(-7000, 7000, size=(n, 3))
        d = tgt - eye
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.concatenate([eye, d], 1).astype(np.float32)
        gpu = wx.hits_to_numpy(g.trace_rays(cuda.from_numpy(rays).cuda(), clip=True, simple=True))
        ora, _ = o.trace_rays(rays, clip=True)
        assert_hits_equal(gpu, ora, f"{len(sp)} roots")
        assert (gpu["state"] == 0).sum() > n // 10
        cam = wx.camera_basis((-5000.0, 3000.0, -6000.0), tuple(sp[0, :3]), (0, 1, 0), 4 / 3, 70.0, 160.0)
        gpu = wx.hits_to_numpy(g.trace_camera(cam, 160, 120, 0.5, clip=True, simple=True))
        ora, _ = o.trace_camera(cam.as_array(), 160, 120, 0.5, clip=True)
        assert_hits_equal(gpu, ora, f"{len(sp)} roots camera")


@pytest.fixture(scope="module")
def cfg2(wx, oracle_mod, cuda):
    say("building the 4096^3 sphere set (2000 spheres)")
    sp = scenes.sphere_set(1)
    lo, hi = scenes.clip_box()
    p = scenes.build_sphere_vdb(sp, lo, hi)
    g = wx.Grid(p)
    k5, v5, k4, v4, v3 = p.masks()
    a5, a4, a3 = p.atlas()
    lay = oracle_mod.Layout.from_arrays(p.origins(), k5, v5, k4, v4, v3, a5, a4, a3)
    del a3
    return p, g, lay


@pytest.mark.parametrize("frame", [0, 21])
def test_config2_orbit_frame_bit_exact(wx, oracle_mod, cuda, cfg2, frame):
    say(f"config2 orbit frame {frame}")
    p, g, lay = cfg2
    spec = scenes.orbit_camera(frame)
    cam = scenes.basis(spec)
    W, H = spec["width"], spec["height"]
    gpu = wx.hits_to_numpy(g.trace_camera(cam, W, H, 0.5, clip=True))
    ora, _ = lay.trace_camera(cam.as_array(), W, H, 0.5, clip=True)
    assert_hits_equal(gpu, ora, f"config2 frame {frame}")
    assert (gpu["state"] == 0).mean() > 0.05


@pytest.mark.parametrize("persistent", [0, 1])
def test_frames_in_flight_two_streams(wx, cuda, cfg2, persistent):
    """Orbit frames alternated over two streams (frame k on stream k % 2, launched
    without waiting, so two frames run at once) give the records of the same frames
    traced one after another on one stream, in the one-lane shape and in the
    persistent shape, whose queue counters each stream holds separately."""
    say(f"frames in flight on two streams, persistent {persistent}")
    p, g, lay = cfg2
    old = wx.tune_get()
    t = wx.tune_get()
    t.persistent = persistent
    wx.tune_set(t)
    try:
        specs = [scenes.orbit_camera(k) for k in range(6)]
        W, H = specs[0]["width"], specs[0]["height"]
        cams = [scenes.basis(s) for s in specs]
        ref = [g.trace_camera(c, W, H, 0.5, clip=True).clone() for c in cams]
        cuda.cuda.synchronize()
        streams = [cuda.cuda.Stream() for _ in range(2)]
        outs = [cuda.empty_like(r) for r in ref]
        for k, c in enumerate(cams):
            g.trace_camera(c, W, H, 0.5, clip=True, out=outs[k], stream=streams[k % 2])
        cuda.cuda.synchronize()
        for k in range(len(cams)):
            assert cuda.equal(outs[k], ref[k]), f"frame {k}"
    finally:
        wx.tune_set(old)


def test_config2_incoherent_sample(wx, oracle_mod, cuda, cfg2):
    say("config2 incoherent 200k")
    p, g, lay = cfg2
    rays = scenes.incoherent_rays(200_000, seed=7)
    gpu = wx.hits_to_numpy(g.trace_rays(cuda.from_numpy(rays).cuda(), clip=True))
    ora, _ = lay.trace_rays(rays, clip=True)
    assert_hits_equal(gpu, ora, "incoherent")
    assert (gpu["state"] == 0).sum() > 2000


def wall_rays():
    """Rays beside a wall of voxels that run into the 1000-step cap, and degenerate
    rays: a zero direction, -0.0 components (idir = -inf, p turns NaN), NaN and
    infinite origins, an origin beyond the bound."""
    r = [(0.5, y, 4.5, 1.0, 0.0, 0.0) for y in (1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5)]
    r += [(0.5, 1.5, 4.5, 0.99999994, 1e-4, 0.0), (4095.5, 2.5, 3.5, -1.0, 0.0, 0.0),
          (0.5, 1.5, 4.5, 0.0, 0.0, 0.0), (0.5, 1.5, 4.5, 1.0, -0.0, 0.0), (10.5, 3.5, 2.5, -0.0, -0.0, 1.0),
          (float("nan"), 1.5, 4.5, 1.0, 0.0, 0.0), (0.5, 1.5, 4.5, float("nan"), 0.0, 0.0),
          (float("inf"), 1.5, 4.5, -1.0, 0.0, 0.0), (0.5, -float("inf"), 4.5, 0.0, 1.0, 0.0),
          (5000.0, 1.5, 4.5, -1.0, 0.0, 0.0), (-4096.5, 1.5, 4.5, 1.0, 0.0, 0.0)]
    return np.array(r, np.float32)


def test_step_cap_and_degenerate_rays(wx, oracle_mod, cuda):
    """HDDA_MAX_RAY_STEPS and degenerate inputs, bit for bit (with and without the
    entry clip)."""
    say("step cap and degenerate rays")
    org = np.array([(x, 0, 0) for x in range(0, 4096, 8)], np.int32)
    masks = np.full((org.shape[0], 8), np.uint64(0xFF))
    p = wx.Vdb345()
    p.add_leaf_masks(org >> 3, masks)
    p.compute_sdf(threads=2)
    g = wx.Grid(p)
    o = oracle_mod.Vdb()
    o.add_leaf_masks(org, masks)
    o.compute_sdf()
    rays = wall_rays()
    t = cuda.from_numpy(rays).cuda()
    for clip in (False, True):
        for simple in (False, True):
            gpu = wx.hits_to_numpy(g.trace_rays(t, clip=clip, simple=simple))
            ora, _ = o.trace_rays(rays, clip=clip)
            assert_hits_equal(gpu, ora, f"wall rays clip={clip} simple={simple}")
    assert (gpu["state"][:4] == wx.MAXSTEPS).all()


def clean_extreme_rays(rng):
    """Waves of 64 rays whose direction, reciprocal and origin are all finite (the
    kernels' NaN-free step) but extreme, and one wave with a single degenerate lane."""
    groups = []
    # beside the wall, nearly axis-aligned: 1000-step cap reached by the wave-counted iterations
    y = 1.5 + np.arange(64) % 7
    groups.append(np.stack([np.full(64, 0.5), y, np.full(64, 4.5), np.ones(64),
                            rng.choice([-1, 1], 64) * 1e-7 * (1 + np.arange(64) % 5),
                            rng.choice([-1, 1], 64) * 1e-7], 1))
    # tiny components: reciprocals near the f32 limit, t overflowing to +-inf
    o = rng.uniform(-50, 350, (64, 3))
    d = np.stack([rng.choice([-1, 1], 64) * rng.uniform(0.5, 1, 64),
                  rng.choice([-1, 1], 64) * 10.0 ** rng.uniform(-38, -30, 64),
                  rng.choice([-1, 1], 64) * 10.0 ** rng.uniform(-37.5, -20, 64)], 1)
    groups.append(np.concatenate([o, d[:, rng.permutation(3)]], 1))
    # huge origins (clipped to the box, or out of bounds at once)
    o = rng.choice([-1, 1], (64, 3)) * 10.0 ** rng.uniform(3, 30, (64, 3))
    d = 128.0 - o
    groups.append(np.concatenate([o, d / np.linalg.norm(d, axis=1, keepdims=True)], 1))
    # origins on voxel and node borders
    o = rng.integers(0, 33, (64, 3)) * 8.0
    d = rng.standard_normal((64, 3))
    groups.append(np.concatenate([o, d], 1))
    # clean rays at the sphere, one lane with a -0.0 component: the whole wave takes the general step
    o = rng.uniform(-100, 350, (64, 3))
    d = 128.0 - o + rng.uniform(-60, 60, (64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[17, 1] = -0.0
    groups.append(np.concatenate([o, d], 1))
    return np.concatenate(groups).astype(np.float32)


def test_clean_waves_with_extreme_rays(wx, oracle_mod, cuda, cfg1):
    """The NaN-free step (waves whose rays are all finite) on extreme but finite inputs,
    the wave-counted step cap, and the general step in a wave with one degenerate lane:
    bit for bit against the oracle, on the wall grid and the config-1 sphere."""
    say("clean waves, extreme rays")
    rng = np.random.default_rng(23)
    rays = clean_extreme_rays(rng)
    org = np.array([(x, 0, 0) for x in range(0, 4096, 8)], np.int32)
    masks = np.full((org.shape[0], 8), np.uint64(0xFF))
    p = wx.Vdb345()
    p.add_leaf_masks(org >> 3, masks)
    p.compute_sdf(threads=2)
    wall = wx.Grid(p)
    ow = oracle_mod.Vdb()
    ow.add_leaf_masks(org, masks)
    ow.compute_sdf()
    t = cuda.from_numpy(rays).cuda()
    _, g1, o1 = cfg1
    for name, g, o in (("wall", wall, ow), ("sphere", g1, o1)):
        for clip in (False, True):
            for simple in (False, True):
                gpu = wx.hits_to_numpy(g.trace_rays(t, clip=clip, simple=simple))
                ora, _ = o.trace_rays(rays, clip=clip)
                assert_hits_equal(gpu, ora, f"{name} clip={clip} simple={simple}")
            if name == "wall":
                assert (gpu["state"][:64] == wx.MAXSTEPS).sum() >= 32
