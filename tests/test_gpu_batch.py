"""GPU: batched persistent launches (many cameras / grids / shadow rays in one
launch) produce exactly what one-request launches and the oracle produce, for
every launch shape."""
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu

LIGHT = (8000.0, 8000.0, 8000.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


@pytest.fixture(scope="module")
def grids(wx, oracle_mod, cuda):
    """config-1 sphere + a 1024^3 sphere set, product grids and oracle trees."""
    out = []
    sp1 = scenes.single_sphere()
    sp2 = scenes.sphere_set(5, count=300, domain=1024, rmin=3, rmax=30)
    for sp, lo, hi in ((sp1, None, None), (sp2, (0, 0, 0), (1024, 1024, 1024))):
        p = scenes.build_sphere_vdb(sp, lo, hi)
        o = oracle_mod.Vdb()
        for s in sp:
            if lo is None:
                o.add_sphere(s[:3], s[3])
            else:
                o.add_sphere(s[:3], s[3], lo, hi)
        o.compute_sdf()
        out.append((p, wx.Grid(p), o))
    return out


def cameras(wx, n, rng, centre, radius, sizes):
    cams = []
    for k in range(n):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        eye = np.asarray(centre) + radius * v
        w, h = sizes[k % len(sizes)]
        cams.append((wx.camera_basis(eye, centre, (0, 1, 0), w / h, 50.0, float(w)), w, h))
    return cams


def test_simple_and_persistent_shapes_agree(wx, cuda, grids):
    say("simple vs persistent")
    p, g, o = grids[0]
    cam = scenes.basis(scenes.config1_camera())
    rays = cuda.from_numpy(scenes.incoherent_rays(70_001, seed=3, domain=256)).cuda()
    a = g.trace_camera(cam, 640, 480, 0.5, simple=True).clone()
    ar = g.trace_rays(rays, clip=True, simple=True).clone()
    saved = wx.tune_get()
    wx.tune_set(wx.Tune(persistent=1, refill=16, chunk=128, waves_per_cu=16, xcd_split=1))
    try:
        b = g.trace_camera(cam, 640, 480, 0.5)
        br = g.trace_rays(rays, clip=True)
    finally:
        wx.tune_set(saved)
    assert cuda.equal(a, b)
    assert cuda.equal(ar, br)


def test_compact_shape_config1_and_odd_sizes(wx, oracle_mod, cuda, grids):
    """The compact shape against the oracle: config 1 and sizes that leave most
    lanes of the edge workgroups idle from the start."""
    say("compact shape vs oracle")
    p, g, o = grids[0]
    saved = wx.tune_get()
    wx.tune_set(wx.Tune(persistent=2, refill=4, chunk=256, waves_per_cu=16, xcd_split=1))
    try:
        for (w, h) in ((640, 480), (37, 29), (1, 1), (8, 200)):
            cam = wx.camera_basis((400, 300, -200), (128, 128, 128), (0, 1, 0), w / h, 60.0, float(w))
            got = wx.hits_to_numpy(g.trace_camera(cam, w, h, 0.5))
            want, _ = o.trace_camera(cam.as_array(), w, h, 0.5)
            assert same(got, want), (w, h)
    finally:
        wx.tune_set(saved)


def test_lds_shape_config1_and_odd_sizes(wx, oracle_mod, cuda, grids):
    """The LDS shape (node-5 records staged in LDS, 32x32-pixel workgroups) against
    the oracle: config 1 and sizes that leave most of an edge workgroup idle."""
    say("LDS shape vs oracle")
    p, g, o = grids[0]
    saved = wx.tune_get()
    wx.tune_set(wx.Tune(persistent=3, refill=16, chunk=256, waves_per_cu=16, xcd_split=1))
    try:
        for (w, h) in ((640, 480), (37, 29), (1, 1), (8, 200), (33, 1)):
            cam = wx.camera_basis((400, 300, -200), (128, 128, 128), (0, 1, 0), w / h, 60.0, float(w))
            got = wx.hits_to_numpy(g.trace_camera(cam, w, h, 0.5))
            want, _ = o.trace_camera(cam.as_array(), w, h, 0.5)
            assert same(got, want), (w, h)
    finally:
        wx.tune_set(saved)


@pytest.mark.parametrize("tune", [dict(persistent=1, refill=1, chunk=64, waves_per_cu=4, xcd_split=0),
                                  dict(persistent=1, refill=64, chunk=512, waves_per_cu=32, xcd_split=1),
                                  dict(persistent=1, refill=24, chunk=128, waves_per_cu=12, xcd_split=1),
                                  dict(persistent=0, refill=16, chunk=256, waves_per_cu=16, xcd_split=1),
                                  dict(persistent=0, refill=16, chunk=256, waves_per_cu=16, xcd_split=2),
                                  dict(persistent=2, refill=1, chunk=256, waves_per_cu=16, xcd_split=1),
                                  dict(persistent=2, refill=8, chunk=256, waves_per_cu=16, xcd_split=1),
                                  dict(persistent=2, refill=64, chunk=256, waves_per_cu=16, xcd_split=1),
                                  dict(persistent=3, refill=16, chunk=256, waves_per_cu=16, xcd_split=1)])
def test_launch_shapes_identical(wx, cuda, grids, tune):
    say(f"tune {tune}")
    p, g, o = grids[1]
    cam = wx.camera_basis((2000, 900, -700), (512, 512, 512), (0, 1, 0), 16 / 9, 50.0, 800.0)
    ref = g.trace_camera(cam, 800, 450, 0.5, clip=True, simple=True).clone()
    t = wx.Tune(**tune)
    out = cuda.empty_like(ref)
    wx.trace_batch([g], [wx.camera_request(0, cam, 800, 450, out, 0.5, True)], tune=t)
    assert cuda.equal(ref, out)


@pytest.mark.parametrize("persistent", [0, 1, 2, 3])
def test_batch_mixed_requests_match_single_launches(wx, oracle_mod, cuda, grids, persistent):
    say(f"mixed batch persistent={persistent}")
    rng = np.random.default_rng(0)
    gl = [grids[0][1], grids[1][1]]
    sizes = [(64, 64), (200, 120), (33, 17), (320, 240), (1, 1), (129, 65)]
    reqs, singles, outs = [], [], []
    for k, (cam, w, h) in enumerate(cameras(wx, 10, rng, (128, 128, 128), 450.0, sizes)):
        gi = k % 2
        if gi == 1:
            cam = wx.camera_basis(np.array([512, 512, 512]) + 1500 * rng.standard_normal(3) / 1.7, (512, 512, 512),
                                  (0, 1, 0), w / h, 50.0, float(w))
        out = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        reqs.append(wx.camera_request(gi, cam, w, h, out, 0.5, True))
        outs.append(out)
        singles.append(gl[gi].trace_camera(cam, w, h, 0.5, clip=True).clone())
    rays = cuda.from_numpy(scenes.incoherent_rays(5000, seed=2, domain=1024)).cuda()
    rout = cuda.empty((5000, 32), dtype=cuda.uint8, device="cuda")
    reqs.append(wx.rays_request(1, rays, rout, True))
    wx.trace_batch(gl, reqs, tune=wx.Tune(persistent=persistent, refill=16, chunk=128, waves_per_cu=16, xcd_split=1))
    cuda.cuda.synchronize()
    for a, b in zip(outs, singles):
        assert cuda.equal(a, b)
    ora, _ = grids[1][2].trace_rays(rays.cpu().numpy(), clip=True)
    assert same(wx.hits_to_numpy(rout), ora)


@pytest.mark.parametrize("persistent", [0, 1, 3])
def test_shadow_rays_match_oracle(wx, oracle_mod, cuda, grids, persistent):
    say(f"shadow rays persistent={persistent}")
    p, g, o = grids[1]
    w, h = 320, 200
    cam = wx.camera_basis((2500, 1800, -900), (512, 512, 512), (0, 1, 0), w / h, 45.0, float(w))
    prim = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
    shad = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
    wx.trace_batch([g], [wx.camera_request(0, cam, w, h, prim, 0.5, True, shadow_out=shad, light=LIGHT)],
                   tune=wx.Tune(persistent=persistent, refill=16, chunk=128, waves_per_cu=16, xcd_split=1))
    gp, gs = wx.hits_to_numpy(prim), wx.hits_to_numpy(shad)
    ora, diag = o.trace_camera(cam.as_array(), w, h, 0.5, clip=True)
    assert same(gp, ora)
    hit = ora["state"] == oracle_mod.HIT
    assert hit.sum() > 500
    assert (gs["state"][~hit] == wx.NO_RAY).all()
    dirs = np.stack([oracle_mod.camera_ray(cam.as_array(), i % w, i // w, 0.5) for i in np.nonzero(hit)[0]])
    srays = []
    for k, i in enumerate(np.nonzero(hit)[0]):
        so, sd = oracle_mod.shadow_ray(diag["p"][i], dirs[k], ora["normal"][i], LIGHT)
        srays.append(np.concatenate([so, sd]))
    sora, _ = o.trace_rays(np.array(srays, np.float32), clip=True)
    assert same(gs[hit], sora)
    occluded = (sora["state"] == oracle_mod.HIT).mean()
    say(f"occluded fraction {occluded:.3f}")


def test_request_validation(wx, cuda, grids):
    p, g, o = grids[0]
    out = cuda.empty((16, 32), dtype=cuda.uint8, device="cuda")
    cam = scenes.basis(scenes.config1_camera())
    bad = wx.camera_request(3, cam, 4, 4, out)
    with pytest.raises(wx.WoxelError) as e:
        wx.trace_batch([g], [bad])
    assert e.value.code == wx.WX_ERR_INVALID
    with pytest.raises(wx.WoxelError):
        wx.trace_batch([g], [wx.camera_request(0, cam, 4, 4, out)],
                       tune=wx.Tune(persistent=1, refill=0, chunk=64, waves_per_cu=8, xcd_split=1))


def test_graph_replay_of_per_request_launches(wx, cuda, grids):
    """Per-request launches captured in a HIP graph: every replay, with eager work and
    synchronisations between replays, equals the eager launches."""
    say("graph capture of per-request launches")
    rng = np.random.default_rng(11)
    reqs, outs = [], []
    for k, (cam, w, h) in enumerate(cameras(wx, 4, rng, (128, 128, 128), 450.0, [(64, 48), (97, 33), (160, 120)])):
        gi = k % 2
        if gi == 1:
            cam = wx.camera_basis(np.array([512, 512, 512]) + 1500 * rng.standard_normal(3) / 1.7, (512, 512, 512),
                                  (0, 1, 0), w / h, 50.0, float(w))
        o = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        s = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        reqs.append(([grids[gi][1]], [wx.camera_request(0, cam, w, h, o, 0.5, True, shadow_out=s, light=LIGHT)]))
        outs.append((o, s))
    t = wx.Tune(persistent=0, refill=16, chunk=256, waves_per_cu=16, xcd_split=1)

    def launch(st):
        for g1, r1 in reqs:
            wx.trace_batch(g1, r1, tune=t, stream=st)

    launch(cuda.cuda.current_stream())
    cuda.cuda.synchronize()
    eager = [(o.clone(), s.clone()) for o, s in outs]
    graph = cuda.cuda.CUDAGraph()
    with cuda.cuda.graph(graph):
        launch(cuda.cuda.current_stream())
    other = grids[0][1].trace_camera(scenes.basis(scenes.config1_camera()), 320, 240, 0.5)
    for rep in range(3):
        for o, s in outs:
            o.zero_()
            s.zero_()
        other = grids[0][1].trace_camera(scenes.basis(scenes.config1_camera()), 320, 240, 0.5)  # eager work between
        cuda.cuda.synchronize()
        graph.replay()
        cuda.cuda.synchronize()
        for (o, s), (eo, es) in zip(outs, eager):
            assert cuda.equal(o, eo) and cuda.equal(s, es), rep
    assert int((wx.hits_to_numpy(eager[0][0])["state"] == wx.HIT).sum()) > 100
    del other
