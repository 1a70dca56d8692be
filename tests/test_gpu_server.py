"""GPU: the serving loop (resident grids, HBM-budget admission, continuous
batching). Every served request is byte-identical to the same request launched
alone, whether it ran in a batch of many, alone, or overlapped with other
launches on several streams; one is also checked against the oracle."""
import sys
import threading

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu

LIGHT = (8000.0, 8000.0, 8000.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(wx, oracle_mod, cuda):
    """config-1 sphere (with its oracle tree) + a 768^3 sphere set."""
    sp1 = scenes.single_sphere()
    p1 = scenes.build_sphere_vdb(sp1)
    o1 = oracle_mod.Vdb()
    o1.add_sphere(sp1[0, :3], sp1[0, 3])
    o1.compute_sdf()
    sp2 = scenes.sphere_set(8, count=200, domain=768, rmin=2, rmax=40)
    p2 = scenes.build_sphere_vdb(sp2, (0, 0, 0), (768, 768, 768))
    return [wx.Grid(p1), wx.Grid(p2)], o1


def requests(wx, n, seed):
    """(grid, camera, w, h, shadow) with mixed grids, sizes and shadow flags."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 48), (160, 120), (97, 131), (256, 256), (33, 17)]
    out = []
    for k in range(n):
        gi = k % 2
        c, r = ((128.0, 128.0, 128.0), 420.0) if gi == 0 else ((384.0, 384.0, 384.0), 1100.0)
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        w, h = sizes[k % len(sizes)]
        cam = wx.camera_basis(np.asarray(c) + r * v, c, (0, 1, 0), w / h, 55.0, float(w))
        out.append((gi, cam, w, h, k % 3 != 2))
    return out


def alone(wx, cuda, grids, gi, cam, w, h, shadow):
    o = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
    s = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda") if shadow else None
    wx.trace_batch(grids, [wx.camera_request(gi, cam, w, h, o, 0.5, True, shadow_out=s, light=LIGHT)])
    cuda.cuda.synchronize()
    return o, s


def served(wx, srv, tk, w, h, shadow):
    r = srv.wait(tk, 60.0)
    assert r.status == wx.WX_OK
    a = wx.device_copy(r.primary, w * h)
    s = wx.device_copy(r.shadow, w * h) if shadow else None
    if not shadow:
        assert not r.shadow
    return r, a, s


def test_batched_mix_identical_to_alone(wx, cuda, setup):
    """A blocker holds the whole arena, so the next requests queue; releasing it
    launches them as ONE batch. Each record matches its solo launch."""
    grids, _ = setup
    reqs = requests(wx, 10, seed=1)
    cap = sum(-(-(w * h * 32 * (2 if sh else 1)) // 256) * 256 for _, _, w, h, sh in reqs)
    # a 64-wide primary + shadow blocker leaves under 4 KiB free: less than any request needs
    bw, bh = 64, cap // (64 * 64)
    assert min(w * h * 32 for _, _, w, h, _ in reqs) > cap - bw * bh * 64 and 1 <= bh <= 65536
    srv = wx.Server(hbm_budget=cap, streams=2, max_batch_requests=64)
    try:
        gids = [srv.add_grid(g) for g in grids]
        say(f"blocker {bw}x{bh}, arena {cap} B")
        blocker = srv.submit(gids[1], reqs[1][1], bw, bh, 0.5, True, shadow=True, light=LIGHT)
        srv.wait(blocker, 60.0)
        tks = [srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT) for gi, cam, w, h, sh in reqs]
        srv.release(blocker)
        for (gi, cam, w, h, sh), tk in zip(reqs, tks):
            r, a, s = served(wx, srv, tk, w, h, sh)
            assert r.batch_requests == len(reqs)
            assert r.submit_s <= r.launch_s <= r.done_s
            srv.release(tk)
            o, so = alone(wx, cuda, grids, gi, cam, w, h, sh)
            assert cuda.equal(a, o), (gi, w, h)
            if sh:
                assert cuda.equal(s, so), (gi, w, h)
        st = srv.stats()
        assert st.completed == len(reqs) + 1 and st.batches == 2
        assert st.budget_peak == cap and st.budget_in_use == 0
        assert 0 < st.p50_ms <= st.p99_ms
    finally:
        srv.close()


def test_served_request_matches_oracle(wx, cuda, oracle_mod, setup):
    grids, o1 = setup
    cam = wx.camera_basis((400, 300, -200), (128, 128, 128), (0, 1, 0), 4 / 3, 60.0, 160.0)
    srv = wx.Server(hbm_budget=64 << 20, streams=1, max_batch_requests=4)
    try:
        gid = srv.add_grid(grids[0])
        srv.add_grid(grids[1])
        tk = srv.submit(gid, cam, 160, 120, 0.5, True, shadow=False)
        r, a, _ = served(wx, srv, tk, 160, 120, False)
        srv.release(tk)
    finally:
        srv.close()
    ora, _ = o1.trace_camera(cam.as_array(), 160, 120, 0.5, clip=True)
    got = wx.hits_to_numpy(a)
    assert np.array_equal(got.view(np.uint8), ora.view(np.uint8))
    assert int((got["state"] == wx.HIT).sum()) > 1000


@pytest.mark.parametrize("streams,max_batch", [(1, 1), (3, 8)])
def test_concurrent_submitters_identical(wx, cuda, setup, streams, max_batch):
    """Two threads submit while launches are in flight on several streams (the
    continuous case); plain one-request-per-launch serving is the (1, 1) row."""
    grids, _ = setup
    reqs = requests(wx, 24, seed=2)
    srv = wx.Server(hbm_budget=48 << 20, streams=streams, max_batch_requests=max_batch)
    results = [None] * len(reqs)
    errs = []
    try:
        gids = [srv.add_grid(g) for g in grids]

        def worker(part):
            try:
                for k in part:
                    gi, cam, w, h, sh = reqs[k]
                    tk = srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT)
                    r, a, s = served(wx, srv, tk, w, h, sh)
                    srv.release(tk)
                    results[k] = (r.batch_requests, a, s)
            except Exception as e:  # noqa: BLE001
                errs.append(repr(e))

        th = [threading.Thread(target=worker, args=(range(i, len(reqs), 2),)) for i in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join(120)
        assert not errs, errs
        st = srv.stats()
        assert st.completed == len(reqs) and st.budget_in_use == 0
        say(f"streams {streams} max_batch {max_batch}: {st.batches} batches, mean {st.mean_batch_requests:.2f}")
    finally:
        srv.close()
    for (gi, cam, w, h, sh), res in zip(reqs, results):
        n, a, s = res
        assert 1 <= n <= max_batch
        o, so = alone(wx, cuda, grids, gi, cam, w, h, sh)
        assert cuda.equal(a, o)
        if sh:
            assert cuda.equal(s, so)


def test_budget_rejects_and_queues(wx, cuda, setup):
    grids, _ = setup
    srv = wx.Server(hbm_budget=1 << 20, streams=1, max_batch_requests=4)
    try:
        gid = srv.add_grid(grids[0])
        cam = scenes.basis(scenes.config1_camera())
        with pytest.raises(wx.WoxelError) as e:
            srv.submit(gid, cam, 1024, 1024, 0.5, True)
        assert e.value.code == wx.WX_ERR_BUDGET
        # two 128x128 primary-only requests (512 KiB each) fill the arena; a third waits for a release
        tks = [srv.submit(gid, cam, 128, 128, 0.5, True) for _ in range(3)]
        r0 = srv.wait(tks[0], 60.0)
        srv.wait(tks[1], 60.0)
        with pytest.raises(wx.WoxelError):
            srv.wait(tks[2], 0.2)
        srv.release(tks[0])
        r2 = srv.wait(tks[2], 60.0)
        assert r2.status == wx.WX_OK and r0.status == wx.WX_OK
        assert srv.stats().rejected == 1
    finally:
        srv.close()
