"""GPU: the serving loop with admission into a running kernel (WX_SERVER_RESIDENT).
Every served request is byte-identical to the same request launched alone: in a
burst that joins one running kernel, across kernels that ended when idle and were
started again, and when the kernel's own waits end on a wait limit set so short
(from the host) that kernels end between requests and unfinished requests run
again in the next one. Every wait of the kernel is bounded, so each of these ends."""
import sys
import threading
import time

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu

LIGHT = (8000.0, 8000.0, 8000.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def grids(wx, cuda):
    """config-1 sphere, a 768^3 sphere set and a 512^3 noise isosurface."""
    p1 = scenes.build_sphere_vdb(scenes.single_sphere())
    sp2 = scenes.sphere_set(8, count=200, domain=768, rmin=2, rmax=40)
    p2 = scenes.build_sphere_vdb(sp2, (0, 0, 0), (768, 768, 768))
    p3 = wx.Vdb345()
    p3.add_noise(3, (0, 0, 0), (512, 512, 512), 8.0 / 512, 0.1, device=True)
    p3.compute_sdf(threads=4, device=True)
    return [wx.Grid(p1), wx.Grid(p2), wx.Grid(p3)]


CENTRES = [((128.0, 128.0, 128.0), 420.0), ((384.0, 384.0, 384.0), 1100.0), ((256.0, 256.0, 256.0), 800.0)]


def requests(wx, n, seed):
    """(grid, camera, w, h, shadow): mixed grids, sizes (some not multiples of 8)
    and shadow flags."""
    rng = np.random.default_rng(seed)
    sizes = [(64, 48), (160, 120), (97, 131), (256, 256), (33, 17), (640, 480), (1, 1), (1023, 9)]
    out = []
    for k in range(n):
        gi = k % 3
        c, r = CENTRES[gi]
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        w, h = sizes[k % len(sizes)]
        cam = wx.camera_basis(np.asarray(c) + r * v, c, (0, 1, 0), w / h, 55.0, float(w))
        out.append((gi, cam, w, h, k % 3 != 2))
    return out


def budget(reqs):
    return sum(-(-(w * h * 32 * (2 if sh else 1)) // 256) * 256 for _, _, w, h, sh in reqs) + (1 << 20)


def check_served(wx, cuda, srv, grids, reqs, tks):
    for (gi, cam, w, h, sh), tk in zip(reqs, tks):
        r = srv.wait(tk, 60.0)
        assert r.status == wx.WX_OK
        assert r.submit_s <= r.launch_s <= r.done_s
        a = wx.device_copy(r.primary, w * h)
        s = wx.device_copy(r.shadow, w * h) if sh else None
        assert bool(r.shadow) == sh
        srv.release(tk)
        o = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        so = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda") if sh else None
        wx.trace_batch(grids, [wx.camera_request(gi, cam, w, h, o, 0.5, True, shadow_out=so, light=LIGHT)])
        cuda.cuda.synchronize()
        assert cuda.equal(a, o), (gi, w, h)
        if sh:
            assert cuda.equal(s, so), (gi, w, h)


def test_burst_joins_one_kernel(wx, cuda, grids):
    reqs = requests(wx, 40, seed=1)
    srv = wx.Server(hbm_budget=budget(reqs), resident=True, idle_stop_us=5000)
    try:
        gids = [srv.add_grid(g) for g in grids]
        say(f"resident burst: {len(reqs)} requests")
        tks = [srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT) for gi, cam, w, h, sh in reqs]
        check_served(wx, cuda, srv, grids, reqs, tks)
        st = srv.stats()
        assert st.completed == len(reqs) and st.budget_in_use == 0
        say(f"sessions {st.batches}, p50 {st.p50_ms:.2f} ms, p99 {st.p99_ms:.2f} ms")
    finally:
        srv.close()


def test_idle_stop_and_restart(wx, cuda, grids):
    reqs = requests(wx, 12, seed=2)
    srv = wx.Server(hbm_budget=budget(reqs), resident=True, idle_stop_us=2000)
    try:
        gids = [srv.add_grid(g) for g in grids]
        for part in (reqs[:6], reqs[6:]):
            say(f"resident: {len(part)} requests after an idle stop")
            tks = [srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT) for gi, cam, w, h, sh in part]
            check_served(wx, cuda, srv, grids, part, tks)
            time.sleep(0.05)  # > idle_stop: the kernel ends
        st = srv.stats()
        assert st.completed == len(reqs) and st.batches >= 2
    finally:
        srv.close()


def test_short_wait_limit_ends_kernels_between_requests(wx, cuda, grids):
    """wait_limit_us = 300 (set from the host): with arrivals ~1-3 ms apart the
    kernel's waits end between requests, kernels end and are started again, and a
    request admitted into a kernel that was ending runs again in the next one."""
    reqs = requests(wx, 30, seed=3)
    srv = wx.Server(hbm_budget=budget(reqs), resident=True, wait_limit_us=300, idle_stop_us=1000000)
    rng = np.random.default_rng(4)
    gaps = rng.uniform(0.001, 0.003, len(reqs))
    try:
        gids = [srv.add_grid(g) for g in grids]
        say(f"resident, wait limit 300 us: {len(reqs)} requests")
        tks = []
        for (gi, cam, w, h, sh), gap in zip(reqs, gaps):
            tks.append(srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT))
            time.sleep(gap)
        check_served(wx, cuda, srv, grids, reqs, tks)
        st = srv.stats()
        say(f"sessions {st.batches}")
        assert st.completed == len(reqs) and st.batches >= 2
    finally:
        srv.close()


def test_concurrent_submitters_and_close_with_work_in_flight(wx, cuda, grids):
    reqs = requests(wx, 24, seed=5)
    srv = wx.Server(hbm_budget=budget(reqs), resident=True, idle_stop_us=5000)
    gids = [srv.add_grid(g) for g in grids]
    tks = [None] * len(reqs)

    def submit(lo):
        for k in range(lo, len(reqs), 3):
            gi, cam, w, h, sh = reqs[k]
            tks[k] = srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT)

    th = [threading.Thread(target=submit, args=(i,)) for i in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    try:
        check_served(wx, cuda, srv, grids, reqs, tks)
        # more work, then close while it is queued or running: close returns
        for gi, cam, w, h, sh in reqs[:8]:
            srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=sh, light=LIGHT)
    finally:
        t0 = time.perf_counter()
        srv.close()
        say(f"closed with work in flight in {time.perf_counter() - t0:.3f} s")


def test_grid_limit(wx, cuda, grids):
    srv = wx.Server(hbm_budget=1 << 20, resident=True)
    try:
        for _ in range(64):
            srv.add_grid(grids[0])
        with pytest.raises(wx.WoxelError) as e:
            srv.add_grid(grids[0])
        assert e.value.code == wx.WX_ERR_UNSUPPORTED
    finally:
        srv.close()
