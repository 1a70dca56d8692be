"""GPU: the batched and served configurations at their real sizes.

Config 3: the 4096^3 sphere set, 64 cameras (poses on the r=6000 sphere,
log-uniform 256^2..3840x2160, one of them 3840x2160) with one shadow ray per hit
toward (8000, 8000, 8000), all in one launch: every request equals its own launch
byte for byte, small requests equal the oracle whole, large ones on sampled 8x8
tiles (primary and shadow records).

Configs 4 and 5: the 32 resident grids (16 sphere sets, 16 value-noise
isosurfaces up to 4096^3). 256 cameras bound to random grids in one launch equal
their per-request launches; on the two smallest sphere-set grids whole requests
and sampled tiles equal the oracle; on two noise grids every sampled hit lies on an
active voxel of the host tree. A Poisson stream through the serving loop gives each
request exactly its solo launch's records.
"""
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu

LIGHT = scenes.LIGHT


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def oracle_layout(oracle_mod, v):
    k5, v5, k4, v4, v3 = v.masks()
    a5, a4, a3 = v.atlas()
    return oracle_mod.Layout.from_arrays(v.origins(), k5, v5, k4, v4, v3, a5, a4, a3)


def check_request_vs_oracle(wx, oracle_mod, lay, cam, w, h, prim, shad, rng, tiles=4, whole_limit=70_000):
    """Primary and shadow records of one camera request against the oracle: the
    whole image when it is small, else `tiles` random 8x8 tiles."""
    gp = wx.hits_to_numpy(prim).reshape(h, w)
    gs = wx.hits_to_numpy(shad).reshape(h, w) if shad is not None else None
    camw = cam.as_array()
    if w * h <= whole_limit:
        pix = [(x, y) for y in range(h) for x in range(w)]
    else:  # random tiles, half of them among the tiles that hold hits
        pix = []
        hy, hx = np.nonzero(gp["state"] == wx.HIT)
        for k in range(tiles):
            if k % 2 and len(hx):
                j = int(rng.integers(len(hx)))
                tx, ty = int(hx[j]) // 8, int(hy[j]) // 8
            else:
                tx, ty = int(rng.integers((w + 7) // 8)), int(rng.integers((h + 7) // 8))
            pix += [(x, y) for y in range(ty * 8, min(h, ty * 8 + 8)) for x in range(tx * 8, min(w, tx * 8 + 8))]
    dirs = np.stack([oracle_mod.camera_ray(camw, x, y, 0.5) for x, y in pix]).astype(np.float32)
    eye = np.broadcast_to(camw[:3].astype(np.float32), dirs.shape)
    ora, diag = lay.trace_rays(np.ascontiguousarray(np.concatenate([eye, dirs], 1)), clip=True)
    got = np.stack([gp[y, x] for x, y in pix])
    assert same(got, ora), "primary records differ from the oracle"
    if gs is None:
        return int((ora["state"] == 0).sum())
    hit = np.nonzero(ora["state"] == oracle_mod.HIT)[0]
    gsh = np.stack([gs[y, x] for x, y in pix])
    assert (gsh["state"][ora["state"] != oracle_mod.HIT] == wx.NO_RAY).all()
    if len(hit):
        srays = []
        for i in hit:
            so, sd = oracle_mod.shadow_ray(diag["p"][i], dirs[i], ora["normal"][i], LIGHT)
            srays.append(np.concatenate([so, sd]))
        sora, _ = lay.trace_rays(np.array(srays, np.float32), clip=True)
        assert same(gsh[hit], sora), "shadow records differ from the oracle"
    return len(hit)


# ---------------------------------------------------------------- config 3

@pytest.fixture(scope="module")
def cfg2(wx, oracle_mod, cuda):
    say("config 3: building the 4096^3 sphere set (2000 spheres)")
    lo, hi = scenes.clip_box()
    p = scenes.build_sphere_vdb(scenes.sphere_set(1), lo, hi)
    g = wx.Grid(p)
    lay = oracle_layout(oracle_mod, p)
    return p, g, lay


def config3_specs():
    specs = scenes.batch_cameras(64)
    if not any(s["width"] == 3840 and s["height"] == 2160 for s in specs):
        big = max(range(len(specs)), key=lambda i: specs[i]["width"] * specs[i]["height"])
        s = specs[big]
        specs[big] = scenes.look_at_spec(s["eye"], s["target"], 3840, 2160)
    return specs


@pytest.mark.parametrize("persistent", [0, 1])
def test_config3_batch_equals_per_request_and_oracle(wx, oracle_mod, cuda, cfg2, persistent):
    p, g, lay = cfg2
    specs = config3_specs()
    say(f"config 3 (shape {persistent}): {len(specs)} cameras, "
        f"{sum(s['width'] * s['height'] for s in specs)} pixels, largest "
        f"{max((s['width'], s['height']) for s in specs)}")
    tune = wx.Tune(persistent=persistent, refill=16, chunk=256, waves_per_cu=16, xcd_split=1)
    reqs, outs, cams = [], [], []
    for sp in specs:
        w, h = sp["width"], sp["height"]
        cam = scenes.basis(sp)
        o = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        s = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        reqs.append(wx.camera_request(0, cam, w, h, o, 0.5, True, shadow_out=s, light=LIGHT))
        outs.append((o, s))
        cams.append((cam, w, h))
    wx.trace_batch([g], reqs, tune=tune)
    cuda.cuda.synchronize()
    for k, ((cam, w, h), (o, s)) in enumerate(zip(cams, outs)):  # each request alone
        o1 = cuda.empty_like(o)
        s1 = cuda.empty_like(s)
        wx.trace_batch([g], [wx.camera_request(0, cam, w, h, o1, 0.5, True, shadow_out=s1, light=LIGHT)], tune=tune)
        cuda.cuda.synchronize()
        assert cuda.equal(o, o1) and cuda.equal(s, s1), (k, w, h)
        del o1, s1
    rng = np.random.default_rng(persistent)
    order = sorted(range(len(specs)), key=lambda i: specs[i]["width"] * specs[i]["height"])
    checked = order[:8] + order[-4:]  # the smallest requests whole, the largest on sampled tiles
    hits = 0
    for k in checked:
        cam, w, h = cams[k]
        hits += check_request_vs_oracle(wx, oracle_mod, lay, cam, w, h, outs[k][0], outs[k][1], rng, tiles=16)
    say(f"config 3: {len(checked)} requests against the oracle, {hits} hits with shadow rays")
    assert hits > 100


# ---------------------------------------------------------------- configs 4 and 5

@pytest.fixture(scope="module")
def grid_set(wx, oracle_mod, cuda):
    """The 32 resident grids; host trees kept for the two smallest sphere sets
    (oracle layouts) and the two smallest noise grids (active-voxel checks)."""
    import time

    specs = scenes.config4_specs()
    grids, keep, t0 = [], {}, time.time()
    sizes = []
    for i, sp in enumerate(specs):
        v = scenes.build_grid(sp, device=True)
        grids.append(wx.Grid(v))
        n = v.count_nodes()
        sizes.append(n[2])
        say(f"grid {i:2d} {sp['kind']:7s} side {sp['side']:4d}: nodes {n}")
        keep[i] = v
        # keep at most a few host trees alive: the smallest sphere sets and noise grids
        sph = sorted([j for j in keep if specs[j]["kind"] == "spheres"], key=lambda j: sizes[j])
        noi = sorted([j for j in keep if specs[j]["kind"] == "noise"], key=lambda j: sizes[j])
        for j in sph[2:] + noi[2:]:
            del keep[j]
    say(f"32 grids resident: {sum(x.nbytes() for x in grids) / 2**30:.2f} GiB in {time.time() - t0:.0f} s")
    layouts = {j: oracle_layout(oracle_mod, keep[j]) for j in keep if specs[j]["kind"] == "spheres"}
    leaves = {}
    for j in keep:
        if specs[j]["kind"] == "noise":
            org, masks = keep[j].leaf_masks()
            key = ((org[:, 0].astype(np.int64) >> 3) << 42) | ((org[:, 1].astype(np.int64) >> 3) << 21) | \
                  (org[:, 2].astype(np.int64) >> 3)
            o = np.argsort(key)
            leaves[j] = (key[o], masks[o])
    return specs, grids, layouts, leaves


def on_active_voxel(leaves, vox):
    """Whether each voxel (n x 3, non-negative) is an active voxel of the host tree."""
    key_s, masks = leaves
    v = vox.astype(np.int64)
    key = ((v[:, 0] >> 3) << 42) | ((v[:, 1] >> 3) << 21) | (v[:, 2] >> 3)
    i = np.clip(np.searchsorted(key_s, key), 0, len(key_s) - 1)
    found = key_s[i] == key
    off = ((v[:, 0] & 7) << 6) | ((v[:, 1] & 7) << 3) | (v[:, 2] & 7)
    bits = (masks[i, off >> 6] >> (off & 63).astype(np.uint64)) & np.uint64(1)
    return found & (bits == 1)


def test_config4_batch_equals_per_request(wx, oracle_mod, cuda, grid_set):
    specs, grids, layouts, leaves = grid_set
    cams = scenes.config4_cameras(specs, 256)
    say(f"config 4: {len(cams)} cameras over {len(set(g for g, _ in cams))} grids, "
        f"{sum(sp['width'] * sp['height'] for _, sp in cams)} pixels")
    reqs, outs, bases = [], [], []
    for gi, sp in cams:
        w, h = sp["width"], sp["height"]
        cam = scenes.basis(sp)
        o = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        s = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
        reqs.append(wx.camera_request(gi, cam, w, h, o, 0.5, True, shadow_out=s, light=LIGHT))
        outs.append((o, s))
        bases.append(cam)
    for persistent in (0, 1):
        tune = wx.Tune(persistent=persistent, refill=16, chunk=256, waves_per_cu=16, xcd_split=1)
        for o, s in outs:
            o.fill_(0x5A)
            s.fill_(0x5A)
        wx.trace_batch(grids, reqs, tune=tune)
        cuda.cuda.synchronize()
        for k, ((gi, sp), cam, (o, s)) in enumerate(zip(cams, bases, outs)):
            w, h = sp["width"], sp["height"]
            o1, s1 = cuda.empty_like(o), cuda.empty_like(s)
            wx.trace_batch([grids[gi]], [wx.camera_request(0, cam, w, h, o1, 0.5, True, shadow_out=s1, light=LIGHT)],
                           tune=tune)
            cuda.cuda.synchronize()
            assert cuda.equal(o, o1) and cuda.equal(s, s1), (persistent, k, gi, w, h)
        say(f"config 4 shape {persistent}: batch equals the 256 per-request launches")
    rng = np.random.default_rng(4)
    checked = hits = 0
    for k, ((gi, sp), cam, (o, s)) in enumerate(zip(cams, bases, outs)):
        w, h = sp["width"], sp["height"]
        if gi in layouts:
            hits += check_request_vs_oracle(wx, oracle_mod, layouts[gi], cam, w, h, o, s, rng, tiles=6)
            checked += 1
        elif gi in leaves:
            rec = wx.hits_to_numpy(o)
            hv = rec["voxel"][rec["state"] == wx.HIT]
            if len(hv):
                sel = hv[rng.choice(len(hv), size=min(len(hv), 20000), replace=False)]
                assert on_active_voxel(leaves[gi], sel).all(), (k, gi)
                hits += len(sel)
            checked += 1
    say(f"config 4: {checked} requests checked against the oracle or the host tree, {hits} hits")
    assert checked >= 4 and hits > 1000


@pytest.mark.parametrize("resident", [False, True], ids=["batched", "resident"])
def test_config5_served_stream_equals_solo(wx, cuda, grid_set, resident):
    """A Poisson stream (rate 1000/s) over the 32 grids through the serving loop
    (3 launch slots with continuous batching, or admission into one running kernel):
    every served request equals its solo launch, and some requests shared a launch
    or a running kernel with others."""
    import threading
    import time

    specs, grids, _, _ = grid_set
    items = [(ta, gi, scenes.basis(sp), sp["width"], sp["height"])
             for ta, gi, sp in scenes.serve_requests(specs, 40, 1000.0, seed=21)]
    budget = sum(w * h * 64 for *_, w, h in items) + (1 << 20)
    if resident:
        srv = wx.Server(hbm_budget=budget, resident=True, idle_stop_us=5000)
    else:
        srv = wx.Server(hbm_budget=budget, streams=3, max_batch_requests=64, max_batch_rays=1 << 24)
    gids = [srv.add_grid(g) for g in grids]
    say(f"config 5: {len(items)} requests, budget {budget / 2**30:.1f} GiB")
    tickets = [None] * len(items)

    def submit():
        t0 = time.perf_counter()
        for k, (ta, gi, cam, w, h) in enumerate(items):
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets[k] = srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=True, light=LIGHT)

    th = threading.Thread(target=submit)
    th.start()
    th.join()
    sizes = []
    try:
        for (ta, gi, cam, w, h), tk in zip(items, tickets):
            r = srv.wait(tk, 60.0)
            assert r.status == wx.WX_OK
            sizes.append(r.batch_requests)
            a, s = wx.device_copy(r.primary, w * h), wx.device_copy(r.shadow, w * h)
            srv.release(tk)
            o = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
            so = cuda.empty((w * h, 32), dtype=cuda.uint8, device="cuda")
            wx.trace_batch([grids[gi]], [wx.camera_request(0, cam, w, h, o, 0.5, True, shadow_out=so, light=LIGHT)])
            cuda.cuda.synchronize()
            assert cuda.equal(a, o) and cuda.equal(s, so), (gi, w, h)
            del a, s, o, so
        st = srv.stats()
    finally:
        srv.close()
    say(f"config 5: batch sizes {sorted(set(sizes))}, p50 {st.p50_ms:.1f} ms, p99 {st.p99_ms:.1f} ms")
    assert st.completed == len(items)
