"""GPU: wx_trace_batch_static, the allocation-free variant of wx_trace_batch (its
tables and queue counters in a static device array per GPU), gives the same hit
records as wx_trace_batch through the C ABI: two calls in a row, the second with a
smaller batch, in the one-lane and the persistent launch shapes."""
import ctypes as C
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu

LIGHT = (8000.0, 8000.0, 8000.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def two_grids(wx, cuda):
    a = scenes.build_sphere_vdb(scenes.single_sphere())
    sp = scenes.sphere_set(5, count=300, domain=1024, rmin=3, rmax=30)
    b = scenes.build_sphere_vdb(sp, (0, 0, 0), (1024, 1024, 1024))
    return [wx.Grid(a), wx.Grid(b)]


def batch(wx, cuda, n_cams, n_rays, seed):
    """n_cams camera requests (odd sizes, shadow rays on half of them) alternating
    between the two grids, plus n_rays ray-list requests; returns (requests, outputs)."""
    rng = np.random.default_rng(seed)
    reqs, outs = [], []
    for k in range(n_cams):
        gi = k % 2
        centre, radius = ((128.0, 128.0, 128.0), 400.0) if gi == 0 else ((512.0, 512.0, 512.0), 1500.0)
        v = rng.standard_normal(3)
        eye = np.asarray(centre) + radius * v / np.linalg.norm(v)
        w, h = int(rng.integers(1, 300)), int(rng.integers(1, 200))
        cam = wx.camera_basis(eye, centre, (0, 1, 0), w / h, 55.0, float(w))
        o = cuda.full((w * h, 32), 0xAB, dtype=cuda.uint8, device="cuda")
        s = cuda.full((w * h, 32), 0xCD, dtype=cuda.uint8, device="cuda") if k % 2 else None
        reqs.append(wx.camera_request(gi, cam, w, h, o, 0.5, True, shadow_out=s, light=LIGHT))
        outs.append((o, s))
    for k in range(n_rays):
        n = int(rng.integers(1, 5000))
        rays = cuda.from_numpy(scenes.incoherent_rays(n, seed=seed * 7 + k, domain=1024)).cuda()
        o = cuda.full((n, 32), 0xEF, dtype=cuda.uint8, device="cuda")
        r = wx.rays_request(1, rays, o, True)
        r._keep = rays
        reqs.append(r)
        outs.append((o, None))
    return reqs, outs


def snapshot(outs):
    return [(o.clone(), None if s is None else s.clone()) for o, s in outs]


def equal(a, b, cuda):
    return all(cuda.equal(x, y) and ((xs is None and ys is None) or cuda.equal(xs, ys))
               for (x, xs), (y, ys) in zip(a, b))


@pytest.mark.parametrize("shape", [dict(persistent=0, refill=16, chunk=256, waves_per_cu=16, xcd_split=1),
                                   dict(persistent=1, refill=16, chunk=128, waves_per_cu=16, xcd_split=1)])
def test_static_equals_allocating_batch(wx, cuda, two_grids, shape):
    tune = wx.Tune(**shape)
    for call, (n_cams, n_rays) in enumerate(((40, 9), (5, 2))):  # the second call is smaller
        say(f"static batch: shape {shape['persistent']}, call {call}: {n_cams} cameras + {n_rays} ray lists")
        reqs, outs = batch(wx, cuda, n_cams, n_rays, seed=11 + call)
        wx.trace_batch(two_grids, reqs, tune=tune)
        cuda.cuda.synchronize()
        ref = snapshot(outs)
        for o, s in outs:  # poison the outputs, then the allocation-free call
            o.fill_(0x5A)
            if s is not None:
                s.fill_(0x5A)
        wx.trace_batch_static(two_grids, reqs, tune=tune)
        cuda.cuda.synchronize()
        assert equal(ref, outs, cuda), (shape, call)


def test_static_single_request_and_limits(wx, cuda, two_grids):
    say("static batch: lone requests and the batch limits")
    for shape in (0, 1):
        tune = wx.Tune(persistent=shape, refill=16, chunk=256, waves_per_cu=16, xcd_split=1)
        reqs, outs = batch(wx, cuda, 1, 0, seed=5)
        wx.trace_batch([two_grids[0]], reqs, tune=tune)
        cuda.cuda.synchronize()
        ref = snapshot(outs)
        outs[0][0].fill_(0)
        wx.trace_batch_static([two_grids[0]], reqs, tune=tune)
        cuda.cuda.synchronize()
        assert equal(ref, outs, cuda), shape
    reqs, _ = batch(wx, cuda, 1, 0, seed=6)
    with pytest.raises(wx.WoxelError) as e:
        wx.trace_batch_static([two_grids[0]] * 65, reqs)
    assert e.value.code == wx.WX_ERR_INVALID
    with pytest.raises(wx.WoxelError) as e:
        wx.trace_batch_static([two_grids[0]], reqs * 1025)
    assert e.value.code == wx.WX_ERR_INVALID
