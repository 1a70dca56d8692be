"""Frozen oracle outputs (tests/golden/make_hit_fixtures.py): the committed bytes
are intact (SHA-256), the oracle still reproduces them, and the product (-m gpu)
matches them bit for bit."""
import hashlib
import importlib.util
import json
import os

import numpy as np
import pytest

from woxel_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLDEN, "hit_fixtures.json")))


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def maker():
    spec = importlib.util.spec_from_file_location("make_hit_fixtures", os.path.join(GOLDEN, "make_hit_fixtures.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("name", sorted(META))
def test_fixture_bytes_intact(name):
    assert hashlib.sha256(open(os.path.join(GOLDEN, name), "rb").read()).hexdigest() == META[name]


def test_oracle_reproduces_fixtures(wx, oracle_mod):
    m = maker()
    lay = m.cfg1_tree().layout()
    cam = wx.camera_basis((400.0, 300.0, -200.0), (128.0, 128.0, 128.0), (0, 1, 0), 64 / 48, 60.0, 64.0)
    for off in (0.5, 0.001):
        hits, _ = lay.trace_camera(cam.as_array(), 64, 48, off)
        assert np.array_equal(hits.view(np.uint8).reshape(-1, 32), load(f"cfg1_cam_64x48_off{off}.npy"))
    rays = load("cfg1_rays_512_in.npy")
    assert np.array_equal(rays, m.fixture_rays())
    hits, _ = lay.trace_rays(rays)
    assert np.array_equal(hits.view(np.uint8).reshape(-1, 32), load("cfg1_rays_512.npy"))
    cam_r = wx.camera_basis((400.0, 300.0, -200.0), (128.0, 128.0, 128.0), (0, 1, 0), 32 / 24, 60.0, 32.0)
    for mode in (3, 4):
        _, rgba8 = lay.render_camera(cam_r.as_array(), 32, 24, wx.shade(mode), 0.5)
        assert np.array_equal(np.ascontiguousarray(rgba8).view(np.uint32).reshape(-1),
                              load(f"cfg1_render_32x24_mode{mode}.npy"))


@pytest.mark.gpu
def test_product_matches_fixtures(wx, cuda):
    g = wx.Grid(scenes.build_sphere_vdb(scenes.single_sphere()))
    cam = wx.camera_basis((400.0, 300.0, -200.0), (128.0, 128.0, 128.0), (0, 1, 0), 64 / 48, 60.0, 64.0)
    for off in (0.5, 0.001):
        got = g.trace_camera(cam, 64, 48, off).cpu().numpy()
        assert np.array_equal(got, load(f"cfg1_cam_64x48_off{off}.npy")), off
    rays = cuda.from_numpy(load("cfg1_rays_512_in.npy")).cuda()
    assert np.array_equal(g.trace_rays(rays).cpu().numpy(), load("cfg1_rays_512.npy"))
    cam_r = wx.camera_basis((400.0, 300.0, -200.0), (128.0, 128.0, 128.0), (0, 1, 0), 32 / 24, 60.0, 32.0)
    for mode in (3, 4):
        got = g.render_camera(cam_r, 32, 24, wx.shade(mode), 0.5).cpu().numpy().view(np.uint32).reshape(-1)
        assert np.array_equal(got, load(f"cfg1_render_32x24_mode{mode}.npy")), mode
