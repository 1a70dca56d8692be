"""GPU: the shading stage (wx_render_camera) against the oracle's restatement of
ray_trace / reflect_ray2 / reflect_ray1 (raycast.comp.wgsl:144-342): float
colours and rgba8 words bit for bit, every render mode, grid overlays on and
off, on the config-1 sphere, a sphere set seen from outside the bound box
(entry clip) and a level set read from a .vdb file."""
import sys

import numpy as np
import pytest

from woxel_amd import scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sphere(wx, oracle_mod, cuda):
    sp = scenes.single_sphere()
    p = scenes.build_sphere_vdb(sp)
    o = oracle_mod.Vdb()
    o.add_sphere(sp[0, :3], sp[0, 3])
    o.compute_sdf()
    return wx.Grid(p), o.layout()


@pytest.fixture(scope="module")
def sphere_set(wx, oracle_mod, cuda):
    sp = scenes.sphere_set(5, count=300, domain=1024, rmin=3, rmax=30)
    lo, hi = (0, 0, 0), (1024, 1024, 1024)
    p = scenes.build_sphere_vdb(sp, lo, hi)
    o = oracle_mod.Vdb()
    for s in sp:
        o.add_sphere(s[:3], s[3], lo, hi)
    o.compute_sdf()
    return wx.Grid(p), o.layout()


def check(wx, grid, lay, cam, w, h, shade, offset=0.5, clip=False):
    f = grid.render_camera(cam, w, h, shade, offset, clip=clip, f32=True).cpu().numpy()
    u8 = grid.render_camera(cam, w, h, shade, offset, clip=clip).cpu().numpy().view(np.uint32)
    rf, r8 = lay.render_camera(cam.as_array(), w, h, shade, offset, clip=clip)
    bad = np.flatnonzero(np.any(f.view(np.uint32) != rf.view(np.uint32), axis=1))
    assert bad.size == 0, (bad[:5], f[bad[:3]], rf[bad[:3]])
    assert np.array_equal(u8, r8)
    return f


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("show", [(0, 0, 0), (1, 1, 1)])
def test_render_config1_bit_exact(wx, sphere, mode, show):
    say(f"config-1 render mode {mode} grid {show}")
    g, lay = sphere
    cam = scenes.basis(scenes.config1_camera())
    f = check(wx, g, lay, cam, 640, 480, wx.shade(mode, show))
    assert np.unique(f[:, :3], axis=0).shape[0] > 3


@pytest.mark.parametrize("mode", [3, 4])
def test_render_sphere_set_clip_bit_exact(wx, sphere_set, mode):
    say(f"sphere set render mode {mode}, camera outside the bound box")
    g, lay = sphere_set
    cam = wx.camera_basis((-2600.0, 3100.0, -4700.0), (512.0, 512.0, 512.0), (0, 1, 0), 4 / 3, 30.0, 320.0)
    check(wx, g, lay, cam, 320, 240, wx.shade(mode, sun_dir=(0.3, -0.8, 0.52), intensity=0.9), clip=True)


def test_render_vdb_file_glossy(wx, oracle_mod, cuda):
    from oracle import vdb_file
    from test_vdb_file import fixture

    say("cube.vdb glossy render")
    p = wx.Vdb345()
    p.read_file(fixture("cube"))
    p.compute_sdf(threads=4)
    g = wx.Grid(p)
    o = oracle_mod.Vdb.from_file_topology(vdb_file.read_grid(fixture("cube")))
    o.compute_sdf()
    cam = wx.camera_basis((260.0, 180.0, -300.0), (4.0, 4.0, 4.0), (0, 1, 0), 4 / 3, 50.0, 256.0)
    for mode in (3, 4):
        check(wx, g, o.layout(), cam, 256, 192, wx.shade(mode, (1, 0, 1)))


def test_render_rejects_bad_arguments(wx, sphere):
    g, _ = sphere
    cam = scenes.basis(scenes.config1_camera())
    with pytest.raises(wx.WoxelError):
        g.render_camera(cam, 16, 16, wx.shade(5))
