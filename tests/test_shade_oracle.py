"""Pins for the oracle's shading (ray_trace / reflect_ray2 / reflect_ray1,
raycast.comp.wgsl:144-342) on scenes whose answers follow from geometry:

* flat faces of voxel boxes seen head-on, where the hit normal is known, give
  the gray / rgb colours and grid overlays in closed form;
* ray-length colours use the hit's own iteration (sum of its step counts - 1);
* diffuse light is lit or ambient according to the sign of -sun . N, and
  shadowed exactly where the shadow ray meets an occluder box (slab test in f64);
* a glossy floor reflects by the mirror law; a reflection that leaves the
  bound box shows the wall of the face it leaves through.

Colours are formed with the reference's constants in float32; the geometric
facts (normals, shadowing, exit faces) come from the scene, not from the
oracle's code.
"""
import numpy as np
import pytest

import woxel_amd as wx

F = np.float32


def box_vdb(oracle_mod, boxes):
    """Solid voxel boxes [lo, hi) with 8-aligned corners (whole leaves)."""
    o = oracle_mod.Vdb()
    org = []
    for lo, hi in boxes:
        for x in range(lo[0], hi[0], 8):
            for y in range(lo[1], hi[1], 8):
                for z in range(lo[2], hi[2], 8):
                    org.append((x, y, z))
    o.add_leaf_masks(np.array(org, np.int32), np.full((len(org), 8), np.uint64(0xFFFFFFFFFFFFFFFF)))
    o.compute_sdf()
    return o.layout()


def render(lay, cam, w, h, shade):
    rgba, _ = lay.render_camera(cam.as_array(), w, h, shade, 0.5)
    hits, diag = lay.trace_camera(cam.as_array(), w, h, 0.5)
    return rgba[:, :3], hits, diag


HEAD_ON = dict(eye=(128.0, 128.0, -300.0), target=(128.0, 128.0, 128.0))


def head_on_camera(target_z=128.0):
    return wx.camera_basis((128.0, 128.0, -300.0), (128.0, 128.0, target_z), (0, 1, 0), 1.0, 12.0, 64.0)


def test_gray_and_rgb_on_a_face_seen_head_on(oracle_mod):
    lay = box_vdb(oracle_mod, [((96, 96, 96), (160, 160, 160))])
    cam = head_on_camera()
    c0, hits, _ = render(lay, cam, 64, 64, wx.shade(0))
    hit = hits["state"] == 0
    assert hit.sum() > 1000 and np.all(hits["normal"][hit] == (0, 0, -1))  # every hit is on the -z face
    assert np.all(c0[hit] == F(0.3))                                      # gray: dot(mask * (.2,.2,.3), 1)
    c1, _, _ = render(lay, cam, 64, 64, wx.shade(1))
    assert np.all(c1[hit] == np.array([F(0.1), F(0.1), F(0.1) + F(0.4)], np.float32))


@pytest.mark.parametrize("show", [(1, 0, 0), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 0, 1)])
@pytest.mark.parametrize("face", [96, 128])
def test_grid_overlay_priorities(oracle_mod, show, face):
    """Overlay when any component of floor(p) is a multiple of 4096 / 128 / 8,
    checked in that order (node-5 first)."""
    lay = box_vdb(oracle_mod, [((96, 96, face), (160, 160, face + 64))])
    c, hits, diag = render(lay, head_on_camera(), 64, 64, wx.shade(0, show))
    hit = hits["state"] == 0
    assert hit.sum() > 1000
    fp = np.floor(diag["p"][hit].astype(np.float64)).astype(np.int64)
    want = np.zeros((fp.shape[0], 3), np.float32)
    done = np.zeros(fp.shape[0], bool)
    for on, size, col in ((show[2], 4096, (-0.3, -0.3, 1.0)), (show[1], 128, (0.6, -0.2, -0.2)),
                          (show[0], 8, (-0.1, 0.5, 0.3))):
        sel = ~done & bool(on) & np.any(fp % size == 0, axis=1)
        want[sel] = np.array(col, np.float32)
        done |= sel
    assert np.array_equal(c[hit], want + F(0.3))
    if show[0] and face == 96:
        assert done.all()  # the face z = 96 lies on a node-3 boundary


def test_ray_length_uses_the_hit_iteration(oracle_mod):
    lay = box_vdb(oracle_mod, [((96, 96, 96), (160, 160, 160))])
    c, hits, _ = render(lay, head_on_camera(), 64, 64, wx.shade(2))
    hit = hits["state"] == 0
    i = hits["steps"].astype(np.int64).sum(axis=1) - 1
    t = (i.astype(np.float32) / F(200.0)).astype(np.float32)
    c1, c2 = np.array([0.72, 1.0, 0.99], np.float32), np.array([1.0, 0.0, 0.0], np.float32)
    want = (c1[None, :] * (F(1.0) - t)[:, None] + c2[None, :] * t[:, None]).astype(np.float32)
    assert np.array_equal(c[hit], want[hit])
    assert len(np.unique(i[hit])) > 1


def diffuse(LN, sun_rgb):
    base, amb = np.array([0.4, 0.2, 0.2], np.float32), np.array([0.4, 0.4, 0.3], np.float32)
    Ia = (F(0.3) * amb) * base
    Id = ((F(0.7) * np.asarray(sun_rgb, np.float32)) * base) * F(LN)
    return Ia.astype(np.float32), (Ia + Id).astype(np.float32)


def test_diffuse_lit_and_unlit_faces(oracle_mod):
    lay = box_vdb(oracle_mod, [((96, 96, 96), (160, 160, 160))])
    cam = head_on_camera()
    s = wx.shade(3, sun_dir=(0.0, 0.0, -1.0))  # sunlight travelling -z: the sun is behind the box
    c, hits, _ = render(lay, cam, 64, 64, s)
    hit = hits["state"] == 0
    ia, _ = diffuse(0.0, s["sun_color"])
    assert np.all(c[hit] == ia)
    s = wx.shade(3, sun_dir=(0.0, 0.0, 1.0))   # sunlight travelling +z onto the -z face: LN = 1, no occluder
    c, _, _ = render(lay, cam, 64, 64, s)
    _, lit = diffuse(1.0, s["sun_color"])
    assert np.all(c[hit] == lit)


def slab_hit(o, d, lo, hi):
    t0, t1 = 0.0, np.inf
    for a in range(3):
        if d[a] == 0:
            if not (lo[a] <= o[a] <= hi[a]):
                return False
            continue
        u, v = (lo[a] - o[a]) / d[a], (hi[a] - o[a]) / d[a]
        t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
    return t0 <= t1


def test_diffuse_shadow_of_an_occluder(oracle_mod):
    floor = ((0, 64, 0), (512, 128, 512))   # top face at y = 128
    occ = ((160, 256, 160), (288, 320, 288))
    lay = box_vdb(oracle_mod, [floor, occ])
    sun = np.array([0.35, -0.85, 0.2], np.float64)
    sun /= np.linalg.norm(sun)
    s = wx.shade(3, sun_dir=tuple(float(x) for x in sun.astype(np.float32)))
    cam = wx.camera_basis((240.0, 700.0, -150.0), (240.0, 128.0, 200.0), (0, 1, 0), 1.0, 55.0, 128.0)
    c, hits, diag = render(lay, cam, 128, 128, s)
    top = (hits["state"] == 0) & np.all(hits["normal"] == (0, 1, 0), axis=1)
    assert top.sum() > 3000
    d = -np.asarray(s["sun_dir"], np.float64)
    ia, _ = diffuse(0.0, s["sun_color"])
    LN = np.float32(np.float32(s["intensity"]) * np.float32(
        (np.float32(-s["sun_dir"][0]) * np.float32(0.0) + np.float32(-s["sun_dir"][1]) * np.float32(1.0))
        + np.float32(-s["sun_dir"][2]) * np.float32(0.0)))
    _, lit = diffuse(LN, s["sun_color"])
    agree, total, shadowed = 0, 0, 0
    for k in np.flatnonzero(top):
        o = diag["p"][k].astype(np.float64) - np.array([0.0, -4e-2, 0.0])   # hit.p - 4e-2 * step * mask
        inner = slab_hit(o, d, np.array(occ[0], float) + 1, np.array(occ[1], float) - 1)
        outer = slab_hit(o, d, np.array(occ[0], float) - 1, np.array(occ[1], float) + 1)
        if inner != outer:
            continue  # within a voxel of the occluder's silhouette
        total += 1
        shadowed += inner
        agree += np.array_equal(c[k], ia if inner else lit)
    assert shadowed > 300 and total - shadowed > 300, (shadowed, total)
    assert agree == total


def test_glossy_floor_reflects_to_the_exit_wall(oracle_mod):
    """Unlit glossy floor: colour = mix(BASE, walls(N'), 0.9), where the reflected
    ray starts at hit.p + 0.04 y, runs along the mirror image of the view ray and
    leaves the bound box through the face whose normal is N' (the traversal's
    own exit, pinned by the traversal tests: a nudge can land exactly on +-4096,
    which the strict bound test lets through to the next face)."""
    lay = box_vdb(oracle_mod, [((0, 64, 0), (512, 128, 512))])
    s = wx.shade(4, sun_dir=(0.0, 1.0, 0.0))  # sunlight travelling up: the floor top is unlit (I = 0)
    cam = wx.camera_basis((256.0, 400.0, -300.0), (256.0, 128.0, 256.0), (0, 1, 0), 1.0, 40.0, 64.0)
    c, hits, diag = render(lay, cam, 64, 64, s)
    top = (hits["state"] == 0) & np.all(hits["normal"] == (0, 1, 0), axis=1)
    assert top.sum() > 1000
    W = np.float32(0.1)
    walls = {  # walls(N) of reflect_ray2 (raycast.comp.wgsl:262-274) for the exit faces
        (-1, 0, 0): np.array([W, W, 0], np.float32), (0, -1, 0): np.array([0, W, W], np.float32),
        (0, 0, -1): np.array([W, 0, W], np.float32), (1, 0, 0): np.array([W, 0, 0], np.float32),
        (0, 0, 1): np.array([0, 0, W], np.float32)}
    m0 = np.array([0.4, 0.2, 0.2], np.float32)  # BASE + 0 * sun
    faces = set()
    for k in np.flatnonzero(top):
        d = camera_dir(cam, k % 64, k // 64)
        r = np.array([d[0], -d[1], d[2]], np.float32)  # mirror image about the y = const floor
        r = (r / np.sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2])).astype(np.float32)
        o = diag["p"][k].copy()
        o[1] = np.float32(o[1] + np.float32(0.04))
        h, _ = lay.hdda_ray(o, r)
        assert h["state"] == 1
        n = tuple(int(v) for v in h["normal"])
        faces.add(n)
        want = (m0 * (F(1.0) - F(0.9)) + walls[n] * F(0.9)).astype(np.float32)
        assert np.array_equal(c[k], want), (k, c[k], want, n)
    assert len(faces) >= 2


def camera_dir(cam, px, py, off=0.5):
    a = cam.as_array().astype(np.float32)
    u, mv, wp = a[4:7], a[8:11], a[12:15]
    fx, fy = np.float32(px) + np.float32(off), np.float32(py) + np.float32(off)
    v = (fx * u + fy * mv) + wp
    return v / np.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
