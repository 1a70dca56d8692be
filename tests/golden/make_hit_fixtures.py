"""Freeze small outputs of the oracle as fixtures: hit records and shaded pixels
on the config-1 sphere, so that a change to the oracle (or to the product, which
must match them bit for bit) shows up against committed bytes.

* cfg1_cam_64x48_off0.5.npy / _off0.001.npy: config-1 camera (eye (400,300,-200),
  target (128,128,128), vfov 60), 64x48 pixels, pixel offset 0.5 and the
  reference's 0.001 (raycast.comp.wgsl:62); 32-byte hit records.
* cfg1_rays_512.npy / cfg1_rays_512_in.npy: 512 seeded rays (origins in
  [-60, 320)^3, directions on S^2, 64 of them axis-aligned, 64 exact-tie
  diagonals from voxel centres) and their hit records, no clip.
* cfg1_render_32x24_mode3.npy / _mode4.npy: rgba8 words of render modes 3 and 4
  with the reference's default sun.

Run from the repository root: python tests/golden/make_hit_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

DST = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(DST)))

import oracle  # noqa: E402
import woxel_amd as wx  # noqa: E402  (camera basis, shade defaults, seeded scenes)
from woxel_amd import scenes  # noqa: E402


def cfg1_tree():
    o = oracle.Vdb()
    sp = scenes.single_sphere()
    o.add_sphere(sp[0, :3], sp[0, 3])
    o.compute_sdf()
    return o


def fixture_rays():
    rng = np.random.default_rng(20)
    n = 512
    o = rng.uniform(-60.0, 320.0, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.concatenate([o, d], 1).astype(np.float32)
    r[:64, 3:] = 0.0
    r[np.arange(64), 3 + np.arange(64) % 3] = np.where(np.arange(64) % 2, 1.0, -1.0)
    r[64:128, 3:] = np.float32(1 / np.sqrt(3))
    r[64:128, :3] = np.floor(r[64:128, :3]) + 0.5
    return r


def main():
    o = cfg1_tree()
    lay = o.layout()
    cam = wx.camera_basis((400.0, 300.0, -200.0), (128.0, 128.0, 128.0), (0, 1, 0), 64 / 48, 60.0, 64.0)
    files = {}
    for off in (0.5, 0.001):
        hits, _ = lay.trace_camera(cam.as_array(), 64, 48, off)
        files[f"cfg1_cam_64x48_off{off}.npy"] = hits.view(np.uint8).reshape(-1, 32)
    rays = fixture_rays()
    hits, _ = lay.trace_rays(rays)
    files["cfg1_rays_512_in.npy"] = rays
    files["cfg1_rays_512.npy"] = hits.view(np.uint8).reshape(-1, 32)
    cam_r = wx.camera_basis((400.0, 300.0, -200.0), (128.0, 128.0, 128.0), (0, 1, 0), 32 / 24, 60.0, 32.0)
    for mode in (3, 4):
        _, rgba8 = lay.render_camera(cam_r.as_array(), 32, 24, wx.shade(mode), 0.5)
        files[f"cfg1_render_32x24_mode{mode}.npy"] = np.ascontiguousarray(rgba8).view(np.uint32).reshape(-1)
    meta = {}
    for name, arr in files.items():
        path = os.path.join(DST, name)
        np.save(path, arr, allow_pickle=False)
        meta[name] = hashlib.sha256(open(path, "rb").read()).hexdigest()
    json.dump(meta, open(os.path.join(DST, "hit_fixtures.json"), "w"), indent=1)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
