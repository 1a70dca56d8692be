"""Synthetic workloads: grids, cameras and ray sets of the benchmark configurations.

Every generator is seeded and pure (same seed -> same bytes), so the product and
the oracle build their grids from identical inputs.
"""
from __future__ import annotations

import math

import numpy as np

DOMAIN = 4096
LIGHT = (8000.0, 8000.0, 8000.0)


# ---------------- grids ----------------

def single_sphere():
    """Config 1: one solid sphere, radius 100 voxels, centred at (128, 128, 128)."""
    return np.array([[128.0, 128.0, 128.0, 100.0]], np.float64)


def sphere_set(seed: int = 1, count: int = 2000, domain: int = DOMAIN, rmin: float = 4.0, rmax: float = 40.0):
    """Config 2: centres uniform in [0, domain)^3, radii uniform in [rmin, rmax] (f64)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, float(domain), size=(count, 3))
    r = rng.uniform(rmin, rmax, size=(count, 1))
    return np.concatenate([c, r], axis=1).astype(np.float64)


def clip_box(domain: int = DOMAIN):
    return (0, 0, 0), (domain, domain, domain)


def build_sphere_vdb(spheres, lo=None, hi=None, threads=None):
    """Product-side grid build (woxel_amd.Vdb345) from a sphere list."""
    from . import Vdb345

    v = Vdb345()
    if lo is None:
        v.add_spheres(spheres, threads=threads)
    else:
        v.add_spheres(spheres, lo, hi, threads=threads)
    v.compute_sdf(threads=threads)
    return v


# ---------------- cameras ----------------

def config1_camera():
    """Config 1: pinhole at (400, 300, -200) looking at (128, 128, 128), 60 deg vfov, 640x480."""
    return dict(eye=(400.0, 300.0, -200.0), target=(128.0, 128.0, 128.0), up=(0.0, 1.0, 0.0),
                aspect=640.0 / 480.0, fovy_deg=60.0, width=640, height=480)


def orbit_camera(frame: int = 0, frames: int = 64, radius: float = 6000.0, width: int = 1920, height: int = 1080,
                 elevation_deg: float = 20.0, fovy_deg: float = 45.0, domain: int = DOMAIN):
    """Config 2: a pinhole orbiting the domain centre at `radius`. Frame k of an
    orbit of `frames` frames sits at azimuth 2*pi*k/frames and the given
    elevation; vfov 45 deg is the reference's default camera (camera.rs:27)."""
    c = domain / 2.0
    az = 2.0 * math.pi * frame / frames
    el = math.radians(elevation_deg)
    eye = (c + radius * math.cos(el) * math.cos(az), c + radius * math.sin(el), c + radius * math.cos(el) * math.sin(az))
    return dict(eye=eye, target=(c, c, c), up=(0.0, 1.0, 0.0), aspect=width / height, fovy_deg=fovy_deg,
                width=width, height=height)


def basis(cam_spec):
    """wx_camera words for a camera spec (product implementation)."""
    from . import camera_basis

    return camera_basis(cam_spec["eye"], cam_spec["target"], cam_spec["up"], cam_spec["aspect"],
                        cam_spec["fovy_deg"], float(cam_spec["width"]))


# ---------------- ray sets ----------------

def incoherent_rays(n: int, seed: int = 7, domain: int = DOMAIN):
    """Config 2, second pass: origins uniform on the bounding sphere of the domain
    (centre domain/2, radius domain/2*sqrt(3)), directions uniform on S^2
    (normalised Gaussian triples). Returns float32 (n, 6)."""
    rng = np.random.default_rng(seed)
    c = domain / 2.0
    R = domain / 2.0 * math.sqrt(3.0)
    o = rng.standard_normal((n, 3))
    o /= np.linalg.norm(o, axis=1, keepdims=True)
    o = c + R * o
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    # keep every direction component away from +-0 (a -0 component is a degenerate ray)
    dd = rays[:, 3:]
    tiny = np.abs(dd) < 1e-7
    dd[tiny] = np.where(dd[tiny] < 0, -1e-7, 1e-7).astype(np.float32)
    return rays


def look_at_spec(eye, target, width, height, fovy_deg=45.0):
    """A camera spec looking from eye at target; up is +y unless nearly parallel."""
    f = np.asarray(target, float) - np.asarray(eye, float)
    f /= np.linalg.norm(f)
    up = (0.0, 1.0, 0.0) if abs(f[1]) < 0.99 else (0.0, 0.0, 1.0)
    return dict(eye=tuple(float(x) for x in eye), target=tuple(float(x) for x in target), up=up,
                aspect=width / height, fovy_deg=fovy_deg, width=int(width), height=int(height))


def batch_cameras(n: int = 64, seed: int = 3, radius: float = 6000.0, domain: int = DOMAIN,
                  lo=(256, 256), hi=(3840, 2160), fovy_deg: float = 45.0):
    """Config 3: n cameras with poses uniform on a sphere of `radius` around the
    domain centre, looking at the centre, resolutions log-uniform between lo and
    hi (one uniform draw u per camera: W = lo_w (hi_w/lo_w)^u, H likewise)."""
    rng = np.random.default_rng(seed)
    c = domain / 2.0
    out = []
    for _ in range(n):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        u = rng.uniform()
        w = int(round(lo[0] * (hi[0] / lo[0]) ** u))
        h = int(round(lo[1] * (hi[1] / lo[1]) ** u))
        out.append(look_at_spec(c + radius * v, (c, c, c), w, h, fovy_deg))
    return out


# ---------------- configs 4 and 5: many resident grids ----------------

def config4_specs(seed: int = 4, n_spheres: int = 16, n_noise: int = 16, noise_max: int = 4096):
    """Config 4 grid list: sphere sets of 200-5000 spheres (radii U[2,60]) in
    domains of side log-uniform in [1024, 4096], and value-noise isosurfaces in
    domains of side log-uniform in [512, noise_max] thresholded to U[1%, 20%]
    occupancy. A smaller noise_max keeps a host-only build short."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n_spheres):
        side = int(round(1024 * 4 ** rng.uniform() / 8)) * 8
        count = int(rng.integers(200, 5001))
        specs.append(dict(kind="spheres", side=side, count=count, seed=int(rng.integers(1 << 30))))
    for i in range(n_noise):
        side = int(round(512 * (noise_max / 512) ** rng.uniform() / 8)) * 8
        occ = float(rng.uniform(0.01, 0.20))
        specs.append(dict(kind="noise", side=side, occupancy=occ, seed=int(rng.integers(1 << 30))))
    return specs


def build_grid(spec, threads=None, device=False):
    """Product-side Vdb345 for one config-4 grid spec. device=True decides the noise
    voxels and replays the SDF on the current GPU (the same grid)."""
    from . import Vdb345

    side = spec["side"]
    v = Vdb345()
    if spec["kind"] == "spheres":
        sp = sphere_set(spec["seed"], count=spec["count"], domain=side, rmin=2.0, rmax=60.0)
        v.add_spheres(sp, (0, 0, 0), (side, side, side), threads=threads)
    else:
        v.add_noise(spec["seed"], (0, 0, 0), (side, side, side), 8.0 / side, spec["occupancy"], threads=threads,
                    device=device)
    v.compute_sdf(threads=threads, device=device)
    return v


def layout_key(spec) -> str:
    """Layout-cache key of a grid spec: what builds it (the spec, with its seed, as
    canonical JSON) and the library version whose builder made the layout."""
    import json

    from . import lib

    return f"woxel grid layout, library {int(lib.wx_version())}: " + json.dumps(spec, sort_keys=True)


def cached_grid(spec, cache_dir: str, threads=None, device=True, bound: float = 4096.0):
    """The device grid of `spec`: loaded from cache_dir when a layout saved under its
    key is there, else built (build_grid) and saved there. Returns (grid, "loaded" or
    "built")."""
    import hashlib
    import os

    from . import Grid, probe_layout

    key = layout_key(spec)
    path = os.path.join(cache_dir, hashlib.sha256(key.encode()).hexdigest()[:24] + ".wxgrid")
    if probe_layout(path, key) is not None:
        return Grid.load(path, key), "loaded"
    v = build_grid(spec, threads=threads, device=device)
    g = Grid(v, bound)
    del v
    os.makedirs(cache_dir, exist_ok=True)
    g.save(path, key)
    return g, "built"


def config4_cameras(specs, n: int = 256, seed: int = 5, lo=(128, 128), hi=(1920, 1080)):
    """Config 4 requests: n cameras, each bound to a uniformly drawn grid, on a sphere
    of radius 1.46 x side around the grid's centre (6000 / 4096 as in config 2),
    resolutions log-uniform between lo and hi. Returns [(grid index, camera spec)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        gi = int(rng.integers(len(specs)))
        side = specs[gi]["side"]
        c = side / 2.0
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        u = rng.uniform()
        w = int(round(lo[0] * (hi[0] / lo[0]) ** u))
        h = int(round(lo[1] * (hi[1] / lo[1]) ** u))
        out.append((gi, look_at_spec(c + 1.46 * side * v, (c, c, c), w, h)))
    return out


SERVE_RESOLUTIONS = ((512, 512), (1024, 1024), (1920, 1080), (3840, 2160))


def serve_requests(specs, n: int, rate: float, seed: int = 6):
    """Config 5 stream: Poisson arrivals at `rate` per second; each request picks a
    uniform grid, an eye uniform on the grid's bounding sphere (centre side/2,
    radius side*sqrt(3)/2) looking at the centre, and a resolution from
    SERVE_RESOLUTIONS. Returns [(arrival time s, grid index, camera spec)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        gi = int(rng.integers(len(specs)))
        side = specs[gi]["side"]
        c = side / 2.0
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        w, h = SERVE_RESOLUTIONS[int(rng.integers(len(SERVE_RESOLUTIONS)))]
        out.append((t, gi, look_at_spec(c + side * math.sqrt(3.0) / 2.0 * v, (c, c, c), w, h)))
    return out
