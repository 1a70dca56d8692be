#!/usr/bin/env python3
"""Benchmark: primary-ray throughput of the HDDA + grid-SDF traversal on MI355X.

Workload (default): the 4096^3 sparse sphere set (2000 solid spheres, centres
uniform in [0,4096)^3, radii uniform in [4,40], seed 1, clipped to the domain,
grid SDF rebuilt) rendered by a 1920x1080 pinhole camera orbiting the domain
centre at radius 6000 (elevation 20 deg, vfov 45 deg, 64-frame orbit). One step
= one frame = one launch of the traversal kernel over 2,073,600 primary rays;
step k renders orbit frame k. Inputs (grid, camera words) are resident on the
GPU before the timed region; each frame writes its own 32-byte hit records.

Prints ONE JSON line (rank 0). See DESIGN.md for the roofline accounting.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(hits_u8, ray_in_bytes: int):
    """Per-launch algorithmic bytes: sum over rays of 64 B x (3 L5 + 3 L4 + 2 L3)
    + ray-in bytes + 32 B hit record out (L5/L4/L3 = node-5/4/3 lookups)."""
    import torch

    lk = hits_u8.view(torch.int32).reshape(-1, 8)[:, 7].to(torch.int64)
    l5 = lk & 1023
    l4 = (lk >> 10) & 1023
    l3 = (lk >> 20) & 1023
    words = (3 * l5 + 3 * l4 + 2 * l3).sum().item()
    n = lk.numel()
    return 64 * words + n * (ray_in_bytes + 32), 4 * words + n * (ray_in_bytes + 32)


def measured_traffic():
    """HBM bytes per launch of the frame kernel from the committed rocprofv3 PMC
    summary, used only while the kernel source is the one that was profiled."""
    import glob
    import hashlib

    src = os.path.join(ROOT, "woxel_amd", "csrc", "trace.hip")
    sha = hashlib.sha256(open(src, "rb").read()).hexdigest()[:16]
    for f in glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json")):
        rec = json.load(open(f))
        if rec.get("trace_hip_sha256_prefix") == sha and "1920x1080" in rec.get("workload", ""):
            return rec["bytes_per_launch"], os.path.relpath(f, ROOT)
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--threads", type=int, default=0, help="host threads for build / CPU baseline")
    ap.add_argument("--simple", action="store_true", help="one-lane-per-ray launch shape instead of persistent")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default persistent launch shape")
    ap.add_argument("--tune", type=str, default="",
                    help="persistent,refill,chunk,waves_per_cu,xcd_split (skips autotune)")
    ap.add_argument("--workload", choices=["frame", "batch"], default="frame",
                    help="frame: one 1920x1080 frame per step (the headline); batch: 64 cameras + shadow rays "
                         "in one launch vs one launch per request")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 10 plain device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n, "launches": 10}), flush=True)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    if not torch.cuda.is_available():
        log("bench.py: no HIP device visible")
        sys.exit(1)
    threads = args.threads or min(16, os.cpu_count() or 1)

    if args.workload == "batch":
        return bench_batch(args, threads)

    # ---- workload: grid + cameras (untimed) ----
    t0 = time.time()
    spheres = scenes.sphere_set(args.seed)
    lo, hi = scenes.clip_box()
    vdb = scenes.build_sphere_vdb(spheres, lo, hi, threads=threads)
    grid = wx.Grid(vdb)
    n5, n4, n3 = vdb.count_nodes()
    log(f"grid built in {time.time() - t0:.1f}s: N5={n5} N4={n4} N3={n3} "
        f"active={vdb.active_voxels()} device={grid.nbytes() / 2**20:.1f} MiB")
    W, H = args.width, args.height
    nframes = args.warmup + args.steps
    cams = [scenes.basis(scenes.orbit_camera(k % 64, width=W, height=H)) for k in range(nframes)]
    outs = [torch.empty((W * H, 32), dtype=torch.uint8, device="cuda") for _ in range(args.steps)]
    warm = torch.empty((W * H, 32), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    tune = None
    if args.tune:
        f = [int(x) for x in args.tune.split(",")]
        tune = wx.Tune(*f)
        wx.tune_set(tune)
    elif not args.simple and not args.no_autotune:
        t1 = time.time()
        tune, tms = wx.autotune([grid], [wx.camera_request(0, cams[0], W, H, warm, 0.5, True)], stream=stream)
        log(f"autotune {time.time() - t1:.1f}s: {tune.as_dict()} {tms:.3f} ms")
    for k in range(args.warmup):
        grid.trace_camera(cams[k], W, H, 0.5, clip=True, out=warm, stream=stream, simple=args.simple)
    torch.cuda.synchronize()

    # ---- timed region: exactly K launches, events on the launch stream ----
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        grid.trace_camera(cams[args.warmup + k], W, H, 0.5, clip=True, out=outs[k], stream=stream,
                          simple=args.simple)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    launch_ms = np.array([a.elapsed_time(b) for a, b in ev])
    ms_per_step = wall * 1e3 / args.steps
    rays = W * H
    value = rays * args.steps / wall

    # ---- roofline: algorithmic bytes per launch / mean launch duration ----
    alg = [algorithmic_bytes(o, 0) for o in outs]
    bytes_per_launch = float(np.mean([a[0] for a in alg]))
    useful_per_launch = float(np.mean([a[1] for a in alg]))
    mean_launch_s = float(launch_ms.mean()) / 1e3
    achieved = bytes_per_launch / mean_launch_s
    hits = sum(int((o.view(torch.int32).reshape(-1, 8)[:, 6] & 0xFF).eq(0).sum().item()) for o in outs)

    one_lane = args.simple or (tune or wx.tune_get()).persistent == 0  # the profiled launch shape
    traffic, traffic_src = measured_traffic() if one_lane and (W, H) == (1920, 1080) else (None, None)

    # ---- CPU baseline: the oracle's traversal on the host cores ----
    cpu = None
    if not args.no_cpu:
        import oracle

        k5, v5, k4, v4, v3 = vdb.masks()
        a5, a4, a3 = vdb.atlas()
        lay = oracle.Layout.from_arrays(vdb.origins(), k5, v5, k4, v4, v3, a5, a4, a3)
        del a3, a4, a5
        done, frames, t = 0, 0, time.perf_counter()
        while time.perf_counter() - t < args.cpu_seconds and frames < 64:
            cam = cams[frames % nframes].as_array()
            lay.trace_camera(cam, W, H, 0.5, clip=True, threads=threads)
            done += W * H
            frames += 1
        cpu_s = time.perf_counter() - t
        cpu = {"value": done / cpu_s / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": f"{frames} full {W}x{H} orbit frames (oracle C traversal, {threads} threads, "
                         f"{cpu_s:.1f}s)"}
        lay.close()

    line = {
        "metric": "Grays/s primary rays, 1920x1080 on 4096^3 sphere-set VDB",
        "value": value / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"sparse sphere set 4096^3 (2000 spheres r in [4,40], seed {args.seed}) + grid SDF; "
                        f"{W}x{H} pinhole orbiting at r=6000 (elev 20 deg, vfov 45 deg), primary rays",
            "rays_per_step": rays,
            "nodes": [n5, n4, n3],
            "hit_fraction": hits / (rays * args.steps),
            "launch": "simple" if args.simple else (tune or wx.tune_get()).as_dict(),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved / 1e9,
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel_ms": float(launch_ms.mean()),
            "alg_bytes_per_launch": bytes_per_launch,
            "useful_4B_frac": useful_per_launch / mean_launch_s / HBM_PEAK,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_batch(args, threads):
    """64 cameras (poses on the r=6000 sphere, log-uniform 256^2..3840x2160) with
    one shadow ray per primary hit toward (8000, 8000, 8000): all requests in one
    launch, against one launch per request, identical outputs."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    spheres = scenes.sphere_set(args.seed)
    lo, hi = scenes.clip_box()
    vdb = scenes.build_sphere_vdb(spheres, lo, hi, threads=threads)
    grid = wx.Grid(vdb)
    specs = scenes.batch_cameras(64)
    stream = torch.cuda.current_stream()
    reqs, outs = [], []
    for sp in specs:
        w, h = sp["width"], sp["height"]
        o = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        s = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        reqs.append(wx.camera_request(0, scenes.basis(sp), w, h, o, 0.5, True, shadow_out=s, light=scenes.LIGHT))
        outs.append((o, s))
    pixels = sum(sp["width"] * sp["height"] for sp in specs)
    log(f"batch: {len(specs)} cameras, {pixels} pixels")
    tune, tms = wx.autotune([grid], reqs, stream=stream)
    log(f"autotune: {tune.as_dict()} {tms:.3f} ms")

    def run_batched():
        wx.trace_batch([grid], reqs, tune=tune, stream=stream)

    def run_single():
        for r in reqs:
            wx.trace_batch([grid], [r], tune=tune, stream=stream)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t_single = timed(run_single, args.steps)
    ref = [(o.clone(), s.clone()) for o, s in outs]
    t_batch = timed(run_batched, args.steps)
    same = all(torch.equal(a, c) and torch.equal(b, d) for (a, b), (c, d) in zip(ref, outs))
    shadow_rays = sum(int((o.view(torch.int32).reshape(-1, 8)[:, 6] & 0xFF).eq(0).sum().item()) for o, _ in outs)
    rays = pixels + shadow_rays
    line = {
        "metric": "Grays/s primary + shadow rays, 64 cameras in one launch (config 3)",
        "value": rays / t_batch / 1e9, "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "64 cameras on the r=6000 sphere around the 4096^3 sphere set, log-uniform "
                               "256x256..3840x2160, one shadow ray per hit toward (8000,8000,8000)",
                   "primary_rays": pixels, "shadow_rays": shadow_rays, "tune": tune.as_dict()},
        "per_request_launches": {"value": rays / t_single / 1e9, "ms": t_single * 1e3},
        "speedup_vs_per_request": t_single / t_batch,
        "identical_output": same,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
