#!/usr/bin/env python3
"""Benchmark: primary-ray throughput of the HDDA + grid-SDF traversal on MI355X.

Workload (default): the 4096^3 sparse sphere set (2000 solid spheres, centres
uniform in [0,4096)^3, radii uniform in [4,40], seed 1, clipped to the domain,
grid SDF rebuilt) rendered by a 1920x1080 pinhole camera orbiting the domain
centre at radius 6000 (elevation 20 deg, vfov 45 deg, 64-frame orbit). One step
= one frame = one launch of the traversal kernel over 2,073,600 primary rays;
step k renders orbit frame k. Inputs (grid, camera words) are resident on the
GPU before the timed region; each frame writes its own 32-byte hit records.

Prints ONE JSON line (rank 0). The traversal is a chain of dependent lookups and
~120 FP32/integer VALU instructions per step, with the gather working set served from
the caches, so the bound it runs against is VALU issue: roofline.achieved is wave64
VALU instructions per launch (the rocprofv3 counter SQ_INSTS_VALU of this kernel
source, committed under profiles/) divided by the live mean kernel time, against the
issue peak of 1024 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction. roofline.hbm is
the measured HBM traffic (FETCH_SIZE, WRITE_SIZE, calibrated against a plain copy) over
the same time against 8 TB/s. roofline.effective_gather keeps the gather model
64 B x (3 L5 + 3 L4 + 2 L3) + 32 B per ray (node-5/4/3 lookups of the reference
algorithm, counted in every hit record), which exceeds HBM peak because the lookups
hit the caches; it is a throughput figure, not a bound. latency_ms gives the
p50 / p99 / max of the per-frame device time (the hipEvent span of each launch).
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
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD every 4 cycles
# (16 lanes wide) at the 2.4 GHz peak engine clock
SIMDS, CLOCK_HZ, CYCLES_PER_WAVE64_VALU = 1024, 2.4e9, 4
VALU_PEAK = SIMDS * CLOCK_HZ / CYCLES_PER_WAVE64_VALU  # wave64 VALU instructions / s
# the fastest issue cost measured on gfx950 (v_mov / v_and repeated, profiles/issue_rate_gfx950.md)
FASTEST_ISSUE_CYCLES = 2.15


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


INCOHERENT_SHAPE = "1,16,256,24,1"  # persistent, refill, chunk, waves per CU, XCD split


def kernel_sources_sha():
    import hashlib

    h = hashlib.sha256()
    for f in ("trace.hip", "lane.hpp", "request.hpp"):
        h.update(open(os.path.join(ROOT, "woxel_amd", "csrc", f), "rb").read())
    return h.hexdigest()[:16]


def counter_record(scene="frame", size="1920x1080", tune=None):
    """Per-dispatch rocprofv3 counters of the dominant kernel for this scene, size and
    launch shape (profiles/counters_*.json), used only while the kernel sources are
    the ones that were profiled. Returns (record, path) or (None, None)."""
    import glob

    sha = kernel_sources_sha()
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "counters_*.json"))):
        rec = json.load(open(f))
        if (rec.get("kernel_sources_sha256_prefix") == sha and rec.get("scene") == scene
                and size in rec.get("workload", "") and rec.get("tune", tune) == tune):
            return rec, os.path.relpath(f, ROOT)
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
    ap.add_argument("--no-batch", action="store_true",
                    help="frame workload: skip the config-3 batched-vs-per-request comparison after the timed region")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--threads", type=int, default=0, help="host threads for build / CPU baseline")
    ap.add_argument("--simple", action="store_true", help="one-lane-per-ray launch shape instead of persistent")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default persistent launch shape")
    ap.add_argument("--tune", type=str, default="",
                    help="persistent,refill,chunk,waves_per_cu,xcd_split (skips autotune)")
    ap.add_argument("--workload", choices=["frame", "config1", "incoherent", "batch", "multigrid", "serve", "render",
                                           "build"],
                    default="frame",
                    help="frame: one 1920x1080 frame per step (the headline); render: shaded 1920x1080 "
                         "frames (--mode); config1: the 640x480 config-1 "
                         "frame; incoherent: 16M incoherent rays per step; batch: 64 cameras + shadow rays "
                         "in one launch vs one launch per request; multigrid: 256 cameras over 32 resident "
                         "grids in one launch vs one launch per request; serve: Poisson request stream "
                         "through the serving loop vs plain per-request serving")
    ap.add_argument("--mode", type=int, default=3, help="render: 0 gray, 1 rgb, 2 ray length, 3 diffuse, 4 glossy")
    ap.add_argument("--layout-cache", type=str, default="",
                    help="multigrid / serve: load the 32 grids from layout files in this directory, building "
                         "and saving the missing ones; build: where the save/load timing writes its file")
    ap.add_argument("--noise-max", type=int, default=4096,
                    help="largest noise-isosurface domain side of the 32-grid set (4096: the configuration's)")
    ap.add_argument("--rates", type=str, default="50,100,200,500,1000,2000",
                    help="serve: Poisson arrival rates (requests/s)")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="serve: arrival window per rate")
    ap.add_argument("--budget-gib", type=float, default=24.0, help="serve: HBM budget for request outputs")
    ap.add_argument("--streams", type=int, default=3, help="serve: launch slots (HIP streams)")
    ap.add_argument("--servers", type=str, default="batched,plain,resident",
                    help="serve: which servers to run (batched always runs)")
    ap.add_argument("--max-batch-rays", type=int, default=1 << 24, help="serve: primary rays per batched launch")
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
    threads = args.threads or host_threads()

    if args.workload == "batch":
        return bench_batch(args, threads)
    if args.workload == "multigrid":
        specs, grids = build_grid_set(args, threads)
        return bench_multigrid(args, specs, grids)
    if args.workload == "serve":
        specs, grids = build_grid_set(args, threads)
        return bench_serve(args, specs, grids)

    if args.workload == "incoherent":
        return bench_incoherent(args, threads)
    if args.workload == "render":
        return bench_render(args, threads)
    if args.workload == "build":
        return bench_build(args, threads)
    return bench_frames(args, threads, args.workload)


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_layout(vdb):
    """The oracle's reference-format layout built from the product tree's exports
    (identical to the oracle's own build: checked by the layout parity tests)."""
    import oracle

    k5, v5, k4, v4, v3 = vdb.masks()
    a5, a4, a3 = vdb.atlas()
    return oracle.Layout.from_arrays(vdb.origins(), k5, v5, k4, v4, v3, a5, a4, a3)


def roofline(outs, launch_ms, ray_in_bytes, rec, rec_src):
    """Bounds of the dominant kernel over its live mean launch time: VALU issue (from
    the kernel's SQ_INSTS_VALU per launch) and HBM bytes (FETCH_SIZE / WRITE_SIZE per
    launch), the larger fraction naming the bound; plus the gather model's effective
    GB/s and the ray iterations per second (both from the hit records)."""
    import torch

    kernel_s = float(np.mean(launch_ms)) / 1e3
    alg = [algorithmic_bytes(o, ray_in_bytes) for o in outs]
    gather = float(np.mean([a[0] for a in alg]))
    useful = float(np.mean([a[1] for a in alg]))  # 4 B per gathered word instead of a 64-B request
    def iterations(o):  # sum of the four 16-bit step counters of every record = iterations
        w = o.view(torch.int32).reshape(-1, 8)[:, 4:6].to(torch.int64) & 0xFFFFFFFF
        return int(((w & 0xFFFF) + (w >> 16)).sum().item())

    iters = float(np.mean([iterations(o) for o in outs]))
    out = {"bound": None, "achieved": None, "peak": VALU_PEAK / 1e9, "unit": "G wave64 VALU instr/s",
           "frac": None, "traffic": None, "kernel_ms": kernel_s * 1e3, "counters_source": rec_src,
           "ray_iterations_per_s": iters / kernel_s,
           "effective_gather": {"GBps": gather / kernel_s / 1e9, "alg_bytes_per_launch": gather,
                                "useful_GBps": useful / kernel_s / 1e9, "useful_bytes_per_launch": useful,
                                "model": "64 B x (3 L5 + 3 L4 + 2 L3) + ray in + 32 B out per ray; lookups hit the "
                                         "caches, so this is a throughput figure, not a bound"}}
    if rec is None:
        return out
    c = rec["per_dispatch"]
    valu = float(c["SQ_INSTS_VALU"])
    out["achieved"] = valu / kernel_s / 1e9
    out["frac"] = valu / kernel_s / VALU_PEAK
    out["issue_model"] = ("4 cycles per wave64 VALU instruction; gfx950 issues 2.2-4.3 by instruction and mix "
                          "(profiles/issue_rate_gfx950.md), so frac overstates VALU saturation")
    # a ceiling no instruction mix exceeds: every instruction at the fastest measured issue cost
    out["frac_at_fastest_issue"] = valu * FASTEST_ISSUE_CYCLES / (SIMDS * CLOCK_HZ) / kernel_s
    util = rec.get("derived", {}).get("VALUUtilization")
    if util is not None:
        out["valu_utilization"] = util / 100.0
        out["lane_frac"] = out["frac"] * util / 100.0
    out["valu_insts_per_launch"] = valu
    traffic = float(c["FETCH_SIZE"]) * 1024 * rec["fetch_scale"] + float(c["WRITE_SIZE"]) * 1024 * rec["write_scale"]
    out["traffic"] = traffic
    out["hbm"] = {"achieved": traffic / kernel_s / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": traffic / kernel_s / HBM_PEAK}
    out["bound"] = "valu" if out["frac"] >= out["hbm"]["frac"] else "hbm"
    return out


def pipelined_frames(grid, cams, W, H, clip, ref_outs, simple, nstreams=2, min_frames=60):
    """The timed frames again with `nstreams` frames in flight, frame i on stream i % K %
    nstreams (default priority), as a renderer that submits frame k+1 before frame k
    completes: the next frame's waves fill the CUs that the previous frame's last waves
    leave idle. The K frames are cycled to at least `min_frames` launches; buffer b is
    written only on stream b % nstreams (never by two streams at once) and ends holding
    frame b, which must equal the single-stream frame b byte for byte."""
    import torch

    K = len(ref_outs)
    if K < nstreams:
        return None
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    outs = [torch.empty_like(o) for o in ref_outs]
    for j, s in enumerate(streams):  # first launch on each stream (its scratch)
        grid.trace_camera(cams[j], W, H, 0.5, clip=clip, out=outs[j], stream=s, simple=simple)
    torch.cuda.synchronize()
    n = max(min_frames, K)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n)]
    t = time.perf_counter()
    for i in range(n):
        b = i % K
        s = streams[b % nstreams]
        ev[i][0].record(s)
        grid.trace_camera(cams[b], W, H, 0.5, clip=clip, out=outs[b], stream=s, simple=simple)
        ev[i][1].record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    same = all(torch.equal(a, b) for a, b in zip(outs, ref_outs))
    log(f"{nstreams} frames in flight: {wall * 1e3 / n:.4f} ms per frame, {W * H * n / wall / 1e9:.2f} Grays/s, "
        f"identical {same}")
    return {"value": W * H * n / wall / 1e9, "unit": "Grays/s", "streams": nstreams, "frames": n,
            "ms_per_frame": wall * 1e3 / n,
            # each frame's hipEvent span on its stream, time sharing the CUs with its neighbour included
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "max": float(lat.max())},
            "identical": bool(same)}


def chain_bound(grid, cam, W, H, clip, hits, kernel_ms, stream):
    """A frame cannot finish before its longest wave: the 64 rays of the 8x8 tile with
    the most iterations (from the frame's hit records), traced alone as one wave on
    the whole GPU, give that chain's time. Directions are computed on the host in the
    kernel's FP32 order (camera_ray), so the lone rays repeat the frame's iterations;
    the figure is reported against the frame's mean kernel time."""
    import torch

    import woxel_amd as wx

    w = hits.view(torch.int32).reshape(-1, 8)[:, 4:6].to(torch.int64) & 0xFFFFFFFF
    it = ((w & 0xFFFF) + (w >> 16)).sum(1).cpu().numpy().reshape(H, W)
    Ht, Wt = (H + 7) // 8, (W + 7) // 8
    pad = np.zeros((Ht * 8, Wt * 8), np.int64)
    pad[:H, :W] = it
    wmax = pad.reshape(Ht, 8, Wt, 8).max(axis=(1, 3))
    ty, tx = (int(v) for v in np.unravel_index(np.argmax(wmax), wmax.shape))
    c = cam.as_array().astype(np.float32)
    eye, u, mv, wp = c[0:3], c[4:7], c[8:11], c[12:15]
    rays, want = [], []
    for lane in range(64):
        px, py = tx * 8 + (lane & 7), ty * 8 + (lane >> 3)
        if px >= W or py >= H:
            continue
        fx, fy = np.float32(px) + np.float32(0.5), np.float32(py) + np.float32(0.5)
        v = ((fx * u + fy * mv) + wp).astype(np.float32)
        ln = np.float32(np.sqrt(np.float32((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])))
        rays.append(np.concatenate([eye, v / ln]).astype(np.float32))
        want.append(it[py, px])
    r = torch.from_numpy(np.stack(rays)).cuda()
    out = torch.empty((len(rays), 32), dtype=torch.uint8, device="cuda")
    miss = torch.tensor([[1e6, 1e6, 1e6, 1.0, 0.0, 0.0]], dtype=torch.float32, device="cuda")
    mout = torch.empty((1, 32), dtype=torch.uint8, device="cuda")

    def med(fn, reps):
        for _ in range(5):
            fn()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    lone = med(lambda: grid.trace_rays(r, clip=clip, out=out, stream=stream, simple=True), 100)
    empty = med(lambda: grid.trace_rays(miss, clip=True, out=mout, stream=stream, simple=True), 100)
    got = out.view(torch.int32).reshape(-1, 8)[:, 4:6].to(torch.int64) & 0xFFFFFFFF
    got_it = ((got & 0xFFFF) + (got >> 16)).sum(1).cpu().numpy()
    iters = int(wmax.max())
    return {"longest_wave_iterations": iters, "tile": [tx, ty], "same_iterations": bool(np.array_equal(got_it, want)),
            "lone_wave_ms": lone, "empty_launch_ms": empty,
            "per_iteration_cycles": (lone - empty) * 1e-3 * CLOCK_HZ / iters,
            "wave_iterations_p50": float(np.percentile(wmax, 50)), "wave_iterations_p99": float(np.percentile(wmax, 99)),
            "frac": lone / kernel_ms}


def host_threads():
    """Threads for the host legs: every core this process may run on (its CPU
    affinity), capped by OMP_NUM_THREADS when the environment sets one (a GPU host's
    share of cores for one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def host_desc(threads):
    return {"cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "threads_used": threads}


def pick_tune(args, grids, req, stream):
    import woxel_amd as wx

    if args.tune:
        tune = wx.Tune(*[int(x) for x in args.tune.split(",")])
        wx.tune_set(tune)
        return tune
    if args.simple or args.no_autotune:
        return None
    t1 = time.time()
    tune, tms = wx.autotune(grids, [req], stream=stream)
    log(f"autotune {time.time() - t1:.1f}s: {tune.as_dict()} {tms:.3f} ms")
    return tune


def bench_frames(args, threads, scene):
    """scene "frame": config-2 orbit frames (entry clip on); scene "config1": the
    config-1 sphere and camera at 640x480, reference bound semantics (no clip)."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    t0 = time.time()
    if scene == "config1":
        vdb = scenes.build_sphere_vdb(scenes.single_sphere(), threads=threads)
        W, H, clip = 640, 480, False
    else:
        lo, hi = scenes.clip_box()
        vdb = scenes.build_sphere_vdb(scenes.sphere_set(args.seed), lo, hi, threads=threads)
        W, H, clip = args.width, args.height, True
    grid = wx.Grid(vdb)
    n5, n4, n3 = vdb.count_nodes()
    log(f"grid built in {time.time() - t0:.1f}s: N5={n5} N4={n4} N3={n3} "
        f"active={vdb.active_voxels()} device={grid.nbytes() / 2**20:.1f} MiB")
    nframes = args.warmup + args.steps
    if scene == "config1":
        cams = [scenes.basis(scenes.config1_camera())] * nframes
    else:
        cams = [scenes.basis(scenes.orbit_camera(k % 64, width=W, height=H)) for k in range(nframes)]
    outs = [torch.empty((W * H, 32), dtype=torch.uint8, device="cuda") for _ in range(args.steps)]
    warm = torch.empty((W * H, 32), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    tune = pick_tune(args, [grid], wx.camera_request(0, cams[0], W, H, warm, 0.5, clip), stream)
    for k in range(args.warmup):
        grid.trace_camera(cams[k], W, H, 0.5, clip=clip, out=warm, stream=stream, simple=args.simple)
    torch.cuda.synchronize()

    # ---- timed region: exactly K launches, events on the launch stream ----
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        grid.trace_camera(cams[args.warmup + k], W, H, 0.5, clip=clip, out=outs[k], stream=stream,
                          simple=args.simple)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    launch_ms = np.array([a.elapsed_time(b) for a, b in ev])
    rays = W * H
    value = rays * args.steps / wall
    hits = sum(hit_count(o) for o in outs)

    one_lane = args.simple or (tune or wx.tune_get()).persistent == 0  # the profiled launch shape
    rec, rec_src = counter_record(scene, f"{W}x{H}") if one_lane else (None, None)
    roof = roofline(outs, launch_ms, 0, rec, rec_src)
    # the frame's chain bound: its longest 8x8 wave tile traced alone
    roof["chain"] = chain_bound(grid, cams[args.warmup], W, H, clip, outs[0], float(np.mean(launch_ms)), stream)
    roof["valu"] = {k: roof[k] for k in ("achieved", "peak", "unit", "frac", "issue_model", "frac_at_fastest_issue")
                    if k in roof}
    if roof["chain"]["frac"] > (roof["frac"] or 0.0):
        # the binding bound is the frame's longest dependent chain: frames/s achieved
        # against the rate that chain alone allows (its wave traced by itself)
        ch = roof["chain"]
        roof.update(bound="chain", achieved=1e3 / roof["kernel_ms"], peak=1e3 / ch["lone_wave_ms"],
                    unit="frames/s (peak: the longest 8x8 tile traced alone)", frac=ch["frac"])
    # after the timed region: the same frames with two in flight (the headline stays one stream)
    piped = pipelined_frames(grid, cams[args.warmup:], W, H, clip, outs, args.simple)
    if piped and roof.get("valu_insts_per_launch"):
        # the counter record's VALU instructions per frame over the pipelined wall time per frame;
        # above 1 in the 4-cycle model, which is no ceiling, against the fastest issue cost
        s = piped["ms_per_frame"] * 1e-3
        insts = roof["valu_insts_per_launch"]
        piped["valu"] = {"achieved": insts / s / 1e9, "unit": "G wave64 VALU instr/s",
                         "frac_4_cycle_model": insts / s / VALU_PEAK,
                         "frac_at_fastest_issue": insts * FASTEST_ISSUE_CYCLES / (SIMDS * CLOCK_HZ) / s}

    # ---- CPU baseline: the oracle's traversal on the host cores ----
    cpu = None
    if not args.no_cpu:
        lay = oracle_layout(vdb)
        done, frames, t = 0, 0, time.perf_counter()
        while time.perf_counter() - t < args.cpu_seconds and frames < 1000:
            lay.trace_camera(cams[frames % nframes].as_array(), W, H, 0.5, clip=clip, threads=threads)
            done += W * H
            frames += 1
        cpu_s = time.perf_counter() - t
        # one thread over the first rows of frame 0 (about a second of work)
        rows = max(1, min(H, int(H * (0.6e6 if scene == "frame" else 4e6) / (W * H))))
        t = time.perf_counter()
        lay.trace_camera(cams[0].as_array(), W, rows, 0.5, clip=clip, threads=1)
        one_s = time.perf_counter() - t
        cpu = {"value": done / cpu_s / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": f"{frames} full {W}x{H} frames (oracle C traversal, {threads} threads, {cpu_s:.1f}s); "
                         f"1 thread: {W}x{rows} rows of frame 0 in {one_s:.2f}s",
               "single_thread_value": W * rows / one_s / 1e9, **host_desc(threads)}
        lay.close()

    if scene == "config1":
        metric = "Grays/s primary rays, 640x480 on the config-1 sphere (r=100 at (128,128,128))"
        workload = ("one solid sphere r=100 voxels at (128,128,128) + grid SDF; 640x480 pinhole at "
                    "(400,300,-200), vfov 60 deg, pixel centres, reference bound semantics (no clip)")
    else:
        metric = "Grays/s primary rays, 1920x1080 on 4096^3 sphere-set VDB; p50/p99 latency"
        workload = (f"sparse sphere set 4096^3 (2000 spheres r in [4,40], seed {args.seed}) + grid SDF; "
                    f"{W}x{H} pinhole orbiting at r=6000 (elev 20 deg, vfov 45 deg), primary rays")
    line = {
        "metric": metric,
        "value": value / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "rays_per_step": rays,
            "nodes": [n5, n4, n3],
            "hit_fraction": hits / (rays * args.steps),
            "launch": "simple" if args.simple else (tune or wx.tune_get()).as_dict(),
        },
        # per-frame latency: the hipEvent span of each timed launch on its stream
        "latency_ms": {"p50": float(np.percentile(launch_ms, 50)), "p99": float(np.percentile(launch_ms, 99)),
                       "max": float(launch_ms.max())},
        "roofline": roof,
        "cpu_baseline": cpu,
        # the same frames, two in flight on two streams (a renderer's pipelined frames)
        "pipelined_2_streams": piped,
    }
    if scene == "frame" and not args.no_batch:
        # after the timed region: config 3 (64 cameras + shadow rays) on the same grid, one
        # launch against one launch per request; the headline metric above is unchanged
        del outs, warm
        torch.cuda.empty_cache()
        line["batched_config3"] = batch_compare(grid, 3, "config 3")
    print(json.dumps(line), flush=True)
    return line


def bench_render(args, threads):
    """Shaded config-2 orbit frames (render mode --mode, the reference's default sun):
    primary rays plus, per mode, shadow rays and two reflection bounces with
    their shadow rays, colours stored as rgba8. One step = one frame."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    lo, hi = scenes.clip_box()
    vdb = scenes.build_sphere_vdb(scenes.sphere_set(args.seed), lo, hi, threads=threads)
    grid = wx.Grid(vdb)
    W, H = args.width, args.height
    shade = wx.shade(args.mode)
    cams = [scenes.basis(scenes.orbit_camera(k % 64, width=W, height=H)) for k in range(args.warmup + args.steps)]
    out = torch.empty(W * H, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    for k in range(args.warmup):
        grid.render_camera(cams[k], W, H, shade, 0.5, clip=True, out=out, stream=stream)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        grid.render_camera(cams[args.warmup + k], W, H, shade, 0.5, clip=True, out=out, stream=stream)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    cpu = None
    if not args.no_cpu:
        lay = oracle_layout(vdb)
        rows = max(1, min(H, int(H * 4e6 / (W * H) / max(1, args.mode))))
        t = time.perf_counter()
        lay.render_camera(cams[0].as_array(), W, rows, shade, 0.5, clip=True, threads=threads)
        cs = time.perf_counter() - t
        cpu = {"value": W * rows / cs / 1e6, "unit": "Mpixel/s", "cores": threads, "kind": "port",
               "sample": f"{W}x{rows} rows of frame 0 (oracle ray_trace, {threads} threads, {cs:.1f}s)",
               **host_desc(threads)}
        lay.close()
    # handoff to the presentation path: the rgba8 frame copied into pinned host memory
    # (the staging of a texture upload) and, for comparison, into another device buffer
    pinned = torch.empty(W * H, dtype=torch.int32, pin_memory=True)
    dev2 = torch.empty_like(out)

    def ev_ms(fn, reps=20):
        fn()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            fn()
            e1.record(stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    d2h = ev_ms(lambda: pinned.copy_(out, non_blocking=True))
    d2d = ev_ms(lambda: dev2.copy_(out))
    same_px = bool(torch.equal(pinned, out.cpu()))
    handoff = {"frame_bytes": W * H * 4, "d2h_pinned_ms": d2h, "d2h_GBps": W * H * 4 / d2h / 1e6,
               "d2d_ms": d2d, "d2d_GBps": W * H * 4 / d2d / 1e6, "copied_frame_identical": same_px}
    names = ["gray", "rgb", "ray length", "diffuse", "glossy"]
    line = {
        "metric": f"Mpixel/s shaded 1920x1080 frames ({names[args.mode]}) on 4096^3 sphere-set VDB",
        "value": W * H * args.steps / wall / 1e6, "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"sparse sphere set 4096^3 (2000 spheres r in [4,40], seed {args.seed}); {W}x{H} "
                               f"orbit frames, render mode {args.mode} ({names[args.mode]}), default sun, rgba8 out",
                   "kernel_ms": kms},
        "handoff": handoff,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return line


def build_frame(wx, g):
    """A 320x240 frame over the 4096^3 domain (build checks: identical grids)."""
    cam = wx.camera_basis((7500.0, 5000.0, 6000.0), (2048.0, 2048.0, 2048.0), (0, 1, 0), 320 / 240, 40.0, 320.0)
    return g.trace_camera(cam, 320, 240, 0.5, clip=True).cpu().numpy()


def layout_cache_times(wx, g, name, cache_dir):
    """Writes grid g to a layout file (wx_grid_save) and loads it back
    (wx_grid_load); the load comes from the page cache the write just filled."""
    import shutil
    import tempfile

    base = cache_dir or tempfile.gettempdir()
    need = g.nbytes() * 2 + (1 << 30)
    if shutil.disk_usage(base).free < need:
        return {"skipped": f"less than {need / 2**30:.1f} GiB free under {base}"}
    path = os.path.join(base, f"bench_layout_{os.getpid()}.wxgrid")
    key = f"bench build: {name}"
    try:
        t = time.perf_counter()
        g.save(path, key)
        save_s = time.perf_counter() - t
        size = os.path.getsize(path)
        t = time.perf_counter()
        g2 = wx.Grid.load(path, key)
        load_s = time.perf_counter() - t
        same = np.array_equal(build_frame(wx, g), build_frame(wx, g2))
        del g2
    finally:
        if os.path.exists(path):
            os.remove(path)
    return {"file_bytes": size, "save_s": save_s, "load_s": load_s, "load_GBps": size / load_s / 1e9,
            "identical_frames": same, "where": base}


def bench_build(args, threads):
    """Grid build (topology, grid SDF, device layout) with the SDF replayed on the
    host threads and on the GPU, on the 4096^3 sphere set and value-noise
    isosurfaces at 20 % occupancy of 1024^3 and 4096^3 (the largest config-4 size,
    more than 2^24 leaves) whose voxels are decided on the GPU (and, at 1024^3, on
    the host as well); layouts must be identical."""
    import woxel_amd as wx
    from woxel_amd import scenes

    out = {}
    lo, hi = scenes.clip_box()
    sp = scenes.sphere_set(args.seed)

    def sphere(device):
        v = wx.Vdb345()
        t = time.perf_counter()
        v.add_spheres(sp, lo, hi, threads=threads)
        return v, time.perf_counter() - t

    def noise(side):
        def make(device):
            v = wx.Vdb345()
            t = time.perf_counter()
            v.add_noise(7, (0, 0, 0), (side, side, side), 8.0 / side, 0.2, threads=threads, device=device)
            return v, time.perf_counter() - t
        return make

    # the first device call of a process initialises the HIP runtime (about 0.15 s): an untimed
    # small device build first, so that no case below is charged for it
    w = wx.Vdb345()
    w.add_spheres(np.array([[64.0, 64.0, 64.0, 20.0]]), threads=threads)
    w.compute_sdf(threads=threads, device=True)
    del w

    # the host generator would take minutes over 4096^3 (identical voxels are shown at 1024^3):
    # both trees of that case take the device's voxels; spheres are rasterised on the host only
    cases = (("sphere set 4096^3 (2000 spheres)", sphere, True, False),
             ("value noise 1024^3, 20 %", noise(1024), True, True),
             ("value noise 4096^3, 20 %", noise(4096), False, True))
    for name, make, host_voxels, device_voxels in cases:
        log(f"build {name}: voxels on the {'host' if host_voxels else 'device'} ...")
        a, gen_h = make(not host_voxels)
        if not host_voxels:
            gen_h = None
        log(f"build {name}: voxels on the {'device' if device_voxels else 'host'} ...")
        b, gen_d = make(device_voxels)
        if not device_voxels:
            gen_d = None
        t = time.perf_counter()
        a.compute_sdf(threads=threads)
        th = time.perf_counter() - t
        t = time.perf_counter()
        launches = b.compute_sdf(threads=threads, device=True)
        td = time.perf_counter() - t
        # the device grid: leaf words copied where the replay left them, against the
        # host round trip (download for the host layout, upload by wx_grid_create)
        t = time.perf_counter()
        gb = wx.Grid(b)
        grid_dev_s = time.perf_counter() - t
        t = time.perf_counter()
        d = b.desc()
        d.bound = 4096.0
        gh = wx.Grid.from_desc(d)
        grid_host_s = time.perf_counter() - t
        same_grid = np.array_equal(build_frame(wx, gb), build_frame(wx, gh))
        del gh
        la, lb = a.layout_arrays(), b.layout_arrays()
        same = all(np.array_equal(la[k], lb[k]) for k in ("root", "n5", "n4", "leaves"))
        out[name] = {"nodes": a.count_nodes(), "voxels_host_s": gen_h, "voxels_device_s": gen_d,
                     "finalize_host_s": th, "finalize_device_s": td, "device_launches": launches,
                     "device_breakdown": b.build_info, "identical": same,
                     "grid_create_s": {"device_words": grid_dev_s, "host_round_trip": grid_host_s,
                                       "identical_frames": same_grid},
                     "layout_cache": layout_cache_times(wx, gb, name, args.layout_cache)}
        log(f"build {name}: {out[name]}")
        del a, b, la, lb, gb
    line = {"metric": "grid finalize time (topology + grid SDF + layout), host SDF replay vs device SDF replay",
            "value": out["value noise 4096^3, 20 %"]["finalize_device_s"], "unit": "s", "n_gpus": 1,
            "higher_is_better": False, "data": "synthetic", "config": {"workload": "build", "threads": threads},
            "results": out}
    print(json.dumps(line), flush=True)
    return line


def bench_incoherent(args, threads):
    """Config 2, second pass: 16,777,216 rays with origins uniform on the domain's
    bounding sphere and directions uniform on S^2, entry clip on. One step = one
    launch over all rays (ray buffer resident on the device)."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    lo, hi = scenes.clip_box()
    vdb = scenes.build_sphere_vdb(scenes.sphere_set(args.seed), lo, hi, threads=threads)
    grid = wx.Grid(vdb)
    n = 1 << 24
    rays_np = scenes.incoherent_rays(n, seed=7)
    rays = torch.from_numpy(rays_np).cuda()
    outs = [torch.empty((n, 32), dtype=torch.uint8, device="cuda") for _ in range(2)]
    stream = torch.cuda.current_stream()
    if not args.tune and not args.simple and not args.no_autotune:
        # the shape wx_autotune selects for this workload (it picks 24 or 32 waves per CU,
        # within noise of each other), fixed so the launch matches its traffic record
        args.tune = INCOHERENT_SHAPE
    tune = pick_tune(args, [grid], wx.rays_request(0, rays, outs[0], True), stream)
    for _ in range(args.warmup):
        grid.trace_rays(rays, clip=True, out=outs[0], stream=stream, simple=args.simple)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        ev[k][0].record(stream)
        grid.trace_rays(rays, clip=True, out=outs[k % 2], stream=stream, simple=args.simple)
        ev[k][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    launch_ms = np.array([a.elapsed_time(b) for a, b in ev])
    same = torch.equal(outs[0], outs[1]) if args.steps > 1 else None
    shape = "simple" if args.simple else (tune or wx.tune_get()).as_dict()
    rec, rec_src = counter_record("incoherent", "16777216 rays", shape)
    roof = roofline(outs[:1], launch_ms, 24, rec, rec_src)
    roof["valu"] = {k: roof[k] for k in ("achieved", "peak", "unit", "frac", "issue_model", "frac_at_fastest_issue")
                    if k in roof}
    hits = hit_count(outs[0])
    cpu = None
    if not args.no_cpu:
        lay = oracle_layout(vdb)
        m = 1 << 20
        done, t = 0, time.perf_counter()
        while time.perf_counter() - t < args.cpu_seconds and done + m <= n:
            lay.trace_rays(rays_np[done:done + m], clip=True, threads=threads)
            done += m
        cpu_s = time.perf_counter() - t
        cpu = {"value": done / cpu_s / 1e9, "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": f"first {done} of the {n} rays (oracle C traversal, {threads} threads, {cpu_s:.1f}s)",
               **host_desc(threads)}
        lay.close()
    line = {
        "metric": "Grays/s incoherent rays on 4096^3 sphere-set VDB",
        "value": n * args.steps / wall / 1e9, "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"sparse sphere set 4096^3 (2000 spheres r in [4,40], seed {args.seed}); 16,777,216 "
                               "rays, origins uniform on the bounding sphere (centre 2048^3, r = 2048 sqrt 3), "
                               "directions uniform on S^2, entry clip",
                   "hit_fraction": hits / n, "repeat_identical": same,
                   "launch": shape},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return line


def batch_compare(grid, steps, log_prefix="batch"):
    """Config 3 on a built 4096^3 sphere-set grid: 64 cameras (poses on the r=6000
    sphere, log-uniform 256^2..3840x2160), one shadow ray per primary hit toward
    (8000, 8000, 8000). All requests in one launch against one launch per request
    (back to back on one stream, and waiting for each), identical outputs; per-request
    latency = the hipEvent span of each request's own launch."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

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
    log(f"{log_prefix}: {len(specs)} cameras, {pixels} pixels")
    t1 = time.time()
    tune, tms = wx.autotune([grid], reqs, stream=stream)
    log(f"{log_prefix}: autotune {time.time() - t1:.1f}s {tune.as_dict()} {tms:.3f} ms")

    def run_batched():
        wx.trace_batch([grid], reqs, tune=tune, stream=stream)

    def run_single():
        for r in reqs:
            wx.trace_batch([grid], [r], tune=tune, stream=stream)

    def run_single_sync():  # a request loop that waits for each result before the next
        for r in reqs:
            wx.trace_batch([grid], [r], tune=tune, stream=stream)
            stream.synchronize()

    t_sync = timed_reps(run_single_sync, steps)
    t_single = timed_reps(run_single, steps)
    ref = [(o.clone(), s.clone()) for o, s in outs]
    t_batch = timed_reps(run_batched, steps)
    same = all(torch.equal(a, c) and torch.equal(b, d) for (a, b), (c, d) in zip(ref, outs))
    del ref
    # each request's own launch, event-timed (its latency when served alone)
    lat = []
    for r in reqs:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        wx.trace_batch([grid], [r], tune=tune, stream=stream)
        e1.record(stream)
        e1.synchronize()
        lat.append(e0.elapsed_time(e1))
    bl = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run_batched()
        e1.record(stream)
        e1.synchronize()
        bl.append(e0.elapsed_time(e1))
    shadow_rays = sum(hit_count(o) for o, _ in outs)
    rays = pixels + shadow_rays
    # the same requests without shadow rays: what the in-lane shadow continuation costs
    plain = []
    for r in reqs:
        q = wx.Request.from_buffer_copy(r)
        q.flags &= ~wx.TRACE_SHADOW
        q.shadow_out = None
        plain.append(q)
    t_batch_ns = timed_reps(lambda: wx.trace_batch([grid], plain, tune=tune, stream=stream), steps)
    t_single_ns = timed_reps(lambda: [wx.trace_batch([grid], [q], tune=tune, stream=stream) for q in plain], steps)
    return {
        "workload": "64 cameras on the r=6000 sphere around the 4096^3 sphere set, log-uniform 256x256..3840x2160, "
                    "one shadow ray per hit toward (8000,8000,8000)",
        "primary_rays": pixels, "shadow_rays": shadow_rays, "tune": tune.as_dict(),
        "batched": {"value": rays / t_batch / 1e9, "unit": "Grays/s", "ms": t_batch * 1e3,
                    "launch_ms_p50": float(np.percentile(bl, 50)), "launch_ms_p99": float(np.percentile(bl, 99))},
        "per_request_launches": {"value": rays / t_single / 1e9, "unit": "Grays/s", "ms": t_single * 1e3},
        "per_request_synchronous": {"value": rays / t_sync / 1e9, "unit": "Grays/s", "ms": t_sync * 1e3},
        "speedup_vs_per_request": t_single / t_batch,
        "speedup_vs_per_request_synchronous": t_sync / t_batch,
        "per_request_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                   "max": float(np.max(lat))},
        "identical_output": same,
        "without_shadow_rays": {"batched_ms": t_batch_ns * 1e3, "per_request_ms": t_single_ns * 1e3,
                                "shadow_cost_batched_ms": (t_batch - t_batch_ns) * 1e3,
                                "shadow_cost_per_request_ms": (t_single - t_single_ns) * 1e3},
    }


def bench_batch(args, threads):
    """64 cameras + shadow rays in one launch against one launch per request (config 3)."""
    import woxel_amd as wx
    from woxel_amd import scenes

    lo, hi = scenes.clip_box()
    vdb = scenes.build_sphere_vdb(scenes.sphere_set(args.seed), lo, hi, threads=threads)
    grid = wx.Grid(vdb)
    r = batch_compare(grid, args.steps)
    line = {
        "metric": "Grays/s primary + shadow rays, 64 cameras in one launch (config 3)",
        "value": r["batched"]["value"], "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
        "ms_per_step": r["batched"]["ms"], "higher_is_better": True, "dtype": "f32", "data": "synthetic",
        "config": {"workload": r["workload"], "primary_rays": r["primary_rays"], "shadow_rays": r["shadow_rays"],
                   "tune": r["tune"]},
        **{k: r[k] for k in ("per_request_launches", "per_request_synchronous", "speedup_vs_per_request",
                             "speedup_vs_per_request_synchronous", "per_request_latency_ms", "identical_output",
                             "without_shadow_rays")},
        "batched_launch_ms": {"p50": r["batched"]["launch_ms_p50"], "p99": r["batched"]["launch_ms_p99"]},
    }
    print(json.dumps(line), flush=True)


def hit_count(t):
    import torch

    return int((t.view(torch.int32).reshape(-1, 8)[:, 6] & 0xFF).eq(0).sum().item())


def build_grid_set(args, threads):
    """The 32 resident grids: 16 sphere sets and 16 value-noise isosurfaces."""
    import woxel_amd as wx
    from woxel_amd import scenes

    specs = scenes.config4_specs(noise_max=args.noise_max)
    grids, t0 = [], time.time()
    for i, sp in enumerate(specs):
        t = time.time()
        if args.layout_cache:
            g, how = scenes.cached_grid(sp, args.layout_cache, threads=threads)
            grids.append(g)
        else:
            v = scenes.build_grid(sp, threads=threads, device=True)
            grids.append(wx.Grid(v))
            how = "built"
            del v
        log(f"grid {i:2d} {sp['kind']:7s} side {sp['side']:4d}: nodes {grids[-1].counts} "
            f"{grids[-1].nbytes() / 2**20:.0f} MiB, {how} in {time.time() - t:.1f}s")
    log(f"32 grids resident: {sum(g.nbytes() for g in grids) / 2**30:.2f} GiB, ready in {time.time() - t0:.0f}s")
    return specs, grids


def set_desc(specs, noise_max):
    sides = [s["side"] for s in specs]
    return (f"32 resident grids: 16 sphere sets (200-5000 spheres, r in [2,60], sides 1024-4096) + 16 value-noise "
            f"isosurfaces (3 octaves, 1-20% occupancy, sides 512-{noise_max}); sides {min(sides)}-{max(sides)}")


def bench_multigrid(args, specs, grids):
    """256 cameras, each bound to a random grid of the 32, log-uniform 128^2..1920x1080,
    primary + one shadow ray per hit: one launch against one launch per request."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    cams = scenes.config4_cameras(specs, 256)
    stream = torch.cuda.current_stream()
    reqs, outs, singles = [], [], []
    for gi, sp in cams:
        w, h = sp["width"], sp["height"]
        o = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        s = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        reqs.append(wx.camera_request(gi, scenes.basis(sp), w, h, o, 0.5, True, shadow_out=s, light=scenes.LIGHT))
        # the same request on its own grid: a plain launch of the one-request kernel
        singles.append(([grids[gi]], [wx.camera_request(0, scenes.basis(sp), w, h, o, 0.5, True, shadow_out=s,
                                                         light=scenes.LIGHT)]))
        outs.append((o, s))
    pixels = sum(sp["width"] * sp["height"] for _, sp in cams)
    log(f"multigrid: {len(cams)} cameras over {len(set(g for g, _ in cams))} grids, {pixels} pixels")
    t1 = time.time()
    tune, tms = wx.autotune(grids, reqs, stream=stream)
    log(f"autotune {time.time() - t1:.1f}s: {tune.as_dict()} {tms:.3f} ms")

    def run_batched():
        wx.trace_batch(grids, reqs, tune=tune, stream=stream)

    def run_single(st=stream):
        for g1, r1 in singles:
            wx.trace_batch(g1, r1, tune=tune, stream=st)

    def run_single_sync():  # a request loop that waits for each result before the next
        for g1, r1 in singles:
            wx.trace_batch(g1, r1, tune=tune, stream=stream)
            stream.synchronize()

    t_sync = timed_reps(run_single_sync, args.steps)
    t_single = timed_reps(run_single, args.steps)
    ref = [(o.clone(), s.clone()) for o, s in outs]
    # the 256 per-request launches captured once in a HIP graph and replayed
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        run_single(torch.cuda.current_stream())
    t_graph = timed_reps(graph.replay, args.steps)
    same_graph = all(torch.equal(a, c) and torch.equal(b, d) for (a, b), (c, d) in zip(ref, outs))
    t_batch = timed_reps(run_batched, args.steps)
    same = same_graph and all(torch.equal(a, c) and torch.equal(b, d) for (a, b), (c, d) in zip(ref, outs))
    del ref, graph
    shadow_rays = sum(hit_count(o) for o, _ in outs)
    rays = pixels + shadow_rays
    res = resident_burst(args, grids, cams, outs)
    line = {
        "metric": "Grays/s primary + shadow rays, 256 cameras over 32 resident grids in one launch (config 4)",
        "value": rays / t_batch / 1e9, "unit": "Grays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 1, "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": set_desc(specs, args.noise_max) + "; 256 cameras on spheres of radius 1.46 x side, "
                               "log-uniform 128x128..1920x1080, one shadow ray per hit toward (8000,8000,8000)",
                   "primary_rays": pixels, "shadow_rays": shadow_rays, "tune": tune.as_dict(),
                   "grid_bytes": sum(g.nbytes() for g in grids)},
        "per_request_launches": {"value": rays / t_single / 1e9, "ms": t_single * 1e3},
        "speedup_vs_per_request": t_single / t_batch,
        "per_request_synchronous": {"value": rays / t_sync / 1e9, "ms": t_sync * 1e3},
        "speedup_vs_per_request_synchronous": t_sync / t_batch,
        "per_request_graph": {"value": rays / t_graph / 1e9, "ms": t_graph * 1e3},
        "speedup_vs_per_request_graph": t_graph / t_batch,
        "identical_output": same,
        "resident_burst": dict(res, value=rays / res["s"] / 1e9) if res else None,
    }
    print(json.dumps(line), flush=True)
    return line


def resident_burst(args, grids, cams, outs):
    """The same 256 requests submitted at once to a server that admits them into one
    running kernel: seconds from the first submit to the last completion (best of
    args.steps bursts), and whether every record equals the batched launch's."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    need = sum(-(-(sp["width"] * sp["height"] * 64) // 256) * 256 for _, sp in cams) + (1 << 20)
    srv = wx.Server(hbm_budget=need, resident=True, idle_stop_us=200000)
    try:
        gids = [srv.add_grid(g) for g in grids]
        best, same = None, True
        for rep in range(max(1, args.steps) + 1):  # the first burst starts the kernel: not timed
            tks = [srv.submit(gids[gi], scenes.basis(sp), sp["width"], sp["height"], 0.5, True, shadow=True,
                              light=scenes.LIGHT) for gi, sp in cams]
            res = [srv.wait(tk, 120.0) for tk in tks]
            t = max(r.done_s for r in res) - min(r.submit_s for r in res)
            if rep:
                best = t if best is None else min(best, t)
            if rep == 1:
                for r, (gi, sp), (o, s) in zip(res, cams, outs):
                    n = sp["width"] * sp["height"]
                    same = same and torch.equal(wx.device_copy(r.primary, n), o) and \
                        torch.equal(wx.device_copy(r.shadow, n), s)
            for tk in tks:
                srv.release(tk)
        st = srv.stats()
        log(f"resident burst: {best * 1e3:.1f} ms, sessions {st.batches}, identical {same}")
        return {"s": best, "ms": best * 1e3, "identical_output": same, "sessions": int(st.batches)}
    finally:
        srv.close()


def timed_reps(fn, reps):
    import torch

    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def serve_stream(srv, gids, items):
    """Submit (arrival, grid, camera, w, h) items at their arrival times; a waiter
    thread waits for each ticket in order and releases it. Returns per-request
    (submit, launch, done, batch size, primary rays) in server clock seconds."""
    import queue
    import threading

    from woxel_amd import scenes

    q, res, err = queue.Queue(), [], []

    def waiter():
        while True:
            it = q.get()
            if it is None:
                return
            tk, rays = it
            try:
                r = srv.wait(tk, 60.0)
                res.append((r.submit_s, r.launch_s, r.done_s, r.batch_requests, rays, r.status))
                srv.release(tk)
            except Exception as e:  # noqa: BLE001 - reported after the run
                err.append(repr(e))

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    try:
        for ta, gi, cam, w, h in items:
            dt = ta - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            q.put((srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=True, light=scenes.LIGHT), w * h))
    finally:
        q.put(None)
        th.join()
    if err:
        raise RuntimeError(f"serving errors: {err[:3]}")
    return res


def serve_summary(res, rate):
    sub = np.array([r[0] for r in res])
    done = np.array([r[2] for r in res])
    lat = (done - sub) * 1e3
    span = done.max() - sub.min()
    rays = sum(r[4] for r in res)
    return {"rate": rate, "requests": len(res), "primary_rays_per_s": rays / span,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_batch": float(np.mean([r[3] for r in res])), "span_s": float(span),
            "failed": int(sum(1 for r in res if r[5] != 0))}


def bench_serve(args, specs, grids):
    """Poisson request streams (rates swept) through the continuous-batching server
    (HBM-budget admission, several launch slots) and through plain per-request
    serving (one slot, one request per launch), same arrivals; p50/p99 and rays/s."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    budget = int(args.budget_gib * 2**30)
    rates = [float(x) for x in args.rates.split(",") if x]

    def items_for(rate, seed):
        n = max(32, int(rate * args.serve_seconds))
        return [(ta, gi, scenes.basis(sp), sp["width"], sp["height"])
                for ta, gi, sp in scenes.serve_requests(specs, n, rate, seed=seed)]

    # the served launch shape: tuned once on a representative mix, kept for both servers
    mix = items_for(1000.0, 99)[:16]
    bufs, reqs = [], []
    for _, gi, cam, w, h in mix:
        o = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        s = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        bufs.append((o, s))
        reqs.append(wx.camera_request(gi, cam, w, h, o, 0.5, True, shadow_out=s, light=scenes.LIGHT))
    tune, tms = wx.autotune(grids, reqs, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    log(f"serve: tune {tune.as_dict()} ({tms:.2f} ms for 16 mixed requests)")

    servers = {
        "batched": dict(hbm_budget=budget, streams=args.streams, max_batch_requests=64,
                        max_batch_rays=args.max_batch_rays),
        "plain": dict(hbm_budget=budget, streams=1, max_batch_requests=1),
        # admission into one running kernel (requests join while it runs)
        "resident": dict(hbm_budget=budget, resident=True),
    }
    want = [x for x in args.servers.split(",") if x]
    servers = {k: v for k, v in servers.items() if k in want or k == "batched"}
    out = {k: [] for k in servers}
    for name, cfg in servers.items():
        srv = wx.Server(**cfg)
        gids = [srv.add_grid(g) for g in grids]
        serve_stream(srv, gids, items_for(200.0, 98)[:24])  # warm-up
        for i, rate in enumerate(rates):
            res = serve_stream(srv, gids, items_for(rate, 7 + i))
            out[name].append(serve_summary(res, rate))
            log(f"serve {name:7s} rate {rate:6.0f}/s: {json.dumps(out[name][-1])}")
        srv.close()

    # the same requests alone and in one served mix must give identical records
    same = served_mix_identical(grids, items_for(1000.0, 97)[:12])
    b, p, rs = out["batched"], out.get("plain"), out.get("resident")
    peak_b = max(x["primary_rays_per_s"] for x in b)
    peak_p = max(x["primary_rays_per_s"] for x in p) if p else None
    line = {
        "metric": "Grays/s primary rays served (each with shadow rays), Poisson arrivals over 32 resident grids; "
                  "p50/p99 latency (config 5)",
        "value": peak_b / 1e9, "unit": "Grays/s", "n_gpus": 1, "steps": len(rates), "warmup": 1,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": set_desc(specs, args.noise_max) + "; requests: uniform grid, eye uniform on the "
                               "grid's bounding sphere, resolution from {512^2, 1024^2, 1920x1080, 3840x2160}, "
                               "primary + shadow rays; Poisson arrivals, rates " + args.rates + " req/s, "
                               f"{args.serve_seconds:g} s per rate",
                   "budget_bytes": budget, "streams": args.streams, "max_batch_requests": 64,
                   "max_batch_rays": args.max_batch_rays,
                   "tune": tune.as_dict()},
        "batched": b, "plain_per_request": p, "resident_session": rs,
        "peak_speedup_vs_plain": peak_b / peak_p if p else None,
        "resident_peak_vs_batched": (max(x["primary_rays_per_s"] for x in rs) / peak_b) if rs else None,
        "identical_output": same,
    }
    print(json.dumps(line), flush=True)
    return line


def served_mix_identical(grids, items):
    """Hold a request that fills the whole arena so the next requests queue up,
    release it so they launch as ONE batch, and compare every served record with
    the same request launched alone."""
    import torch

    import woxel_amd as wx
    from woxel_amd import scenes

    # arena = exactly the mix's outputs; the blocker (4096 wide, primary + shadow)
    # leaves less than 256 KiB free, smaller than any request of the mix
    cap = sum(-(-(w * h * 64) // 256) * 256 for _, _, _, w, h in items)
    bw, bh = 4096, cap // (4096 * 64)
    assert 1 <= bh <= 65536 and min(w * h * 64 for *_, w, h in items) > cap - bw * bh * 64
    srv = wx.Server(hbm_budget=cap, streams=2, max_batch_requests=64)
    gids = [srv.add_grid(g) for g in grids]
    _, g0, cam0, _, _ = items[0]
    blocker = srv.submit(gids[g0], cam0, bw, bh, 0.5, True, shadow=True, light=scenes.LIGHT)
    srv.wait(blocker, 60.0)
    tks = [srv.submit(gids[gi], cam, w, h, 0.5, True, shadow=True, light=scenes.LIGHT) for _, gi, cam, w, h in items]
    srv.release(blocker)
    ok, sizes = True, set()
    for (_, gi, cam, w, h), tk in zip(items, tks):
        r = srv.wait(tk, 60.0)
        sizes.add(r.batch_requests)
        a, s = wx.device_copy(r.primary, w * h), wx.device_copy(r.shadow, w * h)
        srv.release(tk)
        o = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        so = torch.empty((w * h, 32), dtype=torch.uint8, device="cuda")
        wx.trace_batch(grids, [wx.camera_request(gi, cam, w, h, o, 0.5, True, shadow_out=so, light=scenes.LIGHT)])
        torch.cuda.synchronize()
        ok = ok and torch.equal(a, o) and torch.equal(s, so)
    srv.close()
    log(f"served mix: batch sizes {sorted(sizes)}, identical alone and in the mix: {ok}")
    return ok and sizes == {len(items)}


if __name__ == "__main__":
    main()
