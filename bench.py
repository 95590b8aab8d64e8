"""Benchmark: exact FP32 trilinear sampling of a 512^3 structuredRegular
volume at 2^26 uniform random points (vklComputeSampleN, device-resident
inputs), on one MI355X.

Prints one JSON line: throughput in Msamples/s, the roofline of the sampling
kernel (algorithmic bytes / kernel time vs 8 TB/s HBM), and the CPU oracle
timed on this host's cores as a reported baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at fixed RMSE: exact vs batched vs neural-field; % HBM roofline"
HBM_PEAK_GBS = 8000.0
SEED = 20261202


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_volume(torch, n):
    """sin(x) cos(y) sin(z) wavelet (object coords = voxel index * 2pi/n) plus
    seeded U(0, 0.1) noise, float32, generated on the GPU."""
    g = torch.Generator(device="cuda").manual_seed(SEED)
    c = torch.arange(n, device="cuda", dtype=torch.float32) * (8.0 * np.pi / n)
    v = (torch.sin(c)[None, None, :] * torch.cos(c)[None, :, None] * torch.sin(c)[:, None, None])
    v = v + 0.1 * torch.rand((n, n, n), device="cuda", generator=g)
    return v.contiguous()


def distinct_corner_voxels(torch, pts, n):
    """U = number of distinct voxels the linear filter reads for these points."""
    mask = torch.zeros(n * n * n, dtype=torch.bool, device="cuda")
    for chunk in torch.split(pts, 1 << 22):
        ic = torch.floor(chunk).to(torch.int64)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    x = (ic[:, 0] + dx).clamp(0, n - 1)
                    y = (ic[:, 1] + dy).clamp(0, n - 1)
                    z = (ic[:, 2] + dz).clamp(0, n - 1)
                    mask[(z * n + y) * n + x] = True
    return int(mask.sum().item())


def cpu_baseline(data_np, pts_np, count):
    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    grid = O.Grid.dense(data_np)
    sub = np.ascontiguousarray(pts_np[:count])
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        grid.sample(sub, threads=threads)
        best = min(best, time.perf_counter() - t0)
    return {"value": round(count / best / 1e6, 2), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"{count} of the 2^26 points (first {count}), best of 3, C++ restatement "
                      f"(oracle/vkl_oracle.cpp, -O2, no FMA) on {threads} threads"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--points-log2", type=int, default=26)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--brick", default="0,0,0", help="device voxel layout: brick shifts sx,sy,sz (0,0,0 = x-fastest)")
    ap.add_argument("--binned", choices=("auto", "on", "off"), default="auto",
                    help="spatially binned sampling (auto: on for large volumes and batches)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device-to-device copy (a known-size kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a measured path).", file=sys.stderr)
        sys.exit(2)

    import torch

    import openvkl_amd as vkl

    n, npts = args.n, 1 << args.points_log2
    dev = vkl.Device("hip")
    log(f"building {n}^3 volume ...")
    vol_t = make_volume(torch, n)
    brick = tuple(int(v) for v in args.brick.split(","))
    vol = dev.structured_regular(vol_t, brick_shift=brick)
    sampler = vol.sampler(filter=vkl.VKL_FILTER_LINEAR, binned={"auto": None, "on": True, "off": False}[args.binned])
    binned = args.binned == "on" or (args.binned == "auto" and npts >= (1 << 20) and n ** 3 * 4 >= (64 << 20))
    g = torch.Generator(device="cuda").manual_seed(SEED + 1)
    pts = torch.rand((npts, 3), device="cuda", generator=g, dtype=torch.float32) * float(n - 1)
    out = torch.empty(npts, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        sampler.sample(pts, out=out)
    torch.cuda.synchronize()

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        sampler.sample(pts, out=out)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_event = start.elapsed_time(stop) / args.steps
    value = npts * args.steps / wall / 1e6

    if args.copy_calibration:
        src = torch.empty(1 << 28, device="cuda", dtype=torch.float32).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1 GiB read + 1 GiB written")
        del src, dst

    U = distinct_corner_voxels(torch, pts, n)
    bytes_per_launch = npts * (12 + 4) + 4 * U
    achieved = bytes_per_launch / (ms_event * 1e-3) / 1e9
    log(f"U = {U} distinct voxels ({U / n ** 3:.3f} of V); {bytes_per_launch / npts:.2f} B/sample; "
        f"kernel avg {ms_event:.3f} ms; {achieved:.0f} GB/s")

    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded wavelet + U(0,0.1) noise volume; seeded uniform points)",
        "config": {"workload": f"structuredRegular {n}^3 float32 vertex-centered, VKL_FILTER_LINEAR, "
                               f"vklComputeSampleN, 2^{args.points_log2} uniform random points in bounds",
                   "volume_dims": [n, n, n], "points": npts, "filter": "linear", "exact": True,
                   "device_layout": "x-fastest" if brick == (0, 0, 0) else "bricks %dx%dx%d" % tuple(1 << b for b in brick),
                   "sampling_path": "binned (bin scatter + LDS-brick sample)" if binned else "direct gather"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "algorithmic_bytes_per_launch": bytes_per_launch, "kernel_ms": round(ms_event, 4)},
    }
    if not args.no_cpu_baseline:
        log("timing the CPU oracle ...")
        data_np = vol_t.cpu().numpy()
        pts_np = pts[: 1 << 22].cpu().numpy()
        result["cpu_baseline"] = cpu_baseline(data_np, pts_np, 1 << 22)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
