"""Benchmark: exact FP32 trilinear sampling of a 512^3 structuredRegular
volume at 2^26 uniform random points (vklComputeSampleN, device-resident
inputs), on one MI355X. The volume uses the device's corner-packed layout
(volume parameter "cornerPacked"; --layout linear selects the caller's
x-fastest array).

Prints one JSON line: throughput in Msamples/s, the roofline of the sampling
kernel (algorithmic bytes / kernel time vs 8 TB/s HBM, plus the measured HBM
traffic from the committed rocprofv3 counters), the CPU oracle timed on this
host's cores as a reported baseline, and the other workloads (other layouts,
ray marching, batched multi-volume sampling, the neural-field sampler) as
"variants".
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


def timed(torch, fn, steps, warmup):
    """Average ms per call of fn over `steps` calls (HIP events on the null
    stream, which the library uses by default)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def variant_layouts(torch, vkl, dev, vol_t, pts, out, steps, warmup, alg_bytes):
    """The same workload on the other device layouts of the same voxels: the
    caller's x-fastest array (direct gather, and spatially binned), 4x4x2
    bricks (direct gather)."""
    res = {}
    for name, kw, binned in (("x_fastest_direct", {}, False), ("x_fastest_binned", {}, True),
                             ("bricks_4x4x2_direct", {"brick_shift": (2, 2, 1)}, False)):
        v = dev.structured_regular(vol_t, **kw)
        s = v.sampler(filter=vkl.VKL_FILTER_LINEAR, binned=binned)
        ms = timed(torch, lambda: s.sample(pts, out=out), steps, warmup)
        gbs = alg_bytes / (ms * 1e-3) / 1e9
        res[name] = {"msamples_s": round(pts.shape[0] / ms / 1e3, 1), "ms": round(ms, 4),
                     "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
        s.release()
        v.release()
    return res


def variant_dense128(torch, vkl, dev, steps, warmup, cpu=False):
    """The reference's own CPU-runnable case: structuredRegular 128^3 float32
    (the same wavelet + noise generator), 2^20 uniform random points,
    linear, vklComputeSampleN on the x-fastest volume (8 MiB: cache
    resident), and on its corner-packed form (one record per sample). With
    `cpu`, the oracle on every usable host core beside it."""
    n, npts = 128, 1 << 20
    vol = make_volume(torch, n)
    v = dev.structured_regular(vol)
    s = v.sampler(filter=vkl.VKL_FILTER_LINEAR)
    g = torch.Generator(device="cuda").manual_seed(SEED + 30)
    pts = torch.rand((npts, 3), device="cuda", generator=g) * float(n - 1)
    out = torch.empty(npts, device="cuda")
    ms = timed(torch, lambda: s.sample(pts, out=out), steps, warmup)
    vp = dev.structured_regular(vol, corner_packed=True)
    sp = vp.sampler(filter=vkl.VKL_FILTER_LINEAR)
    out_p = torch.empty_like(out)
    ms_p = timed(torch, lambda: sp.sample(pts, out=out_p), steps, warmup)
    same_p = bool(torch.equal(out_p.view(torch.int32), out.view(torch.int32)))
    sp.release()
    vp.release()
    res = {"volume": "structuredRegular 128^3 float32", "points": npts, "ms": round(ms, 4),
           "msamples_s": round(npts / ms / 1e3, 1),
           "corner_packed": {"ms": round(ms_p, 4), "msamples_s": round(npts / ms_p / 1e3, 1),
                             "bitwise_equal_to_x_fastest": same_p}}
    if cpu:
        import oracle as O
        O.use_native()
        threads, _, _ = host_cpu()
        grid = O.Grid.dense(vol.cpu().numpy())
        p_np = pts.cpu().numpy()
        got = grid.sample(p_np, threads=threads)
        best = float("inf")
        for _ in range(5):
            t0 = time.perf_counter()
            got = grid.sample(p_np, threads=threads)
            best = min(best, time.perf_counter() - t0)
        want = out.cpu().numpy()
        res["cpu_msamples_s"] = round(npts / best / 1e6, 2)
        res["cpu_threads"] = threads
        res["cpu_bitwise_equal_to_gpu"] = bool(np.all((got.view(np.uint32) == want.view(np.uint32)) |
                                                      (np.isnan(got) & np.isnan(want))))
    s.release()
    v.release()
    return res


def variant_dense512_matrix(torch, vkl, dev, vol_t, pts, steps, warmup, U):
    """The rest of the 512^3 / 2^26-point matrix on corner-packed volumes:
    half and uint8 (raw codes; the scale is applied by the caller) linear,
    float32 cubic, float32 linear gradient; half / uint8 linear on their raw
    x-fastest arrays (256 / 128 MiB, within the Infinity Cache) and in
    128-byte bricks (direct gathers, binned path off); half cubic; float32
    and half cubic samples and float32 cubic gradients with cubic records
    and with whole cubic stencils (wavefront-cooperative stencil loads)
    beside the same on the corner-packed records alone."""
    npts = pts.shape[0]
    out = torch.empty(npts, device="cuda")
    gout = torch.empty((npts, 3), device="cuda")
    res = {}
    half = vol_t.half()
    u8 = (vol_t * 100.0 + 100.0).clamp(0, 255).to(torch.uint8)
    packed = {"corner_packed": True}
    cases = (("half_linear", half, vkl.VKL_FILTER_LINEAR, False, 2, packed),
             ("uint8_linear", u8, vkl.VKL_FILTER_LINEAR, False, 1, packed),
             ("float32_cubic", vol_t, vkl.VKL_FILTER_CUBIC, False, 4, packed),
             ("float32_linear_gradient", vol_t, vkl.VKL_FILTER_LINEAR, True, 4, packed),
             ("half_linear_x_fastest", half, vkl.VKL_FILTER_LINEAR, False, 2, {}),
             ("half_linear_bricks_4x4x4", half, vkl.VKL_FILTER_LINEAR, False, 2, {"brick_shift": (2, 2, 2)}),
             ("uint8_linear_x_fastest", u8, vkl.VKL_FILTER_LINEAR, False, 1, {}),
             ("uint8_linear_bricks_8x4x4", u8, vkl.VKL_FILTER_LINEAR, False, 1, {"brick_shift": (3, 2, 2)}),
             ("half_cubic", half, vkl.VKL_FILTER_CUBIC, False, 2, packed),
             ("float32_cubic_gradient", vol_t, vkl.VKL_FILTER_CUBIC, True, 4, packed),
             ("float32_cubic_records", vol_t, vkl.VKL_FILTER_CUBIC, False, 4, dict(packed, cubic_records=True)),
             ("float32_cubic_gradient_records", vol_t, vkl.VKL_FILTER_CUBIC, True, 4,
              dict(packed, cubic_records=True)),
             ("half_cubic_records", half, vkl.VKL_FILTER_CUBIC, False, 2, dict(packed, cubic_records=True)),
             ("float32_cubic_stencils", vol_t, vkl.VKL_FILTER_CUBIC, False, 4, dict(packed, cubic_stencils=True)),
             ("float32_cubic_gradient_stencils", vol_t, vkl.VKL_FILTER_CUBIC, True, 4,
              dict(packed, cubic_stencils=True)),
             ("half_cubic_stencils", half, vkl.VKL_FILTER_CUBIC, False, 2, dict(packed, cubic_stencils=True)))
    for name, data, filt, grad, es, layout in cases:
        v = dev.structured_regular(data.contiguous(), **layout)
        sm = v.sampler(filter=filt, binned=False)
        vox = es * (U if filt != vkl.VKL_FILTER_CUBIC else data.numel())
        if grad:
            ms = timed(torch, lambda: sm.gradient(pts, out=gout), steps, warmup)
            alg = npts * (12 + 12) + vox
        else:
            ms = timed(torch, lambda: sm.sample(pts, out=out), steps, warmup)
            alg = npts * (12 + 4) + vox
        gbs = alg / (ms * 1e-3) / 1e9
        res[name] = {"msamples_s": round(npts / ms / 1e3, 1), "ms": round(ms, 4), "algorithmic_gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"  {name}: {res[name]}")
        sm.release()
        v.release()
        del data
    del half, u8
    return res


def variant_batched(torch, vkl, dev, steps, warmup, packed=False):
    """64 structuredRegular volumes, cube sides seeded in [32, 384], float32 and
    half alternating, point counts log-uniform in [1e4, 1e7]: one persistent,
    graph-captured launch with an autotuned shape (vklNewSampleBatchPlan)."""
    rng = np.random.default_rng(SEED + 7)
    g = torch.Generator(device="cuda").manual_seed(SEED + 8)
    keep, samplers, pts, outs = [], [], [], []
    alg = 0.0
    for i in range(64):
        side = int(rng.integers(32, 385))
        dt = torch.float32 if i % 2 == 0 else torch.float16
        data = torch.rand((side, side, side), device="cuda", generator=g).to(dt)
        v = dev.structured_regular(data, corner_packed=packed or None)
        s = v.sampler(filter=vkl.VKL_FILTER_LINEAR)
        count = int(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        p = torch.rand((count, 3), device="cuda", generator=g) * float(side - 1)
        o = torch.empty(count, device="cuda")
        vox = float(side) ** 3
        alg += count * 16 + (4 if dt == torch.float32 else 2) * vox * (1 - np.exp(-8.0 * count / vox))
        keep.append((data, v))
        samplers.append(s)
        pts.append(p)
        outs.append(o)
    plan = vkl.SampleBatchPlan(dev, samplers, pts, outs, autotune=True)
    total = sum(int(p.shape[0]) for p in pts)
    ms = timed(torch, plan.execute, steps, warmup)
    blocks, threads, tile = plan.launch_shape()
    launches = plan.launch_count()
    plan.release()
    # the two groupings at the fixed default shape: one launch per element
    # type, and both element types in one launch
    fixed = {}
    for name, mode in (("per_type", 0), ("merged", 2)):
        pl = vkl.SampleBatchPlan(dev, samplers, pts, outs, autotune=mode)
        fixed[name] = {"ms": round(timed(torch, pl.execute, steps, warmup), 4), "launches": pl.launch_count()}
        pl.release()
    gbs = alg / (ms * 1e-3) / 1e9
    return {"layout": "corner-packed" if packed else "x-fastest",
            "msamples_s": round(total / ms / 1e3, 1), "ms": round(ms, 4), "points": total, "volumes": 64,
            "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "launch": {"blocks": blocks, "threads": threads, "tile": tile, "launches": launches},
            "fixed_shape_groupings": fixed}


def shell_vdb(torch, res, radius=0.35, width=0.0095, seed=SEED + 9):
    """Sphere shell |r - R| < w modulated by 4-octave sine fBm, as 8^3 float32
    leaves on the GPU. Returns (origins int32 [n,3] numpy, leaves [n*512] cuda)."""
    nl, c = res // 8, res / 2.0
    centres = torch.arange(nl, device="cuda", dtype=torch.float32) * 8 + 4.0
    yy, xx = torch.meshgrid(centres, centres, indexing="ij")
    margin = width + 4.0 * 3 ** 0.5 / res
    orgs = []
    for z in range(nl):
        r = torch.sqrt((xx - c) ** 2 + (yy - c) ** 2 + (centres[z] - c) ** 2) / res
        ky, kx = torch.nonzero((r - radius).abs() < margin, as_tuple=True)
        orgs.append(torch.stack([kx * 8, ky * 8, torch.full_like(kx, z * 8)], 1))
    origins = torch.cat(orgs).to(torch.int32)
    rng = np.random.default_rng(seed)
    freq = [0.02 * 2 ** k for k in range(4)]
    phase = rng.uniform(0, 2 * np.pi, (4, 3))
    l8 = torch.arange(8, device="cuda", dtype=torch.float32)
    lx, ly, lz = torch.meshgrid(l8, l8, l8, indexing="ij")  # leaf order x*64 + y*8 + z
    lx, ly, lz = lx.reshape(1, -1), ly.reshape(1, -1), lz.reshape(1, -1)
    leaves = torch.empty((origins.shape[0], 512), device="cuda", dtype=torch.float32)
    for a in range(0, origins.shape[0], 1 << 17):
        o = origins[a:a + (1 << 17)].float()
        px, py, pz = o[:, 0:1] + lx + 0.5, o[:, 1:2] + ly + 0.5, o[:, 2:3] + lz + 0.5
        rr = torch.sqrt((px - c) ** 2 + (py - c) ** 2 + (pz - c) ** 2) / res
        shell = torch.clamp(1.0 - (rr - radius).abs() / width, 0.0, 1.0)
        noise = torch.zeros_like(px)
        for k in range(4):
            noise += 0.5 ** k * torch.sin(px * freq[k] + phase[k, 0]) * torch.sin(py * freq[k] + phase[k, 1]) \
                * torch.sin(pz * freq[k] + phase[k, 2])
        leaves[a:a + (1 << 17)] = shell * (1.0 + 0.5 * noise)
    return origins.cpu().numpy(), leaves.reshape(-1)


def leaf_points(torch, origins_np, count, seed):
    """Uniform points inside the active leaves (the data-carrying region)."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    o = torch.from_numpy(origins_np).cuda().float()
    k = torch.randint(0, o.shape[0], (count,), device="cuda", generator=g)
    return (o[k] + torch.rand((count, 3), device="cuda", generator=g) * 7.999).contiguous()


def variant_neural(torch, vkl, dev, res, steps, warmup, train_log2=28, infer_log2=26, batch_log2=16,
                   max_resolution=256):
    """Neural-field stand-in for vdb sampling: train on 2^28 exact samples,
    per-leaf bounds with exact fallback, infer 2^26 leaf points; compared with
    the exact vdb sampler on the same points (normalised RMSE = RMSE / value
    range of the volume). max_resolution: the finest hash-grid level (256
    measured fastest at the same accuracy as 512, 1024 and 2048)."""
    origins, leaves = shell_vdb(torch, res)
    n = origins.shape[0]
    vol = dev.vdb(np.full(n, 3, np.uint32), origins, np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32),
                  packed_dense=[leaves])
    s = vol.sampler(filter=vkl.VKL_FILTER_LINEAR)
    span = float(np.diff(vol.value_range())[0])
    log(f"  vdb {res}^3 committed ({n} leaves); training")
    nf = vkl.NeuralField(s, seed=1, max_resolution=max_resolution)
    train_steps = 1 << (train_log2 - batch_log2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = nf.train(train_steps, batch=1 << batch_log2, lr=1e-2, leaf_fraction=0.9, seed=2)
    train_s = time.perf_counter() - t0
    log(f"  trained {train_steps} steps in {train_s:.2f} s")
    pts = leaf_points(torch, origins, 1 << infer_log2, SEED + 10)
    exact = torch.empty(pts.shape[0], device="cuda")
    approx = torch.empty_like(exact)
    ws = torch.empty(pts.shape[0] + 1, device="cuda", dtype=torch.int32)
    ms_exact = timed(torch, lambda: s.sample(pts, out=exact), steps, warmup)
    # tricubic on the same points (the stencil's leaf cells looked up once)
    s_c = vol.sampler(filter=vkl.VKL_FILTER_CUBIC)
    ms_cubic = timed(torch, lambda: s_c.sample(pts, out=approx), steps, warmup)
    s_c.release()
    # the exact sampler on the corner-packed form of the same volume
    vol_p = dev.vdb(np.full(n, 3, np.uint32), origins, np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32),
                    packed_dense=[leaves], corner_packed=True)
    s_p = vol_p.sampler(filter=vkl.VKL_FILTER_LINEAR)
    exact_p = torch.empty_like(exact)
    ms_exact_p = timed(torch, lambda: s_p.sample(pts, out=exact_p), steps, warmup)
    same = bool(((exact_p == exact) | (torch.isnan(exact_p) & torch.isnan(exact))).all())
    s_p.release()
    vol_p.release()
    fb = [0]

    def run_nf():
        fb[0] = nf.sample(pts, fallback=True, out=approx, workspace=ws)[1]

    # Points whose exact value is undefined (filter taps on the background
    # outside the domain) are left out of the error.
    defined = torch.isfinite(exact)
    # Per-leaf tolerances, loosest first: tighter ones bring the normalised
    # RMSE under 1e-3 as more leaves fall back to the exact sampler.
    by_tol = []
    for rel_tol in (1e-2, 5e-3, 2.5e-3):
        tol = rel_tol * span
        approved = nf.finalize(tol, points_per_leaf=64, safety=1.0)
        ms_nf = timed(torch, run_nf, steps, warmup)
        err = (approx - exact)[defined].double()
        by_tol.append({"leaf_tolerance_rel": rel_tol, "approved_leaves": int(approved),
                       "fallback_fraction": round(fb[0] / pts.shape[0], 4),
                       "neural_msamples_s": round(pts.shape[0] / ms_nf / 1e3, 1),
                       "speedup_vs_exact": round(ms_exact / ms_nf, 3),
                       "speedup_vs_exact_corner_packed": round(ms_exact_p / ms_nf, 3),
                       "nrmse": float(torch.sqrt((err ** 2).mean())) / span,
                       "max_abs_err_rel": float(err.abs().max()) / span})
        log(f"  tolerance {rel_tol}: {by_tol[-1]}")
    net_only = torch.empty_like(exact)
    ms_net = timed(torch, lambda: nf.sample(pts, fallback=False, out=net_only), steps, warmup)
    nrmse_net = float(torch.sqrt(((net_only - exact)[defined].double() ** 2).mean())) / span
    # Point order: the same points sorted by 8^3 leaf cell, so that a
    # wavefront's points share leaves, table entries and cache lines; for the
    # network and the exact sampler alike, at the loosest tolerance. The sort
    # is timed on its own.
    nf.finalize(by_tol[0]["leaf_tolerance_rel"] * span, points_per_leaf=64, safety=1.0)
    cell = pts.floor().long() >> 3
    cells = (res + 7) // 8
    key = (cell[:, 2] * cells + cell[:, 1]) * cells + cell[:, 0]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sp = pts[torch.argsort(key)].contiguous()
    torch.cuda.synchronize()
    sort_ms = (time.perf_counter() - t1) * 1e3
    del cell, key
    out_s = torch.empty_like(exact)
    ms_exact_s = timed(torch, lambda: s.sample(sp, out=out_s), steps, warmup)
    fb_s = [0]

    def run_nf_sorted():
        fb_s[0] = nf.sample(sp, fallback=True, out=out_s, workspace=ws)[1]

    ms_nf_s = timed(torch, run_nf_sorted, steps, warmup)
    ms_net_s = timed(torch, lambda: nf.sample(sp, fallback=False, out=out_s), steps, warmup)
    sorted_line = {"sort_ms": round(sort_ms, 2), "exact_msamples_s": round(pts.shape[0] / ms_exact_s / 1e3, 1),
                   "neural_msamples_s": round(pts.shape[0] / ms_nf_s / 1e3, 1),
                   "fallback_fraction": round(fb_s[0] / pts.shape[0], 4),
                   "network_only_msamples_s": round(pts.shape[0] / ms_net_s / 1e3, 1),
                   "speedup_vs_exact": round(ms_exact_s / ms_nf_s, 3),
                   "what": f"the same points sorted by 8^3 leaf cell (torch.argsort, sort_ms), tolerance "
                           f"{by_tol[0]['leaf_tolerance_rel']} of the value range"}
    log(f"  leaf-sorted points: {sorted_line}")
    del sp, out_s
    nf.release()
    return {"volume": f"vdb shell+fBm, {res}^3 index domain, {n} float32 8^3 leaves ({n * 2048 / 2 ** 30:.2f} GiB)",
            "max_resolution": max_resolution,
            "exact_msamples_s": round(pts.shape[0] / ms_exact / 1e3, 1),
            "exact_cubic_msamples_s": round(pts.shape[0] / ms_cubic / 1e3, 1),
            "exact_corner_packed_msamples_s": round(pts.shape[0] / ms_exact_p / 1e3, 1),
            "exact_corner_packed_bitexact": same,
            "neural_msamples_s": by_tol[0]["neural_msamples_s"], "speedup_vs_exact": by_tol[0]["speedup_vs_exact"],
            "nrmse": by_tol[0]["nrmse"], "by_tolerance": by_tol,
            "network_only_msamples_s": round(pts.shape[0] / ms_net / 1e3, 1), "network_only_nrmse": nrmse_net,
            "leaf_sorted_points": sorted_line,
            "undefined_fraction": round(1 - float(defined.float().mean()), 5),
            "train_samples": train_steps << batch_log2, "train_s": round(train_s, 2),
            "final_train_mse": float(hist[-16:].mean()), "points": int(pts.shape[0])}


def variant_raymarch(torch, vkl, vol, n, steps, warmup, width=1920, height=1080, march=512):
    """Ray marching: 1920x1080 orthographic rays along +z through the
    512^3 volume, 512 fixed steps each (samples at step midpoints), fused in
    one kernel (vklComputeRayMarchN)."""
    s = vol.sampler(filter=vkl.VKL_FILTER_LINEAR)
    ys, xs = torch.meshgrid(torch.arange(height, device="cuda", dtype=torch.float32),
                            torch.arange(width, device="cuda", dtype=torch.float32), indexing="ij")
    r = width * height
    org = torch.stack([(xs.reshape(-1) + 0.5) * ((n - 1) / width), (ys.reshape(-1) + 0.5) * ((n - 1) / height),
                       torch.zeros(r, device="cuda")], 1).contiguous()
    d = torch.zeros((r, 3), device="cuda")
    d[:, 2] = 1.0
    tr = torch.zeros((r, 2), device="cuda")
    tr[:, 1] = float(n - 1)
    out = torch.empty(r, device="cuda")
    ms = timed(torch, lambda: s.ray_march(org, d, tr, march, out=out), steps, warmup)
    samples = r * march
    alg = r * (12 + 12 + 8 + 4) + n ** 3 * 4
    res = {"rays": r, "steps_per_ray": march, "ms": round(ms, 4), "msamples_s": round(samples / ms / 1e3, 1),
           "mrays_s": round(r / ms / 1e3, 2), "algorithmic_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "note": "coherent gathers hit L1/L2; bound by vector ALU and L1 issue, not HBM"}
    # The ray-march iterator renderer: intervals, sub-steps of nominalDeltaT,
    # transfer-function compositing to opacity 0.99 (vklComputeRayMarchIntervalsN).
    ctx = s.interval_context()
    lo, hi = vol.value_range()
    tf = torch.rand((32, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(SEED))
    tf[:, 3] = torch.linspace(0.0, 0.02, 32, device="cuda")
    got = {}

    def run():
        got["rgba"], got["cnt"] = ctx.ray_march(org, d, tf, (lo, hi), t_ranges=tr, sampling_rate=1.0)

    ms_i = timed(torch, run, steps, warmup)
    taken = float(got["cnt"].double().sum())
    res["intervals_compositing"] = {
        "ms": round(ms_i, 4), "mrays_s": round(r / ms_i / 1e3, 2), "samples_per_ray": round(taken / r, 1),
        "msamples_s": round(taken / ms_i / 1e3, 1),
        "opaque_rays": round(float((got["rgba"][:, 3] >= 0.99).float().mean()), 4)}
    # The hit-iterator renderer: hits of one iso value, gradient normals, two
    # shadow hit queries per hit (vklComputeHitRenderN).
    iso = float(lo + 0.6 * (hi - lo))
    hctx = s.hit_context([iso])

    def run_h():
        got["rgba"], got["cnt"] = hctx.render_hits(org, d, tf, (lo, hi), t_ranges=tr)

    ms_h = timed(torch, run_h, steps, warmup)
    hits = float(got["cnt"].double().sum())
    res["hit_render"] = {"iso": round(iso, 4), "ms": round(ms_h, 4), "mrays_s": round(r / ms_h / 1e3, 2),
                         "hits_per_ray": round(hits / r, 3)}
    return res


def variant_woodcock(torch, vkl, vol, n, steps, warmup, width=1920, height=1080, sigma=0.5):
    """Free-flight sampling: 1920x1080 rays along +z through the
    512^3 volume, Woodcock tracking with majorant `sigma` per voxel length and
    an opacity ramp 0 -> 0.05 over the volume's value range
    (vklComputeWoodcockN; TEA states continue across steps)."""
    s = vol.sampler(filter=vkl.VKL_FILTER_LINEAR)
    ys, xs = torch.meshgrid(torch.arange(height, device="cuda", dtype=torch.float32),
                            torch.arange(width, device="cuda", dtype=torch.float32), indexing="ij")
    r = width * height
    org = torch.stack([(xs.reshape(-1) + 0.5) * ((n - 1) / width), (ys.reshape(-1) + 0.5) * ((n - 1) / height),
                       torch.zeros(r, device="cuda")], 1).contiguous()
    d = torch.zeros((r, 3), device="cuda")
    d[:, 2] = 1.0
    tr = torch.zeros((r, 2), device="cuda")
    tr[:, 1] = float(n - 1)
    rng = torch.zeros((r, 2), dtype=torch.int32, device="cuda")
    rng[:, 1] = torch.arange(r, dtype=torch.int32, device="cuda")
    lo, hi = vol.value_range()
    tf = torch.zeros((16, 4), device="cuda")
    tf[:, 3] = torch.linspace(0.0, 0.05, 16, device="cuda")
    res = {}

    def run():
        res["ev"], _, res["cnt"] = s.woodcock(org, d, tr, rng, sigma, tf, (lo, hi))

    ms = timed(torch, run, steps, warmup)
    coll = float(res["cnt"].double().sum())
    return {"rays": r, "sigma_max": sigma, "ms": round(ms, 4), "mrays_s": round(r / ms / 1e3, 2),
            "collisions_per_ray": round(coll / r, 2), "gcollisions_s": round(coll / ms / 1e6, 2),
            "escaped": round(float(torch.isinf(res["ev"]).float().mean()), 4),
            "note": "each tentative collision: two 8-round TEA draws + one trilinear gather; the lanes of a "
                    "wavefront (adjacent pixels) advance through similar depths, so their gathers share cache lines"}


def boundary_rays(torch, lo, hi, count, seed):
    """Origins uniform on the faces of the box [lo, hi], directions towards a
    uniform point inside it (unit length)."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    lo_t = torch.tensor(lo, device="cuda", dtype=torch.float32)
    ext = torch.tensor(hi, device="cuda", dtype=torch.float32) - lo_t
    org = lo_t + torch.rand((count, 3), device="cuda", generator=g) * ext
    axis = torch.randint(0, 3, (count,), device="cuda", generator=g)
    side = torch.randint(0, 2, (count,), device="cuda", generator=g).float()
    rows = torch.arange(count, device="cuda")
    org[rows, axis] = lo_t[axis] + side * ext[axis]
    tgt = lo_t + torch.rand((count, 3), device="cuda", generator=g) * ext
    d = tgt - org
    d = d / torch.linalg.norm(d, dim=1, keepdim=True)
    return org.contiguous(), d.contiguous()


def iterator_cpu_baseline(origins, leaves, org, d, gpu_i, gpu_h, max_i, max_h, seconds=2.0):
    """The oracle's interval and hit iteration (C++ restatement of the
    reference's iterators, -O3 -march=native, every usable host core) over the
    first k of the same rays, k doubled from 2^12 until one pass takes
    `seconds`; best of 3 passes at that k. Its results are checked bit for bit
    against the GPU's for those rays."""
    import oracle as O

    O.use_native()
    threads, nproc, model = host_cpu()
    n = origins.shape[0]
    t0 = time.perf_counter()
    grid = O.Grid.vdb(np.full(n, 3, np.uint32), origins, np.full(n, 1, np.uint32), packed_dense=[leaves])
    commit_s = time.perf_counter() - t0
    org_np, d_np = org.cpu().numpy(), d.cpu().numpy()
    out = {"cores": threads, "nproc": nproc, "cpu_model": model, "kind": "port",
           "oracle_commit_s": round(commit_s, 2)}
    for name, values, cap, gpu in (("interval", None, max_i, gpu_i), ("hit", [0.5], max_h, gpu_h)):
        k = 1 << 12
        while True:
            t0 = time.perf_counter()
            res, cnt = grid.iterate_many(org_np[:k], d_np[:k], cap, values=values, threads=threads)
            el = time.perf_counter() - t0
            if el >= seconds or 2 * k > org_np.shape[0]:
                break
            k *= 2
        best = el
        for _ in range(2):
            t0 = time.perf_counter()
            res, cnt = grid.iterate_many(org_np[:k], d_np[:k], cap, values=values, threads=threads)
            best = min(best, time.perf_counter() - t0)
        g_out, g_cnt = gpu[0][:k].cpu().numpy(), gpu[1][:k].cpu().numpy().astype(np.uint32)
        live = np.arange(cap)[None, :] < cnt[:, None]
        same = bool(np.array_equal(cnt, g_cnt) and np.array_equal(res.view(np.uint32)[live],
                                                                  g_out.view(np.uint32)[live]))
        out[f"{name}_mrays_s"] = round(k / best / 1e6, 3)
        out[f"{name}_rays"] = k
        out[f"{name}_bitwise_equal_to_gpu"] = same
    out["sample"] = (f"the first {out['interval_rays']} / {out['hit_rays']} of the 2^21 rays (interval / hit), "
                     f"best of 3; oracle/vkl_oracle.cpp (g++ -O3 -march=native -ffp-contract=off) on "
                     f"{threads} threads")
    del grid
    return out


def variant_iterators(torch, vkl, dev, res, steps, warmup, rays_log2=21, cpu=False):
    """Sparse vdb (sphere shell + fBm, `res`^3 index domain, 8^3 float32
    leaves): vklIterateIntervalsBatch (default hint) and vklIterateHitsBatch
    (iso 0.5) over 2^21 rays entering the bounding box; with `cpu`, the
    oracle's iterators timed beside them on a bounded sample of the rays."""
    origins, leaves = shell_vdb(torch, res)
    n = origins.shape[0]
    log(f"  vdb {res}^3: {n} leaves generated; committing")
    vol = dev.vdb(np.full(n, 3, np.uint32), origins, np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32),
                  packed_dense=[leaves])
    log("  committed")
    s = vol.sampler(filter=vkl.VKL_FILTER_LINEAR)
    bb = vol.bounding_box()
    org, d = boundary_rays(torch, bb[:3], bb[3:], 1 << rays_log2, SEED + 11)
    ictx = s.interval_context()
    hctx = s.hit_context([0.5])
    res_i, res_h = {}, {}

    def run_i():
        res_i["out"], res_i["cnt"] = ictx.batch(org, d, max_per_ray=256)
        torch.cuda.synchronize()
        log("  interval batch done")

    def run_h():
        res_h["out"], res_h["cnt"] = hctx.batch(org, d, max_per_ray=16)
        torch.cuda.synchronize()
        log("  hit batch done")

    ms_i = timed(torch, run_i, max(1, steps // 2), 1)
    ms_h = timed(torch, run_h, max(1, steps // 2), 1)
    # the same rays at the finest interval resolution (hint 1: depth 3)
    ictx1 = s.interval_context(hint=1.0)
    cnt1 = []
    ms_i1 = timed(torch, lambda: cnt1.append(ictx1.batch(org, d, max_per_ray=1024)[1]), max(1, steps // 2), 1)
    per_ray1 = float(cnt1[-1].double().mean())
    del cnt1
    ci, ch = res_i["cnt"].double(), res_h["cnt"].double()
    r = org.shape[0]
    # fused ray march over the same rays: 256 samples between box entry (the
    # origin lies on a face) and exit
    lo_t = torch.tensor(bb[:3], device="cuda", dtype=torch.float32)
    hi_t = torch.tensor(bb[3:], device="cuda", dtype=torch.float32)
    inv = 1.0 / d
    t_exit = torch.maximum((lo_t - org) * inv, (hi_t - org) * inv).amin(dim=1).clamp_min(0.0)
    tr = torch.stack([torch.zeros_like(t_exit), t_exit], 1).contiguous()
    march = 256
    acc = torch.empty(r, device="cuda")
    ms_m = timed(torch, lambda: s.ray_march(org, d, tr, march, out=acc), max(1, steps // 2), 1)
    log("  ray march done")
    # sampling along the rays at nominalDeltaT inside the intervals, with
    # transfer-function compositing (vklComputeRayMarchIntervalsN; a faint
    # opacity so that most rays cross the whole shell)
    lo_v, hi_v = vol.value_range()
    tf = torch.rand((32, 4), device="cuda", generator=torch.Generator(device="cuda").manual_seed(SEED + 12))
    tf[:, 3] = torch.linspace(0.0, 0.002, 32, device="cuda")
    rmi = {}

    def run_rmi():
        rmi["rgba"], rmi["cnt"] = ictx.ray_march(org, d, tf, (lo_v, hi_v), sampling_rate=1.0)

    ms_rmi = timed(torch, run_rmi, max(1, steps // 2), 1)
    rmi_samples = float(rmi["cnt"].double().sum())
    rmi_opaque = float((rmi["rgba"][:, 3] >= 0.99).double().mean())
    log("  interval ray march done")
    # ragged ray batches of four contexts (two interval depths, two iso sets)
    # in one plan, against one batch call per context
    ctxs = [ictx, ictx1, hctx, s.hit_context([0.25])]
    sizes, caps = [1 << 20, 1 << 18, 1 << 19, 1 << 17], [256, 1024, 16, 16]
    batches, off = [], 0
    for c, k, cap in zip(ctxs, sizes, caps):
        batches.append((c, org[off:off + k], d[off:off + k], cap))
        off += k
    plan = vkl.RayBatchPlan(dev, batches)
    ms_plan = timed(torch, plan.execute, max(1, steps // 2), 1)
    ms_calls = timed(torch, lambda: [c.batch(o, dd, max_per_ray=cap) for c, o, dd, cap in batches],
                     max(1, steps // 2), 1)
    plan.release()
    log("  ray batch plan done")
    cpu_line = None
    if cpu:
        log("  oracle iterators on the host ...")
        cpu_line = iterator_cpu_baseline(origins, leaves.cpu().numpy(), org, d, (res_i["out"], res_i["cnt"]),
                                         (res_h["out"], res_h["cnt"]), 256, 16)
        log(f"  host: {cpu_line['interval_mrays_s']} / {cpu_line['hit_mrays_s']} Mrays/s")
    return {"volume": f"vdb shell+fBm, {res}^3 index domain, {n} float32 8^3 leaves ({n * 2048 / 2 ** 30:.2f} GiB)",
            "cpu_baseline": cpu_line,
            "rays": r, "interval_mrays_s": round(r / ms_i / 1e3, 2), "interval_ms": round(ms_i, 3),
            "intervals_per_ray": round(float(ci.mean()), 2), "max_intervals_per_ray": int(ci.max()),
            "interval_hint1_mrays_s": round(r / ms_i1 / 1e3, 2), "interval_hint1_ms": round(ms_i1, 3),
            "intervals_per_ray_hint1": round(per_ray1, 2),
            "hit_mrays_s": round(r / ms_h / 1e3, 2), "hit_ms": round(ms_h, 3),
            "hits_per_ray": round(float(ch.mean()), 3),
            "ray_batch_plan": {"batches": len(batches), "rays": off, "plan_ms": round(ms_plan, 3),
                               "one_call_per_batch_ms": round(ms_calls, 3),
                               "what": "intervals at hint 0.5 and 1.0, hits at iso 0.5 and 0.25; 2^20, 2^18, "
                                       "2^19, 2^17 rays; one graph with a branch per batch"},
            "interval_raymarch": {"ms": round(ms_rmi, 3), "mrays_s": round(r / ms_rmi / 1e3, 2),
                                  "samples_per_ray": round(rmi_samples / r, 1),
                                  "msamples_s": round(rmi_samples / ms_rmi / 1e3, 1),
                                  "opaque_rays": round(rmi_opaque, 4),
                                  "what": "samples at nominalDeltaT inside the intervals (hint 0.5) with "
                                          "transfer-function compositing, vklComputeRayMarchIntervalsN"},
            "raymarch_steps_per_ray": march, "raymarch_mrays_s": round(r / ms_m / 1e3, 2),
            "raymarch_msamples_s": round(r * march / ms_m / 1e3, 1)}


def host_cpu():
    """(threads this process may use, nproc, CPU model)."""
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return usable, os.cpu_count() or usable, model


def cpu_baseline(data_np, pts_np, gpu_out_np):
    """The oracle (a C++ restatement of the reference's CPU algorithm, built
    here with -O3 -march=native and without FMA contraction) on every usable
    host core, points split into contiguous chunks; best of 5 after a warm-up
    of at least 1 s (the reference benchmark harness's protocol,
    testing/apps/benchmark_suite/utility.h:20-40). Its samples are checked
    against the GPU's for the same points."""
    import oracle as O

    O.use_native()
    threads, nproc, model = host_cpu()
    grid = O.Grid.dense(data_np)
    sub = np.ascontiguousarray(pts_np)
    t_end = time.perf_counter() + 1.0
    while time.perf_counter() < t_end:
        out = grid.sample(sub, threads=threads)
    best = float("inf")
    for _ in range(5):
        t0 = time.perf_counter()
        out = grid.sample(sub, threads=threads)
        best = min(best, time.perf_counter() - t0)
    same = bool(np.all((out.view(np.uint32) == gpu_out_np.view(np.uint32)) | (np.isnan(out) & np.isnan(gpu_out_np))))
    count = sub.shape[0]
    return {"value": round(count / best / 1e6, 2), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "nproc": nproc, "cpu_model": model, "bitwise_equal_to_gpu": same,
            "sample": f"the first {count} of the 2^26 points, best of 5 after a >= 1 s warm-up; C++ restatement "
                      f"of the reference algorithm (oracle/vkl_oracle.cpp, g++ -O3 -march=native "
                      f"-ffp-contract=off) on {threads} threads"}


def copy_peak(torch, steps=20):
    """Sustained device-to-device copy rate (1 GiB read + 1 GiB written per
    call): the measured HBM denominator beside the 8 TB/s nominal peak."""
    n = 1 << 28
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    ms = timed(torch, lambda: dst.copy_(src), steps, 3)
    del src, dst
    return {"gbs": round(2 * 4 * n / (ms * 1e-3) / 1e9, 1), "bytes_per_call": 2 * 4 * n, "ms": round(ms, 4)}


def random_line_ceiling(torch, npts, steps=10, rows=1 << 27):
    """The measured ceiling of a sampler that reads one random 32-byte record
    per point in random order: torch.take of one 4-byte element from each of
    npts uniformly random 32-byte records of a rows * 32 B array (4 GiB by
    default: 8 B index in, 4 B value out, one random cache line per point)."""
    buf = torch.zeros(rows * 8, device="cuda")
    g = torch.Generator(device="cuda").manual_seed(SEED + 20)
    idx = torch.randint(0, rows, (npts,), device="cuda", generator=g) * 8
    out = torch.empty(npts, device="cuda")
    ms = timed(torch, lambda: torch.take(buf, idx, out=out), steps, 2)
    del buf, idx, out
    return {"gsamples_s": round(npts / ms / 1e6, 2), "ms": round(ms, 4), "records": rows, "record_bytes": 32}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--points-log2", type=int, default=26)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--brick", default="0,0,0", help="device voxel layout: brick shifts sx,sy,sz (0,0,0 = x-fastest)")
    ap.add_argument("--layout", choices=("linear", "packed"), default="packed",
                    help="device voxel layout: the caller's x-fastest array (with --brick) or corner-packed cells")
    ap.add_argument("--binned", choices=("auto", "on", "off"), default="auto",
                    help="spatially binned sampling (auto: on for large volumes and batches)")
    ap.add_argument("--variants", default="dense128,layouts,dense512_matrix,raymarch,woodcock,batched,neural,iterators",
                    help="comma list of extra measurements: dense128, layouts, dense512_matrix, raymarch, woodcock, batched, "
                         "neural, "
                         "iterators "
                         "('' = none)")
    ap.add_argument("--nf-res", type=int, default=4096, help="index-domain edge of the neural-field vdb")
    ap.add_argument("--vdb-res", type=int, default=4096, help="index-domain edge of the iterator vdb")
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
    packed = args.layout == "packed"
    vol = dev.structured_regular(vol_t, brick_shift=None if packed else brick, corner_packed=packed or None)
    sampler = vol.sampler(filter=vkl.VKL_FILTER_LINEAR, binned={"auto": None, "on": True, "off": False}[args.binned])
    binned = args.binned == "on" or (args.binned == "auto" and not packed and npts >= (1 << 20)
                                     and n ** 3 * 4 >= (64 << 20))
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

    # HBM traffic of the headline kernel: not measurable from inside the
    # process; it is the stored result of the rocprofv3 FETCH_SIZE /
    # WRITE_SIZE runs (copy-calibrated) over this same command, committed
    # under profiles/, for the configuration they measured.
    traffic = None
    tname = "kSample_dense512_f32_linear_packed.traffic.txt"
    tfile = os.path.join(ROOT, "profiles", tname)
    if packed and n == 512 and npts == 1 << 26 and os.path.exists(tfile):
        kv = dict(line.split()[:2] for line in open(tfile) if line.strip() and not line.startswith("#"))
        traffic = int(kv["fetch_bytes_per_launch"]) + int(kv["write_bytes_per_launch"])
    # The same call on host arrays (pinned): the library stages the points in
    # and the samples out over PCIe; wall time of the synchronous call.
    log("timing the call on host (pinned) arrays ...")
    pts_h = pts.cpu().pin_memory()
    out_h = torch.empty(npts, dtype=torch.float32).pin_memory()
    sampler.sample(pts_h, out=out_h)
    t0 = time.perf_counter()
    for _ in range(3):
        sampler.sample(pts_h, out=out_h)
    ms_host = (time.perf_counter() - t0) / 3 * 1e3
    host_same = bool(torch.equal(out_h.view(torch.int32), out.cpu().view(torch.int32)))
    del pts_h, out_h
    # and on pageable host arrays (numpy)
    pts_p = pts.cpu().numpy()
    out_p = np.empty(npts, np.float32)
    sampler.sample(pts_p, out=out_p)
    t0 = time.perf_counter()
    for _ in range(3):
        sampler.sample(pts_p, out=out_p)
    ms_pageable = (time.perf_counter() - t0) / 3 * 1e3
    host_same = host_same and bool(np.array_equal(out_p.view(np.int32), out.cpu().numpy().view(np.int32)))
    del pts_p, out_p
    log("measuring the copy peak and the random-record gather ceiling ...")
    peak = copy_peak(torch)
    ceiling = random_line_ceiling(torch, npts)
    # the same random gathers over smaller footprints (Infinity Cache / L2)
    by_footprint = {f"{(r * 32) >> 20} MiB": random_line_ceiling(torch, npts, rows=r)["gsamples_s"]
                    for r in (1 << 19, 1 << 24)}
    by_footprint[f"{(1 << 27) * 32 >> 20} MiB"] = ceiling["gsamples_s"]
    log(f"copy {peak['gbs']} GB/s; random 32-B record gather {ceiling['gsamples_s']} Gsamples/s")
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
                   "device_layout": "corner-packed cells (8 x 4 B per cell)" if packed else
                   "x-fastest" if brick == (0, 0, 0) else "bricks %dx%dx%d" % tuple(1 << b for b in brick),
                   "sampling_path": "binned (bin scatter + LDS-brick sample)" if binned else "direct gather",
                   "seed": SEED},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": f"profiles/{tname} (stored rocprofv3 FETCH_SIZE + WRITE_SIZE of this command)"
                     if traffic is not None else None,
                     "algorithmic_bytes_per_launch": bytes_per_launch, "kernel_ms": round(ms_event, 4),
                     "measured_copy_peak_gbs": peak["gbs"],
                     "frac_of_measured_copy_peak": round(achieved / peak["gbs"], 4),
                     "random_record_gather_ceiling": {
                         "gsamples_s": ceiling["gsamples_s"],
                         "what": "torch.take of one element from each of 2^26 uniformly random 32-byte records "
                                 "of a 4 GiB array: one random cache line per point, the floor of any sampler "
                                 "that reads random points in random order",
                         "kernel_frac_of_ceiling": round(npts / (ms_event * 1e-3) / 1e9 / ceiling["gsamples_s"], 4),
                         "gsamples_s_by_footprint": by_footprint}},
        "host_pointers": {"ms": round(ms_host, 3), "msamples_s": round(npts / ms_host / 1e3, 1),
                          "pageable_ms": round(ms_pageable, 3),
                          "bitwise_equal_to_device_call": host_same,
                          "what": "the same vklComputeSampleN on pinned (ms) and pageable (pageable_ms) host "
                                  "arrays: points staged in and samples out over PCIe inside the call, in "
                                  "overlapped chunks (wall time); not the headline value"},
    }
    variants = [v for v in args.variants.split(",") if v]
    extra = {}
    for v in variants:
        log(f"variant {v} ...")
        if v == "dense128":
            extra["dense128"] = variant_dense128(torch, vkl, dev, args.steps, args.warmup, cpu=not args.no_cpu_baseline)
        elif v == "layouts":
            extra["layouts"] = variant_layouts(torch, vkl, dev, vol_t, pts, out, args.steps, args.warmup,
                                               bytes_per_launch)
        elif v == "raymarch":
            rv = dev.structured_regular(vol_t, corner_packed=True)
            extra["raymarch"] = variant_raymarch(torch, vkl, rv, n, args.steps, args.warmup)
            extra["raymarch"]["layout"] = "corner-packed"
            rv.release()
            rv = dev.structured_regular(vol_t)
            extra["raymarch_x_fastest"] = variant_raymarch(torch, vkl, rv, n, args.steps, args.warmup)
            rv.release()
        elif v == "woodcock":
            rv = dev.structured_regular(vol_t, corner_packed=True)
            extra["woodcock"] = variant_woodcock(torch, vkl, rv, n, args.steps, args.warmup)
            extra["woodcock"]["layout"] = "corner-packed"
            rv.release()
            rv = dev.structured_regular(vol_t, corner_packed=False)
            extra["woodcock_x_fastest"] = variant_woodcock(torch, vkl, rv, n, args.steps, args.warmup)
            rv.release()
        elif v == "dense512_matrix":
            extra["dense512_matrix"] = variant_dense512_matrix(torch, vkl, dev, vol_t, pts, args.steps, args.warmup, U)
        elif v == "batched":
            extra["batched"] = variant_batched(torch, vkl, dev, args.steps, args.warmup, packed=True)
            extra["batched_x_fastest"] = variant_batched(torch, vkl, dev, args.steps, args.warmup)
        elif v == "neural":
            extra["neural_field"] = variant_neural(torch, vkl, dev, args.nf_res, args.steps, args.warmup)
        elif v == "iterators":
            extra["vdb_iterators"] = variant_iterators(torch, vkl, dev, args.vdb_res, args.steps, args.warmup,
                                                       cpu=not args.no_cpu_baseline)
        log(f"  {json.dumps(extra[list(extra)[-1]])}")
    if extra:
        result["variants"] = extra
    if not args.no_cpu_baseline:
        log("timing the CPU oracle ...")
        data_np = vol_t.cpu().numpy()
        count = 1 << 23
        pts_np = pts[:count].cpu().numpy()
        sampler.sample(pts, out=out)
        result["cpu_baseline"] = cpu_baseline(data_np, pts_np, out[:count].cpu().numpy())
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
