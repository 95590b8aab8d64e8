"""GPU parity for the reference's special cases and the configurations no
other test covers: the torus regression rays through vklIterateInterval, the
VDB wavelet value range, error voxels with the nearest filter, 64-bit dense
addressing (more than 2^32 voxel elements), the headline corner-packed
512^3 float32 layout at full size, and the batched multi-volume plan at the
shape of its benchmark. Each compares the C-ABI results with the oracle."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same
from tests.test_oracle_pins import (DIMS, assert_special_case_intervals, dims_for_64_32bit_addressing,
                                    torus_rays)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def log(msg):
    print(msg, flush=True)


def test_torus_rays_scalar_api(device):
    origins, leaves = V.torus_vdb_leaves()
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    g = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    s = vol.sampler()
    ctx = s.interval_context()
    for org, d in torus_rays():
        log(f"torus ray dir.z bits {np.float32(d[2]).view(np.uint32):#010x}")
        got = ctx.intervals(org, d)
        assert_special_case_intervals(got)
        want = g.intervals(org, d)
        assert got.shape == want.shape
        assert_same(got, want)
    ctx.release()
    s.release()
    vol.release()


@pytest.mark.parametrize("dtype", [np.float32, np.float16])
@pytest.mark.parametrize("packed", [True, False])
def test_vdb_wavelet_value_range(device, dtype, packed):
    origins, (leaves,) = V.dense_vdb_leaves(128, [V.wavelet])
    leaves = leaves.astype(dtype)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    if packed:
        vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
        g = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    else:
        vol = device.vdb(lv, origins, fm, leaf_data=[[leaves[i]] for i in range(n)])
        g = O.Grid.vdb(lv, origins, fm, leaf_data=[[leaves[i]] for i in range(n)])
    lo, hi = vol.value_range()
    assert abs((hi - lo) - 6.0) <= 0.001 * 6.0
    assert (np.float32(lo), np.float32(hi)) == tuple(np.float32(x) for x in g.value_range())
    vol.release()


@pytest.mark.parametrize("depth", [0, 1, 2])
def test_error_voxels_nearest(device, depth):
    rng = np.random.default_rng(91)
    nx, ny, nz = DIMS
    arr = (rng.standard_normal((nz, ny, nx)) * 3).astype(np.float32)
    origins, leaves = V.leaves_from_dense(arr)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    g = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    pts = V.uniform_points(20000, (-1, -1, -1), (nx + 1, ny + 1, nz + 1), seed=93)
    s = vol.sampler(filter=vkl.VKL_FILTER_NEAREST, max_sampling_depth=depth)
    g.set_sampler(O.NEAREST, max_sampling_depth=depth)
    assert_same(s.sample(torch.from_numpy(pts).cuda()).cpu().numpy(), g.sample(pts))
    s.release()
    vol.release()


# ---- 64 / 32-bit addressing ---------------------------------------------------------------

def u8_ramp(nx, ny, nz):
    """(x + 3 y + 7 z) mod 256 as uint8, built slab by slab."""
    out = np.empty((nz, ny, nx), np.uint8)
    base = ((np.arange(nx, dtype=np.int64)[None, :] + 3 * np.arange(ny, dtype=np.int64)[:, None]) & 255)
    base = base.astype(np.uint8)
    for z in range(nz):
        np.add(base, np.uint8((7 * z) & 255), out=out[z])  # uint8 arithmetic wraps mod 256
    return out


def vertex_check(s, dims, fn, step=32, dev_data=None):
    """structured_regular_volume.h:96-130 / sampling_utility.h:35-119: samples
    at the vertices on a `step` lattice equal the generating function."""
    nx, ny, nz = dims
    gx, gy, gz = np.meshgrid(np.arange(0, nx, step), np.arange(0, ny, step), np.arange(0, nz, step), indexing="ij")
    v = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    got = s.sample(torch.from_numpy(v.astype(np.float32)).cuda()).cpu().numpy()
    want = fn(v[:, 0], v[:, 1], v[:, 2]).astype(np.float32)
    bad = np.nonzero(got != want)[0]
    if len(bad) and dev_data is not None:
        # where do the wrong samples come from: the voxels in device memory
        # or the sampling?
        vb = v[bad[:8]]
        on_dev = dev_data[torch.from_numpy(vb[:, 2]).cuda(), torch.from_numpy(vb[:, 1]).cuda(),
                          torch.from_numpy(vb[:, 0]).cuda()].cpu().numpy()
        log(f"{len(bad)} bad vertices; first {vb.tolist()} got {got[bad[:8]].tolist()} want "
            f"{want[bad[:8]].tolist()} device voxels {on_dev.tolist()}; "
            f"bad z values {np.unique(v[bad, 2]).tolist()[:20]} y {np.unique(v[bad, 1]).tolist()[:60]}")
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("case", ["float32_reference_dims", "uint8_corner_packed_820", "uint8_raw_1626"])
def test_64_32bit_addressing(device, case):
    """float32 at the reference's 64 / 32-bit dimension (813^3, 2.1 GB) and
    two uint8 volumes past 2^32 voxel elements, where this device switches
    to 64-bit offsets: corner-packed 820^3 (8 elements per cell, 4.4e9) and
    x-fastest 1626^3 (4.3e9)."""
    if case == "float32_reference_dims":
        n = dims_for_64_32bit_addressing(4)
        x = np.arange(n, dtype=np.float32)
        data = (x[None, None, :] + np.float32(2) * x[None, :, None] + np.float32(0.5) * x[:, None, None])
        data = data.astype(np.float32)

        def fn(a, b, c):
            return np.float32(a) + np.float32(2) * np.float32(b) + np.float32(0.5) * np.float32(c)
        packed = False
    else:
        n = 820 if "820" in case else 1626
        data = u8_ramp(n, n, n)

        def fn(a, b, c):
            return ((a + 3 * b + 7 * c) & 255).astype(np.float32)
        packed = "packed" in case
    log(f"{case}: {n}^3 {data.dtype}, corner packed {packed}")
    dev_data = torch.from_numpy(data).cuda()
    vol = device.structured_regular(dev_data, corner_packed=packed or None)
    s = vol.sampler()
    vertex_check(s, (n, n, n), lambda a, b, c: fn(a, b, c), dev_data=dev_data)
    grid = O.Grid.dense(data)
    pts = V.uniform_points(200000, (-1, -1, -1), (n, n, n), seed=94)
    # points near the far corner, whose offsets pass 2^32 elements
    pts[:20000] = V.uniform_points(20000, (n - 40, n - 40, n - 40), (n, n, n), seed=95)
    assert_same(s.sample(torch.from_numpy(pts).cuda()).cpu().numpy(), grid.sample(pts, threads=16))
    s.release()
    vol.release()
    del grid
    torch.cuda.empty_cache()


def test_full_size_512_corner_packed(device):
    """The headline layout (corner-packed 512^3 float32, 4 GiB of cell
    records) at full size: linear samples and gradients vs the oracle."""
    data = V.wavelet_volume(512)
    vol = device.structured_regular(torch.from_numpy(data).cuda(), corner_packed=True)
    s = vol.sampler(filter=vkl.VKL_FILTER_LINEAR)
    grid = O.Grid.dense(data)
    pts = V.uniform_points(1 << 20, (-1, -1, -1), (512, 512, 512), seed=96)
    p = torch.from_numpy(pts).cuda()
    assert_same(s.sample(p).cpu().numpy(), grid.sample(pts, threads=16))
    sub = pts[:200000]
    assert_same(s.gradient(torch.from_numpy(sub).cuda()).cpu().numpy(), grid.gradient(sub))
    s.release()
    vol.release()


def test_batch_plan_benchmark_shape(device):
    """The batched plan at its benchmark shape: 64 volumes with seeded cube
    sides in [32, 384], float32 and half alternating, log-uniform point
    counts in [1e4, 1e7]; every volume's samples vs the oracle on a
    subsample, and vs vklComputeSampleN in full."""
    rng = np.random.default_rng(97)
    g = torch.Generator(device="cuda").manual_seed(98)
    keep, samplers, pts, outs, sides = [], [], [], [], []
    for i in range(64):
        side = int(rng.integers(32, 385))
        dt = torch.float32 if i % 2 == 0 else torch.float16
        data = torch.rand((side, side, side), device="cuda", generator=g).to(dt)
        v = device.structured_regular(data)
        s = v.sampler(filter=vkl.VKL_FILTER_LINEAR)
        count = int(np.exp(rng.uniform(np.log(1e4), np.log(1e7))))
        p = torch.rand((count, 3), device="cuda", generator=g) * float(side + 1) - 1.0
        keep.append((data, v))
        samplers.append(s)
        pts.append(p)
        outs.append(torch.empty(count, device="cuda"))
        sides.append(side)
    log(f"64 volumes, {sum(int(p.shape[0]) for p in pts)} points")
    plan = vkl.SampleBatchPlan(device, samplers, pts, outs, autotune=True)
    plan.execute()
    torch.cuda.synchronize()
    for i in range(64):
        got = outs[i].cpu().numpy()
        assert_same(got, samplers[i].sample(pts[i]).cpu().numpy())
        data_np = keep[i][0].cpu().numpy()
        sub = pts[i][:4000].cpu().numpy()
        assert_same(got[:4000], O.Grid.dense(data_np).sample(sub))
    plan.release()
    for s in samplers:
        s.release()
    for _, v in keep:
        v.release()
