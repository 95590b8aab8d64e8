"""Fused ray marching (vklComputeRayMarchN) equals summing the oracle's samples
at the same march points in step order: dense (every filter) and vdb."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def march_reference(grid, org, d, tr, steps):
    """Oracle samples at t_s = t0 + (s + 0.5) dt, binary32 like the kernel
    (numpy float32 arithmetic has no contraction), summed in step order."""
    org, d, tr = (np.asarray(a, np.float32) for a in (org, d, tr))
    t0 = tr[:, 0]
    dt = ((tr[:, 1] - t0) / np.float32(steps)).astype(np.float32)
    acc = np.zeros(org.shape[0], np.float32)
    for k in range(steps):
        t = (t0 + (np.float32(k) + np.float32(0.5)) * dt).astype(np.float32)
        p = (org + t[:, None] * d).astype(np.float32)
        acc = (acc + grid.sample(p)).astype(np.float32)
    return acc


def rays(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, np.float32), np.asarray(hi, np.float32)
    org = (lo + rng.random((count, 3), dtype=np.float32) * (hi - lo)).astype(np.float32)
    d = rng.standard_normal((count, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tr = np.stack([rng.random(count, dtype=np.float32) * 2, 20 + rng.random(count, dtype=np.float32) * 30], 1)
    return org, d.astype(np.float32), tr.astype(np.float32)


@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR])
def test_dense_ray_march_bitexact(device, filt):
    rng = np.random.default_rng(71)
    data = rng.standard_normal((40, 48, 56)).astype(np.float32)
    vol = device.structured_regular(data, origin=(-1.0, 0.5, 2.0), spacing=(0.75, 1.0, 1.25))
    grid = O.Grid.dense(data, origin=(-1.0, 0.5, 2.0), spacing=(0.75, 1.0, 1.25))
    grid.set_sampler(filt)
    org, d, tr = rays(3000, (-3, -1, 0), (42, 50, 52), 72)
    steps = 37
    got = vol.sampler(filter=filt).ray_march(torch.from_numpy(org).cuda(), torch.from_numpy(d).cuda(),
                                            torch.from_numpy(tr).cuda(), steps)
    assert_same(got.cpu().numpy(), march_reference(grid, org, d, tr, steps))
    # host pointers give the same
    assert_same(vol.sampler(filter=filt).ray_march(org, d, tr, steps), got.cpu().numpy())


def test_vdb_ray_march_bitexact(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    org, d, tr = rays(2000, (10, 10, 10), (118, 118, 118), 73)
    tr[:, 1] *= 3
    got = vol.sampler().ray_march(org, d, tr, 64)
    assert_same(got, march_reference(grid, org, d, tr, 64))


def test_ray_march_rejects_zero_steps(device):
    data = np.zeros((8, 8, 8), np.float32)
    vol = device.structured_regular(data)
    org, d, tr = rays(4, (0, 0, 0), (7, 7, 7), 74)
    s = vol.sampler()
    vkl.lib().vklComputeRayMarchN(vkl.C.byref(s.h), 4, org.ctypes.data, d.ctypes.data, tr.ctypes.data, 0,
                                  np.zeros(4, np.float32).ctypes.data, 0)
    code, msg = device.last_error()
    device.errors.clear()
    assert code == vkl.VKL_INVALID_ARGUMENT and "numSteps" in msg
