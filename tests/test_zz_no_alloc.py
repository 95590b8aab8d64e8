"""vklComputeSampleNNoAlloc (the allocation-free variant of vklComputeSampleN)
returns vklComputeSampleN's samples bit for bit, called through the C ABI:
two calls in a row on the binned path (the second with fewer points, so stale
scratch from the first would show), a direct-kernel volume, a call whose
scratch exceeds the fixed array (rejected with VKL_INVALID_ARGUMENT), and
host pointers (rejected)."""
import numpy as np
import pytest

import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _volume(device, dims, seed):
    rng = np.random.default_rng(seed)
    data = (rng.standard_normal((dims[2], dims[1], dims[0])) * 10).astype(np.float32)
    return device.structured_regular(data)


def test_no_alloc_matches_sample_n_twice(device):
    vol = _volume(device, (96, 80, 70), 71)
    binned = vol.sampler(binned=True)
    pts_a = V.uniform_points(300000, (-2, -2, -2), (97, 81, 71), seed=72)
    pts_b = V.uniform_points(70001, (-1, -1, -1), (96, 80, 70), seed=73)
    for pts in (pts_a, pts_b):
        print(f"no-alloc binned call: {len(pts)} points", flush=True)
        p = torch.from_numpy(pts).cuda()
        got = binned.sample_noalloc(p).cpu().numpy()
        ref = binned.sample(p).cpu().numpy()
        assert_same(got, ref)
    binned.release()
    vol.release()


def test_no_alloc_direct_kernel_volume(device):
    """A sampler whose calls take the direct kernel (binned off) gives the
    same samples through both entry points."""
    vol = _volume(device, (40, 33, 29), 74)
    s = vol.sampler(binned=False)
    pts = torch.from_numpy(V.uniform_points(50000, (-1, -1, -1), (41, 34, 30), seed=75)).cuda()
    assert_same(s.sample_noalloc(pts).cpu().numpy(), s.sample(pts).cpu().numpy())
    s.release()
    vol.release()


def test_no_alloc_rejects_oversized_scratch_and_host_pointers(device):
    vol = _volume(device, (96, 80, 70), 76)
    s = vol.sampler(binned=True)
    # 8 M points need ~200 MiB of binned scratch, more than the fixed 64 MiB
    n = 8 << 20
    pts = torch.zeros((n, 3), device="cuda")
    with pytest.raises(vkl.VklError) as e:
        s.sample_noalloc(pts)
    assert e.value.code == vkl.VKL_INVALID_ARGUMENT and "scratch" in e.value.msg
    with pytest.raises(vkl.VklError) as e:
        s.sample_noalloc(np.zeros((16, 3), np.float32))
    assert e.value.code == vkl.VKL_INVALID_ARGUMENT and "device pointers" in e.value.msg
    # the entry point still works after the rejected calls
    p = torch.from_numpy(V.uniform_points(1000, (0, 0, 0), (95, 79, 69), seed=77)).cuda()
    assert_same(s.sample_noalloc(p).cpu().numpy(), s.sample(p).cpu().numpy())
    s.release()
    vol.release()
