"""Host-pointer calls of many points (vklComputeSampleN / GradientN on host
arrays): the library stages them through the GPU in overlapped chunks.
Results must equal the same call on device arrays bit for bit, for pageable
and pinned host memory, host input with device output and the reverse, a
ragged last chunk, and temporal volumes with host times."""
import numpy as np
import pytest

import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N = (1 << 23) + 12345  # two full chunks of 2^22 points and a ragged third


@pytest.fixture(scope="module")
def dense_case(device):
    rng = np.random.default_rng(101)
    data = rng.standard_normal((70, 60, 50)).astype(np.float32)
    vol = device.structured_regular(data, spacing=(0.5, 1.0, 2.0))
    pts = V.uniform_points(N, (-1, -1, -1), (26, 61, 141), seed=102)
    return vol, pts


@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_CUBIC])
def test_host_arrays_equal_device_arrays(dense_case, filt):
    vol, pts = dense_case
    s = vol.sampler(filter=filt, binned=False)
    dp = torch.from_numpy(pts).cuda()
    want = s.sample(dp).cpu().numpy()
    assert_same(s.sample(pts), want)                                  # pageable in, pageable out
    pin_in = torch.from_numpy(pts).pin_memory()
    pin_out = torch.empty(N, dtype=torch.float32).pin_memory()
    s.sample(pin_in, out=pin_out)                                     # pinned in and out
    assert_same(pin_out.numpy(), want)
    dev_out = torch.empty(N, device="cuda")
    s.sample(pts, out=dev_out)                                        # host in, device out
    assert_same(dev_out.cpu().numpy(), want)
    host_out = np.empty(N, np.float32)
    s.sample(dp, out=host_out)                                        # device in, host out
    assert_same(host_out, want)
    if filt != vkl.VKL_FILTER_NEAREST:
        gw = s.gradient(dp).cpu().numpy()
        assert_same(s.gradient(pts), gw)
    s.release()


def test_temporal_host_times(device):
    rng = np.random.default_rng(103)
    steps = 3
    data = rng.standard_normal(20 * 21 * 22 * steps).astype(np.float32)
    vol = device.structured_regular(data, dims=(22, 21, 20), structured_steps=steps)
    pts = V.uniform_points(N, (0, 0, 0), (21, 20, 19), seed=104)
    times = rng.random(N, dtype=np.float32)
    s = vol.sampler()
    want = s.sample(torch.from_numpy(pts).cuda(), times=torch.from_numpy(times).cuda()).cpu().numpy()
    assert_same(s.sample(pts, times=times), want)
    s.release()


def test_sample_m_host_arrays(device):
    rng = np.random.default_rng(105)
    a0 = rng.standard_normal((30, 31, 32)).astype(np.float32)
    a1 = rng.integers(0, 255, (30, 31, 32)).astype(np.uint8)
    vol = device.structured_regular([a0, a1])
    pts = V.uniform_points(N, (-1, -1, -1), (32, 31, 30), seed=106)
    s = vol.sampler(filter=vkl.VKL_FILTER_LINEAR, binned=False)
    want = s.sample_m(torch.from_numpy(pts).cuda(), [1, 0]).cpu().numpy()
    assert_same(s.sample_m(pts, [1, 0]), want)
    s.release()
