"""Spatially binned stream sampling (dense volumes, linear filter) returns the
direct kernel's samples bit for bit, and the oracle's: every element type,
brick layouts, transforms, points outside the domain, NaN points, and inputs
so skewed that bin segments overflow into the direct fallback."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def both(vol, pts):
    got = vol.sampler(binned=True).sample(pts)
    ref = vol.sampler(binned=False).sample(pts)
    return got, ref


@pytest.mark.parametrize("dtype", [np.float32, np.float16, np.uint8, np.int16, np.float64])
def test_binned_equals_direct_and_oracle(device, dtype):
    rng = np.random.default_rng(61)
    dims = (150, 97, 71)  # x, y, z: ragged against every brick shape
    data = rng.standard_normal((dims[2], dims[1], dims[0])) * 40
    data = data.astype(dtype) if np.dtype(dtype).kind == "f" else np.abs(data).astype(dtype)
    vol = device.structured_regular(data)
    grid = O.Grid.dense(data)
    pts = V.uniform_points(300000, (-2, -2, -2), (dims[0] + 1, dims[1] + 1, dims[2] + 1))
    pts[:7] = [[np.nan, 1, 1], [1, np.nan, 1], [1, 1, np.inf], [-np.inf, 0, 0], [1e30, 5, 5], [0, 0, 0],
               [dims[0] - 1, dims[1] - 1, dims[2] - 1]]
    got, ref = both(vol, torch.from_numpy(pts).cuda())
    assert_same(got.cpu().numpy(), ref.cpu().numpy())
    assert_same(got.cpu().numpy(), grid.sample(pts))


@pytest.mark.parametrize("shift", [(2, 2, 1), (1, 1, 1)])
def test_binned_brick_layout_and_transform(device, shift):
    rng = np.random.default_rng(62)
    data = rng.standard_normal((64, 80, 130)).astype(np.float32)
    vol = device.structured_regular(data, origin=(-3.5, 2.0, 0.25), spacing=(0.5, 1.25, 2.0), brick_shift=shift)
    grid = O.Grid.dense(data, origin=(-3.5, 2.0, 0.25), spacing=(0.5, 1.25, 2.0))
    lo, hi = np.array([-4.0, 1.0, -1.0]), np.array([-3.5 + 0.5 * 130, 2.0 + 1.25 * 80, 0.25 + 2.0 * 64])
    pts = V.uniform_points(400000, lo, hi)
    got, ref = both(vol, torch.from_numpy(pts).cuda())
    assert_same(got.cpu().numpy(), ref.cpu().numpy())
    assert_same(got.cpu().numpy(), grid.sample(pts))


def test_binned_overflow_path(device):
    """All points in a few bins: segments overflow and the rest is sampled by
    the direct fallback inside the binned launch."""
    rng = np.random.default_rng(63)
    data = rng.standard_normal((96, 96, 96)).astype(np.float32)
    vol = device.structured_regular(data)
    grid = O.Grid.dense(data)
    n = 1 << 20
    pts = np.empty((n, 3), np.float32)
    pts[: n // 2] = 10.0 + rng.random((n // 2, 3), dtype=np.float32) * 3.0  # one bin
    pts[n // 2:] = rng.random((n - n // 2, 3), dtype=np.float32) * 95.0
    t = torch.from_numpy(pts).cuda()
    got, ref = both(vol, t)
    assert_same(got.cpu().numpy(), ref.cpu().numpy())
    sub = slice(0, n, 97)
    assert_same(got.cpu().numpy()[sub], grid.sample(pts[sub]))
    # a second call on the same sampler reuses (and re-clears) the scratch
    s = vol.sampler(binned=True)
    a = s.sample(t[: n // 3])
    b = s.sample(t[: n // 3])
    assert torch.equal(a, b)
    assert_same(a.cpu().numpy(), ref[: n // 3].cpu().numpy())


@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_CUBIC])
@pytest.mark.parametrize("cell_centered", [False, True])
def test_binned_filters_and_gradients(device, filt, cell_centered):
    """Every filter's samples and gradients through the binned path equal the
    direct kernels' bit for bit (cubic bricks carry a 1/2-voxel halo)."""
    rng = np.random.default_rng(65)
    data = rng.standard_normal((45, 70, 101)).astype(np.float32)
    vol = device.structured_regular(data, cell_centered=cell_centered)
    pts = torch.from_numpy(V.uniform_points(300000, (-1.5, -1.5, -1.5), (102, 71, 46))).cuda()
    sb = vol.sampler(filter=filt, binned=True)
    sd = vol.sampler(filter=filt, binned=False)
    assert_same(sb.sample(pts).cpu().numpy(), sd.sample(pts).cpu().numpy())
    assert_same(sb.gradient(pts).cpu().numpy(), sd.gradient(pts).cpu().numpy())
    if filt == vkl.VKL_FILTER_LINEAR:
        grid = O.Grid.dense(data, cell_centered=cell_centered)
        sub = pts[::7].contiguous()
        assert_same(sb.gradient(sub).cpu().numpy(), grid.gradient(sub.cpu().numpy()))


def test_binned_host_pointers_and_sample_m(device):
    rng = np.random.default_rng(64)
    data = rng.standard_normal((40, 50, 60)).astype(np.float32)
    vol = device.structured_regular(data)
    grid = O.Grid.dense(data)
    pts = V.uniform_points(50000, (0, 0, 0), (59, 49, 39))
    s = vol.sampler(binned=True)
    assert_same(s.sample(pts), grid.sample(pts))  # numpy in / out: staged copies
    m = s.sample_m(pts, [0, 0])  # strided output, one attribute twice
    assert_same(m[:, 0], grid.sample(pts))
    assert_same(m[:, 1], grid.sample(pts))
