"""Interval ray marching with transfer-function compositing
(vklComputeRayMarchIntervalsN) on the GPU vs the oracle, bit for bit: the
composited colour and opacity and the number of samples per ray, over
structuredRegular volumes of several element types and filters, a sparse vdb
with value-range culling, a temporal volume at a ray time, device and host
pointers, default t ranges and the per-ray sample cap."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same
from tests.test_gpu_iterators import rays_through

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def tf_table(n=7, seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    tf = rng.random((n, 4)).astype(np.float32)
    tf[:, 3] *= scale
    return tf


def compare(ctx, grid, org, d, tf, tf_range, value_ranges=None, **kw):
    """value_ranges: the ones the context was created with."""
    out, cnt = ctx.ray_march(org, d, tf, tf_range, **kw)
    want, wcnt = grid.ray_march_intervals(org, d, tf, tf_range, value_ranges=value_ranges, **kw)
    assert_same(out, want)
    np.testing.assert_array_equal(cnt, wcnt)
    return out, cnt


@pytest.fixture(scope="module")
def wavelet_pair(device):
    dims, origin, spacing = (40, 36, 44), (-1.0, 0.5, 2.0), (0.25, 0.3, 0.2)
    data = V.procedural(dims, V.wavelet, origin, spacing)
    return device.structured_regular(data, origin=origin, spacing=spacing), \
        O.Grid.dense(data, origin=origin, spacing=spacing)


@pytest.mark.parametrize("filt", [(vkl.VKL_FILTER_NEAREST, O.NEAREST), (vkl.VKL_FILTER_LINEAR, O.LINEAR)],
                         ids=["nearest", "linear"])
@pytest.mark.parametrize("rate", [1.0, 3.5])
def test_dense_float_bitexact(wavelet_pair, filt, rate):
    vol, grid = wavelet_pair
    grid.set_sampler(filt[1])
    ctx = vol.sampler(filter=filt[0]).interval_context()
    org, d = rays_through(grid.bounding_box(), 3000, 21)
    out, cnt = compare(ctx, grid, org, d, tf_table(scale=3.0), (-3.0, 3.0), sampling_rate=rate)
    assert (out[:, 3] >= 0.99).any() and (out[:, 3] < 0.99).any() and cnt.max() > 10


def test_dense_cubic_close(wavelet_pair):
    """Tricubic samples match the oracle to ~1e-6 of scale, not bit for bit,
    so compositing agrees to 1e-4 and an early stop may move by a sample."""
    vol, grid = wavelet_pair
    grid.set_sampler(O.CUBIC)
    ctx = vol.sampler(filter=vkl.VKL_FILTER_CUBIC).interval_context()
    org, d = rays_through(grid.bounding_box(), 2000, 22)
    out, cnt = ctx.ray_march(org, d, tf_table(scale=0.6), (-3.0, 3.0), sampling_rate=2.0)
    want, wcnt = grid.ray_march_intervals(org, d, tf_table(scale=0.6), (-3.0, 3.0), sampling_rate=2.0)
    same = cnt == wcnt
    assert same.mean() >= 0.99
    np.testing.assert_allclose(out[same], want[same], atol=1e-4)


@pytest.mark.parametrize("dtype", [np.uint8, np.float16, np.float64], ids=["uint8", "half", "double"])
def test_dense_types_bitexact(device, dtype):
    dims = (30, 25, 20)
    f = V.procedural(dims, V.wavelet, spacing=(0.3, 0.3, 0.3), dtype=np.float64)
    data = ((f + 3) * 40).astype(dtype) if dtype == np.uint8 else f.astype(dtype)
    vol = device.structured_regular(data, spacing=(0.3, 0.3, 0.3))
    grid = O.Grid.dense(data, spacing=(0.3, 0.3, 0.3))
    lo, hi = grid.value_range()
    org, d = rays_through(grid.bounding_box(), 2000, 23)
    compare(vol.sampler().interval_context(), grid, org, d, tf_table(5, 1, 0.8), (lo, hi), sampling_rate=2.0)


def test_vdb_with_value_ranges_bitexact(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid.set_sampler(O.LINEAR)
    org, d = rays_through(grid.bounding_box(), 3000, 24)
    s = vol.sampler()
    out, cnt = compare(s.interval_context(), grid, org, d, tf_table(6, 2, 0.3), (0.0, 1.2), sampling_rate=1.5)
    assert (cnt > 0).any() and (cnt == 0).any()
    vr = [[0.4, 0.9]]
    compare(s.interval_context(value_ranges=vr), grid, org, d, tf_table(6, 2, 0.3), (0.0, 1.2),
            sampling_rate=1.5, value_ranges=vr)


def test_temporal_at_ray_time(device):
    dims = (14, 12, 10)
    data, kw = V.temporal_volume(dims, V.wavelet_time, ("structured", 5))
    vol = device.structured_regular(data, dims=dims, **kw)
    grid = O.Grid.dense_temporal(dims, data, **kw)
    org, d = rays_through(grid.bounding_box(), 1500, 25)
    for t in (0.0, 0.35, 1.0):
        compare(vol.sampler().interval_context(), grid, org, d, tf_table(scale=0.5), (-3.0, 3.0), time=t,
                sampling_rate=2.0)


def test_device_pointers_ranges_and_cap(wavelet_pair):
    vol, grid = wavelet_pair
    grid.set_sampler(O.LINEAR)
    ctx = vol.sampler().interval_context()
    org, d = rays_through(grid.bounding_box(), 2000, 26)
    tf = tf_table(scale=0.2)
    tr = np.tile(np.array([[1.0, 6.0]], np.float32), (2000, 1))
    out, cnt = ctx.ray_march(*[torch.from_numpy(a).cuda() for a in (org, d)], torch.from_numpy(tf).cuda(),
                             (-3.0, 3.0), t_ranges=torch.from_numpy(tr).cuda(), sampling_rate=2.0)
    want, wcnt = grid.ray_march_intervals(org, d, tf, (-3.0, 3.0), t_ranges=tr, sampling_rate=2.0)
    assert_same(out.cpu().numpy(), want)
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), wcnt)
    out, cnt = compare(ctx, grid, org, d, tf, (-3.0, 3.0), sampling_rate=4.0, max_samples=5)
    assert cnt.max() == 5
    compare(ctx, grid, org[:0], d[:0], tf, (-3.0, 3.0))
