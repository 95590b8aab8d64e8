"""The hit-iterator renderer (vklComputeHitRenderN) on the GPU vs the oracle,
bit for bit: composited colour and opacity and hits shaded per ray, with
gradient normals and shadow hit queries, on structuredRegular volumes of
several element types, a sparse vdb, a temporal volume at a ray time, a
nearest gradient filter, device pointers, t ranges and the hit cap."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same
from tests.test_gpu_iterators import rays_through

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def tf_table(n=6, seed=0):
    return np.random.default_rng(seed).random((n, 4)).astype(np.float32)


def compare(ctx, grid, org, d, iso, tf, tf_range, **kw):
    out, cnt = ctx.render_hits(org, d, tf, tf_range, **kw)
    want, wcnt = grid.hit_render(org, d, iso, tf, tf_range, **kw)
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
def test_dense_float_bitexact(wavelet_pair, filt):
    vol, grid = wavelet_pair
    grid.set_sampler(filt[1], gradient_filter=O.LINEAR)
    iso = [-1.0, 0.5, 1.5]
    ctx = vol.sampler(filter=filt[0], gradient_filter=vkl.VKL_FILTER_LINEAR).hit_context(iso)
    org, d = rays_through(grid.bounding_box(), 2000, 31)
    out, cnt = compare(ctx, grid, org, d, iso, tf_table(), (-3.0, 3.0))
    assert cnt.max() >= 3 and (cnt == 0).any()


def test_nearest_gradient_and_cap(wavelet_pair):
    vol, grid = wavelet_pair
    grid.set_sampler(O.LINEAR, gradient_filter=O.NEAREST)
    iso = [0.0, 1.0]
    ctx = vol.sampler(filter=vkl.VKL_FILTER_LINEAR, gradient_filter=vkl.VKL_FILTER_NEAREST).hit_context(iso)
    org, d = rays_through(grid.bounding_box(), 1500, 32)
    compare(ctx, grid, org, d, iso, tf_table(4, 1), (-3.0, 3.0))
    grid.set_sampler(O.LINEAR)
    ctx = vol.sampler().hit_context(iso)
    out, cnt = compare(ctx, grid, org, d, iso, tf_table(4, 1), (-3.0, 3.0), max_hits=2)
    assert cnt.max() == 2


@pytest.mark.parametrize("dtype", [np.uint8, np.float16], ids=["uint8", "half"])
def test_dense_types_bitexact(device, dtype):
    dims = (30, 25, 20)
    f = V.procedural(dims, V.wavelet, spacing=(0.3, 0.3, 0.3), dtype=np.float64)
    data = ((f + 3) * 40).astype(dtype) if dtype == np.uint8 else f.astype(dtype)
    iso = [120.0, 140.0] if dtype == np.uint8 else [0.0, 1.2]
    vol = device.structured_regular(data, spacing=(0.3, 0.3, 0.3))
    grid = O.Grid.dense(data, spacing=(0.3, 0.3, 0.3))
    lo, hi = grid.value_range()
    org, d = rays_through(grid.bounding_box(), 1500, 33)
    compare(vol.sampler().hit_context(iso), grid, org, d, iso, tf_table(5, 2), (lo, hi))


def test_vdb_bitexact(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid.set_sampler(O.LINEAR)
    org, d = rays_through(grid.bounding_box(), 1500, 34)
    out, cnt = compare(vol.sampler().hit_context([0.5]), grid, org, d, [0.5], tf_table(5, 3), (0.0, 1.2))
    assert (cnt > 0).any()


def test_temporal_and_device_pointers(device):
    dims = (14, 12, 10)
    data, kw = V.temporal_volume(dims, V.wavelet_time, ("structured", 5))
    vol = device.structured_regular(data, dims=dims, **kw)
    grid = O.Grid.dense_temporal(dims, data, **kw)
    org, d = rays_through(grid.bounding_box(), 1000, 35)
    ctx = vol.sampler().hit_context([0.3])
    for t in (0.0, 0.6):
        compare(ctx, grid, org, d, [0.3], tf_table(), (-3.0, 3.0), time=t)
    tr = np.tile(np.array([[0.5, 40.0]], np.float32), (1000, 1))
    tf = tf_table()
    out, cnt = ctx.render_hits(*[torch.from_numpy(a).cuda() for a in (org, d)], torch.from_numpy(tf).cuda(),
                               (-3.0, 3.0), t_ranges=torch.from_numpy(tr).cuda(), time=0.6)
    want, wcnt = grid.hit_render(org, d, [0.3], tf, (-3.0, 3.0), t_ranges=tr, time=0.6)
    assert_same(out.cpu().numpy(), want)
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), wcnt)
    compare(ctx, grid, org[:0], d[:0], [0.3], tf, (-3.0, 3.0))
