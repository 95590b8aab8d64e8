"""structuredRegular volumes wider than one level-1 node (4096 voxels) on an
axis: their iterator queries use the general node-table query, not the dense
node numbering, so the interval batches and the fused interval and hit
renderers run their node-table instantiations on a dense source. Bar: bit
for bit against the oracle, as for every other dense volume."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_iterators import compare_intervals, rays_through
from tests import test_gpu_hit_render as HR
from tests import test_gpu_raymarch_intervals as RMI

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def long_pair(device):
    dims, origin, spacing = (4200, 6, 5), (-2.0, 0.5, 1.0), (0.01, 0.3, 0.35)
    data = V.procedural(dims, V.wavelet, origin, spacing)
    return device.structured_regular(data, origin=origin, spacing=spacing), \
        O.Grid.dense(data, origin=origin, spacing=spacing)


@pytest.mark.parametrize("hint", [0.5, 1.0])
def test_interval_batch_general_query(long_pair, hint):
    vol, grid = long_pair
    ctx = vol.sampler().interval_context(hint=hint)
    org, d = rays_through(grid.bounding_box(), 400, seed=71)
    res = ctx.batch(torch.from_numpy(org).cuda(), torch.from_numpy(d).cuda(), max_per_ray=1024)
    out, cnt = res[0].cpu().numpy(), res[1].cpu().numpy()
    assert cnt.max() > 2
    compare_intervals((out, cnt), grid, org, d, O.map_depth(hint), max_per_ray=1024)


def test_interval_renderer_general_query(long_pair):
    vol, grid = long_pair
    grid.set_sampler(O.LINEAR)
    ctx = vol.sampler(filter=vkl.VKL_FILTER_LINEAR).interval_context()
    org, d = rays_through(grid.bounding_box(), 2000, 72)
    out, cnt = RMI.compare(ctx, grid, org, d, RMI.tf_table(scale=3.0), (-3.0, 3.0))
    assert cnt.max() > 10


def test_hit_renderer_general_query(long_pair):
    vol, grid = long_pair
    grid.set_sampler(O.LINEAR, gradient_filter=O.LINEAR)
    iso = [-0.5, 0.5]
    ctx = vol.sampler(filter=vkl.VKL_FILTER_LINEAR, gradient_filter=vkl.VKL_FILTER_LINEAR).hit_context(iso)
    org, d = rays_through(grid.bounding_box(), 2000, 73)
    out, cnt = HR.compare(ctx, grid, org, d, iso, HR.tf_table(), (-3.0, 3.0))
    assert cnt.max() >= 1
