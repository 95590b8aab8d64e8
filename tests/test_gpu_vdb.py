"""Sparse vdb sampling on the GPU vs the CPU oracle: node table, leaf
addressing (packed and per-leaf node.data, float and half), tiles, error
voxels below maxSamplingDepth."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same, check_cubic

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def shell(res=256):
    origins, leaves = V.sparse_shell_leaves(res=res)
    return origins, leaves


@pytest.fixture(scope="module")
def shell_vdb(device):
    origins, leaves = shell()
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    pts = V.uniform_points(400000, (-4, -4, -4), (260, 260, 260))
    return vol, grid, pts, (origins, leaves)


def test_node_table_bitexact(shell_vdb):
    vol, grid, _, _ = shell_vdb
    for l in range(3):
        n, vox, rng_ = vol.level(l)
        org, ovox, orng = grid.level(l)
        assert n == org.shape[0]
        np.testing.assert_array_equal(vox, ovox)
        np.testing.assert_array_equal(rng_.view(np.uint32), orng.view(np.uint32))
    np.testing.assert_array_equal(vol.value_range(), grid.value_range())
    np.testing.assert_array_equal(vol.bounding_box(), grid.bounding_box())


@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR])
def test_sample_gradient_bitexact(shell_vdb, filt):
    vol, grid, pts, _ = shell_vdb
    grid.set_sampler(filt)
    s = vol.sampler(filter=filt)
    assert_same(s.sample(torch.from_numpy(pts).cuda()).cpu().numpy(), grid.sample(pts))
    assert_same(s.gradient(pts[:100000]), grid.gradient(pts[:100000]))


def test_cubic(shell_vdb):
    vol, grid, pts, _ = shell_vdb
    grid.set_sampler(O.CUBIC)
    s = vol.sampler(filter=vkl.VKL_FILTER_CUBIC)
    check_cubic(grid, s.sample(pts[:100000]), pts[:100000])


@pytest.mark.parametrize("depth", [0, 1, 2])
def test_max_sampling_depth_error_voxels(shell_vdb, depth):
    """VdbSampleInner.ih.in:46-54: below maxSamplingDepth the sample is the
    midpoint of the stopping voxel's value range."""
    vol, grid, pts, _ = shell_vdb
    grid.set_sampler(O.LINEAR, max_sampling_depth=depth)
    s = vol.sampler(max_sampling_depth=depth)
    assert_same(s.sample(pts[:100000]), grid.sample(pts[:100000]))


@pytest.mark.parametrize("packed", [False, True])
def test_node_data_half_tiles_multi_attribute(device, packed):
    """Per-leaf node.data (no repack), half + float attributes, tiles on
    levels 1..3, background; also through corner-packed cell records."""
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    rng = np.random.default_rng(43)
    # add tiles: a level-2 tile (128^3) region far away and a few level-3 tiles
    tile_orig = np.array([[256, 0, 0], [384, 128, 0], [0, 0, 4096]], np.int32)
    tile_lvl = np.array([2, 2, 1], np.uint32)
    tile_fmt = np.zeros(3, np.uint32)
    lv = np.concatenate([np.full(n, 3, np.uint32), tile_lvl])
    org = np.concatenate([origins, tile_orig])
    fmt = np.concatenate([np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32), tile_fmt])
    a1 = rng.standard_normal((n, 512)).astype(np.float16)
    leaf_data = [[leaves[i], a1[i]] for i in range(n)]
    leaf_data += [[np.array([0.75 + k], np.float32), np.array([-1.5 - k], np.float16)] for k in range(3)]
    vol = device.vdb(lv, org, fmt, leaf_data=leaf_data, background=0.5, corner_packed=packed)
    grid = O.Grid.vdb(lv, org, fmt, leaf_data=leaf_data, background=[0.5, 0.5])
    box = grid.bounding_box()
    np.testing.assert_array_equal(vol.bounding_box(), box)
    pts = np.concatenate([V.uniform_points(150000, (-2, -2, -2), (130, 130, 130)),
                          V.uniform_points(50000, box[:3] - 2, box[3:] + 2, seed=5)])
    for a in range(2):
        for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
            grid.set_sampler(f)
            s = vol.sampler(filter=f)
            assert_same(s.sample(pts, attr=a), grid.sample(pts, attr=a))
        np.testing.assert_array_equal(vol.value_range(a), grid.value_range(a))
    for l in range(3):
        _, vox, r = vol.level(l)
        _, ovox, orng = grid.level(l)
        np.testing.assert_array_equal(vox, ovox)
        np.testing.assert_array_equal(r.view(np.uint32), orng.view(np.uint32))


def test_corner_packed_vdb_bitexact(device):
    """Corner-packed vdb (leaf index grid + cell records baked at commit):
    samples and gradients equal the oracle's for every filter, points inside
    leaves, at leaf faces, in empty space and outside the domain."""
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)], corner_packed=True)
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    rng = np.random.default_rng(81)
    k = rng.integers(0, n, 150000)
    inside = (origins[k] + rng.random((150000, 3)) * 8.0).astype(np.float32)
    pts = np.concatenate([inside, V.uniform_points(100000, (-3, -3, -3), (131, 131, 131))])
    for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
        grid.set_sampler(f)
        s = vol.sampler(filter=f)
        assert_same(s.sample(pts), grid.sample(pts))
        assert_same(s.gradient(pts), grid.gradient(pts))
    grid.set_sampler(O.CUBIC)
    check_cubic(grid, vol.sampler(filter=vkl.VKL_FILTER_CUBIC).sample(pts[:60000]), pts[:60000])


@pytest.mark.parametrize("volume_depth,sampler_depth", [(2, None), (2, 3), (3, 1)])
def test_corner_packed_vdb_sampling_depth(device, volume_depth, sampler_depth):
    """Cell records hold the volume's maxSamplingDepth values; a sampler with
    another depth reads through the node table. Both equal the oracle."""
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)], corner_packed=True,
                     max_sampling_depth=volume_depth)
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    depth = volume_depth if sampler_depth is None else sampler_depth
    grid.set_sampler(O.LINEAR, max_sampling_depth=depth)
    pts = V.uniform_points(200000, (-2, -2, -2), (130, 130, 130))
    s = vol.sampler(max_sampling_depth=sampler_depth)
    assert_same(s.sample(pts), grid.sample(pts))


def test_index_to_object_transform(device):
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    i2o = np.array([0.5, 0.0, 0.1, 0.0, 0.25, 0.0, 0.0, 0.05, 0.75, -10.0, 5.0, 2.0], np.float32)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, 1, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)], index_to_object=i2o)
    rm = np.array([i2o[0], i2o[3], i2o[6], i2o[1], i2o[4], i2o[7], i2o[2], i2o[5], i2o[8], i2o[9], i2o[10], i2o[11]],
                  np.float32)
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)], i2o=rm)
    box = grid.bounding_box()
    pts = V.uniform_points(100000, box[:3], box[3:])
    assert_same(vol.sampler().sample(pts), grid.sample(pts))
    assert_same(vol.sampler().gradient(pts), grid.gradient(pts))


@pytest.mark.parametrize("index_grid", [0, 1])
def test_leaf_index_grid_on_and_off(device, index_grid):
    """The leaf index grid as a flattened node table (dense-leaf slots, empty
    cells, cells under tiles of levels 1..3 left to the node table) samples
    exactly as the traversal: every filter, gradients, maxSamplingDepth 1..3,
    points in empty space, in leaves, under tiles and outside the domain."""
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    tile_orig = np.array([[256, 0, 0], [144, 8, 8], [0, 0, 4096], [136, 0, 0]], np.int32)
    tile_lvl = np.array([2, 3, 1, 3], np.uint32)
    lv = np.concatenate([np.full(n, 3, np.uint32), tile_lvl])
    org = np.concatenate([origins, tile_orig])
    fmt = np.concatenate([np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32), np.zeros(4, np.uint32)])
    leaf_data = [[leaves[i]] for i in range(n)] + [[np.array([0.75 + k], np.float32)] for k in range(4)]
    vol = device.vdb(lv, org, fmt, leaf_data=leaf_data, leaf_index_grid=index_grid)
    grid = O.Grid.vdb(lv, org, fmt, leaf_data=leaf_data)
    box = grid.bounding_box()
    pts = np.concatenate([V.uniform_points(100000, (-2, -2, -2), (130, 130, 130), seed=81),
                          V.uniform_points(50000, (250, -2, -2), (390, 130, 130), seed=82),
                          V.uniform_points(30000, box[:3] - 2, box[3:] + 2, seed=83)])
    for depth in (3, 2, 1):
        for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
            grid.set_sampler(f, f, max_sampling_depth=depth)
            s = vol.sampler(filter=f, gradient_filter=f, max_sampling_depth=depth)
            assert_same(s.sample(pts), grid.sample(pts))
            assert_same(s.gradient(pts[:20000]), grid.gradient(pts[:20000]))
            s.release()
    grid.set_sampler(vkl.VKL_FILTER_CUBIC)
    s = vol.sampler(filter=vkl.VKL_FILTER_CUBIC)
    got = s.sample(pts[:20000])
    grid.set_cubic_double(True)
    want = grid.sample(pts[:20000])
    grid.set_cubic_double(False)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.abs(got[ok] - want[ok]).max() <= 1e-5 * max(np.abs(want[ok]).max(), 1e-30)
    grid.set_sampler()
    vol.release()


def test_cubic_stencil_gather_equals_per_tap_reads(device):
    """Tricubic stencils through the leaf index grid (the <= 8 leaf cells of
    a stencil looked up once) equal 64 per-tap traversals of the node table
    bit for bit: samples and gradients, stencils across leaf boundaries, in
    empty cells, under tiles of levels 1..3 and outside the domain."""
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    tile_orig = np.array([[256, 0, 0], [144, 8, 8], [0, 0, 4096], [136, 0, 0]], np.int32)
    tile_lvl = np.array([2, 3, 1, 3], np.uint32)
    lv = np.concatenate([np.full(n, 3, np.uint32), tile_lvl])
    org = np.concatenate([origins, tile_orig])
    fmt = np.concatenate([np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32), np.zeros(4, np.uint32)])
    leaf_data = [[leaves[i]] for i in range(n)] + [[np.array([0.75 + k], np.float32)] for k in range(4)]
    grid = O.Grid.vdb(lv, org, fmt, leaf_data=leaf_data)
    box = grid.bounding_box()
    pts = np.concatenate([V.uniform_points(150000, (-2, -2, -2), (130, 130, 130), seed=84),
                          V.uniform_points(50000, (250, -2, -2), (390, 130, 130), seed=85),
                          V.uniform_points(30000, box[:3] - 2, box[3:] + 2, seed=86)])
    p = torch.from_numpy(pts).cuda()
    got = {}
    for index_grid in (1, 0):
        vol = device.vdb(lv, org, fmt, leaf_data=leaf_data, leaf_index_grid=index_grid)
        s = vol.sampler(filter=vkl.VKL_FILTER_CUBIC)
        got[index_grid] = (s.sample(p).cpu().numpy(), s.gradient(p).cpu().numpy())
        s.release()
        vol.release()
    assert_same(got[1][0], got[0][0])
    assert_same(got[1][1], got[0][1])
    grid.set_sampler(O.CUBIC)
    check_cubic(grid, got[1][0][:30000], pts[:30000])
