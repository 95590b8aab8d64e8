"""Motion-blur (temporal) structuredRegular volumes on the GPU vs the oracle:
structured and unstructured time samples (uint32 and uint64 indices), every
element type, per-point times through the stream, multi-attribute and vector
calls; gradients; majorants and interval iteration; hit iteration at the ray
time; the reference's known answer; commit-time validation."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_close_scale, assert_same
from tests.test_gpu_iterators import rays_through
from tests.test_oracle_temporal import tuv_known_answer

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

KINDS = [("structured", 2), ("structured", 5), ("unstructured", [0.0, 0.15, 0.3, 0.65, 0.9, 1.0]),
         ("varying", None)]
KIND_IDS = ["s2", "s5", "u6", "uv"]


def pair(device, dims, fn, kind, dtype=np.float32, index64=False, **geo):
    data, kw = V.temporal_volume(dims, fn, kind, dtype=dtype, index64=index64, **{
        k: v for k, v in geo.items() if k in ("origin", "spacing")})
    vol = device.structured_regular(data, dims=dims, **geo, **kw)
    grid = O.Grid.dense_temporal(dims, data, **geo, **kw)
    return vol, grid, data, kw


def test_reference_known_answer(device):
    """structured_regular_volume_sampling_motion_blur.cpp:84-173: exact."""
    vox, idx, times = tuv_known_answer()
    vol = device.structured_regular(vox, dims=(2, 2, 2), tu_indices=idx, tu_times=times)
    s = vol.sampler(filter=vkl.VKL_FILTER_NEAREST)
    pts = np.array([[x, y, z] for z in range(2) for y in range(2) for x in range(2)], np.float32)
    for t in (0.0, 0.25, 0.5, 0.75, 1.0):
        want = (t * pts.sum(1)).astype(np.float32)
        assert_same(s.sample(pts, times=np.full(8, t, np.float32)), want)
        for i in range(8):  # scalar API with its time argument
            assert s.sample1(pts[i], time=t) == want[i]


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR])
def test_sample_and_gradient_bitexact(device, kind, filt):
    dims = (23, 17, 12)
    vol, grid, _, _ = pair(device, dims, V.wavelet_time, kind, origin=(-1.0, 0.5, 2.0), spacing=(0.5, 0.75, 1.25))
    grid.set_sampler(filt)
    lo, hi = np.array([-2.0, 0.0, 1.0]), np.array([-1 + 0.5 * 23, 0.5 + 0.75 * 17, 2 + 1.25 * 12])
    pts = V.uniform_points(40000, lo, hi, seed=3)
    t = np.random.default_rng(4).random(40000, dtype=np.float32)
    t[:5] = [0.0, 1.0, 0.5, 0.15, 0.9]
    s = vol.sampler(filter=filt)
    got = s.sample(torch.from_numpy(pts).cuda(), times=torch.from_numpy(t).cuda()).cpu().numpy()
    assert_same(got, grid.sample(pts, times=t))
    assert_same(s.sample(pts, times=t), got)  # host pointers
    if filt == vkl.VKL_FILTER_LINEAR:
        assert_same(s.gradient(pts[:8000], times=t[:8000]), grid.gradient(pts[:8000], times=t[:8000]))
    # no times: time 0
    assert_same(s.sample(pts[:3000]), grid.sample(pts[:3000], times=np.zeros(3000, np.float32)))


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.float16, np.float64])
@pytest.mark.parametrize("kind", [("structured", 3), ("varying", None)], ids=["s3", "uv"])
def test_element_types(device, dtype, kind):
    def field(x, y, z, t):
        return np.abs(V.wavelet_time(x, y, z, t)) * 40.0

    vol, grid, _, _ = pair(device, (14, 9, 11), field, kind, dtype=dtype)
    pts = V.uniform_points(20000, (-1, -1, -1), (14, 9, 11), seed=5)
    t = np.random.default_rng(6).random(20000, dtype=np.float32)
    s = vol.sampler()
    assert_same(s.sample(pts, times=t), grid.sample(pts, times=t))


def test_uint64_indices_multi_attribute_and_vector_calls(device):
    dims = (10, 12, 9)
    d0, kw = V.temporal_volume(dims, V.wavelet_time, ("varying", None), index64=True)
    d1, _ = V.temporal_volume(dims, V.xyz_time, ("varying", None), index64=True)
    assert kw["tu_indices"].dtype == np.uint64
    vol = device.structured_regular([d0, d1], dims=dims, **kw)
    grid = O.Grid.dense_temporal(dims, [d0, d1], **kw)
    pts = V.uniform_points(5000, (0, 0, 0), (9, 11, 8), seed=8)
    t = np.random.default_rng(9).random(5000, dtype=np.float32)
    s = vol.sampler()
    m = s.sample_m(pts, [1, 0], times=t)
    assert_same(m[:, 0], grid.sample(pts, attr=1, times=t))
    assert_same(m[:, 1], grid.sample(pts, attr=0, times=t))
    # 8-wide call with per-lane times and a masked lane
    p8 = pts[:8]
    valid = np.ones(8, np.int32)
    valid[3] = 0
    got = s.sample_wide(p8, attr=1, width=8, valid=valid, times=t[:8])
    want = grid.sample(p8, attr=1, times=t[:8])
    assert_same(got[valid == 1], want[valid == 1])


@pytest.mark.parametrize("kind", [("structured", 4), ("varying", None)], ids=["s4", "uv"])
def test_cubic_within_tolerance(device, kind):
    vol, grid, _, _ = pair(device, (16, 15, 14), V.wavelet_time, kind)
    pts = V.uniform_points(20000, (0, 0, 0), (15, 14, 13), seed=10)
    t = np.random.default_rng(11).random(20000, dtype=np.float32)
    grid.set_sampler(O.CUBIC)
    got = vol.sampler(filter=vkl.VKL_FILTER_CUBIC).sample(pts, times=t)
    grid.set_cubic_double(True)
    exact = grid.sample(pts, times=t)
    grid.set_cubic_double(False)
    assert_close_scale(got, exact, rel=1e-5)


@pytest.mark.parametrize("kind", [("structured", 3), ("varying", None)], ids=["s3", "uv"])
def test_majorants_and_intervals(device, kind):
    """Value ranges cover every time sample (node table equals the oracle's),
    so interval iteration is time independent and bit-identical."""
    sp = np.float32(1.0 / 39.0)
    vol, grid, data, _ = pair(device, (40, 40, 40), V.wavelet_time, kind, spacing=(sp, sp, sp))
    for l in range(3):
        n, vox, rng_ = vol.level(l)
        _, ovox, orng = grid.level(l)
        np.testing.assert_array_equal(vox, ovox)
        np.testing.assert_array_equal(rng_.view(np.uint32), orng.view(np.uint32))
    np.testing.assert_array_equal(vol.value_range(), [data.min(), data.max()])
    ctx = vol.sampler().interval_context(value_ranges=[[1.0, 1.5]])
    org, d = rays_through(grid.bounding_box(), 200, seed=12)
    out, cnt = ctx.batch(org, d, max_per_ray=256)
    for r in range(200):
        want = grid.intervals(org[r], d[r], depth=O.map_depth(0.5), value_ranges=[[1.0, 1.5]], max_out=256)
        np.testing.assert_array_equal(out[r, :cnt[r]].view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("kind", [("structured", 3), ("varying", None)], ids=["s3", "uv"])
def test_hits_at_ray_time(device, kind):
    sp = np.float32(1.0 / 31.0)
    vol, grid, _, _ = pair(device, (32, 32, 32), V.wavelet_time, kind, spacing=(sp, sp, sp))
    org, d = rays_through(grid.bounding_box(), 64, seed=13)
    times = np.random.default_rng(14).random(64, dtype=np.float32)
    ctx = vol.sampler().hit_context([0.8, 1.9])
    out, cnt = ctx.batch(org, d, times=times, max_per_ray=16)
    total = 0
    for r in range(64):
        want = grid.hits(org[r], d[r], [0.8, 1.9], time=float(times[r]), max_out=16)
        got = out[r, :cnt[r]]
        assert got.shape == want.shape, r
        assert_same(got, want)
        total += len(want)
    assert total > 20


def expect_error(device, fn, text):
    with pytest.raises(vkl.VklError) as e:
        fn()
    assert text in e.value.msg
    device.errors.clear()


def test_commit_validation(device):
    dims = (2, 2, 2)
    vox, idx, times = tuv_known_answer()
    sr = device.structured_regular
    expect_error(device, lambda: sr(vox, dims=dims, tu_indices=idx[:5], tu_times=times),
                 "temporallyUnstructuredIndices is too small.")
    bad = idx.copy()
    bad[3] = bad[2]
    expect_error(device, lambda: sr(vox, dims=dims, tu_indices=bad, tu_times=times),
                 "temporallyUnstructuredIndices must increase strictly monotonically.")
    t2 = times.copy()
    t2[5], t2[6] = t2[6], t2[5]
    expect_error(device, lambda: sr(vox, dims=dims, tu_indices=idx, tu_times=t2),
                 "temporallyUnstructuredTimes must increase strictly monotonically for every voxel.")
    t3 = times.copy()
    t3[1] = 1.5
    expect_error(device, lambda: sr(vox, dims=dims, tu_indices=idx, tu_times=t3),
                 "Values in temporallyUnstructuredTimes must be bounded by 0.0 and 1.0 for every voxel.")
    expect_error(device, lambda: sr(vox, dims=dims, tu_indices=idx),
                 "but temporallyUnstructuredTimes is not specified.")
    expect_error(device, lambda: sr(vox, dims=dims, structured_steps=2, tu_indices=idx, tu_times=times),
                 "Temporally structured volumes should not have temporallyUnstructuredIndices")
    expect_error(device, lambda: sr(np.zeros(8 * 3 - 1, np.float32), dims=dims, structured_steps=3),
                 "attribute 0 has 23 elements, but expected 24")
    expect_error(device, lambda: sr(np.zeros(16, np.float32), dims=dims, structured_steps=2, corner_packed=True),
                 "cornerPacked and brickShift are not supported for temporal data")


# ---- vdb with temporal leaves --------------------------------------------------------------


def vdb_pair(device, fn=V.wavelet_time, dtype=np.float32, seed=V.SEED + 6):
    nd = V.temporal_vdb_nodes(fn, seed=seed, dtype=dtype)
    vol = device.vdb(nd["level"], nd["origin"], nd["fmt"], leaf_data=nd["leaf_data"],
                     temporal_format=nd["temporal_format"], steps=nd["steps"], tu_indices=nd["tu_indices"],
                     tu_times=nd["tu_times"])
    grid = O.Grid.vdb_temporal(nd["level"], nd["origin"], nd["fmt"], nd["leaf_data"], nd["temporal_format"],
                               nd["steps"], nd["tu_indices"], nd["tu_times"])
    return vol, grid, nd


@pytest.mark.parametrize("dtype", [np.float32, np.float16])
@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR])
def test_vdb_temporal_sample_gradient_bitexact(device, dtype, filt):
    vol, grid, _ = vdb_pair(device, dtype=dtype)
    grid.set_sampler(filt)
    pts = V.uniform_points(40000, (-2, -2, -2), (258, 258, 132), seed=41)
    t = np.random.default_rng(42).random(40000, dtype=np.float32)
    s = vol.sampler(filter=filt)
    assert_same(s.sample(torch.from_numpy(pts).cuda(), times=torch.from_numpy(t).cuda()).cpu().numpy(),
                grid.sample(pts, times=t))
    if filt == vkl.VKL_FILTER_LINEAR:
        assert_same(s.gradient(pts[:8000], times=t[:8000]), grid.gradient(pts[:8000], times=t[:8000]))


def test_vdb_temporal_node_table_intervals_hits(device):
    vol, grid, _ = vdb_pair(device, seed=V.SEED + 7)
    for l in range(3):
        n, vox, rng_ = vol.level(l)
        _, ovox, orng = grid.level(l)
        np.testing.assert_array_equal(vox, ovox)
        np.testing.assert_array_equal(rng_.view(np.uint32), orng.view(np.uint32))
    org, d = rays_through(grid.bounding_box(), 64, seed=43)
    out, cnt = vol.sampler().interval_context().batch(org, d, max_per_ray=256)
    for r in range(64):
        want = grid.intervals(org[r], d[r], depth=O.map_depth(0.5), max_out=256)
        np.testing.assert_array_equal(out[r, :cnt[r]].view(np.uint32), want.view(np.uint32))
    times = np.random.default_rng(44).random(64, dtype=np.float32)
    out, cnt = vol.sampler().hit_context([0.5, 1.5]).batch(org, d, times=times, max_per_ray=16)
    total = 0
    for r in range(64):
        want = grid.hits(org[r], d[r], [0.5, 1.5], time=float(times[r]), max_out=16)
        assert_same(out[r, :cnt[r]], want)
        total += len(want)
    assert total > 10


def test_vdb_temporal_observer_and_errors(device):
    vol, grid, nd = vdb_pair(device, seed=V.SEED + 8)
    s = vol.sampler()
    obs = s.observer("LeafNodeAccess")
    acc = grid.observe_leaf_access()
    pts = V.uniform_points(5000, (0, 0, 0), (200, 200, 128), seed=45)
    t = np.random.default_rng(46).random(5000, dtype=np.float32)
    assert_same(s.sample(pts, times=t), grid.sample(pts, times=t))
    np.testing.assert_array_equal(obs.map(), acc)
    obs.release()
    grid.observe_leaf_access(False)
    n = len(nd["level"])
    dense = np.concatenate([a[0] for a, f in zip(nd["leaf_data"], nd["temporal_format"]) if f == 0])
    expect_error(device, lambda: device.vdb(nd["level"], nd["origin"], nd["fmt"], packed_dense=[dense],
                                            temporal_format=nd["temporal_format"]),
                 "only temporally constant data is supported for packed dense / tile data")
    expect_error(device, lambda: device.vdb(nd["level"], nd["origin"], nd["fmt"], leaf_data=nd["leaf_data"],
                                            temporal_format=nd["temporal_format"], steps=nd["steps"][:-1],
                                            tu_indices=nd["tu_indices"], tu_times=nd["tu_times"]),
                 "If node.temporallyStructuredNumTimesteps is set, it must have the same size as node.data")
    expect_error(device, lambda: device.vdb(nd["level"], nd["origin"], nd["fmt"], leaf_data=nd["leaf_data"],
                                            temporal_format=nd["temporal_format"], steps=nd["steps"],
                                            tu_indices=nd["tu_indices"], tu_times=nd["tu_times"],
                                            corner_packed=True),
                 "cornerPacked is not supported for temporal data")
    assert n > 100
