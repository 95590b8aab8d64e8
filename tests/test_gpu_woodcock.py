"""Woodcock (delta) tracking fused with sampling (vklComputeWoodcockN) on the
GPU vs the oracle, bit for bit: event distances, event samples, collision
counts and the updated TEA states, over structuredRegular volumes of several
element types and every filter, a sparse vdb, temporal structuredRegular and
vdb volumes with a shutter, device and host pointers, the no-majorant and
collision-cap cases, and a subset of a 2^20-ray launch."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same
from tests.test_gpu_iterators import rays_through

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FILTERS = [(vkl.VKL_FILTER_NEAREST, O.NEAREST), (vkl.VKL_FILTER_LINEAR, O.LINEAR), (vkl.VKL_FILTER_CUBIC, O.CUBIC)]


def ramp_tf(n=8, seed=0):
    rng = np.random.default_rng(seed)
    tf = rng.random((n, 4)).astype(np.float32)
    tf[:, 3] = np.linspace(0.0, 1.0, n, dtype=np.float32) ** 2
    return tf


def ray_set(box, count, seed):
    org, d = rays_through(box, count, seed)
    lo, hi = box[:3].astype(np.float64), box[3:].astype(np.float64)
    # entry / exit of the box (rays aim inside it, so the slab test is finite)
    inv = 1.0 / np.where(d == 0, 1e-30, d).astype(np.float64)
    t0 = np.minimum((lo - org) * inv, (hi - org) * inv).max(1)
    t1 = np.maximum((lo - org) * inv, (hi - org) * inv).min(1)
    tr = np.stack([np.maximum(t0, 0.0), t1], 1).astype(np.float32)
    st = np.empty((count, 2), np.uint32)
    st[:, 0] = 5
    st[:, 1] = np.arange(count, dtype=np.uint32)
    return org, d, tr, st


def compare(sampler, grid, org, d, tr, st, sigma, tf, tf_range, max_collisions=1 << 20, **kw):
    st_dev = st.copy()
    ev, sm, cnt = sampler.woodcock(org, d, tr, st_dev, sigma, tf, tf_range, max_collisions=max_collisions, **kw)
    st_ref = st.copy()
    wev, wsm, wcnt = grid.woodcock(org, d, tr, st_ref, sigma, tf, tf_range, max_collisions=max_collisions, **kw)
    assert_same(ev, wev)
    assert_same(sm, wsm)
    np.testing.assert_array_equal(cnt, wcnt)
    np.testing.assert_array_equal(st_dev, st_ref)
    return ev, cnt


def compare_cubic(sampler, grid, org, d, tr, st, sigma, tf, tf_range):
    """The tricubic filter matches the oracle to ~1e-6 of the field scale, not
    bit for bit (test_gpu_dense.check_cubic), so an acceptance test
    u1 < opacity(sample) can flip where the two differ: rays must agree
    exactly except for such rare flips, and event samples to 1e-5 of scale."""
    st_dev, st_ref = st.copy(), st.copy()
    ev, sm, cnt = sampler.woodcock(org, d, tr, st_dev, sigma, tf, tf_range)
    wev, wsm, wcnt = grid.woodcock(org, d, tr, st_ref, sigma, tf, tf_range)
    same = (ev.view(np.uint32) == wev.view(np.uint32)) & (cnt == wcnt) & (st_dev == st_ref).all(1)
    assert same.mean() >= 0.995, same.mean()
    assert np.abs(sm[same].astype(np.float64) - wsm[same]).max() <= 1e-5 * np.abs(wsm).max()
    return ev, cnt


@pytest.fixture(scope="module")
def wavelet_pair(device):
    dims, origin, spacing = (40, 36, 44), (-1.0, 0.5, 2.0), (0.25, 0.3, 0.2)
    data = V.procedural(dims, V.wavelet, origin, spacing)
    vol = device.structured_regular(data, origin=origin, spacing=spacing)
    grid = O.Grid.dense(data, origin=origin, spacing=spacing)
    return vol, grid


@pytest.mark.parametrize("filt", FILTERS, ids=["nearest", "linear", "cubic"])
def test_dense_float_bitexact(wavelet_pair, filt):
    vol, grid = wavelet_pair
    grid.set_sampler(filt[1])
    s = vol.sampler(filter=filt[0])
    box = grid.bounding_box()
    org, d, tr, st = ray_set(box, 3000, 11)
    if filt[1] == O.CUBIC:
        ev, cnt = compare_cubic(s, grid, org, d, tr, st, 2.5, ramp_tf(), (-3.0, 3.0))
    else:
        ev, cnt = compare(s, grid, org, d, tr, st, 2.5, ramp_tf(), (-3.0, 3.0))
    assert np.isfinite(ev).mean() > 0.3 and np.isinf(ev).mean() > 0.005  # both outcomes occur
    assert cnt.mean() > 3


@pytest.mark.parametrize("dtype", [np.uint8, np.float16, np.float64], ids=["uint8", "half", "double"])
def test_dense_types_bitexact(device, dtype):
    dims = (30, 25, 20)
    f = np.float64(V.procedural(dims, V.wavelet, spacing=(0.3, 0.3, 0.3), dtype=np.float64))
    data = ((f + 3) * 40).astype(dtype) if dtype == np.uint8 else f.astype(dtype)
    vol = device.structured_regular(data, spacing=(0.3, 0.3, 0.3))
    grid = O.Grid.dense(data, spacing=(0.3, 0.3, 0.3))
    lo, hi = grid.value_range()
    org, d, tr, st = ray_set(grid.bounding_box(), 2000, 12)
    compare(vol.sampler(), grid, org, d, tr, st, 3.0, ramp_tf(6, 1), (lo, hi))


def test_vdb_bitexact(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    vol = device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    grid = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    for vf, of in FILTERS[:2]:
        grid.set_sampler(of)
        org, d, tr, st = ray_set(grid.bounding_box(), 3000, 13)
        ev, cnt = compare(vol.sampler(filter=vf), grid, org, d, tr, st, 0.08, ramp_tf(5, 2), (0.0, 1.2))
        assert np.isfinite(ev).any() and np.isinf(ev).any()


@pytest.mark.parametrize("shutter", [0.0, 0.5])
def test_temporal_dense_and_vdb_bitexact(device, shutter):
    dims = (14, 12, 10)
    data, kw = V.temporal_volume(dims, V.wavelet_time, ("structured", 5))
    vol = device.structured_regular(data, dims=dims, **kw)
    grid = O.Grid.dense_temporal(dims, data, **kw)
    org, d, tr, st = ray_set(grid.bounding_box(), 2000, 14)
    compare(vol.sampler(), grid, org, d, tr, st, 1.5, ramp_tf(), (-3.0, 3.0), time=0.4, shutter=shutter)

    nd = V.temporal_vdb_nodes(V.wavelet_time, seed=31)
    vv = device.vdb(nd["level"], nd["origin"], nd["fmt"], leaf_data=nd["leaf_data"],
                    temporal_format=nd["temporal_format"], steps=nd["steps"], tu_indices=nd["tu_indices"],
                    tu_times=nd["tu_times"])
    gv = O.Grid.vdb_temporal(nd["level"], nd["origin"], nd["fmt"], nd["leaf_data"], nd["temporal_format"],
                             nd["steps"], nd["tu_indices"], nd["tu_times"])
    org, d, tr, st = ray_set(gv.bounding_box(), 2000, 15)
    compare(vv.sampler(), gv, org, d, tr, st, 0.05, ramp_tf(), (-3.0, 3.0), time=0.7, shutter=shutter)


def test_device_pointers_and_edge_cases(wavelet_pair):
    vol, grid = wavelet_pair
    grid.set_sampler(O.LINEAR)
    s = vol.sampler()
    org, d, tr, st = ray_set(grid.bounding_box(), 1500, 16)
    tf = ramp_tf()
    ev, sm, cnt = s.woodcock(*[torch.from_numpy(a).cuda() for a in (org, d, tr)],
                             _dev_u32(st), 2.5,
                             torch.from_numpy(tf).cuda(), (-3.0, 3.0))
    st_ref = st.copy()
    wev, wsm, wcnt = grid.woodcock(org, d, tr, st_ref, 2.5, tf, (-3.0, 3.0))
    assert_same(ev.cpu().numpy(), wev)
    assert_same(sm.cpu().numpy(), wsm)
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32), wcnt)
    # no majorant: +inf without drawing; collision cap: NaN after max_collisions rejections
    compare(s, grid, org, d, tr, st, 0.0, tf, (-3.0, 3.0))
    tz = np.zeros((4, 4), np.float32)
    compare(s, grid, org, d, tr, st, 50.0, tz, (-3.0, 3.0), max_collisions=7)
    ev, _, cnt = s.woodcock(org, d, tr, st.copy(), 50.0, tz, (-3.0, 3.0), max_collisions=7)
    assert np.isnan(ev).any() and cnt.max() == 7
    compare(s, grid, org[:0], d[:0], tr[:0], st[:0], 2.5, tf, (-3.0, 3.0))


def _dev_u32(st):
    return torch.from_numpy(st.view(np.int32)).cuda()


def test_million_rays_subset(device):
    """2^20 rays through a 128^3 volume in one launch; 3000 of them checked
    against the oracle (rays are independent)."""
    dims = (128, 128, 128)
    data = V.procedural(dims, V.wavelet, spacing=(0.05, 0.05, 0.05))
    vol = device.structured_regular(data, spacing=(0.05, 0.05, 0.05))
    grid = O.Grid.dense(data, spacing=(0.05, 0.05, 0.05))
    n = 1 << 20
    org, d, tr, st = ray_set(grid.bounding_box(), n, 17)
    tf = ramp_tf()
    st_dev = torch.from_numpy(st.view(np.int32)).cuda()
    ev, sm, cnt = vol.sampler().woodcock(*[torch.from_numpy(a).cuda() for a in (org, d, tr)], st_dev, 4.0,
                                         torch.from_numpy(tf).cuda(), (-3.0, 3.0))
    pick = np.random.default_rng(3).choice(n, 3000, replace=False)
    st_ref = st[pick].copy()
    wev, wsm, wcnt = grid.woodcock(org[pick], d[pick], tr[pick], st_ref, 4.0, tf, (-3.0, 3.0))
    assert_same(ev.cpu().numpy()[pick], wev)
    assert_same(sm.cpu().numpy()[pick], wsm)
    np.testing.assert_array_equal(cnt.cpu().numpy().view(np.uint32)[pick], wcnt)
    np.testing.assert_array_equal(st_dev.cpu().numpy().view(np.uint32)[pick], st_ref)
