"""Observers on the HIP device vs the oracle: InnerNode boxes and value ranges
(every maxDepth, several attributes), LeafNodeAccess marks after sampling,
gradients, multi-attribute sampling and hit iteration (per-node and packed
input, every filter, maxSamplingDepth, corner-packed volumes), observer
epochs, release, and unsupported types."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests.test_gpu_dense import assert_same
from tests.test_oracle_observers import (KA_CASES, check_inner_nodes_known_answer, known_answer_volume,
                                         mixed_node_arrays)

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def both(device, seed, packed, nattr=1, corner_packed=None):
    levels, origins, fmt, dense, tile, per_node = mixed_node_arrays(seed, nattr)
    if packed:
        vol = device.vdb(levels, origins, fmt, packed_dense=dense, packed_tile=tile, corner_packed=corner_packed)
        grid = O.Grid.vdb(levels, origins, fmt, packed_dense=dense, packed_tile=tile)
    else:
        vol = device.vdb(levels, origins, fmt, leaf_data=per_node, corner_packed=corner_packed)
        grid = O.Grid.vdb(levels, origins, fmt, leaf_data=per_node)
    return vol, grid, origins


def points(origins, seed, n_near=3000, n_far=3000):
    rng = np.random.default_rng(seed)
    near = origins[rng.choice(len(origins), n_near)] + rng.random((n_near, 3)) * 10 - 1
    far = rng.random((n_far, 3)) * [700, 400, 400] - 2
    return np.concatenate([near, far, [[4100.5, 3.25, 7.75], [-5, -5, -5], [np.nan, 1, 1]]]).astype(np.float32)


def sorted_rows(a):
    return a[np.lexsort(a.T[::-1])]


@pytest.mark.parametrize("max_depth", [None, 0, 1, 2, 3])
def test_inner_node_matches_oracle(device, max_depth):
    vol, grid, _ = both(device, 21, packed=True, nattr=2)
    obs = vol.observer("InnerNode", max_depth=max_depth)
    got = obs.map()
    assert obs.element_type() == vkl.VKL_OBJECT
    assert vkl.lib().vklGetObserverElementSize(obs.h) == 4 * (6 + 2 * 2)
    want = grid.inner_nodes(1 if max_depth is None else max_depth)
    assert got.shape == want.shape
    np.testing.assert_array_equal(sorted_rows(got).view(np.uint32), sorted_rows(want).view(np.uint32))
    obs.release()


def test_inner_node_transform_and_recommit(device):
    levels, origins, fmt, dense, tile, _ = mixed_node_arrays(22)
    i2o = np.array([0.5, 0, 0, 0, 2.0, 0, 0, 0, 0.25, 4.0, -8.0, 1.0], np.float32)
    vol = device.vdb(levels, origins, fmt, packed_dense=dense, packed_tile=tile,
                     index_to_object=[0.5, 0, 0, 0, 2.0, 0, 0, 0, 0.25, 4.0, -8.0, 1.0])
    grid = O.Grid.vdb(levels, origins, fmt, packed_dense=dense, packed_tile=tile, i2o=i2o)
    obs = vol.observer("InnerNode")
    for d in (2, 0, 1):
        obs.set_int("maxDepth", d)
        obs.commit()
        got, want = obs.map(), grid.inner_nodes(d)
        np.testing.assert_array_equal(sorted_rows(got), sorted_rows(want))
    obs.release()


@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_CUBIC])
def test_leaf_access_matches_oracle(device, packed, filt):
    vol, grid, origins = both(device, 23, packed)
    pts = points(origins, 24, 120, 3000)  # reaches some nodes, not all
    s = vol.sampler(filter=filt)
    obs = s.observer("LeafNodeAccess")
    assert obs.map().sum() == 0
    grid.set_sampler(filt)
    acc = grid.observe_leaf_access()
    got = s.sample(torch.from_numpy(pts).cuda()).cpu().numpy()
    want = grid.sample(pts)
    if filt != vkl.VKL_FILTER_CUBIC:
        assert_same(got, want)
    m = obs.map()
    assert m.dtype == np.uint32 and m.shape == (len(origins),)
    assert obs.element_type() == vkl.VKL_UINT
    np.testing.assert_array_equal(m, acc)
    assert 10 < acc.sum() < len(acc)
    # gradients on a fresh observer and oracle buffer
    obs.release()
    if filt != vkl.VKL_FILTER_NEAREST:
        obs = s.observer("LeafNodeAccess")
        acc = grid.observe_leaf_access()
        s.gradient(pts[:2000])
        grid.gradient(pts[:2000])
        np.testing.assert_array_equal(obs.map(), acc)
        obs.release()
    grid.observe_leaf_access(False)


@pytest.mark.parametrize("md", [0, 1, 2])
def test_leaf_access_max_sampling_depth_and_sample_m(device, md):
    vol, grid, origins = both(device, 25, packed=True, nattr=2)
    pts = points(origins, 26, 1500, 1500)
    s = vol.sampler(max_sampling_depth=md)
    obs = s.observer("LeafNodeAccess")
    grid.set_sampler(O.LINEAR, max_sampling_depth=md)
    acc = grid.observe_leaf_access()
    got = s.sample_m(pts, [1, 0])
    want = grid.sample_m(pts, [1, 0])
    assert_same(got, want)
    np.testing.assert_array_equal(obs.map(), acc)
    if md == 0:
        assert acc.sum() == 0
    obs.release()


def test_corner_packed_volume_observed(device):
    """Observation routes corner-packed samplers through the node table (the
    cell records skip it); samples stay bit-exact and the packed path returns
    once the observer is released."""
    vol, grid, origins = both(device, 27, packed=True, corner_packed=True)
    pts = points(origins, 28)
    s = vol.sampler()
    before = s.sample(pts)
    obs = s.observer("LeafNodeAccess")
    acc = grid.observe_leaf_access()
    during = s.sample(pts)
    want = grid.sample(pts)
    assert_same(during, want)
    assert_same(before, want)
    np.testing.assert_array_equal(obs.map(), acc)
    obs.release()
    assert_same(s.sample(pts), want)


def test_observer_epochs_release_and_recommit(device):
    """A new observer starts empty; older observers keep seeing new accesses
    (vdb observers' buffers are zeroed at creation and written by every later
    traversal); a sampler re-commit keeps observation; interval iteration
    samples nothing."""
    vol, grid, origins = both(device, 29, packed=False)
    s = vol.sampler()
    p1, p2 = points(origins, 30, 800, 0), points(origins, 31, 800, 0)
    a = s.observer("LeafNodeAccess")
    s.sample(p1)
    b = s.observer("LeafNodeAccess")
    s.set_int("filter", vkl.VKL_FILTER_LINEAR)
    s.commit()
    s.sample(p2)
    acc1 = grid.observe_leaf_access()
    grid.sample(p1)
    grid.sample(p2)
    acc2 = grid.observe_leaf_access()
    grid.sample(p2)
    np.testing.assert_array_equal(a.map(), acc1)
    np.testing.assert_array_equal(b.map(), acc2)
    assert acc1.sum() > acc2.sum()
    # interval iteration does not sample
    c = s.observer("LeafNodeAccess")
    ctx = s.interval_context()
    org = np.tile(np.array([[-10, 4.5, 4.5]], np.float32), (16, 1))
    org[:, 1] += np.arange(16, dtype=np.float32) * 8
    d = np.tile(np.array([[1, 0, 0]], np.float32), (16, 1))
    ctx.batch(org, d)
    assert c.map().sum() == 0
    for o in (a, b, c):
        o.release()
    # unobserved again: nothing recorded anywhere
    d2 = s.observer("LeafNodeAccess")
    d2.release()
    s.sample(p1)


def test_hit_iteration_marks_like_oracle(device):
    vol, grid, origins = both(device, 32, packed=True)
    s = vol.sampler()
    obs = s.observer("LeafNodeAccess")
    rng = np.random.default_rng(33)
    n = 48
    tgt = origins[rng.choice(len(origins), n)] + 4.0
    org = (tgt + rng.standard_normal((n, 3)) * 40).astype(np.float32)
    d = (tgt - org).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    values = [0.5, 2.0]
    ctx = s.hit_context(values)
    ctx.batch(org, d.astype(np.float32), max_per_ray=16)
    acc = grid.observe_leaf_access()
    for i in range(n):
        grid.hits(org[i], d[i], values, max_out=16)
    np.testing.assert_array_equal(obs.map(), acc)
    assert acc.sum() > 0
    obs.release()


def test_unsupported_observers(device):
    data = np.zeros((8, 8, 8), np.float32)
    vol = device.structured_regular(data)
    for fn, obj in ((vkl.lib().vklNewVolumeObserver, vol), (vkl.lib().vklNewSamplerObserver, vol.sampler())):
        h = fn(obj.h, b"InnerNode" if fn is vkl.lib().vklNewVolumeObserver else b"LeafNodeAccess")
        code, msg = device.last_error()
        device.errors.clear()
        assert not h.host and "unsupported observer type" in msg
    levels, origins, fmt, dense, tile, _ = mixed_node_arrays(34)
    vdb = device.vdb(levels, origins, fmt, packed_dense=dense, packed_tile=tile)
    h = vkl.lib().vklNewVolumeObserver(vdb.h, b"LeafNodeAccess")
    code, msg = device.last_error()
    device.errors.clear()
    assert not h.host and "unsupported observer type: LeafNodeAccess" in msg


@pytest.mark.parametrize("origin,spacing", KA_CASES)
@pytest.mark.parametrize("packed,nattr", [(True, 1), (True, 3), (False, 1)])
def test_inner_nodes_reference_known_answer(device, origin, spacing, packed, nattr):
    """vdb_volume_inner_node_observer.cpp at its own size (256^3), packed and
    per-node input, one and three attributes."""
    level, origins, fmt, dense, i2o = known_answer_volume(origin, spacing, nattr, dims=256)
    i2o_l = [float(v) for v in i2o]
    if packed:
        vol = device.vdb(level, origins, fmt, packed_dense=dense, index_to_object=i2o_l)
    else:
        per_node = [[d[i * 512:(i + 1) * 512] for d in dense] for i in range(len(origins))]
        vol = device.vdb(level, origins, fmt, leaf_data=per_node, index_to_object=i2o_l)
    s = vol.sampler(filter=vkl.VKL_FILTER_NEAREST)

    def inner(d):
        o = vol.observer("InnerNode", max_depth=d)
        try:
            return o.map()
        finally:
            o.release()

    def sample(p, a):
        return s.sample(torch.from_numpy(p).cuda(), attr=a).cpu().numpy()

    check_inner_nodes_known_answer(inner, sample, vol.bounding_box(), 256, nattr)
