"""Pins of the oracle's motion-blur (temporal) sampling of structuredRegular
volumes (oracle/vkl_oracle.cpp: temporallyStructured,
temporallyUnstructured, temporal majorants).

Known answers: the reference's own test of per-voxel varying unstructured
time samples (structured_regular_volume_sampling_motion_blur.cpp:84-173,
tolerance 0). Closed forms: a time-linear field sampled on vertices equals
the field at that time (structured with 2 and 4 steps, unstructured with 2
and with 6 uneven steps: the reference test's TemporalConfigs); time 0 and
time 1 return the first and last samples exactly; one structured time step
equals the temporally constant volume. Brute force: the volume value range
is the min / max over every time sample of every voxel."""
import numpy as np
import pytest

import oracle as O
from tests import volumes as V


def tuv_known_answer():
    """2^3 voxels; x = 0 voxels have times (0, 1), x = 1 voxels (0, 0.5, 1);
    values 0 at time 0 and x + y + z at time 1, linear in between."""
    idx, times, vox, s = [], [], [], 0
    for z in range(2):
        for y in range(2):
            for x in range(2):
                n = 2 if x == 0 else 3
                idx.append(s)
                t, v = [0.0] * n, [0.0] * n
                t[-1], v[-1] = 1.0, x + y + z
                if x == 1:
                    t[1], v[1] = 0.5, 0.5 * (x + y + z)
                times += t
                vox += v
                s += n
    idx.append(s)
    return np.array(vox, np.float32), np.array(idx, np.uint32), np.array(times, np.float32)


def test_reference_varying_unstructured_known_answer():
    vox, idx, times = tuv_known_answer()
    g = O.Grid.dense_temporal((2, 2, 2), vox, tu_indices=idx, tu_times=times)
    g.set_sampler(O.NEAREST)
    pts = np.array([[x, y, z] for z in range(2) for y in range(2) for x in range(2)], np.float32)
    for t in (0.0, 0.25, 0.5, 0.75, 1.0):
        np.testing.assert_array_equal(g.sample(pts, times=np.full(8, t, np.float32)),
                                      (t * pts.sum(1)).astype(np.float32))


CONFIGS = [("structured", 2), ("structured", 4), ("unstructured", [0.0, 1.0]),
           ("unstructured", [0.0, 0.15, 0.3, 0.65, 0.9, 1.0])]


@pytest.mark.parametrize("kind", CONFIGS, ids=["s2", "s4", "u2", "u6"])
def test_time_linear_field_on_vertices(kind):
    """structured_regular_volume_sampling_motion_blur.cpp:16-80 (32^3,
    spacing 1/31): vertex samples of a field linear in time equal the field at
    the sample time (to binary32 rounding of the interpolation)."""
    n, sp = 32, np.float32(1.0 / 31.0)
    data, kw = V.temporal_volume((n, n, n), V.xyz_time, kind, spacing=(sp, sp, sp))
    g = O.Grid.dense_temporal((n, n, n), data, spacing=(sp, sp, sp), **kw)
    k = np.arange(0, n, 3)
    ix, iy, iz = np.meshgrid(k, k, k, indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1)
    pts = (idx.astype(np.float32) * sp).astype(np.float32)
    for filt in (O.NEAREST, O.LINEAR):
        g.set_sampler(filt)
        for t in (0.0, 0.25, 0.5, 0.75, 1.0):
            got = g.sample(pts, times=np.full(len(pts), t, np.float32))
            want = V.xyz_time(pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64),
                              pts[:, 2].astype(np.float64), t)
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-7)


@pytest.mark.parametrize("kind", CONFIGS + [("varying", None)], ids=["s2", "s4", "u2", "u6", "uv"])
def test_end_times_return_end_samples(kind):
    n = 6
    data, kw = V.temporal_volume((n, n, n), V.wavelet_time, kind)
    g = O.Grid.dense_temporal((n, n, n), data, **kw)
    g.set_sampler(O.NEAREST)
    k = np.arange(n)
    iz, iy, ix = np.meshgrid(k, k, k, indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1).astype(np.float32)
    if "structured_steps" in kw:
        T = kw["structured_steps"]
        first, last = data[0::T], data[T - 1::T]
    else:
        idx = kw["tu_indices"].astype(np.int64)
        first, last = data[idx[:-1]], data[idx[1:] - 1]
    np.testing.assert_array_equal(g.sample(pts, times=np.zeros(len(pts), np.float32)), first)
    np.testing.assert_array_equal(g.sample(pts, times=np.ones(len(pts), np.float32)), last)


def test_single_structured_step_equals_constant():
    rng = np.random.default_rng(4)
    data = rng.standard_normal((10, 11, 12)).astype(np.float32)
    g0 = O.Grid.dense(data)
    g1 = O.Grid.dense_temporal((12, 11, 10), data.reshape(-1), structured_steps=1)
    pts = V.uniform_points(5000, (-1, -1, -1), (12, 11, 10))
    t = rng.random(5000, dtype=np.float32)
    for filt in (O.NEAREST, O.LINEAR, O.CUBIC):
        g0.set_sampler(filt)
        g1.set_sampler(filt)
        np.testing.assert_array_equal(g1.sample(pts, times=t), g0.sample(pts))
    np.testing.assert_array_equal(g1.value_range(), g0.value_range())


@pytest.mark.parametrize("kind", [("structured", 3), ("varying", None)], ids=["s3", "uv"])
def test_value_range_covers_every_time_sample(kind):
    data, kw = V.temporal_volume((20, 13, 9), V.wavelet_time, kind, seed=11)
    g = O.Grid.dense_temporal((20, 13, 9), data, **kw)
    np.testing.assert_array_equal(g.value_range(), [data.min(), data.max()])


# ---- vdb with temporal leaves ------------------------------------------------------------


def vdb_pair(fn, **kw):
    nd = V.temporal_vdb_nodes(fn, **kw)
    g = O.Grid.vdb_temporal(nd["level"], nd["origin"], nd["fmt"], nd["leaf_data"], nd["temporal_format"],
                            nd["steps"], nd["tu_indices"], nd["tu_times"])
    return nd, g


def leaf_voxel_points(nd, count, seed):
    """Voxel positions inside the dense leaves (integer index coordinates)."""
    rng = np.random.default_rng(seed)
    dense = np.nonzero(nd["fmt"] == 1)[0]
    pick = rng.choice(dense, count)
    return (nd["origin"][pick] + rng.integers(0, 8, (count, 3))).astype(np.float32)


def expected_time_linear(nd, pts, t):
    """Expected sample of a time-linear field at leaf voxel positions: the
    field at time t, except constant leaves (their values are the field at
    t = 0) and unstructured voxels with a single time sample (that sample)."""
    key = {tuple((o // 8).tolist()): i for i, o in enumerate(nd["origin"]) if nd["fmt"][i] == 1}
    p64 = pts.astype(np.float64)
    want = V.xyz_time(p64[:, 0], p64[:, 1], p64[:, 2], t)
    for k, p in enumerate(pts.astype(np.int64)):
        i = key[tuple((p // 8).tolist())]
        vi = ((p[0] & 7) << 6) | ((p[1] & 7) << 3) | (p[2] & 7)
        if nd["temporal_format"][i] == 0:
            want[k] = V.xyz_time(p64[k, 0], p64[k, 1], p64[k, 2], 0.0)
        elif nd["temporal_format"][i] == 2:
            ix = nd["tu_indices"][i]
            if ix[vi + 1] - ix[vi] == 1:
                want[k] = nd["leaf_data"][i][0][ix[vi]]
    return want


def test_vdb_time_linear_leaves_on_vertices():
    """Nearest samples at voxel positions of a time-linear field equal the
    field at the sample time, whatever each leaf's temporal format."""
    nd, g = vdb_pair(V.xyz_time)
    g.set_sampler(O.NEAREST)
    pts = leaf_voxel_points(nd, 3000, 21)
    for t in (0.0, 0.25, 0.5, 0.75, 1.0):
        got = g.sample(pts, times=np.full(len(pts), t, np.float32))
        np.testing.assert_allclose(got, expected_time_linear(nd, pts, t), rtol=2e-6, atol=1e-3)


def test_vdb_end_times_and_ranges():
    nd, g = vdb_pair(V.wavelet_time, seed=31)
    g.set_sampler(O.NEAREST)
    first, last, pts = [], [], []
    for i in np.nonzero(nd["fmt"] == 1)[0][:40]:
        d, tf = nd["leaf_data"][i][0], nd["temporal_format"][i]
        for vi in (0, 77, 511):
            x, y, z = vi >> 6, (vi >> 3) & 7, vi & 7
            pts.append(nd["origin"][i] + [x, y, z])
            if tf == 0:
                first.append(d[vi]), last.append(d[vi])
            elif tf == 1:
                T = nd["steps"][i]
                first.append(d[vi * T]), last.append(d[vi * T + T - 1])
            else:
                ix = nd["tu_indices"][i]
                first.append(d[ix[vi]]), last.append(d[ix[vi + 1] - 1])
    pts = np.array(pts, np.float32)
    np.testing.assert_array_equal(g.sample(pts, times=np.zeros(len(pts), np.float32)), np.array(first, np.float32))
    np.testing.assert_array_equal(g.sample(pts, times=np.ones(len(pts), np.float32)), np.array(last, np.float32))
    allv = np.concatenate([a[0] for a in nd["leaf_data"]])
    lo, hi = g.value_range()
    assert lo <= allv.min() and hi >= allv.max()
