"""The 4 / 8 / 16-wide calls (openvkl.h:38-152 and 163-434 of the reference's
device header) through the C ABI with `valid` masks, against the oracle:
vklComputeGradient{4,8,16}, vklComputeSampleM{4,8,16} (layout
samples[a * W + lane]) and vklInit/IterateInterval{4,8,16},
vklInit/IterateHit{4,8,16} (every valid lane iterates its own ray; invalid
lanes return 0). Mirrors the reference's vectorized_gradients.cpp,
vectorized_interval_iterator.cpp and vectorized_hit_iterator.cpp."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_close_scale, assert_same
from tests.test_gpu_iterators import rays_through

pytestmark = pytest.mark.gpu
WIDTHS = (4, 8, 16)


def masks(w, rng):
    """All lanes, alternate lanes, one lane, and a random mask."""
    return [[1] * w, [i % 2 for i in range(w)], [1 if i == w - 1 else 0 for i in range(w)],
            list(rng.integers(0, 2, w))]


def soa(pts, w):
    VV = vkl.WIDE[w][0]
    v = VV()
    for i in range(len(pts)):
        v.x[i], v.y[i], v.z[i] = (float(c) for c in pts[i])
    return v


@pytest.fixture(scope="module")
def two_attr(device):
    rng = np.random.default_rng(41)
    a0 = V.wavelet_volume(40)
    a1 = rng.standard_normal((40, 40, 40)).astype(np.float16)
    vol = device.structured_regular([a0, a1], origin=(-2.0, 1.0, 0.5), spacing=(0.5, 0.75, 1.25))
    grid = O.Grid.dense([a0, a1], origin=(-2.0, 1.0, 0.5), spacing=(0.5, 0.75, 1.25))
    return vol, grid


@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_CUBIC])
def test_gradient_wide(two_attr, w, filt):
    vol, grid = two_attr
    s = vol.sampler(filter=filt, gradient_filter=filt)
    grid.set_sampler(filt, filt)
    try:
        check_gradients(s, grid, w, filt)
    finally:
        grid.set_sampler()


def check_gradients(s, grid, w, filt):
    rng = np.random.default_rng(42 + w)
    VV = vkl.WIDE[w][0]
    for valid in masks(w, rng):
        pts = V.uniform_points(w, (-3, 0, 0), (18.5, 31, 50), seed=int(rng.integers(1 << 30)))
        out = VV()
        for i in range(w):
            out.x[i] = out.y[i] = out.z[i] = -7.0
        val = (C.c_int * w)(*[int(v) for v in valid])
        getattr(vkl.lib(), f"vklComputeGradient{w}")(C.addressof(val), C.byref(s.h), C.byref(soa(pts, w)),
                                                      C.byref(out), 1, None)
        s.device.check()
        if filt == vkl.VKL_FILTER_CUBIC:
            # the separable Catmull-Rom form vs the oracle's 64x64 matrix
            # form evaluated in binary64: within 1e-5 of the gradient scale
            grid.set_cubic_double(True)
            want = grid.gradient(pts, attr=1)
            grid.set_cubic_double(False)
        else:
            want = grid.gradient(pts, attr=1)
        got = np.array([[out.x[i], out.y[i], out.z[i]] for i in range(w)], np.float32)
        on = np.array(valid, bool)
        # invalid lanes are not written
        np.testing.assert_array_equal(got[~on], np.full((int((~on).sum()), 3), -7.0, np.float32))
        if filt == vkl.VKL_FILTER_CUBIC:
            assert np.array_equal(np.isnan(got[on]), np.isnan(want[on]))
            if not np.isnan(want[on]).all():
                assert_close_scale(got[on], want[on], rel=1e-5)
        else:
            assert_same(got[on], want[on])


@pytest.mark.parametrize("w", WIDTHS)
def test_sample_m_wide(two_attr, w):
    vol, grid = two_attr
    s = vol.sampler()
    grid.set_sampler()
    rng = np.random.default_rng(52 + w)
    attrs = np.array([1, 0], np.uint32)
    for valid in masks(w, rng):
        pts = V.uniform_points(w, (-3, 0, 0), (18.5, 31, 50), seed=int(rng.integers(1 << 30)))
        out = (C.c_float * (2 * w))(*([-7.0] * (2 * w)))
        val = (C.c_int * w)(*[int(v) for v in valid])
        getattr(vkl.lib(), f"vklComputeSampleM{w}")(C.addressof(val), C.byref(s.h), C.byref(soa(pts, w)),
                                                     C.addressof(out), 2, attrs.ctypes.data, None)
        s.device.check()
        want = grid.sample_m(pts, attrs)  # [w, 2]
        got = np.array(out[:], np.float32).reshape(2, w)  # samples[a * W + lane]
        for i in range(w):
            if valid[i]:
                assert_same(got[:, i], want[i])
            else:
                np.testing.assert_array_equal(got[:, i], [-7.0, -7.0])


def iterate_wide(ctx, w, org, d, valid, hit):
    """Init + iterate the W-wide iterator until every valid lane is done;
    returns per-lane lists of results."""
    L = vkl.lib()
    VV, VR, VI, VH = vkl.WIDE[w]
    kind = "Hit" if hit else "Interval"
    size = getattr(L, f"vklGet{kind}IteratorSize{w}")(C.byref(ctx.h))
    buf = C.create_string_buffer(size)
    tr = VR()
    for i in range(w):
        tr.lower[i], tr.upper[i] = 0.0, float("inf")
    val = (C.c_int * w)(*[int(v) for v in valid])
    times = (C.c_float * w)(*([0.0] * w))
    it = getattr(L, f"vklInit{kind}Iterator{w}")(C.addressof(val), C.byref(ctx.h), C.byref(soa(org, w)),
                                                   C.byref(soa(d, w)), C.byref(tr), C.addressof(times), buf)
    ctx.device.check()
    res = [[] for _ in range(w)]
    live = list(valid)
    for _ in range(10000):
        if not any(live):
            break
        out = VH() if hit else VI()
        result = (C.c_int * w)()
        lv = (C.c_int * w)(*[int(v) for v in live])
        getattr(L, f"vklIterate{kind}{w}")(C.addressof(lv), it, C.byref(out), C.addressof(result))
        ctx.device.check()
        for i in range(w):
            if not live[i]:
                assert result[i] == 0
                continue
            if not result[i]:
                live[i] = 0
                continue
            if hit:
                res[i].append([out.t[i], out.sample[i], out.epsilon[i]])
            else:
                res[i].append([out.tRange.lower[i], out.tRange.upper[i], out.valueRange.lower[i],
                               out.valueRange.upper[i], out.nominalDeltaT[i]])
    return [np.array(r, np.float32).reshape(-1, 3 if hit else 5) for r in res]


@pytest.fixture(scope="module")
def shell_vdb(device):
    origins, leaves = V.sparse_shell_leaves(res=128)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32)
    return (device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)]),
            O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)]))


@pytest.mark.parametrize("w", WIDTHS)
@pytest.mark.parametrize("which", ["dense", "vdb"])
def test_interval_iterator_wide(two_attr, shell_vdb, w, which):
    vol, grid = two_attr if which == "dense" else shell_vdb
    ctx = vol.sampler().interval_context()
    rng = np.random.default_rng(62 + w)
    for valid in masks(w, rng)[:3]:
        org, d = rays_through(grid.bounding_box(), w, seed=int(rng.integers(1 << 30)))
        got = iterate_wide(ctx, w, org, d, valid, hit=False)
        for i in range(w):
            if not valid[i]:
                assert len(got[i]) == 0
                continue
            want = grid.intervals(org[i], d[i])
            assert got[i].shape == want.shape
            assert_same(got[i], want)
    ctx.release()


@pytest.mark.parametrize("w", WIDTHS)
def test_hit_iterator_wide(shell_vdb, w):
    vol, grid = shell_vdb
    ctx = vol.sampler().hit_context([0.5, 0.8])
    rng = np.random.default_rng(72 + w)
    total = 0
    for valid in masks(w, rng)[:3]:
        org, d = rays_through(grid.bounding_box(), w, seed=int(rng.integers(1 << 30)))
        got = iterate_wide(ctx, w, org, d, valid, hit=True)
        for i in range(w):
            if not valid[i]:
                assert len(got[i]) == 0
                continue
            want = grid.hits(org[i], d[i], [0.5, 0.8])
            total += len(want)
            assert got[i].shape == want.shape
            assert_same(got[i], want)
    assert total > 0
    ctx.release()
