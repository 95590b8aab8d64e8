"""Pins the CPU oracle (oracle/vkl_oracle.cpp) against the reference's own
known-answer tests, closed forms, invariants and brute force. CPU only."""
import math
import struct

import numpy as np
import pytest

import oracle as O
from tests import volumes as V

INF = np.float32(np.inf)


@pytest.fixture(scope="module")
def wavelet128():
    sp = np.float32(1.0 / (128.0 - 1.0))
    data = V.procedural((128, 128, 128), V.wavelet, spacing=(sp, sp, sp))
    return O.Grid.dense(data, spacing=(sp, sp, sp)), data


# ---- reference known answers ----------------------------------------------------------

def test_interval_wavelet_known_answer(wavelet128):
    """vectorized_interval_iterator.cpp:26-151: 128^3 wavelet, spacing 1/127,
    rays (x, y, -1) along +z: exactly 16 intervals, first lower == 1, last
    upper == 2, contiguous."""
    g, _ = wavelet128
    rng = np.random.default_rng(7)
    for _ in range(32):
        org = (rng.uniform(0, 1), rng.uniform(0, 1), -1.0)
        iv = g.intervals(org, (0, 0, 1))
        assert len(iv) == 128 // 8
        assert iv[0, 0] == 1.0
        assert iv[-1, 1] == 2.0
        assert np.all(iv[1:, 0] == iv[:-1, 1])


@pytest.mark.parametrize("scale", [1.0, 10.0])
def test_nominal_delta_t_known_answer(scale):
    """interval_iterator.cpp:563-630: gridOrigin -64, spacing (1,2,3)."""
    spacing = np.array([1, 2, 3], np.float32)
    data = V.procedural((128, 128, 128), V.wavelet, origin=(-64, -64, -64), spacing=spacing)
    g = O.Grid.dense(data, origin=(-64, -64, -64), spacing=spacing)
    cases = [((1, 0, 0), 1.0), ((-1, 0, 0), 1.0), ((0, 1, 0), 2.0), ((0, -1, 0), 2.0), ((0, 0, 1), 3.0),
             ((0, 0, -1), 3.0)]
    n = spacing / np.linalg.norm(spacing)
    cases += [(tuple(n), float(np.linalg.norm(spacing))), (tuple(-n), float(np.linalg.norm(spacing)))]
    for d, expected in cases:
        dd = np.asarray(d, np.float32) * np.float32(scale)
        org = -dd * 1000.0  # outside, pointing through the volume
        iv = g.intervals(tuple(org), tuple(dd))
        assert len(iv) > 0
        np.testing.assert_allclose(iv[:, 4], expected / scale, rtol=1e-5)


def test_hit_zramp_known_answer():
    """hit_iterator.cpp:60-120: z-ramp 128^3, spacing 1/127, ray (0.5,0.5,-1)
    along +z, isovalues 0.1..0.9: hit.t == iso + 1 (+-1e-3), sample == iso."""
    sp = np.float32(1.0 / 127.0)
    data = V.procedural((128, 128, 128), V.zramp, spacing=(sp, sp, sp))
    g = O.Grid.dense(data, spacing=(sp, sp, sp))
    iso = []
    f = np.float32(0.1)
    while f < 1.0:
        iso.append(f)
        f = np.float32(f + np.float32(0.1))
    hits = g.hits((0.5, 0.5, -1.0), (0, 0, 1), iso)
    assert len(hits) == len(iso)
    np.testing.assert_array_equal(hits[:, 1], np.array(iso, np.float32))
    np.testing.assert_allclose(hits[:, 0], np.array(iso, np.float32) + 1.0, atol=1e-3)
    assert np.all(hits[:, 2] > 0)  # hit_iterator_epsilon.cpp:54


def test_tricubic_matrix_reference_rows():
    """VdbSampler_tricubic.ih:34-45 (first coefficients) and :385-389 (last)."""
    A = O.tricubic_matrix()
    row = lambda r: {int(c): int(A[r, c]) for c in np.nonzero(A[r])[0]}
    assert row(0) == {0: 1}
    assert row(1) == {24: 1}
    assert row(2) == {0: -3, 1: 3, 24: -2, 25: -1}
    assert row(3) == {0: 2, 1: -2, 24: 1, 25: 1}
    assert row(4) == {16: 1}
    assert row(5) == {48: 1}
    assert list(A[63, 52:64]) == [-2, -2, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1]


def test_background_is_reference_nan():
    """background_undefined.cpp:9-20 / VKLBackgroundUndefined.h:30."""
    g = O.Grid.dense(np.ones((8, 8, 8), np.float32))
    s = g.sample(np.array([[-5.0, 0, 0], [100.0, 3, 3]], np.float32))
    for v in s:
        assert struct.unpack("<I", struct.pack("<f", v))[0] == 0xFFC068B5


# ---- closed forms ------------------------------------------------------------------------

def test_linear_reproduces_trilinear_field():
    f = lambda x, y, z: 0.5 + 0.25 * x - 0.125 * y + 0.0625 * z + 0.01 * x * y - 0.02 * y * z + 0.004 * x * y * z
    data = V.procedural((16, 16, 16), f, dtype=np.float64).astype(np.float32)
    g = O.Grid.dense(data)
    p = V.uniform_points(2000, (0, 0, 0), (15, 15, 15))
    exact = f(p[:, 0].astype(np.float64), p[:, 1].astype(np.float64), p[:, 2].astype(np.float64))
    np.testing.assert_allclose(g.sample(p), exact, rtol=0, atol=2e-5)


def test_cubic_reproduces_triquadratic_field():
    f = lambda x, y, z: 0.3 + 0.1 * x * x - 0.05 * y * y + 0.02 * z * z + 0.01 * x * y * z - 0.03 * x
    data = V.procedural((16, 16, 16), f, dtype=np.float64).astype(np.float32)
    g = O.Grid.dense(data)
    g.set_sampler(O.CUBIC)
    p = V.uniform_points(2000, (1, 1, 1), (13.999, 13.999, 13.999))
    exact = f(p[:, 0].astype(np.float64), p[:, 1].astype(np.float64), p[:, 2].astype(np.float64))
    np.testing.assert_allclose(g.sample(p), exact, rtol=0, atol=5e-4)
    # gradient of the same field
    gx = 0.2 * p[:, 0] + 0.01 * p[:, 1] * p[:, 2] - 0.03
    gy = -0.1 * p[:, 1] + 0.01 * p[:, 0] * p[:, 2]
    gz = 0.04 * p[:, 2] + 0.01 * p[:, 0] * p[:, 1]
    np.testing.assert_allclose(g.gradient(p), np.stack([gx, gy, gz], 1), atol=2e-3)


def test_cubic_equals_separable_catmull_rom():
    """The Lekien-Marsden matrix form equals the tensor product of 1-D
    Catmull-Rom splines (tricubic Hermite on central differences)."""
    rng = np.random.default_rng(3)
    data = rng.standard_normal((12, 12, 12)).astype(np.float32)
    g = O.Grid.dense(data)
    g.set_sampler(O.CUBIC)
    p = V.uniform_points(500, (1, 1, 1), (9.999, 9.999, 9.999))
    got = g.sample(p)

    def w(t):
        return np.array([(-t ** 3 + 2 * t ** 2 - t) / 2, (3 * t ** 3 - 5 * t ** 2 + 2) / 2,
                         (-3 * t ** 3 + 4 * t ** 2 + t) / 2, (t ** 3 - t ** 2) / 2])

    d = data.astype(np.float64)
    ref = []
    for x, y, z in p.astype(np.float64):
        i, j, k = int(np.floor(x)), int(np.floor(y)), int(np.floor(z))
        blk = d[k - 1:k + 3, j - 1:j + 3, i - 1:i + 3]  # [z, y, x]
        ref.append(np.einsum("zyx,x,y,z->", blk, w(x - i), w(y - j), w(z - k)))
    np.testing.assert_allclose(got, np.array(ref), atol=2e-5 * np.abs(d).max())


def test_linear_gradient_of_ramp():
    spacing = (0.5, 2.0, 4.0)
    data = V.procedural((16, 16, 16), lambda x, y, z: 3 * x - 2 * y + z, spacing=spacing, dtype=np.float64)
    g = O.Grid.dense(data.astype(np.float32), spacing=spacing)
    p = V.uniform_points(1000, (0, 0, 0), (7.4, 29.9, 59.9))
    np.testing.assert_allclose(g.gradient(p), np.tile([3, -2, 1], (1000, 1)), atol=1e-4)
    g.set_sampler(O.NEAREST)
    assert np.all(g.gradient(p) == 0)  # VdbSampler_nearest.ih:313-363


@pytest.mark.parametrize("filt", [O.NEAREST, O.LINEAR, O.CUBIC])
def test_vertex_samples_equal_data(filt):
    """sampling_utility.h:35-119: samples at grid vertices equal the voxels."""
    rng = np.random.default_rng(5)
    data = rng.standard_normal((10, 11, 12)).astype(np.float32)
    g = O.Grid.dense(data)
    g.set_sampler(filt)
    ijk = np.stack(np.meshgrid(np.arange(2, 10), np.arange(2, 9), np.arange(2, 8), indexing="ij"), -1).reshape(-1, 3)
    got = g.sample(ijk.astype(np.float32))
    want = data[ijk[:, 2], ijk[:, 1], ijk[:, 0]]
    if filt == O.CUBIC:
        np.testing.assert_allclose(got, want, atol=1e-5)
    else:
        np.testing.assert_array_equal(got, want)


# ---- invariants and brute force -------------------------------------------------------------

def test_leaf_majorants_bruteforce():
    """VdbSampler.ispc:460-557: level-2 voxel range == min/max over the
    [-1, +10) window of its leaf, clipped to the domain."""
    rng = np.random.default_rng(11)
    data = rng.standard_normal((20, 19, 21)).astype(np.float32)
    g = O.Grid.dense(data)
    _, vox, rng2 = g.level(2)
    nz, ny, nx = data.shape
    for lz in range(0, nz, 8):
        for ly in range(0, ny, 8):
            for lx in range(0, nx, 8):
                win = data[max(lz - 1, 0):lz + 10, max(ly - 1, 0):ly + 10, max(lx - 1, 0):lx + 10]
                # level-2 node of the leaf (one node for this volume), voxel index z fastest
                vi = ((lx >> 3) << 8) + ((ly >> 3) << 4) + (lz >> 3)
                assert vox[vi] & 3 == 3
                assert rng2[vi, 0, 0] == win.min() and rng2[vi, 0, 1] == win.max()
    vr = g.value_range()
    assert vr[0] == data.min() and vr[1] == data.max()


def test_dda_visits_every_touched_leaf():
    """Every leaf a ray passes through (dense t sampling) lies in an interval of
    the depth-3 iterator on a sparse volume, and intervals are contiguous."""
    origins, leaves = V.sparse_shell_leaves(res=128, radius=0.3, width=0.08)
    n = len(origins)
    g = O.Grid.vdb(np.full(n, 3), origins, np.full(n, O.FORMAT_DENSE_ZYX), packed_dense=[leaves.reshape(-1)])
    occ = set(map(tuple, (origins // 8).tolist()))
    rng = np.random.default_rng(13)
    for _ in range(30):
        o = rng.uniform(-20, 148, 3).astype(np.float32)
        tgt = rng.uniform(30, 98, 3).astype(np.float32)
        d = (tgt - o).astype(np.float32)
        iv = g.intervals(o, d, depth=3)
        if len(iv) > 1:
            assert np.all(iv[1:, 0] >= iv[:-1, 1])
        ts = np.linspace(0, 2, 20001, dtype=np.float64)
        pts = o.astype(np.float64)[None] + ts[:, None] * d.astype(np.float64)[None]
        cell = np.floor(pts / 8).astype(np.int64)
        inside = np.array([tuple(c) in occ for c in cell.tolist()])
        for t in ts[inside]:
            # skip samples within float noise of a cell boundary
            pt = o.astype(np.float64) + t * d.astype(np.float64)
            if np.min(np.abs(pt - np.round(pt))) < 1e-3:
                continue
            assert np.any((iv[:, 0] <= t + 1e-5) & (t - 1e-5 <= iv[:, 1])), (o, d, t)


def test_interval_value_range_bounds_samples(wavelet128):
    """interval_iterator.cpp:330-386: samples inside an interval lie in its
    value range."""
    g, _ = wavelet128
    rng = np.random.default_rng(17)
    for _ in range(8):
        o = np.array([rng.uniform(0, 1), rng.uniform(0, 1), -1], np.float32)
        d = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 1], np.float32)
        iv = g.intervals(o, d)
        for t0, t1, lo, hi, _ in iv:
            ts = np.linspace(t0, t1, 9, dtype=np.float32)
            s = g.sample(o[None] + ts[:, None] * d[None])
            s = s[~np.isnan(s)]
            assert np.all(s >= lo - 1e-6) and np.all(s <= hi + 1e-6)


def test_value_range_selector_culls():
    """interval_iterator.cpp:449-639: with valueRanges, every returned
    interval's value range overlaps a selected range."""
    sp = np.float32(1.0 / 127.0)
    data = V.procedural((128, 128, 128), V.zramp, spacing=(sp, sp, sp))
    g = O.Grid.dense(data, spacing=(sp, sp, sp))
    sel = np.array([[0.2, 0.3], [0.7, 0.72]], np.float32)
    iv = g.intervals((0.5, 0.5, -1), (0, 0, 1), value_ranges=sel)
    assert len(iv) > 0
    for t0, t1, lo, hi, _ in iv:
        assert any(hi >= a and lo <= b for a, b in sel)
    assert iv[0, 0] > 1.1 and iv[-1, 1] < 1.8


def test_sparse_matches_dense_interior():
    """vdb_volume_dense.cpp:15-29: a vdb built from the leaves of a dense
    array samples like the structuredRegular volume (no half-voxel offset)."""
    rng = np.random.default_rng(19)
    data = rng.standard_normal((24, 24, 24)).astype(np.float32)
    dense = O.Grid.dense(data)
    origins, leaves = [], []
    for k in range(0, 24, 8):
        for j in range(0, 24, 8):
            for i in range(0, 24, 8):
                origins.append((i, j, k))
                leaves.append(np.transpose(data[k:k + 8, j:j + 8, i:i + 8], (2, 1, 0)).reshape(-1))
    n = len(origins)
    sparse = O.Grid.vdb(np.full(n, 3), np.array(origins), np.full(n, O.FORMAT_DENSE_ZYX),
                        packed_dense=[np.concatenate(leaves)])
    p = V.uniform_points(5000, (0, 0, 0), (22.99, 22.99, 22.99))
    for f in (O.NEAREST, O.LINEAR):
        dense.set_sampler(f)
        sparse.set_sampler(f)
        if f == O.NEAREST:
            # sparse (cell-centred) nearest has no +0.5 offset; compare at voxel centres
            q = np.floor(p) + np.float32(0.25)
            np.testing.assert_array_equal(sparse.sample(q), dense.sample(q - np.float32(0.5) + np.float32(0.25)))
        else:
            np.testing.assert_array_equal(sparse.sample(p), dense.sample(p))
    # per-leaf node data gives the same samples as packed data
    sparse2 = O.Grid.vdb(np.full(n, 3), np.array(origins), np.full(n, O.FORMAT_DENSE_ZYX),
                         leaf_data=[[l] for l in leaves])
    np.testing.assert_array_equal(sparse2.sample(p), sparse.sample(p))


def test_depth_mapping():
    """IteratorContext.cpp:65-71."""
    assert [O.map_depth(h) for h in (0, 0.1, 0.2, 0.39, 0.4, 0.5, 0.79, 0.8, 1.0, 2.0, -1.0)] == \
           [0, 0, 1, 1, 2, 2, 2, 3, 3, 3, 0]


def test_half_conversion():
    for bits, val in [(0x3C00, 1.0), (0xC000, -2.0), (0x0001, 2.0 ** -24), (0x7BFF, 65504.0), (0x0000, 0.0)]:
        assert O.half_to_float(bits) == np.float32(val)
    assert math.isinf(O.half_to_float(0x7C00)) and math.isnan(O.half_to_float(0x7E00))
    all_bits = np.arange(0, 0x7C00, dtype=np.uint16)
    want = all_bits.view(np.float16).astype(np.float32)
    got = np.array([O.half_to_float(int(b)) for b in all_bits[::37]], np.float32)
    np.testing.assert_array_equal(got, want[::37])
