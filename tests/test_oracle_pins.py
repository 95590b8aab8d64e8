"""More pins of the CPU oracle: the error-voxel rule of maxSamplingDepth
against a brute-force window min / max, the reference's torus regression rays
(divide_safe edge directions), the VDB wavelet value-range known answer, and
the 64 / 32-bit addressing dimension formula. CPU only."""
import struct

import numpy as np
import pytest

import oracle as O
from tests import volumes as V

FLT_MIN = np.float32(1.17549435e-38)


def f32_bits(u):
    return struct.unpack("<f", struct.pack("<I", u))[0]


# ---- maxSamplingDepth < 3: error voxels ----------------------------------------------

DIMS = (256, 16, 24)  # x, y, z: two level-1 voxels (128^3) along x


@pytest.fixture(scope="module")
def full_vdb():
    rng = np.random.default_rng(91)
    nx, ny, nz = DIMS
    arr = (rng.standard_normal((nz, ny, nx)) * 3).astype(np.float32)
    # a few large values so that different windows have different extremes
    arr.reshape(-1)[rng.choice(arr.size, 40, replace=False)] = rng.uniform(-40, 40, 40).astype(np.float32)
    origins, leaves = V.leaves_from_dense(arr)
    n = len(origins)
    g = O.Grid.vdb(np.full(n, 3, np.uint32), origins, np.full(n, O.FORMAT_DENSE_ZYX, np.uint32),
                   packed_dense=[leaves.reshape(-1)])
    return g, arr


def window_mid(arr, lo, size):
    """1/2 (min + max) over the voxels a level voxel's value range covers:
    its region [lo, lo + size) widened to [lo - 1, lo + size + 2) (every leaf
    range covers [-1, +10) around the leaf, VdbSampler.ispc:460-557), cut to
    the domain (out-of-domain neighbours read the NaN background, which the
    range skips)."""
    nz, ny, nx = arr.shape
    x0, y0, z0 = (max(lo[a] - 1, 0) for a in range(3))
    x1, y1, z1 = min(lo[0] + size + 2, nx), min(lo[1] + size + 2, ny), min(lo[2] + size + 2, nz)
    w = arr[z0:z1, y0:y1, x0:x1]
    return np.float32(0.5) * (np.float32(w.min()) + np.float32(w.max()))


def expected_error_samples(arr, vox, depth):
    """Nearest samples at voxels `vox` [n,3] (x, y, z) when traversal stops at
    level `depth` (VdbSampleInner.ih.in:46-54, VdbSampler_traverseAndSample.ih:
    185-192): 1/2 (lower + upper) of the stopping voxel's range."""
    size = [4096, 128, 8][depth]
    out = np.empty(len(vox), np.float32)
    cache = {}
    for i, v in enumerate(vox):
        lo = tuple(int(c) // size * size for c in v)
        if lo not in cache:
            cache[lo] = window_mid(arr, lo, size)
        out[i] = cache[lo]
    return out


@pytest.mark.parametrize("depth", [0, 1, 2])
def test_error_voxel_rule_matches_bruteforce(full_vdb, depth):
    g, arr = full_vdb
    rng = np.random.default_rng(92 + depth)
    nx, ny, nz = DIMS
    vox = np.stack([rng.integers(1, nx - 1, 3000), rng.integers(1, ny - 1, 3000), rng.integers(1, nz - 1, 3000)], 1)
    # nearest filter on cell data: the voxel is floor(p)
    pts = (vox + np.float32(0.25)).astype(np.float32)
    g.set_sampler(O.NEAREST, max_sampling_depth=depth)
    got = g.sample(pts)
    want = expected_error_samples(arr, vox, depth)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    # The pin has power: the rule evaluated at a neighbouring level (a wrong
    # level, or a leaf range used for a node range) disagrees.
    for other in {0, 1, 2} - {depth}:
        wrong = expected_error_samples(arr, vox, other)
        if {depth, other} == {0, 1}:
            # one level-0 voxel vs two level-1 voxels: they differ on one side
            assert np.mean(wrong != want) > 0.3, other
        else:
            assert np.mean(wrong != want) > 0.9, other
    g.set_sampler()


def test_error_voxel_linear_is_trilinear_of_corner_values(full_vdb):
    """With a linear filter every corner is an error voxel of its own: the
    sample is the trilinear blend of the per-corner values above."""
    g, arr = full_vdb
    rng = np.random.default_rng(95)
    nx, ny, nz = DIMS
    p = np.stack([rng.uniform(1, nx - 2, 2000), rng.uniform(1, ny - 2, 2000), rng.uniform(1, nz - 2, 2000)],
                 1).astype(np.float32)
    g.set_sampler(O.LINEAR, max_sampling_depth=2)
    got = g.sample(p)
    ic = np.floor(p).astype(np.int64)
    d = p - ic.astype(np.float32)
    c = {}
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c[dx, dy, dz] = expected_error_samples(arr, ic + [dx, dy, dz], 2)

    def lerp(t, a, b):
        return (np.float32(1) - t) * a + t * b

    want = lerp(d[:, 0], lerp(d[:, 1], lerp(d[:, 2], c[0, 0, 0], c[0, 0, 1]), lerp(d[:, 2], c[0, 1, 0], c[0, 1, 1])),
                lerp(d[:, 1], lerp(d[:, 2], c[1, 0, 0], c[1, 0, 1]), lerp(d[:, 2], c[1, 1, 0], c[1, 1, 1])))
    np.testing.assert_array_equal(got.view(np.uint32), want.astype(np.float32).view(np.uint32))
    g.set_sampler()


# ---- torus regression rays (vdb_volume.cpp:951-1008) ----------------------------------

TORUS_ORG = [1112900070, 1116163650, 1103628776]   # 53.3769 67.6528 25.0048
TORUS_DIR = [1081551625, 1098411576, 2984533223]   # 3.862 15.5269 -6.64624e-09
TORUS_DIR_Z = [None, 0.0, float(FLT_MIN), float(np.float32(1.1) * FLT_MIN), -0.0, -float(FLT_MIN),
               -float(np.float32(1.1) * FLT_MIN)]


def torus_rays():
    org = np.array([f32_bits(u) for u in TORUS_ORG], np.float32)
    base = np.array([f32_bits(u) for u in TORUS_DIR], np.float32)
    rays = []
    for z in TORUS_DIR_Z:
        d = base.copy()
        if z is not None:
            d[2] = np.float32(z)
        rays.append((org, d))
    return rays


def assert_special_case_intervals(iv):
    """vdb_volume.cpp:896-944: at least one interval, all with non-negative
    t, upper > lower, and each starting where the previous one ended."""
    assert len(iv) >= 1
    assert np.all(iv[:, 0] >= 0) and np.all(iv[:, 1] >= 0)
    assert np.all(iv[:, 1] > iv[:, 0])
    assert np.all(iv[1:, 0] == iv[:-1, 1])


@pytest.fixture(scope="module")
def torus_grid():
    origins, leaves = V.torus_vdb_leaves()
    n = len(origins)
    return O.Grid.vdb(np.full(n, 3, np.uint32), origins, np.full(n, O.FORMAT_DENSE_ZYX, np.uint32),
                      packed_dense=[leaves.reshape(-1)])


def test_torus_rays_divide_safe_edges(torus_grid):
    for org, d in torus_rays():
        iv = torus_grid.intervals(org, d)
        assert_special_case_intervals(iv)
    # the ray direction's signed FLT_MIN cases differ in the sign of the
    # clamped reciprocal only; the intervals themselves agree with dir.z = 0
    ivs = [torus_grid.intervals(o, d) for o, d in torus_rays()[1:]]
    for other in ivs[1:]:
        assert len(other) == len(ivs[0])


# ---- VDB wavelet value range (vdb_volume.cpp:81-127) ------------------------------------

@pytest.mark.parametrize("dtype", [np.float32, np.float16])
@pytest.mark.parametrize("packed", [True, False])
def test_vdb_wavelet_value_range_known_answer(dtype, packed):
    """WaveletVdbVolume 128^3, origin 0, spacing 1 (repackNodes on / off):
    value range width 6 within 0.1 %."""
    origins, (leaves,) = V.dense_vdb_leaves(128, [V.wavelet])
    leaves = leaves.astype(dtype)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, O.FORMAT_DENSE_ZYX, np.uint32)
    if packed:
        g = O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)])
    else:
        g = O.Grid.vdb(lv, origins, fm, leaf_data=[[leaves[i]] for i in range(n)])
    lo, hi = g.value_range()
    assert hi >= lo
    assert abs((hi - lo) - 6.0) <= 0.001 * 6.0
    # and it is the range of the voxel values
    assert lo == np.float32(leaves.min()) and hi == np.float32(leaves.max())


# ---- 64 / 32-bit addressing (structured_regular_volume.h:79-130) -------------------------

def dims_for_64_32bit_addressing(byte_stride):
    """The reference's dimension formula: the smallest cube whose bytes exceed
    the 2^31 limit of 32-bit addressing (Data.ih:45-59)."""
    max32 = 1 << 31
    dim = int(np.cbrt(float(max32) / float(byte_stride))) + 1
    if dim ** 3 * byte_stride <= max32:
        raise RuntimeError("incorrect dimension computed for 64/32-bit addressing mode")
    return dim


@pytest.mark.parametrize("stride,dim", [(1, 1291), (2, 1025), (4, 813), (8, 646)])
def test_64_32bit_addressing_dimensions(stride, dim):
    assert dims_for_64_32bit_addressing(stride) == dim
    assert dim ** 3 * stride > 1 << 31 >= (dim - 1) ** 3 * stride


def test_iterate_many_equals_per_ray_calls():
    """The threaded many-ray entry point (the iterators' CPU baseline) gives
    exactly the per-ray interval and hit iteration results."""
    origins, leaves = V.sparse_shell_leaves(res=64)
    n = len(origins)
    g = O.Grid.vdb(np.full(n, 3, np.uint32), origins, np.full(n, 1, np.uint32), packed_dense=[leaves.reshape(-1)])
    rng = np.random.default_rng(5)
    org = (rng.random((64, 3)) * 80 - 8).astype(np.float32)
    d = rng.standard_normal((64, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out, cnt = g.iterate_many(org, d, 64, threads=3)
    hout, hcnt = g.iterate_many(org, d, 8, values=[0.5], threads=3)
    assert cnt.sum() > 0 and hcnt.sum() > 0
    for r in range(len(org)):
        iv = g.intervals(org[r], d[r], max_out=64)
        np.testing.assert_array_equal(out[r, :cnt[r]].view(np.uint32), iv.view(np.uint32))
        h = g.hits(org[r], d[r], [0.5], max_out=8)
        np.testing.assert_array_equal(hout[r, :hcnt[r]].view(np.uint32), h.view(np.uint32))
