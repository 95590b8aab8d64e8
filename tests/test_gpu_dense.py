"""structuredRegular sampling on the GPU vs the CPU oracle.

Bar: nearest / linear samples and linear gradients bit-identical (NaN ==
NaN); tricubic within 1e-5 of the field scale (separable Catmull-Rom vs the
reference's Lekien-Marsden matrix form)."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def assert_same(got, want):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    both_nan = np.isnan(got) & np.isnan(want)
    bad = ~both_nan & (got.view(np.uint32) != want.view(np.uint32))
    assert not bad.any(), f"{bad.sum()} mismatches, first at {np.argmax(bad.reshape(-1))}: " \
                          f"{got.reshape(-1)[np.argmax(bad.reshape(-1))]} vs {want.reshape(-1)[np.argmax(bad.reshape(-1))]}"


def assert_close_scale(got, want, rel=1e-5, slack=None):
    """|got - want| <= rel * field scale (+ per-point slack)."""
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    m = ~np.isnan(want)
    scale = max(np.abs(want[m]).max(), 1e-30)
    extra = 0.0 if slack is None else np.abs(np.asarray(slack, np.float64)[m])
    err = np.abs(got[m] - want[m]) - extra
    assert err.max() <= rel * scale, (err.max(), rel * scale)


def check_cubic(grid, got, pts, gradient=False, rel=1e-5):
    """Tricubic bar: within rel * scale of the reference algorithm evaluated
    exactly (binary64), and of its literal binary32 form up to that form's own
    rounding |f32 - f64| (large on rough data, ~1e-5 of scale)."""
    fn = grid.gradient if gradient else grid.sample
    grid.set_cubic_double(True)
    exact = fn(pts)
    grid.set_cubic_double(False)
    literal = fn(pts)
    assert_close_scale(got, exact, rel=rel)
    assert_close_scale(got, literal, rel=rel, slack=np.asarray(literal, np.float64) - np.asarray(exact, np.float64))


@pytest.fixture(scope="module")
def wavelet128(device):
    data = V.wavelet_volume(128)
    vol = device.structured_regular(data)
    grid = O.Grid.dense(data)
    # 2^20 points in bounds plus boundary / outside probes
    pts = V.uniform_points(1 << 20, (0, 0, 0), (127, 127, 127))
    edge = np.array([[0, 0, 0], [127, 127, 127], [126.999, 0, 5], [-0.001, 4, 4], [127.001, 3, 3],
                     [63.5, 63.5, 63.5], [1e9, 0, 0], [-1e9, 5, 5]], np.float32)
    return vol, grid, np.concatenate([pts, edge]).astype(np.float32)


@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR])
def test_wavelet128_sample_bitexact(wavelet128, filt):
    vol, grid, pts = wavelet128
    s = vol.sampler(filter=filt)
    grid.set_sampler(filt)
    want = grid.sample(pts)
    assert_same(s.sample(pts), want)  # host pointers (staged)
    got = s.sample(torch.from_numpy(pts).cuda()).cpu().numpy()  # device pointers
    assert_same(got, want)


def test_wavelet128_cubic(wavelet128):
    vol, grid, pts = wavelet128
    s = vol.sampler(filter=vkl.VKL_FILTER_CUBIC)
    grid.set_sampler(O.CUBIC)
    p = pts[:200000]
    check_cubic(grid, s.sample(torch.from_numpy(p).cuda()).cpu().numpy(), p)
    g = s.gradient(torch.from_numpy(p[:50000]).cuda()).cpu().numpy()
    check_cubic(grid, g, p[:50000], gradient=True)


@pytest.mark.parametrize("filt", [vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR])
def test_wavelet128_gradient_bitexact(wavelet128, filt):
    vol, grid, pts = wavelet128
    s = vol.sampler(filter=filt)
    grid.set_sampler(filt)
    p = pts[:300000]
    assert_same(s.gradient(torch.from_numpy(p).cuda()).cpu().numpy(), grid.gradient(p))


def test_scalar_and_wide_calls_match_stream(wavelet128):
    """sampling_utility.h:35-119: scalar, 4/8/16-wide and stream agree."""
    vol, grid, pts = wavelet128
    s = vol.sampler()
    p = pts[:16]
    ref = s.sample(p)
    for i in range(4):
        assert_same(np.float32(s.sample1(p[i])), ref[i])
    for w in (4, 8, 16):
        assert_same(s.sample_wide(p[:w], width=w), ref[:w])
    g = s.gradient(p[:3])
    assert_same(s.gradient1(p[0]), g[0])


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.float16, np.float64])
def test_element_types(device, dtype):
    """Data.ih:66-226: integer types convert without normalisation."""
    rng = np.random.default_rng(23)
    if dtype in (np.float16, np.float64):
        data = rng.standard_normal((40, 37, 33)).astype(dtype)
    else:
        info = np.iinfo(dtype)
        data = rng.integers(info.min, info.max, (40, 37, 33), endpoint=True).astype(dtype)
    vol = device.structured_regular(data, spacing=(0.5, 1.0, 2.0), origin=(1, 2, 3))
    grid = O.Grid.dense(data, spacing=(0.5, 1.0, 2.0), origin=(1, 2, 3))
    pts = V.uniform_points(100000, (0.5, 1.5, 2.0), (18.5, 39.0, 83.0))
    for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
        grid.set_sampler(f)
        s = vol.sampler(filter=f)
        assert_same(s.sample(pts), grid.sample(pts))
        assert_same(s.gradient(pts), grid.gradient(pts))
    grid.set_sampler(O.CUBIC)
    s = vol.sampler(filter=vkl.VKL_FILTER_CUBIC)
    check_cubic(grid, s.sample(pts), pts)


def test_cell_centered_background_transform(device):
    rng = np.random.default_rng(29)
    data = rng.standard_normal((17, 18, 19)).astype(np.float32)
    i2o = np.array([0.9, 0.1, 0.0, -0.2, 1.1, 0.05, 0.0, 0.3, 0.7, 4.0, -2.0, 1.0], np.float32)  # column-major
    vol = device.structured_regular(data, cell_centered=True, background=-3.5, index_to_object=i2o)
    rm = np.array([i2o[0], i2o[3], i2o[6], i2o[1], i2o[4], i2o[7], i2o[2], i2o[5], i2o[8], i2o[9], i2o[10], i2o[11]],
                  np.float32)
    grid = O.Grid.dense(data, i2o=rm, cell_centered=True, background=[-3.5])
    np.testing.assert_array_equal(vol.bounding_box(), grid.bounding_box())
    box = grid.bounding_box()
    pts = V.uniform_points(100000, box[:3] - 1, box[3:] + 1)
    for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
        grid.set_sampler(f)
        s = vol.sampler(filter=f)
        assert_same(s.sample(pts), grid.sample(pts))
        assert_same(s.gradient(pts), grid.gradient(pts))
    np.testing.assert_array_equal(vol.value_range(), grid.value_range())


def test_multi_attribute_sample_m(device):
    rng = np.random.default_rng(31)
    a0 = rng.standard_normal((20, 20, 20)).astype(np.float32)
    a1 = rng.integers(0, 255, (20, 20, 20)).astype(np.uint8)
    a2 = rng.standard_normal((20, 20, 20)).astype(np.float16)
    vol = device.structured_regular([a0, a1, a2])
    grid = O.Grid.dense([a0, a1, a2])
    assert vol.num_attributes() == 3
    pts = V.uniform_points(20000, (0, 0, 0), (19, 19, 19))
    s = vol.sampler()
    got = s.sample_m(pts, [2, 0, 1])
    assert_same(got, grid.sample_m(pts, [2, 0, 1]))
    for a in range(3):
        assert_same(s.sample(pts, attr=a), grid.sample(pts, attr=a))
        np.testing.assert_array_equal(vol.value_range(a), grid.value_range(a))


@pytest.mark.parametrize("shift", [(2, 2, 1), (2, 1, 1), (1, 1, 1), (3, 2, 2)])
def test_brick_layouts_match_reference_order(device, shift):
    """Brick-swizzled device layouts change addressing only: every filter,
    gradient and majorant equals the oracle exactly (cubic: same bar)."""
    rng = np.random.default_rng(53)
    data = rng.standard_normal((37, 29, 45)).astype(np.float32)
    vol = device.structured_regular(data, brick_shift=shift)
    grid = O.Grid.dense(data)
    pts = V.uniform_points(200000, (-1, -1, -1), (45, 29, 37))
    for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
        grid.set_sampler(f)
        s = vol.sampler(filter=f)
        assert_same(s.sample(pts), grid.sample(pts))
        assert_same(s.gradient(pts), grid.gradient(pts))
    grid.set_sampler(O.CUBIC)
    check_cubic(grid, vol.sampler(filter=vkl.VKL_FILTER_CUBIC).sample(pts[:50000]), pts[:50000])
    for l in range(3):
        _, vox, r = vol.level(l)
        _, ovox, orng = grid.level(l)
        np.testing.assert_array_equal(r.view(np.uint32), orng.view(np.uint32))


@pytest.mark.parametrize("dtype", [np.float32, np.float16, np.uint8, np.float64])
def test_corner_packed_layout_matches_reference(device, dtype):
    """Corner-packed cell records (8 corners per cell) change addressing only:
    samples, gradients and majorants equal the oracle's exactly, including
    stencils on the domain's last planes and outside it."""
    rng = np.random.default_rng(57)
    data = rng.standard_normal((33, 26, 41)) * 30
    data = data.astype(dtype) if np.dtype(dtype).kind == "f" else np.abs(data).astype(dtype)
    vol = device.structured_regular(data, corner_packed=True)
    grid = O.Grid.dense(data)
    pts = V.uniform_points(200000, (-1.5, -1.5, -1.5), (41.5, 26.5, 33.5))
    for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
        grid.set_sampler(f)
        s = vol.sampler(filter=f)
        assert_same(s.sample(pts), grid.sample(pts))
        assert_same(s.gradient(pts), grid.gradient(pts))
    grid.set_sampler(O.CUBIC)
    check_cubic(grid, vol.sampler(filter=vkl.VKL_FILTER_CUBIC).sample(pts[:50000]), pts[:50000])
    for l in range(3):
        _, _, r = vol.level(l)
        _, _, orng = grid.level(l)
        np.testing.assert_array_equal(r.view(np.uint32), orng.view(np.uint32))


def test_corner_packed_rejects_brick_shift(device):
    v = device.new_volume("structuredRegular")
    v.set_vec3i("dimensions", 4, 4, 4)
    d = device.data(np.zeros(64, np.float32))
    v.set_data("data", d)
    d.release()
    v.set_vec3i("brickShift", 1, 1, 1)
    v.set_bool("cornerPacked", True)
    vkl.lib().vklCommit(v.h)
    code, msg = device.last_error()
    device.errors.clear()
    assert code == vkl.VKL_INVALID_ARGUMENT and "cornerPacked" in msg


def test_node_table_bitexact(device):
    """Node table (voxel codes + value ranges) equals the oracle's bit for bit."""
    rng = np.random.default_rng(37)
    data = rng.standard_normal((70, 300, 140)).astype(np.float32)  # nz, ny, nx; partial leaves
    for bg in (None, 0.25):
        vol = device.structured_regular(data, background=bg)
        grid = O.Grid.dense(data, background=None if bg is None else [bg])
        for l in range(3):
            n, vox, rng_ = vol.level(l)
            org, ovox, orng = grid.level(l)
            assert n == org.shape[0]
            np.testing.assert_array_equal(vox, ovox)
            np.testing.assert_array_equal(rng_.view(np.uint32), orng.view(np.uint32))
        np.testing.assert_array_equal(vol.value_range(), grid.value_range())


def test_full_size_512_linear(device):
    """512^3 float32 (the benchmark volume): subsample parity + vertex
    reproduction over 2^24 points."""
    n = 512
    rng = np.random.default_rng(41)
    data = rng.random((n, n, n), dtype=np.float32)
    t = torch.from_numpy(data).cuda()
    vol = device.structured_regular(t)  # device-resident shared data
    s = vol.sampler()
    pts = torch.rand((1 << 24, 3), device="cuda", dtype=torch.float32) * 511.0
    out = s.sample(pts)
    # vertex reproduction: integer points sample the voxel exactly
    ip = torch.randint(0, 511, (1 << 20, 3), device="cuda", dtype=torch.int64)
    vv = s.sample(ip.float())
    want = t[ip[:, 2], ip[:, 1], ip[:, 0]]
    assert torch.equal(vv, want)
    # oracle parity on a subsample
    grid = O.Grid.dense(data)
    sub = pts[::64].cpu().numpy()
    assert_same(out[::64].cpu().numpy(), grid.sample(sub))
