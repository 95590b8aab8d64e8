"""Interval and hit iteration on the GPU vs the CPU oracle: batched kernels
and the per-ray C API. Bar: interval t ranges, value ranges and
nominalDeltaT bit-identical; hits bit-identical with the linear filter."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rays_through(box, count, seed):
    rng = np.random.default_rng(seed)
    lo, hi = box[:3].astype(np.float64), box[3:].astype(np.float64)
    c = (lo + hi) / 2
    ext = (hi - lo)
    org = c + (rng.standard_normal((count, 3)) * ext)
    tgt = lo + rng.random((count, 3)) * ext
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def compare_intervals(ctx_batch, grid, org, d, depth, value_ranges=None, max_per_ray=512):
    out, cnt = ctx_batch
    for r in range(len(org)):
        want = grid.intervals(org[r], d[r], depth=depth, value_ranges=value_ranges, max_out=max_per_ray)
        got = out[r, :cnt[r]]
        assert got.shape == want.shape, (r, got.shape, want.shape)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.fixture(scope="module")
def wavelet(device):
    sp = np.float32(1.0 / 127.0)
    data = V.procedural((128, 128, 128), V.wavelet, spacing=(sp, sp, sp))
    return device.structured_regular(data, spacing=(sp, sp, sp)), O.Grid.dense(data, spacing=(sp, sp, sp))


def test_wavelet_known_answer_scalar_api(wavelet):
    """vectorized_interval_iterator.cpp:26-151 through vklIterateInterval."""
    vol, _ = wavelet
    ctx = vol.sampler().interval_context()
    rng = np.random.default_rng(3)
    for _ in range(4):
        iv = ctx.intervals((rng.uniform(0, 1), rng.uniform(0, 1), -1.0), (0, 0, 1))
        assert len(iv) == 16 and iv[0, 0] == 1.0 and iv[-1, 1] == 2.0
        assert np.all(iv[1:, 0] == iv[:-1, 1])


@pytest.mark.parametrize("hint", [0.0, 0.25, 0.5, 1.0])
def test_dense_interval_batch(wavelet, hint):
    vol, grid = wavelet
    ctx = vol.sampler().interval_context(hint=hint)
    org, d = rays_through(grid.bounding_box(), 300, seed=int(hint * 10))
    res = ctx.batch(torch.from_numpy(org).cuda(), torch.from_numpy(d).cuda(), max_per_ray=512)
    compare_intervals((res[0].cpu().numpy(), res[1].cpu().numpy()), grid, org, d, O.map_depth(hint))


def test_scalar_api_equals_batch(wavelet):
    vol, grid = wavelet
    ctx = vol.sampler().interval_context(value_ranges=[[0.5, 1.0], [2.0, 2.5]])
    org, d = rays_through(grid.bounding_box(), 5, seed=99)
    out, cnt = ctx.batch(org, d, max_per_ray=256)
    for r in range(5):
        iv = ctx.intervals(org[r], d[r])
        np.testing.assert_array_equal(iv.view(np.uint32), out[r, :cnt[r]].view(np.uint32))


@pytest.fixture(scope="module")
def shell(device):
    origins, leaves = V.sparse_shell_leaves(res=256)
    n = len(origins)
    lv, fm = np.full(n, 3, np.uint32), np.full(n, 1, np.uint32)
    return (device.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)]),
            O.Grid.vdb(lv, origins, fm, packed_dense=[leaves.reshape(-1)]))


@pytest.mark.parametrize("hint", [0.0, 0.5, 1.0])
def test_vdb_interval_batch(shell, hint):
    vol, grid = shell
    ctx = vol.sampler().interval_context(hint=hint)
    org, d = rays_through(grid.bounding_box(), 400, seed=5 + int(hint * 10))
    res = ctx.batch(torch.from_numpy(org).cuda(), torch.from_numpy(d).cuda(), max_per_ray=2048)
    compare_intervals((res[0].cpu().numpy(), res[1].cpu().numpy()), grid, org, d, O.map_depth(hint),
                      max_per_ray=2048)


def test_vdb_interval_value_selector(shell):
    vol, grid = shell
    sel = [[0.4, 0.6]]
    ctx = vol.sampler().interval_context(value_ranges=sel)
    org, d = rays_through(grid.bounding_box(), 200, seed=77)
    out, cnt = ctx.batch(org, d, max_per_ray=1024)
    compare_intervals((out, cnt), grid, org, d, 2, value_ranges=np.array(sel, np.float32), max_per_ray=1024)


def test_zramp_hits_known_answer(device):
    """hit_iterator.cpp:60-120 through vklIterateHit and the batch."""
    sp = np.float32(1.0 / 127.0)
    data = V.procedural((128, 128, 128), V.zramp, spacing=(sp, sp, sp))
    vol = device.structured_regular(data, spacing=(sp, sp, sp))
    iso, f = [], np.float32(0.1)
    while f < 1.0:
        iso.append(f)
        f = np.float32(f + np.float32(0.1))
    ctx = vol.sampler().hit_context(iso)
    hits = ctx.hits((0.5, 0.5, -1.0), (0, 0, 1))
    assert len(hits) == len(iso)
    np.testing.assert_array_equal(hits[:, 1], np.array(iso, np.float32))
    np.testing.assert_allclose(hits[:, 0], np.array(iso) + 1.0, atol=1e-3)
    grid = O.Grid.dense(data, spacing=(sp, sp, sp))
    want = grid.hits((0.5, 0.5, -1.0), (0, 0, 1), iso)
    np.testing.assert_array_equal(hits.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("which", ["dense", "vdb"])
def test_hit_batch_bitexact(wavelet, shell, which):
    vol, grid = wavelet if which == "dense" else shell
    values = [0.5] if which == "vdb" else [0.3, 1.1]
    ctx = vol.sampler().hit_context(values)
    org, d = rays_through(grid.bounding_box(), 200, seed=123)
    out, cnt = ctx.batch(torch.from_numpy(org).cuda(), torch.from_numpy(d).cuda(), max_per_ray=64)
    out, cnt = out.cpu().numpy(), cnt.cpu().numpy()
    total = 0
    for r in range(len(org)):
        want = grid.hits(org[r], d[r], values, max_out=64)
        got = out[r, :cnt[r]]
        total += len(want)
        assert got.shape == want.shape, r
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert total > 0
