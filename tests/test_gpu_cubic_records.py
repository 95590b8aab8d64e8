"""Tricubic sampling of structuredRegular volumes with cubic records
("cubicRecords": per voxel the 16 elements of its 4x4 (y, z) neighbourhood;
"cubicStencils": per cell its whole stencil), loaded a whole stencil at a
time by the lanes of a wavefront together.

Bar: samples and gradients bit-identical to the per-point tricubic path of
the same volume without records (the same 64 taps, the same arithmetic), for
every element type, base layout, transform and point position (stencils on
the domain's faces and outside it take the per-point path inside the same
kernel); and within 1e-5 of the field scale of the oracle's tricubic (the
reference's Lekien-Marsden form, evaluated in binary64)."""
import numpy as np
import pytest

import oracle as O
import openvkl_amd as vkl
from tests import volumes as V
from tests.test_gpu_dense import assert_same, check_cubic

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

CUBIC = vkl.VKL_FILTER_CUBIC
KINDS = [dict(cubic_records=True), dict(cubic_stencils=True)]
KIND_IDS = ["records", "stencils"]


def random_volume(dtype, shape, seed):
    rng = np.random.default_rng(seed)
    if np.dtype(dtype).kind == "f":
        return (rng.standard_normal(shape) * 30).astype(dtype)
    info = np.iinfo(dtype)
    return rng.integers(info.min, info.max, shape, endpoint=True).astype(dtype)


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
@pytest.mark.parametrize("dtype", [np.float32, np.float16, np.uint8, np.int16, np.uint16, np.float64])
def test_records_match_per_point_path_and_oracle(device, dtype, kind):
    data = random_volume(dtype, (33, 26, 41), 61)
    plain = device.structured_regular(data)
    rec = device.structured_regular(data, **kind)
    grid = O.Grid.dense(data)
    grid.set_sampler(O.CUBIC)
    # points over the domain and a voxel beyond it on every side: interior
    # stencils, stencils on the faces and outside points in the same waves
    pts = torch.from_numpy(V.uniform_points(300000, (-1.5, -1.5, -1.5), (41.5, 26.5, 33.5))).cuda()
    sp, sr = plain.sampler(filter=CUBIC), rec.sampler(filter=CUBIC)
    got = sr.sample(pts)
    assert_same(got.cpu().numpy(), sp.sample(pts).cpu().numpy())
    g = sr.gradient(pts)
    assert_same(g.cpu().numpy(), sp.gradient(pts).cpu().numpy())
    sub = pts[:40000].cpu().numpy()
    check_cubic(grid, got[:40000].cpu().numpy(), sub)
    check_cubic(grid, g[:20000].cpu().numpy(), sub[:20000], gradient=True)


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
@pytest.mark.parametrize("layout", [dict(corner_packed=True), dict(brick_shift=(2, 2, 1))])
def test_records_over_other_base_layouts(device, layout, kind):
    """The records sit beside the volume's own layout, which the per-point
    fallback reads."""
    data = random_volume(np.float32, (29, 37, 45), 67)
    plain = device.structured_regular(data)
    rec = device.structured_regular(data, **kind, **layout)
    pts = torch.from_numpy(V.uniform_points(200000, (-1.5, -1.5, -1.5), (45.5, 37.5, 29.5))).cuda()
    sp, sr = plain.sampler(filter=CUBIC), rec.sampler(filter=CUBIC)
    assert_same(sr.sample(pts).cpu().numpy(), sp.sample(pts).cpu().numpy())
    assert_same(sr.gradient(pts).cpu().numpy(), sp.gradient(pts).cpu().numpy())
    # nearest and linear keep their own kernels on the base layout
    for f in (vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_LINEAR):
        assert_same(rec.sampler(filter=f).sample(pts).cpu().numpy(),
                    plain.sampler(filter=f).sample(pts).cpu().numpy())


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
def test_records_transform_cell_centered_background(device, kind):
    rng = np.random.default_rng(71)
    data = rng.standard_normal((17, 18, 19)).astype(np.float32)
    i2o = np.array([0.9, 0.1, 0.0, -0.2, 1.1, 0.05, 0.0, 0.3, 0.7, 4.0, -2.0, 1.0], np.float32)  # column-major
    kw = dict(cell_centered=True, background=-3.5, index_to_object=i2o)
    plain = device.structured_regular(data, **kw)
    rec = device.structured_regular(data, **kind, **kw)
    rm = np.array([i2o[0], i2o[3], i2o[6], i2o[1], i2o[4], i2o[7], i2o[2], i2o[5], i2o[8], i2o[9], i2o[10], i2o[11]],
                  np.float32)
    grid = O.Grid.dense(data, i2o=rm, cell_centered=True, background=[-3.5])
    grid.set_sampler(O.CUBIC)
    box = grid.bounding_box()
    pts = V.uniform_points(100000, box[:3] - 1, box[3:] + 1)
    sp, sr = plain.sampler(filter=CUBIC), rec.sampler(filter=CUBIC)
    got = sr.sample(pts)
    assert_same(got, sp.sample(pts))
    assert_same(sr.gradient(pts), sp.gradient(pts))
    check_cubic(grid, got, pts)


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
def test_records_sample_m_scalar_and_wide(device, kind):
    """SampleM writes strided outputs through the same kernel; the scalar and
    wide calls equal the stream call."""
    rng = np.random.default_rng(73)
    a0 = rng.standard_normal((20, 21, 22)).astype(np.float32)
    a1 = rng.integers(0, 255, (20, 21, 22)).astype(np.uint8)
    a2 = rng.standard_normal((20, 21, 22)).astype(np.float16)
    plain = device.structured_regular([a0, a1, a2])
    rec = device.structured_regular([a0, a1, a2], **kind)
    pts = V.uniform_points(30000, (-1, -1, -1), (22, 21, 20))
    sp, sr = plain.sampler(filter=CUBIC), rec.sampler(filter=CUBIC)
    assert_same(sr.sample_m(pts, [2, 0, 1]), sp.sample_m(pts, [2, 0, 1]))
    ref = sr.sample(pts[:16])
    assert_same(np.array([sr.sample1(p) for p in pts[:8]], np.float32), ref[:8])
    for w in (4, 8, 16):
        assert_same(sr.sample_wide(pts[:w], width=w), ref[:w])


@pytest.mark.parametrize("autotune", [0, 1, 2])
def test_records_in_batch_plan_equal_stream_call(device, autotune):
    """Batch plans run tricubic batches on cubic records through the
    stencil-cooperative kernel and the rest through the persistent batch
    kernel: every batch equals its stream call bit for bit, an empty batch
    included, and each cubic-records batch is one launch of its own."""
    data = random_volume(np.float32, (40, 40, 40), 79)
    rec = device.structured_regular(data, cubic_records=True)
    whole = device.structured_regular(data.astype(np.float16), cubic_stencils=True)
    plain = device.structured_regular(data, corner_packed=True)
    samplers = [rec.sampler(filter=CUBIC), whole.sampler(filter=CUBIC), plain.sampler(filter=CUBIC),
                rec.sampler(filter=vkl.VKL_FILTER_LINEAR), rec.sampler(filter=CUBIC)]
    counts = [100000, 70000, 50000, 30000, 0]
    pts = [torch.from_numpy(V.uniform_points(max(c, 1), (-1, -1, -1), (40, 40, 40), seed=90 + k)).cuda()[:c]
           for k, c in enumerate(counts)]
    outs = [torch.full((c,), -7.0, device="cuda") for c in counts]
    plan = vkl.SampleBatchPlan(device, samplers, pts, outs, autotune=autotune)
    plan.execute()
    torch.cuda.synchronize()
    assert plan.launch_count() >= 2 + 1  # two records batches + at least one group
    plan.release()
    for s, p, o in zip(samplers, pts, outs):
        if p.shape[0]:
            assert_same(o.cpu().numpy(), s.sample(p).cpu().numpy())
        s.release()


def test_records_rejected_for_temporal_data(device):
    v = device.new_volume("structuredRegular")
    v.set_vec3i("dimensions", 4, 4, 4)
    v.set_int("temporallyStructuredNumTimesteps", 2)
    d = device.data(np.zeros(128, np.float32))
    v.set_data("data", d)
    d.release()
    v.set_bool("cubicRecords", True)
    vkl.lib().vklCommit(v.h)
    code, msg = device.last_error()
    device.errors.clear()
    assert code == vkl.VKL_INVALID_ARGUMENT and "cubicRecords" in msg


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
def test_full_size_512_cubic_records(device, kind):
    """512^3 float32, corner-packed with cubic records (the benchmark's
    cubic volume, 8 GiB of records: element offsets past 2^32 bytes): equal
    to the per-point path on 2^22 points, oracle bar on a subsample."""
    n = 512
    rng = np.random.default_rng(83)
    data = rng.random((n, n, n), dtype=np.float32)
    t = torch.from_numpy(data).cuda()
    rec = device.structured_regular(t, corner_packed=True, **kind)
    sr = rec.sampler(filter=CUBIC)
    pts = torch.rand((1 << 22, 3), device="cuda", dtype=torch.float32) * 513.0 - 1.0
    got = sr.sample(pts)
    gg = sr.gradient(pts[: 1 << 20])
    sr.release()
    rec.release()
    plain = device.structured_regular(t, corner_packed=True)
    sp = plain.sampler(filter=CUBIC)
    assert_same(got.cpu().numpy(), sp.sample(pts).cpu().numpy())
    assert_same(gg.cpu().numpy(), sp.gradient(pts[: 1 << 20]).cpu().numpy())
    sp.release()
    plain.release()
    grid = O.Grid.dense(data)
    grid.set_sampler(O.CUBIC)
    sub = pts[::512].cpu().numpy()
    check_cubic(grid, got[::512].cpu().numpy(), sub)


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
def test_records_beside_64bit_base_layout(device, kind):
    """uint8 corner-packed 820^3 (4.4e9 elements: the per-point path uses
    64-bit offsets) with cubic records (8.8 GB): equal to the per-point path,
    including stencils on the far faces, oracle bar on a subsample."""
    from tests.test_gpu_pins import u8_ramp

    n = 820
    data = u8_ramp(n, n, n)
    t = torch.from_numpy(data).cuda()
    pts = V.uniform_points(200000, (-1, -1, -1), (n, n, n), seed=99)
    pts[:40000] = V.uniform_points(40000, (n - 6, n - 6, n - 6), (n + 0.5, n + 0.5, n + 0.5), seed=100)
    p = torch.from_numpy(pts).cuda()
    rec = device.structured_regular(t, corner_packed=True, **kind)
    sr = rec.sampler(filter=CUBIC)
    got = sr.sample(p)
    gg = sr.gradient(p)
    sr.release()
    rec.release()
    plain = device.structured_regular(t, corner_packed=True)
    sp = plain.sampler(filter=CUBIC)
    assert_same(got.cpu().numpy(), sp.sample(p).cpu().numpy())
    assert_same(gg.cpu().numpy(), sp.gradient(p).cpu().numpy())
    sp.release()
    plain.release()
    grid = O.Grid.dense(data)
    grid.set_sampler(O.CUBIC)
    check_cubic(grid, got[:20000].cpu().numpy(), pts[:20000])


@pytest.mark.parametrize("kind", KINDS, ids=KIND_IDS)
def test_records_tiny_volumes(device, kind):
    """Volumes narrower than a stencil: whole stencils still load their
    records; row records (which need 4 columns) leave them to the per-point
    kernel."""
    rng = np.random.default_rng(107)
    for shape in ((6, 5, 3), (4, 2, 7), (1, 1, 1)):
        data = rng.standard_normal(shape).astype(np.float32)
        plain = device.structured_regular(data)
        rec = device.structured_regular(data, **kind)
        nz, ny, nx = shape
        pts = V.uniform_points(20000, (-1.5, -1.5, -1.5), (nx + 0.5, ny + 0.5, nz + 0.5), seed=108)
        sp, sr = plain.sampler(filter=CUBIC), rec.sampler(filter=CUBIC)
        assert_same(sr.sample(pts), sp.sample(pts))
        assert_same(sr.gradient(pts), sp.gradient(pts))
        sp.release(), sr.release(), plain.release(), rec.release()


def test_records_and_stencils_rejected_together(device):
    v = device.new_volume("structuredRegular")
    v.set_vec3i("dimensions", 4, 4, 4)
    d = device.data(np.zeros(64, np.float32))
    v.set_data("data", d)
    d.release()
    v.set_bool("cubicRecords", True)
    v.set_bool("cubicStencils", True)
    vkl.lib().vklCommit(v.h)
    code, msg = device.last_error()
    device.errors.clear()
    assert code == vkl.VKL_INVALID_ARGUMENT and "cannot be combined" in msg
