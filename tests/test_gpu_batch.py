"""Batched multi-volume sampling (one graph-captured persistent launch) gives
exactly the per-volume vklComputeSampleN results."""
import numpy as np
import pytest

import openvkl_amd as vkl
from tests.volumes import leaves_from_dense

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_batch_plan_matches_per_volume(device):
    rng = np.random.default_rng(47)
    vols, samplers, pts, outs, refs = [], [], [], [], []
    for i in range(12):
        n = int(rng.integers(16, 72))
        dt = np.float32 if i % 2 == 0 else np.float16
        data = rng.standard_normal((n, n + 3, n + 1)).astype(dt)
        filt = [vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_NEAREST, vkl.VKL_FILTER_CUBIC][i % 4]
        v = device.structured_regular(data)
        s = v.sampler(filter=filt)
        count = int(np.exp(rng.uniform(np.log(100), np.log(200000))))
        p = torch.rand((count, 3), device="cuda") * torch.tensor([n + 1.0, n + 3.0, n + 0.0], device="cuda") - 0.5
        vols.append(v)
        samplers.append(s)
        pts.append(p)
        outs.append(torch.full((count,), -7.0, device="cuda"))
        refs.append(s.sample(p))
    plan = vkl.SampleBatchPlan(device, samplers, pts, outs, autotune=True)
    blocks, threads, tile = plan.launch_shape()
    assert blocks > 0 and threads in (256, 512, 1024) and tile >= threads
    for rep in range(3):
        for o in outs:
            o.fill_(-7.0)
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        for o, r in zip(outs, refs):
            a, b = o.cpu().numpy(), r.cpu().numpy()
            same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
            assert same.all(), rep
    plan.release()


def _check_plan(plan, outs, refs):
    for rep in range(2):
        for o in outs:
            o.fill_(-7.0)
        torch.cuda.synchronize()
        plan.execute()
        torch.cuda.synchronize()
        for k, (o, r) in enumerate(zip(outs, refs)):
            a, b = o.cpu().numpy(), r.cpu().numpy()
            same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
            assert same.all(), (rep, k)


def test_batch_plan_every_dense_type_one_launch(device):
    """All six structuredRegular element types with one filter share one
    launch (the element type is chosen per batch); batch sizes from empty to
    large, in an order that interleaves the types."""
    rng = np.random.default_rng(48)
    types = [np.uint8, np.int16, np.uint16, np.float16, np.float32, np.float64]
    vols, samplers, pts, outs, refs = [], [], [], [], []
    for i in range(18):
        dt = types[i % 6]
        n = int(rng.integers(9, 40))
        if np.issubdtype(dt, np.integer):
            info = np.iinfo(dt)
            data = rng.integers(info.min, info.max, (n, n + 2, n + 1), endpoint=True).astype(dt)
        else:
            data = rng.standard_normal((n, n + 2, n + 1)).astype(dt)
        v = device.structured_regular(data)
        s = v.sampler(filter=[vkl.VKL_FILTER_LINEAR, vkl.VKL_FILTER_CUBIC][i // 12])
        count = [0, 1, 63, 65, 300, 5000, 70000][int(rng.integers(0, 7))]
        p = torch.rand((count, 3), device="cuda") * torch.tensor([n + 1.0, n + 2.0, n + 0.0], device="cuda") - 0.5
        vols.append(v)
        samplers.append(s)
        pts.append(p)
        outs.append(torch.full((count,), -7.0, device="cuda"))
        refs.append(s.sample(p) if count else torch.empty(0, device="cuda"))
    for mode, launches in ((0, 11), (2, 2)):
        # per-type groups: 6 linear + 5 cubic element types; merged: one per filter
        plan = vkl.SampleBatchPlan(device, samplers, pts, outs, autotune=mode)
        assert plan.launch_count() == launches
        _check_plan(plan, outs, refs)
        plan.release()


def test_batch_plan_many_small_batches(device):
    """More batches than the workgroup stages in LDS, most smaller than a
    wave (waves span several batches), over a dense and a vdb volume."""
    rng = np.random.default_rng(49)
    data = rng.standard_normal((24, 20, 28)).astype(np.float32)
    dense = device.structured_regular(data)
    origins, leaves = leaves_from_dense(rng.standard_normal((24, 32, 16)).astype(np.float32))
    n = len(origins)
    vdb = device.vdb(np.full(n, 3, np.uint32), origins, np.full(n, vkl.VKL_FORMAT_DENSE_ZYX, np.uint32),
                     packed_dense=[leaves.reshape(-1)])
    sd, sv = dense.sampler(), vdb.sampler()
    samplers, pts, outs, refs = [], [], [], []
    for i in range(2300):
        s = sd if i % 3 else sv
        count = int(rng.integers(0, 40)) if i % 97 else 3000
        p = torch.rand((count, 3), device="cuda") * 30.0 - 1.0
        samplers.append(s)
        pts.append(p)
        outs.append(torch.full((count,), -7.0, device="cuda"))
        refs.append(s.sample(p) if count else torch.empty(0, device="cuda"))
    for mode in (0, 2):
        plan = vkl.SampleBatchPlan(device, samplers, pts, outs, autotune=mode)
        assert plan.launch_count() == 2
        _check_plan(plan, outs, refs)
        plan.release()
