"""Batched multi-volume sampling (one graph-captured persistent launch) gives
exactly the per-volume vklComputeSampleN results."""
import numpy as np
import pytest

import openvkl_amd as vkl

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
